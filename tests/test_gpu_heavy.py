"""The heavy mixed scene (1000 instances of 8 meshes, ~24M instanced
triangles, ~30% alpha-tested leaf quads with stochastic alpha, two
scattering volume cubes, an importance-sampled 2048x1024 sky, 3840x2160,
depth 12): a scaled-down copy against the CPU oracle in both kernel forms,
and the full-size scene through size-independent properties (both forms
bit-identical, finite, non-negative)."""
import numpy as np
import pytest

import oracle
from common import oracle_setup, rel_err
from glslpt import scenes
from glslpt._capi import PT_MODE_PLAIN, PT_MODE_WAVEFRONT
from glslpt.renderer import Renderer, options_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def renderer():
    r = Renderer(0)
    yield r
    r.close()


def small_heavy(**kw):
    args = dict(width=96, height=64, spp=4, max_depth=8, sphere_freqs=(4, 5, 6, 7, 9),
                leaf_quads_per_mesh=(20, 30, 40), counts=(24, 8, 4, 2, 1, 20, 6, 3), extent=20.0, sky=(128, 64))
    args.update(kw)
    return scenes.heavy_mixed(**args)


@pytest.mark.parametrize("mode", [PT_MODE_PLAIN, PT_MODE_WAVEFRONT])
def test_heavy_small_parity(renderer, mode):
    """Within 1e-4 on every pixel (stochastic alpha at bilinear mask edges
    included: filter, sin and cos run the same float operations as the
    oracle's), image means within 1e-4."""
    sc, rs = small_heavy()
    renderer.upload_scene(sc)
    renderer.set_options(options_for(rs, mode=mode, wave_spp=2))
    renderer.set_camera(sc.camera())
    renderer.clear()
    renderer.render(0, rs.spp)
    renderer.synchronize()
    g = renderer.read_accum()
    sa, cam, o = oracle_setup(sc, rs)
    r = oracle.render(sa, cam, o, 0, rs.spp)
    e = rel_err(g[..., :3], r[..., :3], floor=1e-4).max(axis=2)
    within = (e <= 1e-4).mean()
    mg, mr = g[..., :3].astype(np.float64).mean(), r[..., :3].astype(np.float64).mean()
    print(f"heavy mode={mode}: pixels within 1e-4 = {within:.5f}, mean gpu {mg:.6f} oracle {mr:.6f}")
    assert np.isfinite(g).all()
    assert within == 1.0
    assert abs(mg - mr) <= 1e-4 * abs(mr)


def test_heavy_small_plain_equals_wavefront(renderer):
    sc, rs = small_heavy()
    renderer.upload_scene(sc)
    out = []
    for mode, kw in ((PT_MODE_PLAIN, {}), (PT_MODE_WAVEFRONT, dict(wave_spp=3, compact_every=2))):
        renderer.set_options(options_for(rs, mode=mode, **kw))
        renderer.set_camera(sc.camera())
        renderer.clear()
        renderer.render(0, 6)
        renderer.synchronize()
        out.append(renderer.read_accum())
    assert np.array_equal(out[0], out[1])


def test_heavy_full_size_forms_agree(renderer):
    """Full size, 1 spp: the plain and wavefront forms give the same image
    bit for bit; it is finite, non-negative and lit."""
    sc, rs = scenes.heavy_mixed()
    renderer.upload_scene(sc)
    out = []
    for mode in (PT_MODE_WAVEFRONT, PT_MODE_PLAIN):
        renderer.set_options(options_for(rs, mode=mode))
        renderer.set_camera(sc.camera())
        renderer.clear()
        renderer.render(0, 1)
        renderer.synchronize()
        out.append(renderer.read_accum())
        print(f"full heavy mode={mode}: {renderer.stats()['ms_last_render']:.1f} ms, mean {out[-1][..., :3].mean():.4f}")
    a, b = out
    assert a.shape == (rs.height, rs.width, 4)
    assert np.isfinite(a).all() and (a[..., :3] >= 0).all()
    assert a[..., :3].mean() > 0
    assert np.array_equal(a, b)


def test_heavy_full_size_sample3_finite(renderer):
    """Full size, sample 3, both forms with four-wide nodes: pixel (3219,
    1547) scatters in a volume where Henyey-Greenstein sampling's cosine
    rounds a hair above 1; the image is finite (sin theta = 0 there, as the
    reference's clamp gives on a GPU) and the forms agree bit for bit."""
    sc, rs = scenes.heavy_mixed()
    renderer.upload_scene(sc)
    out = []
    for mode in (PT_MODE_WAVEFRONT, PT_MODE_PLAIN):
        renderer.set_options(options_for(rs, mode=mode, bvh_width=4))
        renderer.set_camera(sc.camera())
        renderer.clear()
        renderer.render(3, 1)
        renderer.synchronize()
        out.append(renderer.read_accum())
    a, b = out
    print(f"pixel (3219, 1547): {a[1547, 3219].tolist()}")
    assert np.isfinite(a).all() and (a[..., :3] >= 0).all()
    assert np.array_equal(a, b)


def test_heavy_small_capped_extend_identical(renderer):
    """The capped extend stage on the heavy scene (alpha pass-through, media,
    four-wide nodes): bit-identical to the plain kernel."""
    sc, rs = small_heavy()
    renderer.upload_scene(sc)
    out = []
    for mode, kw in ((PT_MODE_PLAIN, {}), (PT_MODE_WAVEFRONT, dict(extend_cap=8, wave_spp=4, bvh_width=4))):
        renderer.set_options(options_for(rs, mode=mode, **kw))
        renderer.set_camera(sc.camera())
        renderer.clear()
        renderer.render(0, 4)
        renderer.synchronize()
        out.append(renderer.read_accum())
    assert np.array_equal(out[0], out[1])

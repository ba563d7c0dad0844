"""The BLAS traversed as four-wide nodes collapsed from the reference's BVH2
(pt_options.bvh_width = 4) against the oracle, which walks the BVH2:

* closest hit: primitive, instance and light ids identical on >= 99.99% of
  rays, and every ray whose ids differ is a tie (both hits at the same
  distance within 1e-5 relative); hit t within 1e-5 relative; any hit
  identical on >= 99.99% of rays;
* rendered images within 1e-4 relative on every pixel, image means within
  1e-4 relative;
* the plain and wavefront forms bit-identical to each other.
"""
import numpy as np
import pytest

import oracle
from common import camera_rays, oracle_setup, random_dirs, rel_err, small_instanced
from glslpt import scenes
from glslpt._capi import PT_MODE_PLAIN, PT_MODE_WAVEFRONT
from glslpt.renderer import Renderer, options_for
from test_gpu_heavy import small_heavy

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def renderer():
    r = Renderer(0)
    yield r
    r.close()


def _hits(renderer, sc, rs, ro, rd, label):
    renderer.upload_scene(sc)
    renderer.set_options(options_for(rs, bvh_width=4))
    sa, cam, o = oracle_setup(sc, rs)
    g = renderer.closest_hit(ro, rd)
    r = oracle.closest_hit(sa, o, ro, rd)
    same = (g["triangle"] == r["triangle"]) & (g["instance"] == r["instance"]) & (g["light"] == r["light"])
    hit = r["t"] < 1e6
    terr = rel_err(g["t"][hit], r["t"][hit]).max() if hit.any() else 0.0
    print(f"{label} bvh4: rays {len(ro)}, ids identical {same.mean():.6f}, differing {int((~same).sum())}, "
          f"max rel t {terr:.3g}")
    assert same.mean() >= 0.9999
    assert terr <= 1e-5
    assert np.all(rel_err(g["t"][~same], r["t"][~same]) <= 1e-5)  # differences are ties


def test_closest_hit_cornell_bvh4(renderer):
    sc, rs = scenes.cornell()
    rng = np.random.default_rng(21)
    ro, rd = camera_rays(sc.camera(), rs.width, rs.height, rng, 200000)
    ri = rng.uniform([0.05, 0.05, 0.05], [0.5, 0.5, 0.5], (100000, 3)).astype(np.float32)
    ro = np.concatenate([ro, ri]); rd = np.concatenate([rd, random_dirs(rng, 100000)])
    _hits(renderer, sc, rs, ro, rd, "cornell")


def test_closest_hit_instanced_bvh4(renderer):
    sc, rs = small_instanced()
    rng = np.random.default_rng(22)
    ro, rd = camera_rays(sc.camera(), rs.width, rs.height, rng, 200000)
    ri = rng.uniform(-6, 6, (100000, 3)).astype(np.float32)
    ro = np.concatenate([ro, ri]); rd = np.concatenate([rd, random_dirs(rng, 100000)])
    _hits(renderer, sc, rs, ro, rd, "instanced")


def test_closest_hit_headline_bvh4(renderer):
    """The 1080p workload's 64 instances of the 15,680-triangle sphere."""
    sc, rs = scenes.instanced_spheres()
    rng = np.random.default_rng(23)
    ro, rd = camera_rays(sc.camera(), rs.width, rs.height, rng, 100000)
    ri = rng.uniform(-25, 25, (200000, 3)).astype(np.float32)
    ro = np.concatenate([ro, ri]); rd = np.concatenate([rd, random_dirs(rng, 200000)])
    _hits(renderer, sc, rs, ro, rd, "headline")


def test_any_hit_instanced_bvh4(renderer):
    sc, rs = small_instanced()
    renderer.upload_scene(sc)
    renderer.set_options(options_for(rs, bvh_width=4))
    sa, cam, o = oracle_setup(sc, rs)
    rng = np.random.default_rng(24)
    ro = rng.uniform(-6, 6, (200000, 3)).astype(np.float32)
    rd = random_dirs(rng, 200000)
    md = rng.uniform(0.1, 10, 200000).astype(np.float32)
    g = renderer.any_hit(ro, rd, md)
    r = oracle.any_hit(sa, o, ro, rd, md)
    print(f"any-hit bvh4 agreement {(g == r).mean():.6f}")
    assert (g == r).mean() >= 0.9999


def _render(renderer, sc, rs, mode, spp):
    renderer.upload_scene(sc)
    renderer.set_options(options_for(rs, mode=mode, wave_spp=spp, bvh_width=4))
    renderer.set_camera(sc.camera())
    renderer.clear()
    renderer.render(0, spp)
    renderer.synchronize()
    return renderer.read_accum()


# (scene, pixel floor, image-mean tolerance): every pixel, as the BVH2 form
CASES = {
    "cornell": (lambda: scenes.cornell(), 1.0, 1e-4),
    "instanced": (lambda: small_instanced(), 1.0, 1e-4),
    "heavy": (lambda: small_heavy(), 1.0, 1e-4),
    "media": (lambda: scenes.feature_scene("media"), 1.0, 1e-4),
    "envmap": (lambda: scenes.feature_scene("envmap"), 1.0, 1e-4),
    "alpha_mask": (lambda: scenes.feature_scene("alpha_mask"), 1.0, 1e-4),
}


@pytest.mark.parametrize("case", list(CASES))
def test_render_parity_bvh4(renderer, case):
    make, floor, mtol = CASES[case]
    sc, rs = make()
    sa, cam, o = oracle_setup(sc, rs)
    ref = oracle.render(sa, cam, o, 0, rs.spp)
    imgs = []
    for mode in (PT_MODE_PLAIN, PT_MODE_WAVEFRONT):
        g = _render(renderer, sc, rs, mode, rs.spp)
        e = rel_err(g[..., :3], ref[..., :3], floor=1e-4).max(axis=2)
        within = (e <= 1e-4).mean()
        mg, mr = g[..., :3].astype(np.float64).mean(), ref[..., :3].astype(np.float64).mean()
        print(f"{case} bvh4 mode={mode}: pixels within 1e-4 = {within:.6f}, mean {mg:.7f} vs {mr:.7f}")
        assert np.isfinite(g).all()
        assert within >= floor
        assert abs(mg - mr) <= mtol * abs(mr)
        imgs.append(g)
    assert np.array_equal(imgs[0], imgs[1])

"""GPU parity of the option paths beyond the basic scenes (HDR environment
map with importance sampling, textures incl. normal/emission maps,
homogeneous media, volume MIS, alpha mask and stochastic alpha) against the
CPU oracle on the same seeds, in both kernel forms and in the batched form.
Stochastic alpha and volume MIS resolve NEE visibility inside the wavefront
shade stage (the reference's draw order), the other paths through the
shadow stage."""
import numpy as np
import pytest

import oracle
from common import oracle_setup, rel_err
from glslpt import scenes
from glslpt._capi import PT_MODE_PLAIN, PT_MODE_WAVEFRONT
from glslpt.renderer import Renderer, options_for

pytestmark = pytest.mark.gpu

KINDS = ["envmap", "textures", "media", "volmis", "alpha_mask", "alpha_blend"]


@pytest.fixture(scope="module")
def renderer():
    r = Renderer(0)
    yield r
    r.close()


@pytest.mark.parametrize("mode", [PT_MODE_PLAIN, PT_MODE_WAVEFRONT])
@pytest.mark.parametrize("kind", KINDS)
def test_feature_parity(renderer, kind, mode):
    sc, rs = scenes.feature_scene(kind)
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
    print(f"{kind} mode={mode}: pixels within 1e-4 = {within:.5f}, worst {e.max():.3g}, "
          f"mean gpu {mg:.6f} oracle {mr:.6f}")
    assert np.isfinite(g).all()
    assert within == 1.0, within  # every pixel (measured: every scene, both forms)
    assert abs(mg - mr) <= 1e-4 * abs(mr)


SCENE_FILES = ["feature", "camera_matrix"]


@pytest.mark.parametrize("name", SCENE_FILES)
def test_scene_file_render_parity(renderer, name):
    """A .scene file end to end: pts_load_scene_file -> upload -> plain-kernel
    render, against the oracle on the same arrays and options (the fixture
    scene turns on every option the file format has)."""
    import os
    from glslpt.scene import Scene
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes", name + ".scene")
    sc = Scene()
    ro = sc.LoadSceneFile(path)
    sc.ProcessScene()
    rs = scenes.settings_from_render_options(ro, width=80, height=60, spp=4)
    renderer.upload_scene(sc)
    renderer.set_options(options_for(rs, mode=PT_MODE_PLAIN))
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
    print(f"{name}: pixels within 1e-4 = {within:.5f}, mean gpu {mg:.6f} oracle {mr:.6f}")
    assert np.isfinite(g).all()
    assert within == 1.0
    assert abs(mg - mr) <= 1e-4 * abs(mr)


@pytest.mark.parametrize("kind", KINDS)
def test_plain_equals_wavefront_features(renderer, kind):
    """Both kernel forms give bit-identical images on every option path (two
    NEE rays per vertex with an environment map, media, alpha masks,
    textures; stochastic alpha and volume MIS with NEE in the shade stage)."""
    sc, rs = scenes.feature_scene(kind)
    renderer.upload_scene(sc)
    out = []
    for mode in (PT_MODE_PLAIN, PT_MODE_WAVEFRONT):
        renderer.set_options(options_for(rs, mode=mode, wave_spp=3))
        renderer.set_camera(sc.camera())
        renderer.clear()
        renderer.render(0, 6)
        renderer.synchronize()
        out.append(renderer.read_accum())
    d = np.abs(out[0] - out[1]).max()
    print(f"{kind}: plain vs wavefront max abs diff {d}")
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("kind", ["alpha_blend", "volmis"])
def test_batch_equals_single_render_inline_nee(renderer, kind):
    """The batched form with NEE resolved in the shade stage: three jobs of
    the scene from their own cameras and depths, each bit-identical to a
    single render of that job."""
    sc, rs = scenes.feature_scene(kind)
    renderer.upload_scene(sc)
    cam = sc.camera()
    jobs = [dict(camera=cam, width=48, height=32, spp=3, max_depth=4),
            dict(camera=cam, width=40, height=40, spp=2, max_depth=6),
            dict(camera=cam, width=64, height=24, spp=4, max_depth=2)]
    renderer.set_options(options_for(rs, wave_spp=2, use_graph=1))
    outs = renderer.batch_render(jobs)
    for j, (job, out) in enumerate(zip(jobs, outs)):
        renderer.set_options(options_for(rs, width=job["width"], height=job["height"], max_depth=job["max_depth"],
                                         wave_spp=2))
        renderer.set_camera(job["camera"])
        renderer.clear()
        renderer.render(0, job["spp"])
        renderer.synchronize()
        ref = renderer.read_accum()
        print(f"{kind} job {j}: max abs diff {np.abs(out - ref).max()}")
        assert np.array_equal(out, ref), f"{kind} job {j}"

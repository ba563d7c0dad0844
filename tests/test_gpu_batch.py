"""Batched render of uneven jobs (pt_batch_render): every job's image is
identical to rendering that job alone with pt_render (same seeds, same order
of accumulation), whatever the pass packing, launch shape or graph capture,
in the wavefront form and in the one-launch plain form; pt_render itself is
pinned to the CPU oracle by test_gpu_parity.py."""
import numpy as np
import pytest

from glslpt import scenes
from glslpt._capi import PT_MODE_PLAIN, PT_MODE_WAVEFRONT
from glslpt.renderer import Renderer, options_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def renderer():
    r = Renderer(0)
    yield r
    r.close()


def single(r, sc, rs, job, **kw):
    r.set_options(options_for(rs, width=job["width"], height=job["height"], max_depth=job["max_depth"], **kw))
    r.set_camera(job["camera"])
    r.clear()
    r.render(0, job["spp"])
    r.synchronize()
    return r.read_accum()


def small_jobs():
    return scenes.batch_jobs(num_jobs=7, resolutions=((48, 48), (64, 32), (96, 80), (40, 72)),
                             depth_range=(1, 7), spp_range=(1, 6))


@pytest.mark.parametrize("graph,wave_spp,bs,stack,cadence,cap", [(0, 1, 256, 24, 1, 0), (1, 1, 64, 16, 2, 0),
                                                                  (1, 3, 128, 32, 1, 0), (1, 64, 256, 2, 3, 0),
                                                                  (1, 3, 64, 16, 1, 2), (0, 4, 256, 24, 2, 6)])
def test_batch_equals_single_renders(renderer, graph, wave_spp, bs, stack, cadence, cap):
    """cap: pt_options.extend_cap of the batch (its resume launches captured
    in the pass graph); the single renders run uncapped."""
    sc, rs, jobs = small_jobs()
    renderer.upload_scene(sc)
    kw = dict(block_size=bs, stack_depth=stack, compact_every=cadence)
    renderer.set_options(options_for(rs, wave_spp=wave_spp, use_graph=graph, extend_cap=cap, **kw))
    outs = renderer.batch_render(jobs)
    st = renderer.stats()
    print(f"graph={graph} wave_spp={wave_spp}: {st['ms_last_render']:.3f} ms")
    for j, (job, out) in enumerate(zip(jobs, outs)):
        ref = single(renderer, sc, rs, job, wave_spp=2, **kw)
        assert np.isfinite(out).all()
        assert np.array_equal(out, ref), f"job {j}: max diff {np.abs(out - ref).max()}"
        assert out[..., 3].max() == job["spp"]


def test_batch_rejects_invalid_jobs(renderer):
    from glslpt._capi import PtError
    sc, rs, jobs = small_jobs()
    renderer.upload_scene(sc)
    renderer.set_options(options_for(rs))
    bad = dict(jobs[0], spp=0)
    with pytest.raises(PtError):
        renderer.batch_render([jobs[1], bad])


def test_batch_full_size_equals_single_renders(renderer):
    """The full 32-job batch (resolutions up to 1920x1080, depth 2-12, 4-32
    spp, ~3.6e8 samples) in graph-captured passes, against single renders."""
    sc, rs, jobs = scenes.batch_jobs()
    renderer.upload_scene(sc)
    renderer.set_options(options_for(rs, wave_spp=8, use_graph=1, block_size=64, stack_depth=16))
    outs = renderer.batch_render(jobs)
    ms = renderer.stats()["ms_last_render"]
    total = sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    print(f"32 jobs, {total / 1e6:.1f} Msamples in {ms:.1f} ms = {total / ms / 1e3:.0f} Msamples/s")
    for j, (job, out) in enumerate(zip(jobs, outs)):
        ref = single(renderer, sc, rs, job, wave_spp=job["spp"], block_size=64, stack_depth=16)
        assert np.array_equal(out, ref), f"job {j}"


@pytest.mark.parametrize("bs,stack", [(256, 24), (64, 16), (128, 2)])
def test_plain_batch_equals_single_renders(renderer, bs, stack):
    """The one-launch plain form (each thread one pixel, its samples in order)
    against single wavefront renders of each job."""
    sc, rs, jobs = small_jobs()
    renderer.upload_scene(sc)
    renderer.set_options(options_for(rs, mode=PT_MODE_PLAIN, block_size=bs, stack_depth=stack))
    outs = renderer.batch_render(jobs)
    for j, (job, out) in enumerate(zip(jobs, outs)):
        ref = single(renderer, sc, rs, job, mode=PT_MODE_WAVEFRONT, wave_spp=2, block_size=bs, stack_depth=stack)
        assert np.array_equal(out, ref), f"job {j}: max diff {np.abs(out - ref).max()}"


def test_plain_batch_full_size_equals_single_renders(renderer):
    """The full 32-job batch in one plain launch, against single plain renders."""
    sc, rs, jobs = scenes.batch_jobs()
    renderer.upload_scene(sc)
    renderer.set_options(options_for(rs, mode=PT_MODE_PLAIN, block_size=64, stack_depth=16))
    outs = renderer.batch_render(jobs)
    ms = renderer.stats()["ms_last_render"]
    total = sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    print(f"plain batch: 32 jobs, {total / 1e6:.1f} Msamples in {ms:.1f} ms = {total / ms / 1e3:.0f} Msamples/s")
    for j, (job, out) in enumerate(zip(jobs, outs)):
        ref = single(renderer, sc, rs, job, mode=PT_MODE_PLAIN, block_size=64, stack_depth=16)
        assert np.array_equal(out, ref), f"job {j}"

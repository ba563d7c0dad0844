"""Launch-configuration autotuner for the wavefront renderer.

Sweeps workgroup size {64, 128, 256}, LDS traversal-stack depth {16, 24, 32},
samples per wavefront pass, queue-compaction cadence, BLAS node width {2, 4},
where NEE rays are traced (shadow stage or inside shade) and the closest-hit
step cap of later bounces (0 or 32; the 1080p workload only) on the actual
workload, timing it on the GPU, and stores the fastest configuration as JSON
in glslpt/tuning/<key>.json, where bench.py and callers look it up.

    python -m glslpt.autotune c2 [spp] [--quick] [--out DIR]   # 1080p instanced spheres, pt_render
    python -m glslpt.autotune c3 [--quick] [--out DIR]         # 32 uneven Cornell jobs, pt_batch_render
                                                               # with graph-captured passes
--out writes the JSON to DIR instead of glslpt/tuning/.
"""
import itertools
import json
import os
import sys
import time

from . import _capi as api
from .renderer import Renderer, options_for

TUNING_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning")

BLOCK_SIZES = (64, 128, 256)
STACK_DEPTHS = (16, 24, 32)
QUICK = dict(block_size=(64, 128), stack_depth=(16, 24), wave_spp=(16, 32, 64), compact_every=(1, 2),
             bvh_width=(2, 4), nee_inline=(0, 1), extend_cap=(0, 32))
WAVE_SPP = (8, 16, 32, 64)
CADENCE = (1, 2)
BVH_WIDTHS = (2, 4)
NEE_INLINE = (0, 1)
EXTEND_CAP = (0, 32)


def tuned(key):
    """The stored configuration for `key`, or None."""
    path = os.path.join(TUNING_DIR, f"{key}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["best"]


def _time(r, sc, rs, spp, reps, **kw):
    r.set_options(options_for(rs, mode=api.PT_MODE_WAVEFRONT, **kw))
    r.set_camera(sc.camera())
    r.clear()
    r.render(0, spp)  # warm-up (allocations, first touch)
    r.synchronize()
    t0 = time.perf_counter()
    for i in range(reps):
        r.render((i + 1) * spp, spp)
    r.synchronize()
    return (time.perf_counter() - t0) / reps


def sweep(r, sc, rs, spp=16, reps=2, grid=None, log=print):
    grid = grid or dict(block_size=BLOCK_SIZES, stack_depth=STACK_DEPTHS, wave_spp=WAVE_SPP, compact_every=CADENCE,
                        bvh_width=BVH_WIDTHS, nee_inline=NEE_INLINE, extend_cap=EXTEND_CAP)
    keys = list(grid)
    results = []
    for vals in itertools.product(*[grid[k] for k in keys]):
        cfg = dict(zip(keys, vals))
        if cfg.get("wave_spp", 1) > spp:
            continue
        dt = _time(r, sc, rs, spp, reps, **cfg)
        msps = rs.width * rs.height * spp / dt / 1e6
        results.append({"config": cfg, "ms": dt * 1e3, "msamples_per_s": msps})
        log(json.dumps(results[-1]))
    results.sort(key=lambda x: -x["msamples_per_s"])
    return results


def _time_batch(r, rs, jobs, reps, **kw):
    r.set_options(options_for(rs, mode=api.PT_MODE_WAVEFRONT, use_graph=1, **kw))
    r.batch_render(jobs)  # warm-up
    ms = []
    for _ in range(reps):
        r.batch_render(jobs)
        ms.append(r.stats()["ms_last_render"])
    return min(ms) / 1e3


def sweep_batch(r, rs, jobs, reps=1, grid=None, log=print):
    grid = grid or dict(block_size=BLOCK_SIZES, stack_depth=STACK_DEPTHS, wave_spp=(2, 4, 8, 16),
                        compact_every=(1, 2, 3), bvh_width=BVH_WIDTHS, nee_inline=NEE_INLINE)
    keys = list(grid)
    total = sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    results = []
    for vals in itertools.product(*[grid[k] for k in keys]):
        cfg = dict(zip(keys, vals))
        dt = _time_batch(r, rs, jobs, reps, **cfg)
        results.append({"config": cfg, "ms": dt * 1e3, "msamples_per_s": total / dt / 1e6})
        log(json.dumps(results[-1]))
    results.sort(key=lambda x: -x["msamples_per_s"])
    return results


def main(argv):
    if len(argv) < 1 or argv[0] not in ("c2", "c3"):
        sys.exit("usage: python -m glslpt.autotune c2 [spp] [--quick] [--out DIR] | c3 [--quick] [--out DIR]")
    out_dir = TUNING_DIR
    if "--out" in argv:
        out_dir = argv[argv.index("--out") + 1]
        argv = [a for i, a in enumerate(argv) if a != "--out" and (i == 0 or argv[i - 1] != "--out")]
    if argv[0] == "c3":
        from .scenes import batch_jobs
        sc, rs, jobs = batch_jobs()
        r = Renderer(0)
        r.upload_scene(sc)
        grid = dict(block_size=(64, 256), stack_depth=(16, 24), wave_spp=(4, 8, 16), compact_every=(1, 2),
                    nee_inline=NEE_INLINE) if "--quick" in argv else None
        res = sweep_batch(r, rs, jobs, reps=2, grid=grid)
        os.makedirs(out_dir, exist_ok=True)
        with open(os.path.join(out_dir, "c3.json"), "w") as f:
            json.dump({"key": "c3", "best": dict(res[0]["config"], use_graph=1), "results": res}, f, indent=1)
        print("best", json.dumps(res[0]))
        return
    spp = int(argv[1]) if len(argv) > 1 and argv[1].isdigit() else 64
    grid = QUICK if "--quick" in argv else None
    from .scenes import instanced_spheres
    sc, rs = instanced_spheres(num_instances=64, freq=28, seed=0x5EED0002, width=1920, height=1080, spp=64, max_depth=8)
    r = Renderer(0)
    r.upload_scene(sc)
    res = sweep(r, sc, rs, spp=spp, reps=1, grid=grid)
    os.makedirs(out_dir, exist_ok=True)
    out = {"key": "c2", "spp_timed": spp, "best": res[0]["config"], "results": res}
    with open(os.path.join(out_dir, "c2.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("best", json.dumps(res[0]))


if __name__ == "__main__":
    main(sys.argv[1:])

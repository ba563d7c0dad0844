"""Workloads for profiler runs.

    python profiling.py copy NBYTES
    python profiling.py heavy [SPP]
    python profiling.py edits
    python profiling.py l2read [MB] [ROWS] [REPS]
    python profiling.py c2 [SPP]

copy: copies NBYTES device-to-device with hipMemcpy (one blit-kernel dispatch
that reads and writes NBYTES). Profiled with `rocprofv3 --pmc FETCH_SIZE`
(then WRITE_SIZE), its per-dispatch counter values give the factor that turns
the counters of the path's kernels into bytes of HBM traffic.

heavy: one wavefront frame of SPP samples (default 4) of the heavy mixed
scene at 3840x2160, depth 12, then an instrumented frame of the same samples
(k_extend<true, ...> kernels, so a kernel filter on "k_extend<false" selects
the plain frame). Prints k_extend's time and algorithmic bytes (64 B per node,
36 B per triangle, 52 B per instance, 60 B per light record, 56 B per ray).

edits: the device edit paths on the heavy mixed scene: pt_set_env_map of its
2048x1024 sky (k_env_lum, the chunked CDF kernels k_env_chunk_sum, k_env_guess,
k_env_chunk_fun, k_env_carry, k_env_chunk_write, and the k_env_scan fallback)
four times, then twenty pt_rebuild_instances of its 1003 instances with
alternating transforms (k_tlas_prep, k_tlas_build). Prints the wall time of
each call.

l2read: the L2 read rate of this GPU, measured. A float32 tensor of MB MiB
(default 2: inside one XCD's 4 MiB L2) is summed over a stride-0 expansion of
ROWS rows (default 512), so each launch reads the same MB MiB ROWS times and,
after the first touch, every read is served by the L2 (a CU's L1 holds far less
than its share). REPS launches (default 50) timed with HIP events; prints the
read bytes per launch and GB/s. Profiled with `rocprofv3 --pmc TCC_REQ_sum`, the
same run gives the bytes one L2 request carries (read bytes / requests).

c2: one wavefront frame of SPP samples (default 16) of the 1080p headline
workload at the stored launch configuration; a kernel filter on "k_extend"
selects its closest-hit launches.
"""
import ctypes as C
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "glsl-pathtracer_amd"))


def copy(nbytes):
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        rc = hip.hipMalloc(C.byref(p), C.c_size_t(nbytes))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed: {rc}")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed: {rc}")
    print(f"copied {nbytes} bytes device to device")


def heavy(spp):
    import json
    from glslpt import _capi as api
    from glslpt.autotune import tuned
    from glslpt.renderer import Renderer, options_for
    from glslpt.scenes import heavy_mixed
    sc, rs = heavy_mixed(spp=spp)
    cfg = dict(block_size=256, stack_depth=24, wave_spp=4, compact_every=1)
    cfg.update(tuned("c2") or {})
    r = Renderer(0)
    r.upload_scene(sc)
    out = {}
    for count in (0, 1):
        r.set_options(options_for(rs, mode=api.PT_MODE_WAVEFRONT, count_work=count, **cfg))
        r.set_camera(sc.camera())
        r.clear()
        r.reset_stats()
        r.render(0, spp)
        r.synchronize()
        st = r.stats()
        if count == 0:
            out["extend_ms"] = st["ms_stage"]["extend"]
            out["extend_launches"] = st["launches_stage"]["extend"]
            out["frame_ms"] = st["ms_last_render"]
        else:
            out["rays"] = st["rays_closest"]
            out["algorithmic_bytes"] = (64 * st["nodes_visited"] + 36 * st["tris_tested"] + 52 * st["instances_entered"]
                                        + 60 * st["lights_tested"] + 56 * st["rays_closest"])
    r.close()
    out.update(spp=spp, width=rs.width, height=rs.height, launch=cfg)
    print(json.dumps(out))


def edits():
    import json
    import time

    import numpy as np

    from glslpt.renderer import Renderer
    from glslpt.scenes import heavy_mixed
    sc, rs = heavy_mixed()
    a = sc.arrays()
    r = Renderer(0)
    r.upload_scene(sc)
    sky = np.ascontiguousarray(a["env_img"], np.float32).reshape(a["env_size"][1], a["env_size"][0], 3)
    env_ms = []
    for _ in range(4):
        t0 = time.perf_counter()
        r.set_env_map(sky)
        env_ms.append(1e3 * (time.perf_counter() - t0))
    xf0 = np.ascontiguousarray(a["transforms"], np.float32).reshape(-1, 16)
    xf1 = xf0.copy()
    xf1[:, 12:15] += np.random.default_rng(0).uniform(-0.5, 0.5, (xf0.shape[0], 3)).astype(np.float32)
    tlas_ms = []
    for k in range(20):
        t0 = time.perf_counter()
        r.rebuild_instances(xf1 if k % 2 == 0 else xf0)
        tlas_ms.append(1e3 * (time.perf_counter() - t0))
    r.close()
    print(json.dumps({"env": {"texels": int(sky.shape[0] * sky.shape[1]), "set_env_map_ms": env_ms},
                      "tlas": {"instances": int(xf0.shape[0]), "rebuild_instances_ms": tlas_ms}}))


def l2read(mb=2, rows=512, reps=50):
    import json

    import torch
    n = mb * (1 << 20) // 4
    x = torch.ones(n, device="cuda", dtype=torch.float32)
    y = x.unsqueeze(0).expand(rows, n)
    for _ in range(3):
        y.sum()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        y.sum()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = rows * n * 4
    print(json.dumps({"bytes_per_launch": nbytes, "ms_per_launch": ms, "gb_per_s": nbytes / ms / 1e6,
                      "mb": mb, "rows": rows, "reps": reps}))


def c2(spp=16):
    import json
    from glslpt._capi import PT_MODE_WAVEFRONT
    from glslpt.autotune import tuned
    from glslpt.renderer import Renderer, options_for
    from glslpt.scenes import instanced_spheres
    sc, rs = instanced_spheres()
    cfg = dict(block_size=256, stack_depth=24, wave_spp=4, compact_every=1)
    cfg.update(tuned("c2") or {})
    r = Renderer(0)
    r.upload_scene(sc)
    r.set_options(options_for(rs, mode=PT_MODE_WAVEFRONT, **cfg))
    r.set_camera(sc.camera())
    r.clear()
    r.reset_stats()
    r.render(0, spp)
    r.synchronize()
    st = r.stats()
    r.close()
    print(json.dumps({"spp": spp, "launch": cfg, "ms_last_render": st["ms_last_render"],
                      "stage_ms": {k: v for k, v in st["ms_stage"].items() if v > 0},
                      "launches": {k: v for k, v in st["launches_stage"].items() if v > 0}}))


if __name__ == "__main__":
    if len(sys.argv) == 3 and sys.argv[1] == "copy":
        copy(int(sys.argv[2]))
    elif len(sys.argv) in (2, 3) and sys.argv[1] == "heavy":
        heavy(int(sys.argv[2]) if len(sys.argv) == 3 else 4)
    elif len(sys.argv) == 2 and sys.argv[1] == "edits":
        edits()
    elif len(sys.argv) >= 2 and sys.argv[1] == "l2read":
        l2read(*[int(a) for a in sys.argv[2:5]])
    elif len(sys.argv) in (2, 3) and sys.argv[1] == "c2":
        c2(int(sys.argv[2]) if len(sys.argv) == 3 else 16)
    else:
        sys.exit("usage: python profiling.py copy NBYTES | heavy [SPP] | edits | l2read [MB] [ROWS] [REPS] | c2 [SPP]")

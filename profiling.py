"""Workloads for profiler runs.

    python profiling.py copy NBYTES
    python profiling.py heavy [SPP]
    python profiling.py edits
    python profiling.py l2read [MB] [STEPS] [REPS]
    python profiling.py valu [STEPS] [REPS]
    python profiling.py summarize ROOT OUT_JSON
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

l2read: the L2 read rate of this GPU, measured with the probe kernel l2_read
(below): 8192 workgroups of 256 threads, each wave reading STEPS x 4 chunks of
1 KiB scattered over an MB MiB buffer (default 2: inside one XCD's 4 MiB L2,
so after the first touch every read is served by the L2; a CU's 32 KiB L1
holds under 2% of it). REPS launches (default 20) timed with HIP events;
prints the read bytes per launch and GB/s. Profiled with `rocprofv3 --pmc
TCC_REQ_sum`, the same run gives the bytes one L2 request carries.

valu: the VALU issue rate of this GPU, measured with the probe kernel
valu_fma: 8192 x 256 threads, each running eight independent fp32 FMA chains
for STEPS steps (default 4096); prints wave-instructions and lane FMAs per
second. Profiled with the SQ counters, it shows what SQ_INSTS_VALU,
SQ_ACTIVE_INST_VALU and SQ_THREAD_CYCLES_VALU read for a kernel whose VALU
count is known.

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
    cfg.update(tuned("c5") or {"bvh_width": 4, "extend_cap": 0})  # as bench.py's heavy leg
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


# Two probe kernels, compiled at run time with hiprtc for this GPU and launched
# through the HIP module API (they are measuring instruments, not part of the
# renderer):
#   l2_read   each wave reads four consecutive 1 KiB chunks (64 lanes x 16 B)
#             of an L2-resident buffer per step, starting at chunk
#             (w * 7919 + i * 104729) mod nchunks and wrapping; every address is
#             inside the buffer (chunk < nchunks, lane < 64).
#   valu_fma  eight independent chains of fp32 FMA per thread (no packed
#             instructions: built with the packed-FP32 feature off), so the
#             VALU issue rate bounds it.
PROBE_SRC = r"""
extern "C" __global__ void __launch_bounds__(256) l2_read(const float4* __restrict__ buf, unsigned nchunks,
                                                          unsigned stride, int steps, float4* __restrict__ out) {
    // chunk indices are wave-uniform (scalar registers); stride < nchunks
    const unsigned wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const unsigned lane = threadIdx.x & 63u;
    float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), a1 = a0, a2 = a0, a3 = a0;
    unsigned c = (wave * 7919u) % nchunks;
    for (int i = 0; i < steps; ++i) {
        const unsigned c1 = c + 1u < nchunks ? c + 1u : c + 1u - nchunks;
        const unsigned c2 = c1 + 1u < nchunks ? c1 + 1u : c1 + 1u - nchunks;
        const unsigned c3 = c2 + 1u < nchunks ? c2 + 1u : c2 + 1u - nchunks;
        const float4 v0 = buf[c * 64u + lane], v1 = buf[c1 * 64u + lane];
        const float4 v2 = buf[c2 * 64u + lane], v3 = buf[c3 * 64u + lane];
        a0.x += v0.x; a0.y += v0.y; a0.z += v0.z; a0.w += v0.w;
        a1.x += v1.x; a1.y += v1.y; a1.z += v1.z; a1.w += v1.w;
        a2.x += v2.x; a2.y += v2.y; a2.z += v2.z; a2.w += v2.w;
        a3.x += v3.x; a3.y += v3.y; a3.z += v3.z; a3.w += v3.w;
        c = c + stride < nchunks ? c + stride : c + stride - nchunks;
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] =
        make_float4(a0.x + a1.x + a2.x + a3.x, a0.y + a1.y + a2.y + a3.y, a0.z + a1.z + a2.z + a3.z,
                    a0.w + a1.w + a2.w + a3.w);
}
extern "C" __global__ void __launch_bounds__(256) valu_fma(float* __restrict__ out, int steps, float a, float b) {
    float x0 = threadIdx.x * 1e-3f, x1 = x0 + 1.f, x2 = x0 + 2.f, x3 = x0 + 3.f;
    float x4 = x0 + 4.f, x5 = x0 + 5.f, x6 = x0 + 6.f, x7 = x0 + 7.f;
    for (int i = 0; i < steps; ++i) {
        x0 = __builtin_fmaf(x0, a, b); x1 = __builtin_fmaf(x1, a, b);
        x2 = __builtin_fmaf(x2, a, b); x3 = __builtin_fmaf(x3, a, b);
        x4 = __builtin_fmaf(x4, a, b); x5 = __builtin_fmaf(x5, a, b);
        x6 = __builtin_fmaf(x6, a, b); x7 = __builtin_fmaf(x7, a, b);
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = ((x0 + x1) + (x2 + x3)) + ((x4 + x5) + (x6 + x7));
}
"""


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {rc}")


def probe_code(arch="gfx950"):
    """Code object of PROBE_SRC for arch, from hiprtc."""
    rtc = C.CDLL("libhiprtc.so")
    prog = C.c_void_p()
    _check(rtc.hiprtcCreateProgram(C.byref(prog), PROBE_SRC.encode(), b"probe.hip", 0, None, None),
           "hiprtcCreateProgram")
    opts = [f"--offload-arch={arch}".encode(), b"-O3", b"-Xclang", b"-target-feature", b"-Xclang",
            b"-packed-fp32-ops"]
    arr = (C.c_char_p * len(opts))(*opts)
    rc = rtc.hiprtcCompileProgram(prog, len(opts), arr)
    if rc != 0:
        n = C.c_size_t()
        rtc.hiprtcGetProgramLogSize(prog, C.byref(n))
        log = C.create_string_buffer(n.value + 1)
        rtc.hiprtcGetProgramLog(prog, log)
        raise RuntimeError("hiprtc: " + log.value.decode(errors="replace"))
    n = C.c_size_t()
    _check(rtc.hiprtcGetCodeSize(prog, C.byref(n)), "hiprtcGetCodeSize")
    code = C.create_string_buffer(n.value)
    _check(rtc.hiprtcGetCode(prog, code), "hiprtcGetCode")
    rtc.hiprtcDestroyProgram(C.byref(prog))
    return code.raw


class _Probe:
    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.code = probe_code()
        self.mod = C.c_void_p()
        _check(self.hip.hipModuleLoadData(C.byref(self.mod), self.code), "hipModuleLoadData")

    def fn(self, name):
        f = C.c_void_p()
        _check(self.hip.hipModuleGetFunction(C.byref(f), self.mod, name.encode()), "hipModuleGetFunction")
        return f

    def malloc(self, nbytes):
        p = C.c_void_p()
        _check(self.hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)), "hipMalloc")
        return p

    def time(self, f, grid, block, args, reps):
        """Average ms of reps launches of f (after one untimed launch)."""
        params = (C.c_void_p * len(args))(*[C.cast(C.byref(a), C.c_void_p) for a in args])
        launch = lambda: _check(self.hip.hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, None, params, None),
                                "hipModuleLaunchKernel")
        launch()
        _check(self.hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
        e0, e1 = C.c_void_p(), C.c_void_p()
        self.hip.hipEventCreate(C.byref(e0))
        self.hip.hipEventCreate(C.byref(e1))
        self.hip.hipEventRecord(e0, None)
        for _ in range(reps):
            launch()
        self.hip.hipEventRecord(e1, None)
        _check(self.hip.hipEventSynchronize(e1), "hipEventSynchronize")
        ms = C.c_float()
        self.hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        return ms.value / reps


def l2read(mb=2, steps=256, reps=20):
    """L2 read rate: every wave of a full grid reads 1 KiB chunks of an MB MiB
    buffer (default 2: inside one XCD's 4 MiB L2, so after the first touch
    each XCD's L2 serves it), four per step."""
    import json
    pr = _Probe()
    nbytes = mb << 20
    nchunks = nbytes // 1024
    buf = pr.malloc(nbytes)
    _check(pr.hip.hipMemset(buf, 0, C.c_size_t(nbytes)), "hipMemset")
    blocks, threads = 8192, 256
    out = pr.malloc(blocks * threads * 16)
    ms = pr.time(pr.fn("l2_read"), blocks, threads,
                 [buf, C.c_uint(nchunks), C.c_uint(104729 % nchunks), C.c_int(steps), out], reps)
    read = blocks * threads // 64 * steps * 4 * 1024
    print(json.dumps({"kernel": "l2_read", "buffer_bytes": nbytes, "read_bytes_per_launch": read,
                      "ms_per_launch": ms, "gb_per_s": read / ms / 1e6, "blocks": blocks, "threads": threads,
                      "steps": steps, "reps": reps}))


def valu(steps=4096, reps=20):
    """VALU issue rate: eight independent fp32 FMA chains per thread over a
    full grid; prints wave-instructions and lane FMAs per second (the loop's
    own scalar instructions are not VALU)."""
    import json
    pr = _Probe()
    blocks, threads = 8192, 256
    out = pr.malloc(blocks * threads * 4)
    ms = pr.time(pr.fn("valu_fma"), blocks, threads, [out, C.c_int(steps), C.c_float(0.999), C.c_float(1e-3)], reps)
    fma = blocks * threads * steps * 8
    print(json.dumps({"kernel": "valu_fma", "lane_fma_per_launch": fma, "wave_fma_per_launch": fma // 64,
                      "ms_per_launch": ms, "lane_fma_per_s": fma / ms * 1e3, "tflops_fp32": 2 * fma / ms / 1e9,
                      "wave_instr_per_s": fma / 64 / ms * 1e3, "blocks": blocks, "threads": threads,
                      "steps": steps, "reps": reps}))


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


def _counters(path):
    """{kernel name: {dispatch id: {counter: value}}} of a rocprofv3 --pmc csv."""
    import csv
    out = {}
    for row in csv.DictReader(open(path)):
        name = row["Kernel_Name"].split("(")[0]
        d = out.setdefault(name, {}).setdefault(int(row["Dispatch_Id"]), {})
        d[row["Counter_Name"]] = d.get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
    return out


def _one(root, sub):
    import glob
    files = glob.glob(os.path.join(root, sub, "*_counter_collection.csv"))
    if len(files) != 1:
        raise RuntimeError(f"expected one counter csv under {root}/{sub}, found {files}")
    return _counters(files[0])


MI355X_SIMDS = 256 * 4  # 256 CUs of four SIMDs


def summarize(root, out_json):
    """Counter summary of the 1080p frame's closest-hit stage and the probes,
    from the rocprofv3 runs under ROOT:
      fetch, write, tcc, sq   --pmc FETCH_SIZE / WRITE_SIZE / TCC_REQ_sum / SQ
                              counters of `profiling.py c2 64`, k_extend only
      copy_fetch, copy_write  FETCH_SIZE / WRITE_SIZE of `profiling.py copy 1073741824`
      l2req                   TCC_REQ_sum of `profiling.py l2read 2 256 3`
      valusq                  SQ counters and GRBM_GUI_ACTIVE of `profiling.py valu 4096 3`
      l2rate.json, valu.json  the probes' own timed lines (no profiler)
      c2.json                 the line of `profiling.py c2 64` without a profiler
                              (launch configuration, stage times)
    Writes OUT_JSON with the per-step averages (one step = one closest-hit
    stage launch of the renderer, which may run two kernels) the bench line
    uses."""
    import json
    # every k_extend* dispatch of the frame: the camera-ray launch, and for later
    # bounces the (capped) closest-hit launch and, with a step cap, its resume
    # launch; figures are per closest-hit stage step (the stage's launches in
    # the unprofiled frame: one camera-ray step, the rest later bounces)
    frame = json.load(open(os.path.join(root, "c2.json")))
    steps = frame["launches"]["extend"]
    kinds = {"camera": {"kernels": set(), "dispatches": 0, "steps": 1, "sums": {}},
             "later": {"kernels": set(), "dispatches": 0, "steps": steps - 1, "sums": {}}}
    seen = set()
    for sub in ("fetch", "write", "tcc", "sq"):
        for name, ds in _one(root, sub).items():
            kind = "camera" if "k_extend<false, true, false" in name else "later"
            k = kinds[kind]
            k["kernels"].add(name)
            for did, cv in ds.items():
                if (sub, did) in seen:
                    continue
                seen.add((sub, did))
                if sub == "tcc":
                    k["dispatches"] += 1
                for c, v in cv.items():
                    k["sums"][c] = k["sums"].get(c, 0.0) + v
    cf = _one(root, "copy_fetch")["__amd_rocclr_copyBuffer"]
    cw = _one(root, "copy_write")["__amd_rocclr_copyBuffer"]
    nbytes = 1 << 30
    fetch_scale = nbytes / (next(iter(cf.values()))["FETCH_SIZE"] * 1024)
    write_scale = nbytes / (next(iter(cw.values()))["WRITE_SIZE"] * 1024)
    l2 = json.load(open(os.path.join(root, "l2rate.json")))
    vl = json.load(open(os.path.join(root, "valu.json")))
    l2req = _one(root, "l2req")["l2_read"]
    req = sum(d["TCC_REQ_sum"] for d in l2req.values()) / len(l2req)
    bytes_per_req = l2["read_bytes_per_launch"] / req
    vsq = _one(root, "valusq")["valu_fma"]
    vsq = {c: sum(d[c] for d in vsq.values()) / len(vsq) for c in next(iter(vsq.values()))}
    summary = {"workload": "1920x1080, 64 spp, depth 8 (64 instances of a 15680-triangle sphere)",
               "launch": frame["launch"], "stage": "k_extend (closest hit)", "unprofiled_frame": frame,
               "kinds": {}, "frame": {},
               "calibration": {"fetch_scale": fetch_scale, "write_scale": write_scale,
                               "l2_bytes_per_request": bytes_per_req, "l2_read_gb_per_s": l2["gb_per_s"],
                               "l2_probe": l2, "valu_probe": vl, "valu_probe_counters": vsq,
                               "valu_wave_instr_per_s": vl["wave_instr_per_s"],
                               "valu_lane_op_per_s": vl["lane_fma_per_s"]}}
    tot = {}
    for kind, k in kinds.items():
        n, sm = k["steps"], k["sums"]
        per = {c: v / n for c, v in sm.items()}
        per["hbm_bytes"] = per["FETCH_SIZE"] * 1024 * fetch_scale + per["WRITE_SIZE"] * 1024 * write_scale
        per["l2_bytes"] = per["TCC_REQ_sum"] * bytes_per_req
        per["lane_activity"] = sm["SQ_THREAD_CYCLES_VALU"] / (64 * sm["SQ_ACTIVE_INST_VALU"])
        # occupancy: SQ_WAVE_CYCLES counts resident wave-cycles in units of 4 cycles; GRBM_GUI_ACTIVE
        # sums the busy cycles of the 8 XCDs
        per["waves_per_simd"] = 4 * sm["SQ_WAVE_CYCLES"] / (sm["GRBM_GUI_ACTIVE"] / 8 * MI355X_SIMDS)
        per["waves_per_simd_launched"] = sm["SQ_WAVES"] / k["dispatches"] / MI355X_SIMDS
        summary["kinds"][kind] = {"kernels": sorted(k["kernels"]), "dispatches": k["dispatches"], "steps": n,
                                  "per_step": per}
        for c in ("hbm_bytes", "l2_bytes", "SQ_INSTS_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_ACTIVE_INST_VALU"):
            tot[c] = tot.get(c, 0.0) + per[c] * n
    summary["frame"] = {"steps": steps, "per_step": {c: v / steps for c, v in tot.items()},
                        "lane_activity": tot["SQ_THREAD_CYCLES_VALU"] / (64 * tot["SQ_ACTIVE_INST_VALU"])}
    with open(out_json, "w") as f:
        json.dump(summary, f, indent=1)
    print(json.dumps(summary["frame"]))


if __name__ == "__main__":
    if len(sys.argv) == 3 and sys.argv[1] == "copy":
        copy(int(sys.argv[2]))
    elif len(sys.argv) in (2, 3) and sys.argv[1] == "heavy":
        heavy(int(sys.argv[2]) if len(sys.argv) == 3 else 4)
    elif len(sys.argv) == 2 and sys.argv[1] == "edits":
        edits()
    elif len(sys.argv) >= 2 and sys.argv[1] == "l2read":
        l2read(*[int(a) for a in sys.argv[2:5]])
    elif len(sys.argv) == 4 and sys.argv[1] == "summarize":
        summarize(sys.argv[2], sys.argv[3])
    elif len(sys.argv) >= 2 and sys.argv[1] == "valu":
        valu(*[int(a) for a in sys.argv[2:4]])
    elif len(sys.argv) in (2, 3) and sys.argv[1] == "c2":
        c2(int(sys.argv[2]) if len(sys.argv) == 3 else 16)
    else:
        sys.exit("usage: python profiling.py copy NBYTES | heavy [SPP] | edits | l2read [MB] [STEPS] [REPS] | "
                 "valu [STEPS] [REPS] | c2 [SPP] | summarize ROOT OUT_JSON")

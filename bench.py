"""Benchmark of the path-tracing hot path on one MI355X.

Workload (the headline configuration): 64 instances (random rigid transforms,
uniform scale 0.5-2) of a seeded geodesic icosphere with 15,680 triangles
displaced by 3-octave value noise (1,003,520 instanced triangles), 16 Disney
materials drawn from a seed, 2 rect lights + 1 sphere light + a constant
environment, 1920x1080, 64 spp, max depth 8, Russian roulette from depth 3.

A step renders all 64 samples of the frame. The reported value is the
wavefront renderer's throughput in Msamples/s (W*H*spp / step time); the plain
one-kernel-per-frame renderer is timed on the same frame beside it. The scene
is resident in HBM before the timed region starts.

roofline: the dominant kernel is the closest-hit (extend) stage. Its
algorithmic bytes per launch come from an instrumented run of the same frame
(count of internal nodes visited, triangles tested, instances entered, light
records read, rays): 64 B per node (two child boxes + references), 36 B per
triangle (three positions), 52 B per instance (inverse affine + BLAS root),
60 B per light record, 56 B per ray (queue entry, ray, depth, hit record).
achieved = bytes per launch / average launch time (HIP events on the
renderer's stream over the timed region); peak = 8 TB/s HBM.

adjoint: the second half of the metric, backward/forward time ratio, on the
differentiable-materials workload (256 boxes + 8 noise-displaced spheres,
199,072 triangles, 64 Disney materials, 512x512, 16 spp, max depth 4): forward
= one pt_render of the 16 samples, backward = one pt_adjoint of the same
samples for an L2 loss against a target rendered from a second parameter seed
(both device time between HIP events, median of --adjoint-reps). The backward
pass is the whole pt_adjoint call: a primal pass (the forward render, recording
NEE visibility) and the replay pass that accumulates the gradients.

batch: 32 uneven Cornell-box jobs in one pt_batch_render, in the wavefront
form (graph-captured passes) and the one-launch plain form, reported beside
the headline, and the same jobs run one after another (batch.sequential: one
pt_render per job, plain and wavefront).

cornell: the Cornell box (256x256, 4 spp, depth 5) in both kernel forms and
on the CPU oracle.

config.paths: from the instrumented frame, the fraction of camera rays that
hit the scene and the mean number of path vertices (closest hits shaded) per
sample.

heavy: the heavy mixed scene (1000 instances of 8 seeded meshes, ~24.5M
instanced triangles of which ~32% alpha-tested leaf quads, two scattering
volume cubes, an importance-sampled 2048x1024 HDR sky) at 3840x2160, 16 spp,
max depth 12: wavefront and plain throughput and k_extend's algorithmic
bandwidth (bytes from an instrumented 1-spp frame, scaled to the timed
frames) against HBM peak. Its ~260 MB of nodes and triangles exceed L2 and
MALL, yet L2 still serves ~96% of the closest-hit reads (rocprofv3 counters).
heavy.tlas_rebuild: an edit that moves all 1000 instances, as the device
rebuild (pt_rebuild_instances) against the host path it replaces (the scene
builder's RebuildInstances + pt_update_instances with the TLAS range), wall
time per call.
heavy.env_cdf: replacing the heavy scene's 2048x1024 sky, pt_set_env_map (image
upload + CDF built on the GPU) against the host CDF build of the scene library.

cpu_baseline: the CPU oracle (C++ restatement of the reference shaders,
OpenMP) on the host cores, rendering a bounded sample of the same frame
(every k-th row at 1 spp, k chosen for ~15 s of CPU work).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "glsl-pathtracer_amd"))

METRIC = "Msamples/s at 1080p depth 8 (batched vs plain kernel); backward/forward time ratio"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X vector fp32, packed
B_NODE, B_TRI, B_INST, B_LIGHT, B_RAY = 64, 36, 52, 60, 56
# Algorithmic flops of the same work, counted from the kernels' expressions: a BVH2-node
# equivalent is two slab tests (per box 6 subtractions, 6 products, 10 min/max, 2 compares); a
# triangle test 52 (two cross products, four dot products, one reciprocal, 3 products, 2
# subtractions, 5 compares); an instance entry 36 (the affine origin and linear direction
# transforms, 3 reciprocals); a light test 38 (the 1080p scene's mean of 2 rect tests at 44 and
# one sphere test at 25, every branch taken).
F_NODE, F_TRI, F_INST, F_LIGHT = 48, 52, 36, 38


def build_scene(args):
    from glslpt import scenes
    return scenes.instanced_spheres(num_instances=64, freq=28, seed=0x5EED0002, width=args.width,
                                    height=args.height, spp=args.spp, max_depth=8)


def run_mode(r, sc, rs, mode, steps, warmup, spp, opts_kw):
    from glslpt.renderer import options_for
    r.set_options(options_for(rs, mode=mode, **opts_kw))
    r.set_camera(sc.camera())
    r.clear()
    s0 = 0
    for _ in range(warmup):
        r.render(s0, spp)
        s0 += spp
    r.synchronize()
    r.reset_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.render(s0, spp)
        s0 += spp
    r.synchronize()
    t1 = time.perf_counter()
    st = r.stats()
    return (t1 - t0) / steps * 1e3, st


def count_work(r, sc, rs, spp, opts_kw):
    from glslpt._capi import PT_MODE_WAVEFRONT
    from glslpt.renderer import options_for
    r.set_options(options_for(rs, mode=PT_MODE_WAVEFRONT, count_work=1, **opts_kw))
    r.set_camera(sc.camera())
    r.clear()
    r.reset_stats()
    r.render(0, spp)
    r.synchronize()
    return r.stats()


def cpu_baseline(sc, rs, target_s):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle
    from common import oracle_setup
    sa, cam, o = oracle_setup(sc, rs)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    # probe: every 8th row at 1 spp; then size the sample (rows x spp) to ~target_s
    step = 8
    rows = len(range(0, rs.height, step))
    t0 = time.perf_counter()
    oracle.render(sa, cam, o, 0, 1, rows=(0, rs.height, step), threads=threads)
    per_sample = (time.perf_counter() - t0) / (rows * rs.width)
    want = target_s / max(per_sample, 1e-9)
    if want < rows * rs.width:
        step = max(1, int(rs.height * rs.width / want))
        spp = 1
    else:
        step = 1
        spp = max(1, int(want / (rs.height * rs.width)))
    rows = len(range(0, rs.height, step))
    t0 = time.perf_counter()
    oracle.render(sa, cam, o, 1, spp, rows=(0, rs.height, step), threads=threads)
    dt = time.perf_counter() - t0
    samples = rows * rs.width * spp
    return {"value": samples / dt / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port", "cpu": cpu_model(),
            "sample": f"{rows} rows (every {step}th) x {rs.width} px x {spp} spp of the same frame, {dt:.1f} s"}


def adjoint_ratio(reps, wave_spp):
    """Forward and adjoint device time on the differentiable-materials
    workload, L2 loss against a target from a second parameter seed."""
    import numpy as np
    from glslpt import scenes
    from glslpt.renderer import Renderer, options_for
    sc, rs = scenes.material_boxes()
    a = sc.arrays()
    target_mats = scenes.apply_params(a["materials"], scenes.diff_material_params(0xD1FF0005, a["materials"].shape[0]))
    r = Renderer(0)
    r.set_options(options_for(rs, wave_spp=wave_spp))
    r.set_camera(sc.camera())
    r.upload_scene(sc, materials=target_mats)
    r.clear()
    r.render(1 << 20, rs.spp)  # target: other samples, other parameters
    target = r.read_accum()[..., :3] / rs.spp
    r.upload_scene(sc)
    fwd, bwd = [], []
    loss = None
    for i in range(reps + 1):
        r.clear()
        r.render(0, rs.spp)
        r.synchronize()
        f_ms = r.stats()["ms_last_render"]
        img = r.read_accum()[..., :3] / rs.spp
        diff = img - target
        loss = float(np.mean(diff.astype(np.float64) ** 2))
        dl = (2.0 / diff.size) * diff
        g = r.adjoint(0, rs.spp, dl.astype(np.float32))
        st = r.stats()
        b_ms = st["ms_last_render"]
        dropped = st["adjoint_dropped"]
        if i > 0:  # the first round is warm-up
            fwd.append(f_ms)
            bwd.append(b_ms)
    r.close()
    f_med, b_med = float(np.median(fwd)), float(np.median(bwd))
    return {"workload": f"{int(a['vert_indices'].shape[0]) + 12 * 255} triangles (256 boxes + 8 noise-displaced "
                        f"spheres), {a['materials'].shape[0]} Disney materials x 7 parameters, "
                        f"{rs.width}x{rs.height}, {rs.spp} spp, max depth {rs.max_depth}",
            "forward_ms": f_med, "backward_ms": b_med, "backward_over_forward": b_med / f_med,
            "reps": reps, "loss": loss, "grad_norm": float(np.linalg.norm(g)), "dropped_vertices": dropped,
            "launch": {"wave_spp": wave_spp}}


def batch_throughput(reps):
    """The 32 uneven Cornell jobs (resolutions 128^2..1920x1080, depth 2-12,
    4-32 spp) in one pt_batch_render: the wavefront form (graph-captured
    passes, the c3 launch configuration) and the one-launch plain form; device
    time, best of `reps` after a warm-up."""
    from glslpt import scenes
    from glslpt._capi import PT_MODE_PLAIN, PT_MODE_WAVEFRONT
    from glslpt.autotune import tuned
    from glslpt.renderer import Renderer, options_for
    sc, rs, jobs = scenes.batch_jobs()
    cfg = dict(block_size=64, stack_depth=16, wave_spp=8, compact_every=1, use_graph=1)
    cfg.update(tuned("c3") or {})
    r = Renderer(0)
    r.upload_scene(sc)
    total = sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    out = {"workload": f"{len(jobs)} Cornell-box jobs, resolutions 128^2..1920x1080, depth 2-12, 4-32 spp",
           "samples": total}
    for name, mode in (("wavefront", PT_MODE_WAVEFRONT), ("plain", PT_MODE_PLAIN)):
        r.set_options(options_for(rs, mode=mode, **cfg))
        r.batch_render(jobs)
        ms = []
        for _ in range(reps):
            r.batch_render(jobs)
            ms.append(r.stats()["ms_last_render"])
        best = min(ms)
        out[name] = {"ms": best, "msamples_per_s": total / best / 1e3}
    out["launch"] = cfg
    r.close()
    return out


def batch_sequential(reps):
    """The same 32 jobs rendered one after another, one pt_render per job
    (plain: one kernel launch per sample; wavefront: the C3 launch
    configuration without batching). Each job's image is bit-identical to its
    batched image. Device time summed over the jobs, best of `reps`."""
    from glslpt import scenes
    from glslpt._capi import PT_MODE_PLAIN, PT_MODE_WAVEFRONT
    from glslpt.autotune import tuned
    from glslpt.renderer import Renderer, options_for
    sc, rs, jobs = scenes.batch_jobs()
    cfg = dict(block_size=64, stack_depth=16, wave_spp=8, compact_every=1)
    cfg.update({k: v for k, v in (tuned("c3") or {}).items() if k != "use_graph"})
    r = Renderer(0)
    r.upload_scene(sc)
    total = sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    out = {}
    for name, mode in (("plain", PT_MODE_PLAIN), ("wavefront", PT_MODE_WAVEFRONT)):
        best = None
        for _ in range(reps + 1):  # the first round is warm-up
            ms = 0.0
            for j in jobs:
                o = options_for(rs, mode=mode, **cfg)
                o.width, o.height, o.max_depth = j["width"], j["height"], j["max_depth"]
                r.set_options(o)
                r.set_camera(j["camera"])
                r.clear()
                r.render(0, j["spp"])
                r.synchronize()
                ms += r.stats()["ms_last_render"]
            best = ms if best is None else min(best, ms)
        out[name] = {"ms": best, "msamples_per_s": total / best / 1e3}
    r.close()
    out["samples"] = total
    return out


def cornell_throughput(steps):
    """The Cornell box (256x256, 4 spp, depth 5): both kernel forms, device
    time per frame (median of `steps`), and the CPU oracle on the host cores."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle
    from common import oracle_setup
    from glslpt import scenes
    from glslpt._capi import PT_MODE_PLAIN, PT_MODE_WAVEFRONT
    from glslpt.renderer import Renderer, options_for
    sc, rs = scenes.cornell()
    samples = rs.width * rs.height * rs.spp
    r = Renderer(0)
    r.upload_scene(sc)
    r.set_camera(sc.camera())
    out = {"workload": f"Cornell box, 34 triangles, {rs.width}x{rs.height}, {rs.spp} spp, max depth {rs.max_depth}"}
    for name, mode in (("plain", PT_MODE_PLAIN), ("wavefront", PT_MODE_WAVEFRONT)):
        r.set_options(options_for(rs, mode=mode, wave_spp=rs.spp))
        ms = []
        for k in range(steps + 1):
            r.clear()
            r.render(k * rs.spp, rs.spp)
            r.synchronize()
            ms.append(r.stats()["ms_last_render"])
        med = float(np.median(ms[1:]))
        out[name] = {"ms": med, "msamples_per_s": samples / med / 1e3}
    r.close()
    sa, cam, o = oracle_setup(sc, rs)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    t = []
    for k in range(3):
        t0 = time.perf_counter()
        oracle.render(sa, cam, o, 0, rs.spp, threads=threads)
        t.append(time.perf_counter() - t0)
    out["cpu_oracle"] = {"msamples_per_s": samples / float(np.median(t)) / 1e6, "cores": threads}
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def heavy_throughput(steps, warmup, count=True):
    """The heavy mixed scene (1000 instances of 8 meshes, ~24.5M instanced
    triangles, ~32% alpha-tested leaf quads, two scattering volume cubes, an
    importance-sampled 2048x1024 sky) at 3840x2160, 16 spp, depth 12:
    wavefront frames (the C2 launch configuration), one plain frame, and
    k_extend's algorithmic bytes from an instrumented 1-spp frame."""
    from glslpt import scenes
    from glslpt._capi import PT_MODE_PLAIN, PT_MODE_WAVEFRONT
    from glslpt.autotune import tuned
    from glslpt.renderer import Renderer
    t0 = time.perf_counter()
    sc, rs = scenes.heavy_mixed()
    build_s = time.perf_counter() - t0
    a = sc.arrays()
    inst_tris = sum(c * 20 * f * f for c, f in zip(scenes.HEAVY_COUNTS, scenes.HEAVY_SPHERE_FREQS))
    leaf_tris = sum(c * 2 * q for c, q in zip(scenes.HEAVY_COUNTS[len(scenes.HEAVY_SPHERE_FREQS):],
                                              scenes.HEAVY_LEAF_QUADS))
    cfg = dict(block_size=256, stack_depth=24, wave_spp=4, compact_every=1, bvh_width=2)
    cfg.update(tuned("c2") or {})
    # this scene's stored configuration: the 1080p one (four-wide nodes) with a
    # longer closest-hit step cap, 128 (closest-hit 477 -> 432 ms per 4-spp
    # frame; 32 is slower than none)
    cfg.update(tuned("c5") or {"bvh_width": 4, "extend_cap": 0})
    r = Renderer(0)
    r.upload_scene(sc)
    samples = rs.width * rs.height * rs.spp
    ms_w, st_w = run_mode(r, sc, rs, PT_MODE_WAVEFRONT, steps, warmup, rs.spp, cfg)
    ms_p, st_p = run_mode(r, sc, rs, PT_MODE_PLAIN, 1, 0, rs.spp, cfg)
    out = {"workload": f"{a['transforms'].shape[0]} instances of 8 meshes + ground + 2 volume cubes, "
                       f"{inst_tris + leaf_tris} instanced triangles ({leaf_tris} alpha-tested), "
                       f"{a['vert_indices'].shape[0]} unique, 2048x1024 HDR sky, "
                       f"{rs.width}x{rs.height}, {rs.spp} spp, max depth {rs.max_depth}",
           "scene_build_s": build_s, "launch": cfg,
           "wavefront": {"msamples_per_s": samples / ms_w / 1e3, "ms_per_frame": ms_w,
                         "stage_ms": {k: v / steps for k, v in st_w["ms_stage"].items() if v > 0}},
           "plain": {"msamples_per_s": samples / ms_p / 1e3, "ms_per_frame": ms_p}}
    out["wavefront_over_plain"] = out["wavefront"]["msamples_per_s"] / out["plain"]["msamples_per_s"]
    if count:
        cw = count_work(r, sc, rs, 1, cfg)
        scale = rs.spp * steps
        bytes_total = scale * (B_NODE * cw["nodes_visited"] + B_TRI * cw["tris_tested"] +
                               B_INST * cw["instances_entered"] + B_LIGHT * cw["lights_tested"] +
                               B_RAY * cw["rays_closest"])
        ext_ms = st_w["ms_stage"]["extend"]
        achieved = bytes_total / (ext_ms / 1e3) / 1e9
        out["roofline_extend"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": achieved / HBM_PEAK_GBS, "bytes_per_frame": bytes_total / steps,
                                  "extend_ms_per_frame": ext_ms / steps,
                                  "per_ray": {"nodes": cw["nodes_visited"] / max(1, cw["rays_closest"]),
                                              "tris": cw["tris_tested"] / max(1, cw["rays_closest"]),
                                              "instances": cw["instances_entered"] / max(1, cw["rays_closest"])},
                                  "rays_per_sample": cw["rays_closest"] / (rs.width * rs.height)}
    out["adjoint"] = heavy_adjoint_ratio(r, sc, rs, cfg)
    out["tlas_rebuild"] = tlas_rebuild_times(r, sc, a)
    out["env_cdf"] = env_cdf_times(r, a)
    r.close()
    return out


def heavy_adjoint_ratio(r, sc, rs, cfg, spp=4, reps=2):
    """Forward and adjoint device time on the heavy scene with respect to the
    Disney parameters and the two volumes' density and colour
    (pt_adjoint_medium): L2 loss against a target rendered from other samples
    with the volumes 1.5x denser and greyer (colour 0.6)."""
    import numpy as np
    from glslpt._capi import PT_MODE_WAVEFRONT
    from glslpt.renderer import options_for
    mats = sc.arrays()["materials"]
    vols = [m for m in range(mats.shape[0]) if mats[m, 18] != 0]
    tgt = mats.copy()
    tgt[vols, 19] *= 1.5
    tgt[vols, 20:23] = 0.6
    r.set_options(options_for(rs, mode=PT_MODE_WAVEFRONT, **cfg))
    r.set_camera(sc.camera())
    r.upload_scene(sc, materials=tgt)
    r.clear()
    r.render(1 << 20, spp)
    r.synchronize()
    target = r.read_accum()[..., :3] / spp
    r.upload_scene(sc)
    fwd, bwd = [], []
    for i in range(reps + 1):
        r.clear()
        r.render(0, spp)
        r.synchronize()
        f_ms = r.stats()["ms_last_render"]
        diff = r.read_accum()[..., :3] / spp - target
        loss = float(np.mean(diff.astype(np.float64) ** 2))
        g, mg = r.adjoint_medium(0, spp, ((2.0 / diff.size) * diff).astype(np.float32))
        st = r.stats()
        if i > 0:  # the first round is warm-up
            fwd.append(f_ms)
            bwd.append(st["ms_last_render"])
    f_med, b_med = float(np.median(fwd)), float(np.median(bwd))
    return {"workload": f"the heavy scene, {spp} spp; Disney parameters of {mats.shape[0]} materials and the "
                        f"density and colour of {len(vols)} volumes",
            "forward_ms": f_med, "backward_ms": b_med, "backward_over_forward": b_med / f_med, "reps": reps,
            "loss": loss, "volume_gradient": {int(m): [float(x) for x in mg[m]] for m in vols},
            "grad_norm": float(np.linalg.norm(g)), "dropped": st["adjoint_dropped"]}


def env_cdf_times(r, a, reps=4):
    """Environment-map change of the heavy scene's 2048x1024 sky: pt_set_env_map
    (24 MB image up, CDF built on the GPU, bit-identical to BuildCDF) against the
    host build of the scene library (EnvironmentMap::BuildCDF's running sum on one
    core, plus the image copy), which a host path would then upload with the
    8 MB CDF. Wall time per call, best of `reps`."""
    import numpy as np

    from glslpt.scene import Scene
    sky = np.ascontiguousarray(a["env_img"], np.float32).reshape(a["env_size"][1], a["env_size"][0], 3)
    dev, host = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        r.set_env_map(sky)
        dev.append(time.perf_counter() - t0)
    s = Scene()
    for _ in range(reps):
        t0 = time.perf_counter()
        s.AddEnvMap(sky)
        host.append(time.perf_counter() - t0)
    return {"texels": int(sky.shape[0] * sky.shape[1]), "device_set_env_map_ms": 1e3 * min(dev),
            "host_build_ms": 1e3 * min(host)}


def tlas_rebuild_times(r, sc, a, reps=5):
    """Instance edit of the heavy scene (every instance moved): the device
    rebuild (pt_rebuild_instances: transforms up, TLAS and instance records
    built on the GPU) against the host path it replaces (the scene builder's
    RebuildInstances, then pt_update_instances with the TLAS node range, which
    re-lays the node array out on the host and uploads it). Wall time per call."""
    import ctypes as C

    import numpy as np

    from glslpt import _capi as api
    xf0 = np.ascontiguousarray(a["transforms"], np.float32).reshape(-1, 16)
    n = xf0.shape[0]
    rng = np.random.default_rng(0)
    xf1 = xf0.copy()
    xf1[:, 12:15] += rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    dev = []
    for k in range(reps + 1):
        t0 = time.perf_counter()
        r.rebuild_instances(xf1 if k % 2 == 0 else xf0)
        dev.append(time.perf_counter() - t0)
    for i in range(n):
        sc.SetInstanceTransform(i, xf1[i])
    host_build, host_upload = [], []
    for k in range(2):
        t0 = time.perf_counter()
        sc.RebuildInstances()
        host_build.append(time.perf_counter() - t0)
        v = sc.view()
        tli = v.top_level_index
        nodes = C.cast(C.byref(v.nodes.contents, tli * C.sizeof(api.PtBvhNode)), C.POINTER(api.PtBvhNode))
        t0 = time.perf_counter()
        rc = r._lib.pt_update_instances(r._h, v.transforms, v.num_instances, v.materials, v.num_materials,
                                        nodes, tli, v.num_nodes - tli)
        host_upload.append(time.perf_counter() - t0)
        if rc != api.PT_OK:
            raise RuntimeError("pt_update_instances failed")
    d_ms = 1e3 * float(np.median(dev[1:]))
    hb, hu = 1e3 * min(host_build), 1e3 * min(host_upload)
    return {"instances": n, "device_ms": d_ms, "host_rebuild_ms": hb, "host_update_ms": hu,
            "host_total_ms": hb + hu, "host_over_device": (hb + hu) / d_ms}


def counter_profile(cfg):
    """The committed rocprofv3 counter summary of the closest-hit stage on this
    workload (profiles/extend_c2_1080p_counters.json, written by `python
    profiling.py summarize`), when it was taken with the same launch
    configuration; else None."""
    path = os.path.join(ROOT, "profiles", "extend_c2_1080p_counters.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t if t.get("launch") == cfg else None


def counter_rooflines(prof, avg_ms):
    """traffic, l2 and valu entries of the roofline object: the profile's
    per-dispatch counter figures over the closest-hit kernel's average launch
    time measured in this run."""
    if prof is None:
        return None, None, None
    per = prof["frame"]["per_step"]
    cal = prof["calibration"]
    sec = avg_ms / 1e3
    src = "profiles/extend_c2_1080p_counters.json"
    l2_ach = per["l2_bytes"] / sec / 1e9
    l2 = {"achieved": l2_ach, "peak": cal["l2_read_gb_per_s"], "unit": "GB/s",
          "frac": l2_ach / cal["l2_read_gb_per_s"], "bytes_per_launch": per["l2_bytes"],
          "peak_basis": "measured: L2-resident read probe (profiling.py l2read), all CUs", "source": src}
    # lane operations = VALU wave-instructions x active lanes per instruction
    lanes = per["SQ_THREAD_CYCLES_VALU"] / per["SQ_ACTIVE_INST_VALU"]
    lane_ops = per["SQ_INSTS_VALU"] * lanes
    v_ach = lane_ops / sec / 1e12
    v_peak = cal["valu_lane_op_per_s"] / 1e12
    issue = per["SQ_INSTS_VALU"] / sec / cal["valu_wave_instr_per_s"]
    valu = {"achieved": v_ach, "peak": v_peak, "unit": "Tlane-op/s", "frac": v_ach / v_peak,
            "issue_frac": issue, "lane_activity": lanes / 64,
            "lane_activity_camera": prof["kinds"]["camera"]["per_step"]["lane_activity"],
            "lane_activity_later": prof["kinds"]["later"]["per_step"]["lane_activity"],
            "waves_per_simd_camera": prof["kinds"]["camera"]["per_step"].get("waves_per_simd"),
            "waves_per_simd_later": prof["kinds"]["later"]["per_step"].get("waves_per_simd"),
            "waves_per_simd_launched": prof["kinds"]["camera"]["per_step"].get("waves_per_simd_launched"),
            "fp32_peak_tflops": FP32_PEAK_TFLOPS, "frac_of_fp32_peak": 2 * v_ach / FP32_PEAK_TFLOPS,
            "peak_basis": "measured: fp32 FMA probe (profiling.py valu), one wave64 VALU instruction per SIMD "
                          "per 4 cycles; 157.3 TFLOP/s counts packed FP32 (2 per lane), which this build "
                          "does not emit",
            "source": src}
    return per["hbm_bytes"], l2, valu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--plain-steps", type=int, default=None)
    ap.add_argument("--wave-spp", type=int, default=None)
    ap.add_argument("--block-size", type=int, default=None)
    ap.add_argument("--stack-depth", type=int, default=None)
    ap.add_argument("--compact-every", type=int, default=None)
    ap.add_argument("--bvh-width", type=int, default=None, choices=(2, 4))
    ap.add_argument("--nee-inline", type=int, default=None, choices=(0, 1))
    ap.add_argument("--extend-cap", type=int, default=None)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-count", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--adjoint-reps", type=int, default=5)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-heavy", action="store_true")
    ap.add_argument("--no-cornell", action="store_true")
    ap.add_argument("--heavy-steps", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the path shards by sample range "
              "across GPUs, which this build does not launch)", file=sys.stderr)
        sys.exit(2)

    from glslpt._capi import PT_MODE_PLAIN, PT_MODE_WAVEFRONT
    from glslpt.renderer import Renderer

    sc, rs = build_scene(args)
    a = sc.arrays()
    ntri_inst = 0
    r = Renderer(0)
    r.upload_scene(sc)
    # launch configuration: explicit flags, else the autotuner's stored choice
    from glslpt.autotune import tuned
    cfg = dict(block_size=256, stack_depth=24, wave_spp=4, compact_every=1, bvh_width=2)
    cfg.update(tuned("c2") or {})
    for k in cfg:
        if getattr(args, k, None) is not None:
            cfg[k] = getattr(args, k)
    opts_kw = dict(cfg)
    samples_per_step = rs.width * rs.height * args.spp

    ms_w, st_w = run_mode(r, sc, rs, PT_MODE_WAVEFRONT, args.steps, args.warmup, args.spp, opts_kw)
    value = samples_per_step / (ms_w / 1e3) / 1e6

    plain = None
    if not args.no_plain:
        ps = args.plain_steps if args.plain_steps is not None else max(1, args.steps)
        ms_p, st_p = run_mode(r, sc, rs, PT_MODE_PLAIN, ps, min(1, args.warmup), args.spp, opts_kw)
        plain = {"value": samples_per_step / (ms_p / 1e3) / 1e6, "ms_per_step": ms_p,
                 "kernel_ms_per_step": st_p["ms_stage"]["plain"] / ps}

    roofline = None
    path_stats = None
    extend_ms = st_w["ms_stage"]["extend"]
    extend_launches = max(1, st_w["launches_stage"]["extend"])
    if not args.no_count:
        cspp = min(args.spp, 4)
        cw = count_work(r, sc, rs, cspp, opts_kw)
        scale = args.spp * args.steps / cspp
        bytes_total = scale * (B_NODE * cw["nodes_visited"] + B_TRI * cw["tris_tested"] +
                               B_INST * cw["instances_entered"] + B_LIGHT * cw["lights_tested"] +
                               B_RAY * cw["rays_closest"])
        per_launch = bytes_total / extend_launches
        avg_ms = extend_ms / extend_launches
        achieved = per_launch / (avg_ms / 1e3) / 1e9
        traffic, l2, valu = counter_rooflines(counter_profile(cfg), avg_ms)
        flops = scale * (F_NODE * cw["nodes_visited"] + F_TRI * cw["tris_tested"] +
                         F_INST * cw["instances_entered"] + F_LIGHT * cw["lights_tested"]) / extend_launches
        if valu is not None:
            tf = flops / (avg_ms / 1e3) / 1e12
            valu.update({"algorithmic_flops_per_launch": flops, "algorithmic_tflops": tf,
                         "algorithmic_frac_of_fp32_peak": tf / FP32_PEAK_TFLOPS,
                         "algorithmic_frac_of_probe": tf * 1e12 / (2 * valu["peak"] * 1e12),
                         "algorithmic_basis": "flops per node, triangle, instance and light test counted from "
                                              "the kernels' expressions (bench.py F_*) times the instrumented "
                                              "frame's counts, over the measured launch time"})
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "l2": l2, "valu": valu,
                    "achieved_basis": "algorithmic bytes (64 B per node, 36 B per triangle, 52 B per instance, "
                                      "60 B per light record, 56 B per ray) per launch over the measured launch "
                                      "time; most are served by L2, see l2",
                    "kernel": "k_extend (closest hit)", "bytes_per_launch": per_launch,
                    "avg_launch_ms": avg_ms, "launches": extend_launches,
                    "per_ray": {"nodes": cw["nodes_visited"] / max(1, cw["rays_closest"]),
                                "tris": cw["tris_tested"] / max(1, cw["rays_closest"]),
                                "instances": cw["instances_entered"] / max(1, cw["rays_closest"])},
                    "rays_per_sample": cw["rays_closest"] / (rs.width * rs.height * cspp),
                    "shadow_per_ray": {"nodes": cw["shadow_nodes"] / max(1, cw["rays_shadow"]),
                                       "tris": cw["shadow_tris"] / max(1, cw["rays_shadow"]),
                                       "instances": cw["shadow_instances"] / max(1, cw["rays_shadow"])},
                    "shadow_rays_per_sample": cw["rays_shadow"] / (rs.width * rs.height * cspp),
                    "stack_spills": cw["stack_spills"]}
        nsmp = rs.width * rs.height * cspp
        path_stats = {"primary_hit_fraction": cw["bounce_hits"][0] / nsmp,
                      "vertices_per_path": sum(cw["bounce_hits"]) / nsmp,
                      "hits_per_bounce": [h / nsmp for h in cw["bounce_hits"] if h > 0]}
    r.close()

    adjoint = None if args.no_adjoint else adjoint_ratio(args.adjoint_reps, 16)
    batch = None
    if not args.no_batch:
        batch = batch_throughput(3)
        batch["sequential"] = batch_sequential(1)
        seq = batch["sequential"]["plain"]["msamples_per_s"]
        batch["batched_plain_over_sequential_plain"] = batch["plain"]["msamples_per_s"] / seq
        batch["batched_wavefront_over_sequential_plain"] = batch["wavefront"]["msamples_per_s"] / seq
    cornell = None if args.no_cornell else cornell_throughput(5)
    heavy = None if args.no_heavy else heavy_throughput(args.heavy_steps, 1, count=not args.no_count)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sc, rs, args.cpu_seconds)

    out = {
        "metric": METRIC, "value": value, "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_w, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "64 instances of a 15680-triangle noise-displaced icosphere (1003520 instanced "
                               "triangles), 16 Disney materials, 2 rect + 1 sphere light + constant environment, "
                               f"{rs.width}x{rs.height}, {args.spp} spp, max depth 8, RR from depth 3",
                   "renderer": "wavefront", "launch": cfg, "samples_per_step": samples_per_step,
                   "unique_triangles": int(a["vert_indices"].shape[0]), "instances": int(a["transforms"].shape[0]),
                   "paths": path_stats},
        "stage_ms_per_step": {k: v / args.steps for k, v in st_w["ms_stage"].items() if v > 0},
        "plain": plain,
        "wavefront_over_plain": (value / plain["value"]) if plain else None,
        "adjoint": adjoint,
        "batch": batch,
        "heavy": heavy,
        "cornell": cornell,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

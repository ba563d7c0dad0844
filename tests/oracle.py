"""ctypes face of oracle/_ref/liboracle.so — the CPU transcription of the
reference shaders. Test infrastructure: only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg use it, always as the reference the HIP
renderer's results are compared with (or timed beside)."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "_ref", "liboracle.so")
P = C.POINTER
i32, f32 = C.c_int32, C.c_float


class OracleScene(C.Structure):
    _fields_ = [("nodes", P(f32)), ("num_nodes", i32), ("top_level_index", i32),
                ("vert_indices", P(i32)), ("num_tris", i32),
                ("vertices_uvx", P(f32)), ("normals_uvy", P(f32)), ("num_vertices", i32),
                ("materials", P(f32)), ("num_materials", i32),
                ("transforms", P(f32)), ("num_instances", i32),
                ("lights", P(f32)), ("num_lights", i32),
                ("textures", P(C.c_uint8)), ("tex_width", i32), ("tex_height", i32), ("num_textures", i32),
                ("env_img", P(f32)), ("env_cdf", P(f32)), ("env_width", i32), ("env_height", i32),
                ("env_total_sum", f32)]


class OracleCamera(C.Structure):
    _fields_ = [("position", f32 * 3), ("up", f32 * 3), ("right", f32 * 3), ("forward", f32 * 3),
                ("fov", f32), ("focalDist", f32), ("aperture", f32)]


class OracleOptions(C.Structure):
    _fields_ = [("width", i32), ("height", i32), ("max_depth", i32),
                ("opt_lights", i32), ("opt_rr", i32), ("rr_depth", i32),
                ("opt_uniform_light", i32), ("uniform_light_col", f32 * 3),
                ("opt_envmap", i32), ("envmap_intensity", f32), ("envmap_rot", f32),
                ("opt_hide_emitters", i32), ("opt_background", i32), ("opt_transparent_background", i32),
                ("opt_alpha_test", i32), ("opt_medium", i32), ("opt_vol_mis", i32),
                ("opt_roughness_mollification", i32), ("roughness_mollification_amt", f32),
                ("opt_opengl_normalmap", i32)]


class OracleCounters(C.Structure):
    _fields_ = [("rays_closest", C.c_int64), ("rays_shadow", C.c_int64), ("nodes", C.c_int64),
                ("tris", C.c_int64), ("instances", C.c_int64), ("lights", C.c_int64)]


def build():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(PATH):
            build()
        l = C.CDLL(PATH)
        l.oracle_render.argtypes = [i32, P(OracleScene), P(OracleCamera), P(OracleOptions), i32, i32, i32, i32, i32,
                                    P(f32), P(C.c_double), i32, P(OracleCounters)]
        l.oracle_trace_samples.argtypes = [P(OracleScene), P(OracleCamera), P(OracleOptions), i32, P(i32), P(i32),
                                           P(i32), P(f32)]
        l.oracle_camera_rays.argtypes = [P(OracleCamera), P(OracleOptions), i32, P(i32), P(i32), P(i32), P(f32),
                                         P(f32)]
        l.oracle_closest_hit.argtypes = [P(OracleScene), P(OracleOptions), P(f32), P(f32), i32, P(f32), P(i32),
                                         P(i32), P(i32), P(f32), P(OracleCounters)]
        l.oracle_any_hit.argtypes = [P(OracleScene), P(OracleOptions), P(f32), P(f32), P(f32), i32, P(i32)]
        l.oracle_pcg4d.argtypes = [P(C.c_uint32), i32, P(C.c_uint32), P(f32)]
        l.oracle_pcg4d.restype = None
        l.oracle_disney_eval.argtypes = [P(f32), f32, P(f32), P(f32), P(f32), P(f32), P(f32)]
        l.oracle_disney_sample.argtypes = [P(f32), f32, P(f32), P(f32), P(C.c_uint32), P(f32), P(f32), P(f32)]
        l.oracle_sample_light.argtypes = [P(f32), i32, P(f32), P(C.c_uint32), P(f32)]
        l.oracle_eval_envmap.argtypes = [P(OracleScene), P(OracleOptions), P(f32), P(f32)]
        l.oracle_sample_envmap.argtypes = [P(OracleScene), P(OracleOptions), P(C.c_uint32), P(f32), P(f32)]
        l.oracle_eval_fn.argtypes = [i32, i32, P(f32), P(f32)]
        l.oracle_render_frozen.argtypes = [P(OracleScene), P(f32), P(OracleCamera), P(OracleOptions), i32, i32,
                                           P(f32), i32]
        l.oracle_tonemap.argtypes = [P(f32), i32, i32, f32, i32, i32, i32, i32, P(f32), P(C.c_uint8)]
        l.oracle_set_inverse_mode.argtypes = [i32]
        l.oracle_set_inverse_mode.restype = None
        _lib = l
    return _lib


FN = dict(GTR1=0, GTR2Aniso=1, SmithG=2, SmithGAniso=3, DielectricFresnel=4, SchlickWeight=5, PowerHeuristic=6,
          SampleGGXVNDF=7, SampleGTR1=8, CosineSampleHemisphere=9, UniformSampleHemisphere=10, SampleHG=11,
          PhaseHG=12, SphereIntersect=13, RectIntersect=14, AABBIntersect=15, sincos=16)
_FN_OUT = {7: 3, 8: 3, 9: 3, 10: 3, 11: 3, 16: 2}


def fn(name, *cols):
    """Evaluate an oracle primitive; each argument is a scalar, a length-n
    vector or an (n, k) array of k packed values."""
    arrs = [np.asarray(c, np.float64) for c in cols]
    n = max([a.shape[0] for a in arrs if a.ndim >= 1] or [1])
    flat = []
    for a in arrs:
        if a.ndim == 0:
            flat.append(np.full((n, 1), float(a)))
        elif a.ndim == 1:
            flat.append(a.reshape(n, 1))
        else:
            flat.append(a.reshape(n, -1))
    inp = np.ascontiguousarray(np.concatenate(flat, axis=1), np.float32)
    k = FN[name]
    out = np.zeros((n, _FN_OUT.get(k, 1)), np.float32)
    rc = lib().oracle_eval_fn(k, n, fp(inp), fp(out))
    assert rc == 0
    return out if out.shape[1] > 1 else out[:, 0]


def set_inverse_mode(mode):
    lib().oracle_set_inverse_mode(int(mode))


def fp(a):
    return None if a is None else a.ctypes.data_as(P(f32))


def ip(a):
    return None if a is None else a.ctypes.data_as(P(i32))


def up(a):
    return a.ctypes.data_as(P(C.c_uint32))


class SceneArrays:
    """Holds numpy copies of a scene's flat arrays and an OracleScene over them."""

    def __init__(self, arrays, textures=None, tex_size=(0, 0)):
        self.a = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in arrays.items()}
        a = self.a
        s = OracleScene()
        s.nodes, s.num_nodes, s.top_level_index = fp(a["nodes"]), a["nodes"].shape[0], a["top_level_index"]
        s.vert_indices, s.num_tris = ip(a["vert_indices"]), a["vert_indices"].shape[0]
        s.vertices_uvx, s.normals_uvy = fp(a["vertices_uvx"]), fp(a["normals_uvy"])
        s.num_vertices = a["vertices_uvx"].shape[0]
        s.materials, s.num_materials = fp(a["materials"]), a["materials"].shape[0]
        s.transforms, s.num_instances = fp(a["transforms"]), a["transforms"].shape[0]
        s.lights, s.num_lights = fp(a["lights"]), a["lights"].shape[0]
        if textures is not None:
            self.tex = np.ascontiguousarray(textures, np.uint8)
            s.textures = self.tex.ctypes.data_as(P(C.c_uint8))
            s.tex_width, s.tex_height = tex_size
            s.num_textures = self.tex.shape[0]
        if "env_img" in a:
            s.env_img, s.env_cdf = fp(a["env_img"]), fp(a["env_cdf"])
            s.env_width, s.env_height = a["env_size"]
            s.env_total_sum = a["env_total_sum"]
        self.s = s


def camera_struct(c):
    o = OracleCamera()
    for k in ("position", "up", "right", "forward"):
        getattr(o, k)[:] = list(getattr(c, k))
    o.fov, o.focalDist, o.aperture = c.fov, c.focalDist, c.aperture
    return o


def options(width, height, max_depth, num_lights, rr=True, rr_depth=2, uniform_light=False,
            uniform_col=(0.3, 0.3, 0.3), materials=None, **kw):
    o = OracleOptions()
    o.width, o.height, o.max_depth = width, height, max_depth
    o.opt_lights = 1 if num_lights > 0 else 0
    o.opt_rr, o.rr_depth = int(rr), rr_depth
    o.opt_uniform_light = int(uniform_light)
    o.uniform_light_col[:] = list(uniform_col)
    o.opt_opengl_normalmap = 1  # RenderOptions::openglNormalMap default (src/core/Renderer.h:57)
    if materials is not None:  # derived like InitShaders (src/core/Renderer.cpp:441-459)
        o.opt_alpha_test = int(np.any(materials[:, 29].astype(np.int32) != 0))
        o.opt_medium = int(np.any(materials[:, 18].astype(np.int32) != 0))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def render(sa, cam, opts, first_sample, num_samples, precision=32, rows=None, threads=0, counters=False):
    W, H = opts.width, opts.height
    r0, r1, rs = (0, H, 1) if rows is None else rows
    cnt = OracleCounters() if counters else None
    if precision == 64:
        acc = np.zeros((H, W, 4), np.float64)
        rc = lib().oracle_render(64, C.byref(sa.s), C.byref(cam), C.byref(opts), first_sample, num_samples,
                                 r0, r1, rs, None, acc.ctypes.data_as(P(C.c_double)), threads,
                                 C.byref(cnt) if cnt else None)
    else:
        acc = np.zeros((H, W, 4), np.float32)
        rc = lib().oracle_render(32, C.byref(sa.s), C.byref(cam), C.byref(opts), first_sample, num_samples,
                                 r0, r1, rs, fp(acc), None, threads, C.byref(cnt) if cnt else None)
    if rc != 0:
        raise RuntimeError(f"oracle_render failed {rc}")
    return (acc, cnt) if counters else acc


def render_frozen(sa, eval_materials, cam, opts, first_sample, num_samples, threads=0):
    """Frozen-sampling render: paths sampled with the scene's materials, BSDF
    values from eval_materials ((num_materials, 32) float32)."""
    W, H = opts.width, opts.height
    em = np.ascontiguousarray(eval_materials, np.float32)
    assert em.shape == sa.a["materials"].shape
    acc = np.zeros((H, W, 4), np.float32)
    rc = lib().oracle_render_frozen(C.byref(sa.s), fp(em), C.byref(cam), C.byref(opts), first_sample, num_samples,
                                    fp(acc), threads)
    if rc != 0:
        raise RuntimeError(f"oracle_render_frozen failed {rc}")
    return acc


def tonemap(accum, num_samples, tonemap=True, aces=False, simple_aces=False, bg_mode=0, background=(1, 1, 1)):
    a = np.ascontiguousarray(accum, np.float32)
    h, w = a.shape[:2]
    bg = np.asarray(background, np.float32)
    out = np.zeros((h, w, 4), np.uint8)
    rc = lib().oracle_tonemap(fp(a), w, h, np.float32(1.0) / np.float32(num_samples), int(tonemap), int(aces),
                              int(simple_aces), int(bg_mode), fp(bg), out.ctypes.data_as(P(C.c_uint8)))
    if rc != 0:
        raise RuntimeError("oracle_tonemap failed")
    return out


def camera_rays(cam, opts, px, py, sample):
    """Origins and directions (n, 3) of the camera rays of tile.glsl."""
    px, py, sample = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.int32), np.shape(px)), np.int32)
                      for a in (px, py, sample)]
    n = px.shape[0]
    o = np.zeros((n, 3), np.float32)
    d = np.zeros((n, 3), np.float32)
    rc = lib().oracle_camera_rays(C.byref(cam), C.byref(opts), n, ip(px), ip(py), ip(sample), fp(o), fp(d))
    assert rc == 0
    return o, d


def eval_envmap(sa, opts, d):
    """EvalEnvMap for directions d (n, 3): (n, 4) of rgb and pdf."""
    d = np.ascontiguousarray(d, np.float32)
    out = np.zeros((d.shape[0], 4), np.float32)
    for i in range(d.shape[0]):
        lib().oracle_eval_envmap(C.byref(sa.s), C.byref(opts), fp(d[i]), fp(out[i]))
    return out


def sample_envmap(sa, opts, seed):
    """SampleEnvMap from RNG state seed: direction and pdf (4), rgb (3), next state."""
    s = np.asarray(seed, np.uint32).copy()
    dp = np.zeros(4, np.float32)
    rgb = np.zeros(3, np.float32)
    lib().oracle_sample_envmap(C.byref(sa.s), C.byref(opts), up(s), fp(dp), fp(rgb))
    return dp, rgb, s


def closest_hit(sa, opts, o, d, counters=False):
    n = o.shape[0]
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    t = np.zeros(n, np.float32)
    inst = np.zeros(n, np.int32)
    tri = np.zeros(n, np.int32)
    li = np.zeros(n, np.int32)
    uv = np.zeros((n, 2), np.float32)
    cnt = OracleCounters() if counters else None
    lib().oracle_closest_hit(C.byref(sa.s), C.byref(opts), fp(o), fp(d), n, fp(t), ip(inst), ip(tri), ip(li), fp(uv),
                             C.byref(cnt) if cnt else None)
    res = dict(t=t, instance=inst, triangle=tri, light=li, uv=uv)
    return (res, cnt) if counters else res


def any_hit(sa, opts, o, d, maxd):
    n = o.shape[0]
    occ = np.zeros(n, np.int32)
    lib().oracle_any_hit(C.byref(sa.s), C.byref(opts), fp(np.ascontiguousarray(o, np.float32)),
                         fp(np.ascontiguousarray(d, np.float32)), fp(np.ascontiguousarray(maxd, np.float32)), n,
                         ip(occ))
    return occ


def pcg4d(seed, n):
    s = np.array(seed, np.uint32)
    st = np.zeros((n, 4), np.uint32)
    r = np.zeros(n, np.float32)
    lib().oracle_pcg4d(up(s), n, up(st), fp(r))
    return st, r


def trace_samples(sa, cam, opts, px, py, sample):
    n = len(px)
    out = np.zeros((n, 4), np.float32)
    lib().oracle_trace_samples(C.byref(sa.s), C.byref(cam), C.byref(opts), n, ip(np.asarray(px, np.int32)),
                               ip(np.asarray(py, np.int32)), ip(np.asarray(sample, np.int32)), fp(out))
    return out


def disney_eval(mat, V, N, L, eta=-1.0):
    f = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    lib().oracle_disney_eval(fp(np.asarray(mat, np.float32)), float(eta), fp(np.asarray(V, np.float32)),
                             fp(np.asarray(N, np.float32)), fp(np.asarray(L, np.float32)), fp(f), fp(pdf))
    return f, float(pdf[0])


def disney_sample(mat, V, N, seed, eta=-1.0):
    s = np.asarray(seed, np.uint32).copy()
    L = np.zeros(3, np.float32)
    f = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    lib().oracle_disney_sample(fp(np.asarray(mat, np.float32)), float(eta), fp(np.asarray(V, np.float32)),
                               fp(np.asarray(N, np.float32)), up(s), fp(L), fp(f), fp(pdf))
    return L, f, float(pdf[0]), s


def sample_light(light, num_lights, pos, seed):
    s = np.asarray(seed, np.uint32).copy()
    rec = np.zeros(13, np.float32)
    lib().oracle_sample_light(fp(np.asarray(light, np.float32)), int(num_lights), fp(np.asarray(pos, np.float32)),
                              up(s), fp(rec))
    return dict(direction=rec[0:3], normal=rec[3:6], emission=rec[6:9], dist=rec[9], pdf=rec[10]), s

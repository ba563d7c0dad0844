"""Shared helpers for the parity tests: the same scene and settings feed the
HIP renderer and the CPU oracle."""
import numpy as np

import oracle
from glslpt import scenes


# RenderSettings.extra (pt_options fields) -> oracle option fields
_EXTRA_TO_ORACLE = {
    "enable_envmap": ("opt_envmap", int), "envmap_intensity": ("envmap_intensity", float),
    "envmap_rot": ("envmap_rot", lambda deg: float(deg) / 360.0), "hide_emitters": ("opt_hide_emitters", int),
    "enable_background": ("opt_background", int), "transparent_background": ("opt_transparent_background", int),
    "enable_volume_mis": ("opt_vol_mis", int),
    "enable_roughness_mollification": ("opt_roughness_mollification", int),
    "roughness_mollification_amt": ("roughness_mollification_amt", float),
    "opengl_normal_map": ("opt_opengl_normalmap", int),
}


def oracle_setup(sc, rs, **kw):
    a = sc.arrays()
    sa = oracle.SceneArrays(a, textures=a.get("textures"), tex_size=a.get("tex_size", (0, 0)))
    cam = oracle.camera_struct(sc.camera())
    extra = {}
    for k, v in rs.extra.items():
        name, conv = _EXTRA_TO_ORACLE[k]
        extra[name] = conv(v)
    if rs.extra.get("enable_envmap") and "env_img" not in a:
        extra["opt_envmap"] = 0  # no map uploaded: the renderer leaves OPT_ENVMAP off too
    extra.update(kw)
    o = oracle.options(rs.width, rs.height, rs.max_depth, a["lights"].shape[0], rr=rs.enable_rr,
                       rr_depth=rs.rr_depth, uniform_light=rs.enable_uniform_light,
                       uniform_col=rs.uniform_light_col, materials=a["materials"], **extra)
    return sa, cam, o


def random_dirs(rng, n):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d.astype(np.float32)


def camera_rays(cam, width, height, rng, n):
    """Rays through random points of the image plane (no jitter model needed)."""
    u = rng.uniform(-1, 1, n)
    v = rng.uniform(-1, 1, n)
    s = np.tan(cam.fov * 0.5)
    right, up, fwd = [np.array(list(getattr(cam, k)), np.float64) for k in ("right", "up", "forward")]
    d = (u * s)[:, None] * right + (v * s * height / width)[:, None] * up + fwd
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.tile(np.array(list(cam.position), np.float32), (n, 1))
    return o.astype(np.float32), d.astype(np.float32)


def small_instanced(num_instances=16, freq=8, seed=7):
    return scenes.instanced_spheres(num_instances=num_instances, freq=freq, seed=seed, width=96, height=64, spp=2,
                                    max_depth=8, extent=6.0)


def rel_err(a, b, floor=1e-6):
    return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)

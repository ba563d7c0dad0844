"""PathTrace / DirectLight / tile.glsl of the oracle at the level of whole
paths: an analytic direct-lighting case and the float/double instantiations."""
import numpy as np

import oracle
from common import oracle_setup
from glslpt import scenes
from glslpt.scene import Scene, mat4, quad_light
from glslpt.scenes import RenderSettings


def plane_under_light(rough, color, le, light_half=0.25, h=1.0):
    sc = Scene()
    S = 50.0
    quad = np.array([[[-S, 0, -S], [S, 0, S], [S, 0, -S]], [[-S, 0, -S], [-S, 0, S], [S, 0, S]]], np.float32)
    n = np.tile(np.array([0, 1, 0], np.float32), (2, 3, 1))
    m = sc.AddMesh(quad, n, name="plane")
    mid = sc.AddMaterial(baseColor=color, roughness=rough, ior=1.0)  # ior 1: no specular lobe, pure Burley diffuse
    sc.AddMeshInstance(m, mat4(), mid)
    a = light_half
    sc.AddLight(quad_light((-a, h, -a), (a, h, -a), (-a, h, a), le))
    sc.AddCamera((0.0, 3.0, 1.0), (0.0, 0.0, 0.0), 0.05)
    sc.ProcessScene()
    return sc


def burley(c, rough, V, L):
    H = V + L
    H = H / np.linalg.norm(H, axis=-1, keepdims=True)
    cd = np.sum(L * H, -1)
    fd90 = 0.5 + 2 * rough * cd * cd
    return c[None, :] / np.pi * ((1 + (fd90 - 1) * (1 - L[:, 1]) ** 5) * (1 + (fd90 - 1) * (1 - V[1]) ** 5))[:, None]


def test_direct_lighting_matches_quadrature():
    """Depth-1 paths (NEE + BSDF sampling with MIS) estimate the direct
    radiance of a Burley-diffuse plane under a rect light; compare the mean of
    100k samples of one pixel with a 2D quadrature of the rendering equation."""
    rough, color, le, a, h = 0.4, np.array([0.8, 0.6, 0.4]), np.array([5.0, 4.0, 3.0]), 0.25, 1.0
    sc = plane_under_light(rough, color, le, a, h)
    rs = RenderSettings(8, 8, 1, 1, enable_rr=False)
    sa, cam, o = oracle_setup(sc, rs)
    n = 100000
    vals = oracle.trace_samples(sa, cam, o, np.full(n, 4), np.full(n, 4), np.arange(n))
    eye = np.array([0.0, 3.0, 1.0])
    V = eye / np.linalg.norm(eye)
    m = 600
    g = (np.arange(m) + 0.5) / m * 2 * a - a
    X, Z = np.meshgrid(g, g, indexing="ij")
    Pl = np.stack([X.ravel(), np.full(X.size, h), Z.ravel()], 1)
    d2 = np.sum(Pl * Pl, 1)
    L = Pl / np.sqrt(d2)[:, None]
    f = burley(color, rough, V, L)
    integrand = f * le[None, :] * (L[:, 1] * L[:, 1] / d2)[:, None]   # cos at P and at the light (both = L.y)
    want = integrand.sum(0) * (2 * a / m) ** 2
    got = vals[:, :3].mean(0)
    se = vals[:, :3].std(0) / np.sqrt(n)
    assert np.all(np.abs(got - want) < 5 * se + 2e-3 * want), (got, want, se)


def test_float_and_double_paths_agree():
    sc, rs = scenes.cornell(width=48, height=48, spp=8)
    sa, cam, o = oracle_setup(sc, rs)
    f = oracle.render(sa, cam, o, 0, 8)
    d = oracle.render(sa, cam, o, 0, 8, precision=64)
    rel = np.abs(f[..., :3] - d[..., :3]) / np.maximum(np.abs(d[..., :3]), 1e-3)
    assert (rel.max(-1) <= 1e-3).mean() > 0.99
    assert abs(f[..., :3].mean() - d[..., :3].mean()) / d[..., :3].mean() < 1e-3
    assert np.all(f[..., 3] == 8)


def test_options_switch_code_paths():
    """OPT_* switches change the image as the reference's defines do:
    hide_emitters removes directly seen lights, background makes missed
    primary rays transparent, RR changes nothing in expectation."""
    sc, rs = scenes.cornell(width=32, height=32, spp=4)
    sa, cam, o = oracle_setup(sc, rs)
    base = oracle.render(sa, cam, o, 0, 4)
    sa2, cam2, o2 = oracle_setup(sc, rs, opt_hide_emitters=1)
    hid = oracle.render(sa2, cam2, o2, 0, 4)
    assert hid[..., :3].max() < base[..., :3].max()
    si, rsi = scenes.instanced_spheres(num_instances=4, freq=4, width=24, height=16, spp=2, max_depth=3, extent=4.0)
    sa3, cam3, o3 = oracle_setup(si, rsi, opt_background=1)
    bg = oracle.render(sa3, cam3, o3, 0, 2)
    assert bg[..., 3].min() == 0.0 and bg[..., 3].max() == 2.0


def test_frozen_sampling_matches_normal_render():
    """With the evaluation table equal to the scene's materials the
    frozen-sampling render is the ordinary render (up to the last ulp of the
    re-evaluated BSDF)."""
    from glslpt import scenes as S
    sc, rs = S.material_boxes(num_boxes=24, sphere_freq=5, num_spheres=2, num_materials=6, width=32, height=32,
                              spp=2, max_depth=4, extent=5.0)
    sa, cam, o = oracle_setup(sc, rs)
    base = oracle.render(sa, cam, o, 0, 2)
    fr = oracle.render_frozen(sa, sa.a["materials"], cam, o, 0, 2)
    assert np.allclose(fr, base, rtol=1e-5, atol=1e-6)


def test_frozen_sampling_derivative_of_diffuse_albedo():
    """Frozen sampling differentiates only f: for a Burley-diffuse plane (ior
    1: no specular lobe) under a rect light the radiance of every depth-1 path
    is linear in baseColor, so a central difference over the evaluation
    table's baseColor gives radiance / baseColor in that channel and nothing
    in the others."""
    color = np.array([0.8, 0.6, 0.4])
    sc = plane_under_light(0.4, color, np.array([5.0, 4.0, 3.0]))
    rs = RenderSettings(16, 16, 4, 1, enable_rr=False)
    sa, cam, o = oracle_setup(sc, rs)
    m = sa.a["materials"]
    base = oracle.render_frozen(sa, m, cam, o, 0, 4)[..., :3].astype(np.float64)
    eps = 1e-3
    for c in range(3):
        mp, mm = m.copy(), m.copy()
        mp[0, c] += eps
        mm[0, c] -= eps
        d = (oracle.render_frozen(sa, mp, cam, o, 0, 4)[..., :3].astype(np.float64)
             - oracle.render_frozen(sa, mm, cam, o, 0, 4)[..., :3]) / (2 * eps)
        want = base[..., c] / color[c]
        lit = want > 1e-3 * want.max()
        assert np.allclose(d[..., c][lit], want[lit], rtol=2e-3), c
        others = [k for k in range(3) if k != c]
        assert np.abs(d[..., others]).max() <= 1e-3 * base.max()

"""The oracle's camera rays, environment-map lookups, light sampling, media
and alpha-tested AnyHit, pinned by what the mathematics fixes rather than by
retyping the oracle's formulas:

* raygen (tile.glsl:41-75): a pinhole ray starts at the camera and passes
  within one pixel of the pixel centre; the jitter follows the tent density
  (moments and a KS test); a thin-lens ray starts on the aperture disc
  (squared radius uniform on [0, aperture]) and meets the pinhole ray of the
  same sample on the focal plane.
* environment map (envmap.glsl:28-83): the texel SampleEnvMap picks equals a
  linear scan of the CDF for the same random number (from an independent
  PCG4D); its colour is the GL_LINEAR value at a texel corner (the mean of
  four wrapped texels); its pdf equals EvalEnvMap's in the sampled direction;
  EvalEnvMap's pdf integrates to one over the sphere; the rotation is a shift
  of the azimuth.
* light sampling (sampling.glsl:186-248): rect and sphere lights draw two
  random numbers and land on the light (rect: at u*r1 + v*r2 from its
  corner; sphere: on the hemisphere facing the point, uniform in area);
  E[1/pdf] equals the rect's solid angle (Van Oosterom-Strackee) and the
  sphere hemisphere's projected-area integral (quadrature); a distant light
  draws nothing and returns its direction with pdf 1.
* homogeneous media (pathtrace.glsl:370-410), through whole paths across an
  alpha-masked slab: absorption gives exp(-(1 - colour) * density * chord),
  an emissive medium adds colour * density * chord, and in a scattering
  medium the unscattered fraction follows exp(-density * chord).
* alpha-tested AnyHit (anyhit.glsl:118-141): occlusion against a float64
  brute force over every triangle with an analytic block mask, and the
  stochastic (blend) test's comparison direction.
* the float sin and cos every sampling routine calls (GLSL leaves their
  precision to the driver; the path fixes one sequence of float operations
  shared by kernels and oracle): within 1.6 ulp of double sin and cos over
  [-64, 64] and next to every multiple of pi/2 there, odd and even, exact at
  zero.
"""
import numpy as np
import pytest
from scipy import stats

import oracle
from common import oracle_setup
from glslpt.scene import Scene, distant_light, mat4, quad_light, sphere_light
from glslpt.scenes import RenderSettings
from test_oracle_rng import pcg4d_np, rand_np

EPS = 0.0003  # globals.glsl:30, the ray-origin offset after a surface


def rands(seed, n):
    """n successive rand() values from seed and the state after them."""
    v = np.array(seed, np.uint32)
    out = []
    for _ in range(n):
        v = pcg4d_np(v)
        out.append(float(rand_np(v)))
    return out, v


# ------------------------------------------------------------------ raygen
def _camera(aperture=0.0, focal=1.0, fov=0.7, seed=3):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    c = oracle.OracleCamera()
    c.position[:] = [1.0, 2.0, 3.0]
    c.right[:] = list(q[:, 0])
    c.up[:] = list(q[:, 1])
    c.forward[:] = list(q[:, 2])
    c.fov, c.focalDist, c.aperture = fov, focal, aperture
    return c


def _tent_cdf(t):
    t = np.clip(t, -1, 1)
    return np.where(t < 0, 0.5 * (1 + t) ** 2, 1 - 0.5 * (1 - t) ** 2)


def test_pinhole_rays_tent_filter_around_pixel_centre():
    W, H = 37, 23
    cam = _camera()
    o = oracle.options(W, H, 1, 0)
    rng = np.random.default_rng(5)
    n = 20000
    px, py = rng.integers(0, W, n), rng.integers(0, H, n)
    smp = rng.integers(0, 1000, n)
    org, d = oracle.camera_rays(cam, o, px, py, smp)
    assert np.all(org == np.array(cam.position[:], np.float32))
    d = d.astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    R, U, F = (np.array(getattr(cam, k)[:], np.float64) for k in ("right", "up", "forward"))
    scale = np.tan(0.5 * cam.fov)
    xi = (d @ R) / (d @ F) / scale
    yi = (d @ U) / (d @ F) / (scale * H / W)
    jx = (xi + 1) * 0.5 * W - (px + 0.5)
    jy = (yi + 1) * 0.5 * H - (py + 0.5)
    for j in (jx, jy):
        assert np.abs(j).max() <= 1 + 1e-3
        assert abs(j.mean()) < 0.01                  # symmetric
        assert abs((j * j).mean() - 1 / 6) < 0.006   # tent variance (uniform: 1/3)
        assert stats.kstest(j, _tent_cdf).pvalue > 1e-4
    assert abs(np.corrcoef(jx, jy)[0, 1]) < 0.03


def test_thin_lens_rays_meet_pinhole_ray_on_focal_plane():
    W, H = 40, 30
    A, Fd = 0.04, 5.0
    pin, lens = _camera(0.0, Fd), _camera(A, Fd)
    o = oracle.options(W, H, 1, 0)
    rng = np.random.default_rng(6)
    n = 8000
    px, py, smp = rng.integers(0, W, n), rng.integers(0, H, n), rng.integers(0, 1000, n)
    o0, d0 = oracle.camera_rays(pin, o, px, py, smp)
    o1, d1 = oracle.camera_rays(lens, o, px, py, smp)
    P = np.array(pin.position[:], np.float64)
    R, U, F = (np.array(getattr(pin, k)[:], np.float64) for k in ("right", "up", "forward"))
    off = o1.astype(np.float64) - P
    assert np.abs(off @ F).max() < 1e-6                      # on the lens plane
    r2 = (off @ R) ** 2 + (off @ U) ** 2
    assert r2.max() <= A * (1 + 1e-5)
    assert stats.kstest(r2 / A, "uniform").pvalue > 1e-4     # uniform on the disc
    ang = np.arctan2(off @ U, off @ R)
    assert stats.kstest((ang + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-4
    focus = P + Fd * d0.astype(np.float64)
    dist = np.linalg.norm(focus - o1, axis=1)
    hit = o1 + d1.astype(np.float64) * dist[:, None]
    assert np.abs(hit - focus).max() < 2e-5 * Fd


# ------------------------------------------------------------------ environment map
def _env_scene(W, H, seed):
    rng = np.random.default_rng(seed)
    img = rng.uniform(0.05, 2.0, (H, W, 3)).astype(np.float32)
    img[rng.integers(H), rng.integers(W)] *= 30.0
    lum = (np.float32(0.212671) * img[..., 0] + np.float32(0.715160) * img[..., 1]
           + np.float32(0.072169) * img[..., 2]).astype(np.float32).ravel()
    cdf = np.add.accumulate(lum, dtype=np.float32)  # sequential float32 running sum
    a = dict(nodes=np.zeros((1, 9), np.float32), top_level_index=0, vert_indices=np.zeros((1, 3), np.int32),
             vertices_uvx=np.zeros((3, 4), np.float32), normals_uvy=np.zeros((3, 4), np.float32),
             materials=np.zeros((1, 32), np.float32), transforms=np.eye(4, dtype=np.float32).reshape(1, 16),
             lights=np.zeros((0, 15), np.float32), env_img=img.reshape(-1, 3), env_cdf=cdf,
             env_total_sum=float(cdf[-1]), env_size=(W, H))
    return oracle.SceneArrays(a), img, cdf


def _env_opts(rot):
    o = oracle.options(16, 16, 1, 0)
    o.opt_envmap, o.envmap_intensity, o.envmap_rot = 1, 1.0, rot
    return o


@pytest.mark.parametrize("W,H,rot", [(4, 2, 0.0), (7, 5, 0.0), (7, 5, 0.3), (16, 9, 0.125)])
def test_sample_envmap_picks_cdf_texel(W, H, rot):
    sa, img, cdf = _env_scene(W, H, 100 + W)
    o = _env_opts(rot)
    total = np.float32(cdf[-1])
    rng = np.random.default_rng(W * H)
    counts = np.zeros(W * H)
    checked_pdf = 0
    for _ in range(3000):
        seed = rng.integers(0, 2 ** 32, 4, dtype=np.uint64).astype(np.uint32)
        dp, rgb, nxt = oracle.sample_envmap(sa, o, seed)
        (r,), v = rands(seed, 1)
        assert np.array_equal(nxt, v)
        value = np.float32(np.float32(r) * total)
        idx = min(int(np.searchsorted(cdf, value, side="right")), W * H - 1)  # first cdf > value
        x, y = idx % W, idx // W
        counts[idx] += 1
        d = dp[:3].astype(np.float64)
        yr = np.arccos(np.clip(d[1], -1, 1)) / np.pi * H
        assert abs(yr - y) < 1e-3, (yr, y)
        ys, xs = [(y - 1) % H, y], [(x - 1) % W, x]
        corner = np.mean([img[j, i].astype(np.float64) for j in ys for i in xs], axis=0)
        assert np.allclose(rgb, corner, rtol=2e-6, atol=0)
        if y == 0:
            assert dp[3] == 0.0 or not np.isfinite(dp[3])  # sin(theta) == 0: pdf forced to 0
            continue
        xr = ((np.arctan2(-d[2], -d[0]) / (2 * np.pi) + rot) * W) % W
        assert min(abs(xr - x), W - abs(xr - x)) < 1e-3, (xr, x)
        ev = oracle.eval_envmap(sa, o, dp[None, :3])[0]
        assert np.allclose(ev[:3], rgb, rtol=1e-3)
        assert abs(ev[3] - dp[3]) <= 1e-3 * dp[3]
        checked_pdf += 1
    assert checked_pdf > 1000
    # frequencies of the picked texels follow luminance / total
    lum = np.diff(np.concatenate([[0.0], cdf.astype(np.float64)])) / float(total)
    assert stats.chisquare(counts, lum * counts.sum()).pvalue > 1e-4


@pytest.mark.parametrize("W,H", [(4, 2), (7, 5)])
def test_eval_envmap_pdf_integrates_to_one(W, H):
    sa, _, _ = _env_scene(W, H, 7 * W)
    o = _env_opts(0.2)
    nt, nphi = 120, 240
    th = (np.arange(nt) + 0.5) * np.pi / nt
    ph = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    pdf = oracle.eval_envmap(sa, o, d)[:, 3].reshape(nt, nphi).astype(np.float64)
    integral = np.sum(pdf * np.sin(T)) * (np.pi / nt) * (2 * np.pi / nphi)
    assert abs(integral - 1.0) < 3e-3, integral


def test_envmap_rotation_shifts_azimuth():
    sa, _, _ = _env_scene(9, 6, 11)
    rot = 0.23
    rng = np.random.default_rng(12)
    th = rng.uniform(0.2, 2.9, 500)
    a = rng.uniform(-np.pi, np.pi, 500)

    def dirs(theta, az):
        return np.stack([np.sin(theta) * np.cos(az), np.cos(theta), np.sin(theta) * np.sin(az)], -1)

    e1 = oracle.eval_envmap(sa, _env_opts(rot), dirs(th, a))
    e0 = oracle.eval_envmap(sa, _env_opts(0.0), dirs(th, a + 2 * np.pi * rot))
    assert np.allclose(e1, e0, rtol=2e-4)


# ------------------------------------------------------------------ light sampling
def _light_rec(l):
    return np.frombuffer(bytes(l), np.float32).copy()


def _solid_angle_triangle(a, b, c):
    """Van Oosterom & Strackee (1983) solid angle of triangle abc seen from 0."""
    la, lb, lc = (np.linalg.norm(x) for x in (a, b, c))
    num = abs(np.dot(a, np.cross(b, c)))
    den = la * lb * lc + np.dot(a, b) * lc + np.dot(a, c) * lb + np.dot(b, c) * la
    return 2 * np.arctan2(num, den)


def test_rect_light_sample_and_pdf():
    p, v1, v2 = np.array([0.3, 2.0, -0.4]), np.array([1.5, 2.2, -0.1]), np.array([0.1, 2.4, 0.9])
    L = quad_light(p, v1, v2, (3, 2, 1))
    rec = _light_rec(L)
    u, v = v1 - p, v2 - p
    nrm = np.cross(u, v) / np.linalg.norm(np.cross(u, v))
    rng = np.random.default_rng(21)
    for s in (np.array([0.2, -0.5, 0.3]), np.array([1.0, 4.0, 0.5])):  # below and above the light
        inv = []
        for _ in range(20000):
            seed = rng.integers(0, 2 ** 32, 4, dtype=np.uint64).astype(np.uint32)
            ls, nxt = oracle.sample_light(rec, 3, s, seed)
            (r1, r2), st = rands(seed, 2)
            assert np.array_equal(nxt, st)
            q = p + u * np.float32(r1) + v * np.float32(r2)
            assert np.allclose(s + ls["direction"] * ls["dist"], q, atol=2e-5 * (1 + np.abs(q).max()))
            assert np.allclose(ls["normal"], nrm, atol=1e-6)
            assert np.allclose(ls["emission"], [9, 6, 3])
            inv.append(1.0 / ls["pdf"])
        inv = np.array(inv)
        corners = [p - s, p + u - s, p + u + v - s, p + v - s]
        omega = _solid_angle_triangle(corners[0], corners[1], corners[2]) + \
            _solid_angle_triangle(corners[0], corners[2], corners[3])
        se = inv.std() / np.sqrt(inv.size)
        assert abs(inv.mean() - omega) < 5 * se + 1e-6, (inv.mean(), omega, se)


def test_sphere_light_sample_and_pdf():
    c, R = np.array([0.5, 3.0, -1.0]), 0.7
    L = sphere_light(c, R, (1, 1, 1))
    rec = _light_rec(L)
    s = np.array([0.0, 0.0, 0.5])
    axis = (s - c) / np.linalg.norm(s - c)
    rng = np.random.default_rng(22)
    inv, z = [], []
    for _ in range(20000):
        seed = rng.integers(0, 2 ** 32, 4, dtype=np.uint64).astype(np.uint32)
        ls, nxt = oracle.sample_light(rec, 2, s, seed)
        _, st = rands(seed, 2)
        assert np.array_equal(nxt, st)
        q = s + ls["direction"].astype(np.float64) * ls["dist"]
        assert abs(np.linalg.norm(q - c) - R) < 2e-5
        assert np.allclose(ls["normal"], (q - c) / R, atol=2e-5)
        assert np.allclose(ls["emission"], [2, 2, 2])
        z.append(np.dot((q - c) / R, axis))
        inv.append(1.0 / ls["pdf"])
    z, inv = np.array(z), np.array(inv)
    assert z.min() > -1e-5                                    # hemisphere facing the point
    assert stats.kstest(np.clip(z, 0, 1), "uniform").pvalue > 1e-4  # uniform in area
    # E[1/pdf] = integral over the hemisphere of |cos| / d^2 dA (quadrature)
    e1 = np.cross(axis, [0, 0, 1.0]); e1 /= np.linalg.norm(e1)
    e2 = np.cross(axis, e1)
    nt, nphi = 400, 400
    t = (np.arange(nt) + 0.5) * (np.pi / 2) / nt
    ph = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    T, P = np.meshgrid(t, ph, indexing="ij")
    n = (np.sin(T) * np.cos(P))[..., None] * e1 + (np.sin(T) * np.sin(P))[..., None] * e2 + np.cos(T)[..., None] * axis
    q = c + R * n
    w = q - s
    d2 = np.sum(w * w, -1)
    cos = np.abs(np.sum(n * w, -1)) / np.sqrt(d2)
    integral = np.sum(cos / d2 * R * R * np.sin(T)) * (np.pi / 2 / nt) * (2 * np.pi / nphi)
    se = inv.std() / np.sqrt(inv.size)
    assert abs(inv.mean() - integral) < 5 * se, (inv.mean(), integral, se)


def test_distant_light_draws_nothing():
    L = distant_light((0.3, 0.8, -0.2), (2, 3, 4))
    rec = _light_rec(L)
    s = np.array([1.0, -2.0, 0.5])
    seed = np.array([9, 8, 7, 6], np.uint32)
    ls, nxt = oracle.sample_light(rec, 4, s, seed)
    assert np.array_equal(nxt, seed)
    dirn = np.array([0.3, 0.8, -0.2]) / np.linalg.norm([0.3, 0.8, -0.2])
    assert np.allclose(ls["direction"], dirn, atol=1e-7)
    assert ls["pdf"] == 1.0 and ls["dist"] == 1e6
    assert np.allclose(ls["emission"], [8, 12, 16])
    assert np.allclose(ls["normal"], (s - [0.3, 0.8, -0.2]) / np.linalg.norm(s - [0.3, 0.8, -0.2]), atol=1e-6)


# ------------------------------------------------------------------ media
def _slab_scene(medium_type, density, color, backdrop=(0.5, 0.2, 0.1), thickness=1.0, W=48, H=32):
    """Camera on +z looking down -z through a slab between z = 0 and
    z = -thickness whose two faces are fully alpha-masked (every hit passes
    through) and carry the medium; an emissive backdrop behind it."""
    sc = Scene()
    big = 60.0

    def quad(z, nz):
        pts = np.array([[-big, -big, z], [big, -big, z], [big, big, z], [-big, big, z]], np.float32)
        tri = np.array([[pts[0], pts[1], pts[2]], [pts[0], pts[2], pts[3]]], np.float32)
        if nz < 0:
            tri = tri[:, ::-1]
        return tri, np.tile(np.array([0, 0, nz], np.float32), (2, 3, 1))

    f_p, f_n = quad(0.0, 1.0)
    b_p, b_n = quad(-thickness, -1.0)
    slab = sc.AddMesh(np.concatenate([f_p, b_p]), np.concatenate([f_n, b_n]), name="slab")
    back_p, back_n = quad(-thickness - 1.0, 1.0)
    back = sc.AddMesh(back_p, back_n, name="backdrop")
    m_slab = sc.AddMaterial(baseColor=(1, 1, 1), alphaMode=2, opacity=0.0, alphaCutoff=0.5, mediumType=medium_type,
                            mediumDensity=density, mediumColor=color)
    # black and ior 1 (no Fresnel reflection): the backdrop reflects nothing
    m_back = sc.AddMaterial(baseColor=(0, 0, 0), ior=1.0, emission=backdrop)
    sc.AddMeshInstance(slab, mat4(), m_slab)
    sc.AddMeshInstance(back, mat4(), m_back)
    sc.AddCamera((0.0, 0.0, 4.0), (0.0, 0.0, -1.0), 40.0)
    sc.ProcessScene()
    return sc, RenderSettings(W, H, 1, 3, enable_rr=False)


def _per_sample(sc, rs, spp):
    sa, cam, o = oracle_setup(sc, rs)
    assert o.opt_medium == 1 and o.opt_alpha_test == 1 and o.opt_lights == 0
    px, py = np.meshgrid(np.arange(rs.width), np.arange(rs.height))
    px, py = np.tile(px.ravel(), spp), np.tile(py.ravel(), spp)
    smp = np.repeat(np.arange(spp), rs.width * rs.height)
    rad = oracle.trace_samples(sa, cam, o, px, py, smp)[:, :3].astype(np.float64)
    _, d = oracle.camera_rays(cam, o, px, py, smp)
    chord = 1.0 / np.abs(d[:, 2].astype(np.float64)) - EPS  # the second face's hit distance
    return rad, chord


def test_absorbing_medium_transmittance():
    dens, col, E = 1.3, np.array([0.2, 0.6, 0.9]), np.array([0.5, 0.2, 0.1])
    sc, rs = _slab_scene(1, dens, col, E)
    rad, chord = _per_sample(sc, rs, 2)
    want = E * np.exp(-(1 - col)[None, :] * chord[:, None] * dens)
    assert np.allclose(rad, want, rtol=2e-5, atol=0)


def test_emissive_medium_adds_along_chord():
    dens, col, E = 0.7, np.array([0.9, 0.3, 0.05]), np.array([0.5, 0.2, 0.1])
    sc, rs = _slab_scene(3, dens, col, E)
    rad, chord = _per_sample(sc, rs, 2)
    want = col[None, :] * chord[:, None] * dens + E
    assert np.allclose(rad, want, rtol=2e-5, atol=0)


def test_scattering_medium_free_flight():
    """Colour 0: a scattered path carries nothing further, so a sample is the
    backdrop's emission exactly when it crossed the slab unscattered."""
    dens, E = 0.8, np.array([0.5, 0.2, 0.1])
    sc, rs = _slab_scene(2, dens, (0.0, 0.0, 0.0), E, W=64, H=48)
    rad, chord = _per_sample(sc, rs, 12)
    through = np.all(rad == E.astype(np.float32).astype(np.float64), axis=1)
    assert np.all(through | np.all(rad == 0, axis=1))
    p = np.exp(-dens * chord)
    sd = np.sqrt(np.sum(p * (1 - p)))
    assert abs(through.sum() - p.sum()) < 5 * sd, (through.sum(), p.sum(), sd)
    # and per chord length: split the samples at the median chord
    lo = chord < np.median(chord)
    for m in (lo, ~lo):
        sdm = np.sqrt(np.sum(p[m] * (1 - p[m])))
        assert abs(through[m].sum() - p[m].sum()) < 5 * sdm


# ------------------------------------------------------------------ alpha-tested AnyHit
def _block_mask(n=8, block=2, seed=31):
    rng = np.random.default_rng(seed)
    blocks = rng.random((n // block, n // block)) < 0.5
    a = np.kron(blocks, np.ones((block, block), bool))
    img = np.zeros((n, n, 4), np.uint8)
    img[..., 0] = 200
    img[..., 3] = np.where(a, 255, 0)
    return img, a


def _mt(o, d, v0, v1, v2):
    """Moller-Trumbore in float64 for one triangle, many rays."""
    e1, e2 = v1 - v0, v2 - v0
    pv = np.cross(d, e2)
    det = pv @ e1
    tv = o - v0
    b1 = np.sum(tv * pv, 1) / det
    qv = np.cross(tv, e1)
    b2 = np.sum(d * qv, 1) / det
    t = (qv @ e2) / det
    return t, b1, b2


def test_alpha_mask_anyhit_brute_force():
    n = 8
    img, alpha = _block_mask(n)
    rng = np.random.default_rng(32)
    sc = Scene()
    sc.SetTextureArraySize(n, n)
    tex = sc.AddTexture(img, "mask")
    mat = sc.AddMaterial(baseColor=(1, 1, 1), baseColorTexId=tex, alphaMode=2, alphaCutoff=0.5)
    tris, quads = [], []
    for k in range(6):
        c = rng.uniform(-2, 2, 3)
        u, v = rng.normal(size=3), rng.normal(size=3)
        u *= rng.uniform(1.0, 2.0) / np.linalg.norm(u)
        v *= rng.uniform(1.0, 2.0) / np.linalg.norm(v)
        t0 = rng.uniform(-1, 1, 2)
        tu, tv = rng.uniform(0.5, 1.5, 2), rng.uniform(0.5, 1.5, 2)
        P = np.array([c, c + u, c + u + v, c + v], np.float32)
        T = np.array([t0, t0 + tu, t0 + tu + tv, t0 + tv], np.float32)
        pos = np.array([[P[0], P[1], P[2]], [P[0], P[2], P[3]]], np.float32)
        uvs = np.array([[T[0], T[1], T[2]], [T[0], T[2], T[3]]], np.float32)
        nrm = np.tile(np.cross(u, v).astype(np.float32), (2, 3, 1))
        mesh = sc.AddMesh(pos, nrm, uvs, name=f"q{k}")
        sc.AddMeshInstance(mesh, mat4(), mat)
        tris += [(pos[i].astype(np.float64), uvs[i].astype(np.float64)) for i in range(2)]
        quads.append((c, u, v))
    sc.AddCamera((0, 0, 8), (0, 0, 0), 40)
    sc.ProcessScene()
    sa, _, o = oracle_setup(sc, RenderSettings(16, 16, 1, 2))
    assert o.opt_alpha_test == 1 and o.opt_medium == 0
    m = 40000
    org = rng.uniform(-4, 4, (m, 3)) + np.array([0, 0, 6.0])
    # aim at a random point of a random quad (most rays cross one or more)
    qi = rng.integers(0, len(quads), m)
    ab = rng.uniform(0, 1, (m, 2))
    tgt = np.array([quads[q][0] + a * quads[q][1] + b * quads[q][2] for q, (a, b) in zip(qi, ab)])
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    org, d = org.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)
    maxd = rng.uniform(2.0, 14.0, m).astype(np.float32).astype(np.float64)
    occ_bf = np.zeros(m, bool)
    occ_geo = np.zeros(m, bool)
    keep = np.ones(m, bool)
    for (v, t) in tris:
        th, b1, b2 = _mt(org, d, v[0], v[1], v[2])
        b0 = 1 - b1 - b2
        inside = (b0 >= 0) & (b1 >= 0) & (b2 >= 0) & (th >= 0) & (th < maxd)
        near = (np.minimum(np.minimum(b0, b1), b2) > -1e-4) & (np.minimum(np.minimum(b0, b1), b2) < 1e-4)
        near |= np.abs(th - maxd) < 1e-4 * maxd
        uv = t[0] * b0[:, None] + t[1] * b1[:, None] + t[2] * b2[:, None]
        s = uv * n                                    # texel space
        f = np.mod(s, 2.0)                            # position inside a 2-texel block
        clear = np.all((f > 0.55) & (f < 1.45), axis=1)  # all bilinear taps inside one block
        ij = np.floor(np.mod(s, n)).astype(int)
        a = alpha[ij[:, 1], ij[:, 0]]
        keep &= ~(near | (inside & ~clear))
        occ_geo |= inside
        occ_bf |= inside & a
    got = oracle.any_hit(sa, o, org, d, maxd).astype(bool)
    k = keep
    print(f"kept {k.sum()} of {m}; occluded {occ_bf[k].sum()}, geometric hits passed by the mask "
          f"{(occ_geo & ~occ_bf)[k].sum()}")
    assert k.sum() > m // 2
    assert (occ_geo & ~occ_bf)[k].sum() > 500 and occ_bf[k].sum() > 500
    assert np.array_equal(got[k], occ_bf[k])


def test_alpha_blend_anyhit_compares_rand_with_opacity():
    """The oracle's ray query seeds each ray's RNG with (1, 2, 3, 4): the
    first rand() decides a blend hit, which is skipped when rand() > opacity."""
    (r0,), _ = rands([1, 2, 3, 4], 1)
    res = []
    for opacity in (r0 - 2e-3, r0 + 2e-3):
        sc = Scene()
        pos = np.array([[[-1, -1, 0], [1, -1, 0], [1, 1, 0]], [[-1, -1, 0], [1, 1, 0], [-1, 1, 0]]], np.float32)
        mesh = sc.AddMesh(pos, np.tile(np.array([0, 0, 1], np.float32), (2, 3, 1)), name="q")
        sc.AddMeshInstance(mesh, mat4(), sc.AddMaterial(alphaMode=1, opacity=float(opacity)))
        sc.AddCamera((0, 0, 8), (0, 0, 0), 40)
        sc.ProcessScene()
        sa, _, o = oracle_setup(sc, RenderSettings(16, 16, 1, 2))
        org = np.array([[0.1, 0.2, 3.0]], np.float32)
        d = np.array([[0, 0, -1]], np.float32)
        res.append(int(oracle.any_hit(sa, o, org, d, np.array([10.0], np.float32))[0]))
    assert res == [0, 1]


def _ulp_err(got, ref):
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / ulp


def test_sincos_within_ulps_of_double():
    rng = np.random.default_rng(8)
    # every float within 64 ulps of k * pi/2 (|k| <= 40), where the result is
    # smallest and the reduction's error counts most, a dense grid and
    # uniform angles
    near = []
    for k in range(-40, 41):
        c = np.float32(k * np.pi / 2)
        near.append(c + np.arange(-64, 65, dtype=np.float32) * np.spacing(c if c != 0 else np.float32(1)))
    x = np.concatenate(near + [np.linspace(-64, 64, 1_600_001, dtype=np.float32),
                               rng.uniform(0, 2 * np.pi, 400_000).astype(np.float32),
                               np.float32([1e-30, 1e-7, 3e-4, np.pi, 2 * np.pi])])
    sc = oracle.fn("sincos", x)
    xd = x.astype(np.float64)
    es, ec = _ulp_err(sc[:, 0], np.sin(xd)), _ulp_err(sc[:, 1], np.cos(xd))
    print(f"sin max {es.max():.3f} ulp, cos max {ec.max():.3f} ulp; correctly rounded "
          f"{(sc[:, 0] == np.sin(xd).astype(np.float32)).mean():.3f} / {(sc[:, 1] == np.cos(xd).astype(np.float32)).mean():.3f}")
    assert es.max() <= 1.6 and ec.max() <= 1.6
    # a term dropped from either series, or a quadrant mixed up, is off by
    # far more: the mean error stays below one ulp
    assert es.mean() < 0.5 and ec.mean() < 0.5


def test_sincos_symmetry_and_zero():
    x = np.linspace(0, 7, 70_001, dtype=np.float32)
    p, m = oracle.fn("sincos", x), oracle.fn("sincos", -x)
    assert np.array_equal(m[:, 0], -p[:, 0])
    assert np.array_equal(m[:, 1], p[:, 1])
    z = oracle.fn("sincos", np.float32([0.0]))
    assert z[0, 0] == 0.0 and z[0, 1] == 1.0

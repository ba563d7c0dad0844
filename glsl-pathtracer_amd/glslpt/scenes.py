"""Procedural scenes for the workloads (synthetic inputs; no files needed).

* ``cornell(seed)``      Cornell box: five wall quads (2 triangles each) and two
                         blocks (12 triangles each) = 34 triangles, one rect
                         light; diffuse Disney materials with seeded baseColor.
                         Geometry from the published Cornell-box measurements
                         (mm, scaled to metres), each mesh placed by an instance
                         transform the way the reference's .scene files do
                         (translate + scale 0.01, assets/cornell_box_orig.scene).
* ``instanced_spheres``  N instances of a seeded geodesic icosphere displaced by
                         3-octave value noise, random rigid transforms and
                         uniform scale, Disney materials drawn from a seed,
                         rect + sphere lights and a constant environment.
* ``batch_jobs``         32 uneven render jobs of the Cornell scene (own
                         camera, resolution, depth and spp each).
* ``material_boxes``     differentiable-materials workload: 256 boxes and 8
                         noise-displaced spheres (~200k triangles), 64
                         materials with seeded optimisable parameters.
* ``feature_scene``      small scenes exercising one option path each
                         (environment map, textures, media, alpha).
* ``heavy_mixed``        the heavy mixed scene: 1000 instances of 8 seeded
                         meshes (five noise-displaced spheres of 20k-2M
                         triangles, three clusters of alpha-tested leaf quads),
                         ~24M instanced triangles of which ~30% alpha-tested,
                         two scattering volume cubes, a 2048x1024 HDR sky,
                         3840x2160, 16 spp, depth 12.
* ``tiny``               a few triangles, for smoke runs.

Each builder returns (Scene, RenderSettings) where RenderSettings carries the
resolution, depth and option flags of the workload.
"""
from dataclasses import dataclass, field

import numpy as np

from .scene import Scene, material, quad_light, sphere_light, mat4


@dataclass
class RenderSettings:
    width: int
    height: int
    spp: int
    max_depth: int
    enable_rr: bool = True
    rr_depth: int = 2
    enable_uniform_light: bool = False
    uniform_light_col: tuple = (0.3, 0.3, 0.3)
    extra: dict = field(default_factory=dict)


# ---------------------------------------------------------------- Cornell box
# Published Cornell-box geometry (mm). Walls are axis-aligned quads.
_CB_WALLS = {
    "floor":   [(552.8, 0, 0), (0, 0, 0), (0, 0, 559.2), (549.6, 0, 559.2)],
    "ceiling": [(556.0, 548.8, 0), (556.0, 548.8, 559.2), (0, 548.8, 559.2), (0, 548.8, 0)],
    "back":    [(549.6, 0, 559.2), (0, 0, 559.2), (0, 548.8, 559.2), (556.0, 548.8, 559.2)],
    "green":   [(0, 0, 559.2), (0, 0, 0), (0, 548.8, 0), (0, 548.8, 559.2)],
    "red":     [(552.8, 0, 0), (549.6, 0, 559.2), (556.0, 548.8, 559.2), (556.0, 548.8, 0)],
}
_CB_SHORT = [(130.0, 165.0, 65.0), (82.0, 165.0, 225.0), (240.0, 165.0, 272.0), (290.0, 165.0, 114.0)]
_CB_TALL = [(423.0, 330.0, 247.0), (265.0, 330.0, 296.0), (314.0, 330.0, 456.0), (472.0, 330.0, 406.0)]


def _quad_tris(q):
    a, b, c, d = [np.asarray(p, np.float64) for p in q]
    return [np.stack([a, b, c]), np.stack([a, c, d])]


def _block_tris(top):
    top = [np.asarray(p, np.float64) for p in top]
    bot = [np.array([p[0], 0.0, p[2]]) for p in top]
    tris = _quad_tris(top)
    for i in range(4):
        j = (i + 1) % 4
        tris += _quad_tris([bot[i], bot[j], top[j], top[i]])
    tris += _quad_tris([bot[3], bot[2], bot[1], bot[0]])
    return tris


def _face_normals(tris):
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return np.repeat(n[:, None, :], 3, axis=1)


def _orient_inward(tris, center):
    """Flip triangles so their geometric normal faces `center` (walls face the room)."""
    out = []
    for t in tris:
        n = np.cross(t[1] - t[0], t[2] - t[0])
        c = t.mean(axis=0)
        if np.dot(n, center - c) < 0:
            t = t[[0, 2, 1]]
        out.append(t)
    return np.stack(out)


def _orient_outward(tris, center):
    out = []
    for t in tris:
        n = np.cross(t[1] - t[0], t[2] - t[0])
        if np.dot(n, t.mean(axis=0) - center) < 0:
            t = t[[0, 2, 1]]
        out.append(t)
    return np.stack(out)


def cornell(seed=0xC0FFEE01, width=256, height=256, spp=4, max_depth=5):
    rng = np.random.Generator(np.random.PCG64(seed))
    sc = Scene()
    room_center = np.array([278.0, 274.4, 279.6])
    parts = []
    for name, q in _CB_WALLS.items():
        parts.append((name, _orient_inward(np.stack(_quad_tris(q)), room_center)))
    for name, top in (("short", _CB_SHORT), ("tall", _CB_TALL)):
        tr = np.stack(_block_tris(top))
        parts.append((name, _orient_outward(tr, tr.reshape(-1, 3).mean(axis=0))))
    for i, (name, tris) in enumerate(parts):
        # Mesh in centred local units of 0.1 mm-per-unit (cm), placed by translate + scale 0.01.
        centre = tris.reshape(-1, 3).mean(axis=0)
        local = ((tris - centre) / 10.0).astype(np.float32)          # cm
        mesh = sc.AddMesh(local, _face_normals(local.astype(np.float64)).astype(np.float32), name=name)
        bc = rng.random(3)
        mid = sc.AddMaterial(baseColor=bc)
        xf = mat4(translate=(centre / 1000.0), scale=(0.01, 0.01, 0.01))
        sc.AddMeshInstance(mesh, xf, mid)
    sc.AddLight(quad_light((0.343, 0.5478, 0.227), (0.343, 0.5478, 0.332), (0.213, 0.5478, 0.227), (17, 12, 4)))
    sc.AddCamera((0.276, 0.275, -0.75), (0.276, 0.275, 0.10), 40.0)
    sc.ProcessScene()
    return sc, RenderSettings(width, height, spp, max_depth, enable_rr=True, rr_depth=2)


# ------------------------------------------------------------- icospheres
def _icosahedron():
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    return v, f


def geodesic_sphere(freq):
    """Frequency-`freq` geodesic subdivision of the icosahedron: 20*freq^2
    triangles on the unit sphere, shared vertices. Returns (verts, faces).

    Vertices are numbered in order of first appearance over the faces' grids
    (vectorised; identical to `_geodesic_sphere_loop`)."""
    v0, f0 = _icosahedron()
    ii, jj = [], []
    for i in range(freq + 1):
        for j in range(freq + 1 - i):
            ii.append(i)
            jj.append(j)
    ii = np.array(ii)
    jj = np.array(jj)
    kk = freq - ii - jj
    ng = ii.shape[0]
    A, B, C = v0[f0[:, 0]], v0[f0[:, 1]], v0[f0[:, 2]]
    pts = (A[:, None, :] * kk[None, :, None] + B[:, None, :] * ii[None, :, None]
           + C[:, None, :] * jj[None, :, None]) / freq          # (20, ng, 3), face-major grid order
    pts = pts.reshape(-1, 3)
    keys = np.round(pts, 12)
    _, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")                  # unique keys by first appearance
    rank = np.empty_like(order)
    rank[order] = np.arange(order.shape[0])
    gid = rank[inv.reshape(-1)]                                # grid point -> vertex id
    first_pts = pts[first[order]]
    # per-vector norm, computed as np.linalg.norm of one vector rounds (a BLAS dot)
    norms = np.fromiter((np.linalg.norm(p) for p in first_pts), np.float64, first_pts.shape[0])
    verts = first_pts / norms[:, None]
    # grid (i, j) -> position in the per-face grid list
    pos = -np.ones((freq + 2, freq + 2), np.int64)
    pos[ii, jj] = np.arange(ng)
    ti, tj = [], []
    up = []
    for i in range(freq):
        for j in range(freq - i):
            ti.append(i); tj.append(j); up.append(True)
            if i + j < freq - 1:
                ti.append(i); tj.append(j); up.append(False)
    ti = np.array(ti); tj = np.array(tj); up = np.array(up)
    c0 = np.where(up, pos[ti, tj], pos[ti + 1, tj])
    c1 = np.where(up, pos[ti + 1, tj], pos[ti + 1, tj + 1])
    c2 = pos[ti, tj + 1]
    base = (np.arange(20) * ng)[:, None]
    faces = np.stack([gid[base + c0[None, :]], gid[base + c1[None, :]], gid[base + c2[None, :]]], -1).reshape(-1, 3)
    return verts, faces.astype(np.int64)


def _geodesic_sphere_loop(freq):
    """Reference loop form of `geodesic_sphere` (kept for its test)."""
    v0, f0 = _icosahedron()
    verts, index = [], {}

    def vid(p):
        key = tuple(np.round(p, 12))
        if key not in index:
            index[key] = len(verts)
            verts.append(p / np.linalg.norm(p))
        return index[key]

    faces = []
    for a, b, c in f0:
        A, B, C = v0[a], v0[b], v0[c]
        grid = {}
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                k = freq - i - j
                grid[(i, j)] = vid((A * k + B * i + C * j) / freq)
        for i in range(freq):
            for j in range(freq - i):
                faces.append((grid[(i, j)], grid[(i + 1, j)], grid[(i, j + 1)]))
                if i + j < freq - 1:
                    faces.append((grid[(i + 1, j)], grid[(i + 1, j + 1)], grid[(i, j + 1)]))
    return np.array(verts), np.array(faces, np.int64)


def value_noise3(p, seed, octaves=3):
    """Sum of `octaves` octaves of trilinear value noise on an integer lattice
    with hashed values in [-1, 1]; octave k has frequency 2^k and weight 2^-k."""
    out = np.zeros(p.shape[0])
    for k in range(octaves):
        q = p * (2.0 ** k) * 2.0
        i0 = np.floor(q).astype(np.int64)
        fr = q - i0
        w = fr * fr * (3 - 2 * fr)
        acc = np.zeros(p.shape[0])
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    c = i0 + np.array([dx, dy, dz])
                    h = (c[:, 0] * 73856093) ^ (c[:, 1] * 19349663) ^ (c[:, 2] * 83492791) ^ (seed + k * 0x9E3779B1)
                    h = (h * 0x2545F491) & 0xFFFFFFFF
                    h ^= h >> 15
                    val = (h & 0xFFFF) / 32767.5 - 1.0
                    wx = w[:, 0] if dx else 1 - w[:, 0]
                    wy = w[:, 1] if dy else 1 - w[:, 1]
                    wz = w[:, 2] if dz else 1 - w[:, 2]
                    acc += wx * wy * wz * val
        out += acc * (0.5 ** k)
    return out


def displaced_icosphere(freq=28, amplitude=0.1, seed=0x5EED0002):
    """Triangle soup (positions[n,3,3], normals[n,3,3]) of a noise-displaced
    geodesic sphere with smooth vertex normals."""
    v, f = geodesic_sphere(freq)
    d = 1.0 + amplitude * value_noise3(v, seed)
    p = v * d[:, None]
    fn = np.cross(p[f[:, 1]] - p[f[:, 0]], p[f[:, 2]] - p[f[:, 0]])
    vn = np.zeros_like(p)
    for k in range(3):
        np.add.at(vn, f[:, k], fn)
    vn /= np.linalg.norm(vn, axis=1, keepdims=True)
    return p[f].astype(np.float32), vn[f].astype(np.float32)


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    x, y, z, w = q
    # column-per-row Mat4 data layout as Mat4::QuatToMatrix (src/math/Mat4.h:109-139)
    r = np.eye(4)
    r[0, :3] = [1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)]
    r[1, :3] = [2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)]
    r[2, :3] = [2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)]
    return r.astype(np.float32)


def disney_materials(rng, n):
    mats = []
    for _ in range(n):
        mats.append(material(baseColor=rng.random(3), roughness=rng.uniform(0.05, 1.0),
                             metallic=float(rng.random() < 0.5), ior=1.5, clearcoat=rng.uniform(0, 1)))
    return mats


def instanced_spheres(num_instances=64, freq=28, seed=0x5EED0002, width=1920, height=1080, spp=64, max_depth=8,
                      extent=20.0, num_materials=16):
    rng = np.random.Generator(np.random.PCG64(seed))
    sc = Scene()
    pos, nrm = displaced_icosphere(freq, 0.1, seed)
    mesh = sc.AddMesh(pos, nrm, name=f"icosphere{freq}")
    mids = [sc.AddMaterial(m) for m in disney_materials(rng, num_materials)]
    for i in range(num_instances):
        s = rng.uniform(0.5, 2.0)
        xf = mat4(translate=rng.uniform(-extent, extent, 3), scale=(s, s, s), rot=random_rotation(rng))
        sc.AddMeshInstance(mesh, xf, mids[i % num_materials])
    h = extent * 1.5
    sc.AddLight(quad_light((-10, h, -10), (0, h, -10), (-10, h, 0), (6, 6, 6)))
    sc.AddLight(quad_light((5, h, 5), (15, h, 5), (5, h, 15), (5, 6, 8)))
    sc.AddLight(sphere_light((0, extent * 0.6, extent * 1.2), 2.0, (8, 7, 5)))
    sc.AddCamera((0.0, extent * 0.5, extent * 3.0), (0.0, 0.0, 0.0), 45.0)
    sc.ProcessScene()
    return sc, RenderSettings(width, height, spp, max_depth, enable_rr=True, rr_depth=3,
                              enable_uniform_light=True, uniform_light_col=(0.3, 0.3, 0.3))


def settings_from_render_options(o, width=None, height=None, spp=1):
    """RenderSettings for a pts_render_options (from Scene.LoadSceneFile):
    resolution, depth, RR and every path switch the renderer reads; width and
    height override the file's resolution."""
    extra = dict(enable_envmap=int(o.enable_envmap), envmap_intensity=float(o.envmap_intensity),
                 envmap_rot=float(o.envmap_rot), hide_emitters=int(o.hide_emitters),
                 enable_background=int(o.enable_background), transparent_background=int(o.transparent_background),
                 enable_volume_mis=int(o.enable_volume_mis),
                 enable_roughness_mollification=int(o.enable_roughness_mollification),
                 roughness_mollification_amt=float(o.roughness_mollification_amt),
                 opengl_normal_map=int(o.opengl_normal_map))
    return RenderSettings(width or o.render_resolution[0], height or o.render_resolution[1], spp, o.max_depth,
                          enable_rr=bool(o.enable_rr), rr_depth=o.rr_depth,
                          enable_uniform_light=bool(o.enable_uniform_light),
                          uniform_light_col=tuple(o.uniform_light_col), extra=extra)


# ------------------------------------------------------------ batched jobs
BATCH_RESOLUTIONS = ((128, 128), (256, 256), (512, 512), (1024, 1024), (1920, 1080))


def batch_jobs(num_jobs=32, seed=0xBA7C0003, radius=10.0, fov=6.0, resolutions=BATCH_RESOLUTIONS,
               depth_range=(2, 12), spp_range=(4, 32)):
    """Batched uneven jobs over the Cornell scene: job j (seed + j) looks at
    the room centre from a point uniform on a sphere of `radius` around it,
    with a resolution drawn from `resolutions`, max depth uniform in
    depth_range and spp uniform in spp_range (inclusive). The field of view
    (degrees) is narrow enough for the ~0.56-unit room to fill the frame from
    that distance. Returns (scene, settings, jobs); a job is a dict with
    camera, width, height, spp, max_depth."""
    sc, rs = cornell()
    centre = np.array([0.278, 0.2744, 0.2796])
    jobs = []
    for j in range(num_jobs):
        rng = np.random.Generator(np.random.PCG64(seed + j))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        w, h = resolutions[int(rng.integers(len(resolutions)))]
        depth = int(rng.integers(depth_range[0], depth_range[1] + 1))
        spp = int(rng.integers(spp_range[0], spp_range[1] + 1))
        sc.AddCamera(tuple(centre + radius * d), tuple(centre), fov)
        jobs.append(dict(camera=sc.camera(), width=w, height=h, spp=spp, max_depth=depth))
    return sc, rs, jobs


# --------------------------------------------------- differentiable materials
def _unit_box():
    """12 outward-facing triangles of the cube [-1, 1]^3 with flat normals."""
    c = np.array([[x, y, z] for x in (-1.0, 1.0) for y in (-1.0, 1.0) for z in (-1.0, 1.0)])
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for f in faces:
        tris += _quad_tris([c[i] for i in f])
    tris = _orient_outward(np.stack(tris), np.zeros(3))
    return tris.astype(np.float32), _face_normals(tris).astype(np.float32)


def ior_from_specular(specular):
    """Disney `specular` to index of refraction: F0 = 0.08 * specular =
    ((ior - 1) / (ior + 1))^2 (specular 0.5 -> ior 1.5)."""
    r = np.sqrt(0.08 * np.asarray(specular, np.float64))
    return (1.0 + r) / (1.0 - r)


def diff_material_params(seed, n=64):
    """Seeded optimisable parameters of n Disney materials, inside the open
    ranges (no parameter at a lobe's on/off boundary): baseColor U[0.05, 0.95]^3,
    roughness U[0.1, 0.9], metallic U[0.05, 0.95], specular U[0.2, 0.8] mapped
    to ior, clearcoat U[0.05, 0.95]. Returns (n, 7) in pt_adjoint order."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = np.zeros((n, 7), np.float32)
    p[:, 0:3] = rng.uniform(0.05, 0.95, (n, 3))
    p[:, 3] = rng.uniform(0.1, 0.9, n)
    p[:, 4] = rng.uniform(0.05, 0.95, n)
    p[:, 5] = ior_from_specular(rng.uniform(0.2, 0.8, n))
    p[:, 6] = rng.uniform(0.05, 0.95, n)
    return p


def apply_params(materials, params):
    """Material table (n, 32) with the pt_adjoint parameters replaced."""
    from ._capi import ADJOINT_FIELDS
    m = np.array(materials, np.float32, copy=True)
    for k, f in enumerate(ADJOINT_FIELDS):
        m[:, f] = params[:, k]
    return m


def material_boxes(num_boxes=256, sphere_freq=35, num_spheres=8, num_materials=64, seed=0xD1FF0004,
                   param_seed=0xD1FF0004, width=512, height=512, spp=16, max_depth=4, extent=10.0):
    """Differentiable-materials workload: `num_boxes` boxes (12 triangles
    each, random rotation and size, one of them a ground slab) and
    `num_spheres` distinct noise-displaced geodesic spheres of frequency
    `sphere_freq` (20 f^2 triangles each; 8 x 24,500 + 256 x 12 = 199,072
    triangles by default), `num_materials` Disney materials from
    diff_material_params(param_seed), two rect lights and a constant
    environment."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sc = Scene()
    bpos, bnrm = _unit_box()
    box = sc.AddMesh(bpos, bnrm, name="box")
    spheres = []
    v, f = geodesic_sphere(sphere_freq)
    for k in range(num_spheres):
        d = 1.0 + 0.1 * value_noise3(v, seed + 17 * k)
        pv = v * d[:, None]
        fn = np.cross(pv[f[:, 1]] - pv[f[:, 0]], pv[f[:, 2]] - pv[f[:, 0]])
        vn = np.zeros_like(pv)
        for j in range(3):
            np.add.at(vn, f[:, j], fn)
        vn /= np.linalg.norm(vn, axis=1, keepdims=True)
        spheres.append(sc.AddMesh(pv[f].astype(np.float32), vn[f].astype(np.float32), name=f"sphere{k}"))
    params = diff_material_params(param_seed, num_materials)
    mids = []
    for i in range(num_materials):
        p = params[i]
        mids.append(sc.AddMaterial(material(baseColor=p[0:3], roughness=p[3], metallic=p[4], ior=p[5],
                                            clearcoat=p[6], clearcoatGloss=0.5)))
    # ground slab, then boxes scattered over it
    sc.AddMeshInstance(box, mat4(translate=(0, -0.5, 0), scale=(extent * 1.5, 0.5, extent * 1.5)),
                       mids[int(rng.integers(num_materials))])
    for i in range(num_boxes - 1):
        sx, sy, sz = rng.uniform(0.2, 0.9, 3)
        t = (rng.uniform(-extent, extent), rng.uniform(0.2, 3.0), rng.uniform(-extent, extent))
        sc.AddMeshInstance(box, mat4(translate=t, scale=(sx, sy, sz), rot=random_rotation(rng)),
                           mids[int(rng.integers(num_materials))])
    for k, m in enumerate(spheres):
        r = rng.uniform(1.0, 2.0)
        t = (rng.uniform(-0.7, 0.7) * extent, r + 0.2, rng.uniform(-0.7, 0.7) * extent)
        sc.AddMeshInstance(m, mat4(translate=t, scale=(r, r, r), rot=random_rotation(rng)),
                           mids[int(rng.integers(num_materials))])
    h = 12.0
    sc.AddLight(quad_light((-6, h, -6), (0, h, -6), (-6, h, 0), (12, 11, 10)))
    sc.AddLight(quad_light((3, h, 3), (9, h, 3), (3, h, 9), (8, 10, 12)))
    sc.AddCamera((0.0, 9.0, 19.0), (0.0, 1.0, 0.0), 45.0)
    sc.ProcessScene()
    return sc, RenderSettings(width, height, spp, max_depth, enable_rr=True, rr_depth=2,
                              enable_uniform_light=True, uniform_light_col=(0.3, 0.3, 0.3))


# ------------------------------------------------------ feature scenes
def _uv_quad(corner, u, v, tiles=1.0):
    """Two triangles of the parallelogram corner + [0,1]u + [0,1]v with UVs
    in [0, tiles], facing along cross(u, v)."""
    c = np.asarray(corner, np.float64)
    u = np.asarray(u, np.float64)
    v = np.asarray(v, np.float64)
    p = [c, c + u, c + u + v, c + v]
    t = [(0, 0), (tiles, 0), (tiles, tiles), (0, tiles)]
    pos = np.array([[p[0], p[1], p[2]], [p[0], p[2], p[3]]], np.float32)
    uvs = np.array([[t[0], t[1], t[2]], [t[0], t[2], t[3]]], np.float32)
    n = np.cross(u, v)
    n /= np.linalg.norm(n)
    return pos, np.tile(n.astype(np.float32), (2, 3, 1)), uvs


def sky_envmap(width=64, height=32, seed=0xE7A11, sun=(0.3, 0.22), sun_power=400.0):
    """Procedural HDR sky (RGB float, rows top-down like stb's HDR loader):
    a vertical gradient, faint value noise and a bright sun disc, so that
    importance sampling matters."""
    y = (np.arange(height) + 0.5) / height
    x = (np.arange(width) + 0.5) / width
    X, Y = np.meshgrid(x, y)
    sky = np.stack([0.3 + 0.5 * (1 - Y), 0.4 + 0.5 * (1 - Y), 0.6 + 0.6 * (1 - Y)], -1)
    rng = np.random.Generator(np.random.PCG64(seed))
    sky *= 1.0 + 0.2 * rng.random((height, width, 1))
    d2 = (X - sun[0]) ** 2 + ((Y - sun[1]) * height / width) ** 2
    sky += (d2 < 0.0012)[..., None] * np.array([sun_power, 0.9 * sun_power, 0.7 * sun_power])
    return sky.astype(np.float32)


def _checker(n, a, b, cells=8, alpha=None):
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    m = ((i * cells // n) + (j * cells // n)) % 2
    img = np.where(m[..., None] == 0, np.array(a), np.array(b)).astype(np.float64)
    out = np.zeros((n, n, 4), np.uint8)
    out[..., :3] = np.clip(img * 255 + 0.5, 0, 255).astype(np.uint8)
    out[..., 3] = 255 if alpha is None else alpha
    return out


def feature_scene(kind, width=96, height=64, spp=4, max_depth=5, seed=0xFEA70005):
    """Small scenes exercising one option path each (C5 features):

    ``envmap``       importance-sampled HDR sky (OPT_ENVMAP), rotated, scaled
    ``textures``     base-colour, metallic-roughness, normal and emission maps
    ``media``        scatter / absorb / emissive homogeneous media in boxes
                     (volume_cube.scene style: spectrans 1, blend opacity 0)
    ``volmis``       the media scene with volume MIS (plain kernel only)
    ``alpha_mask``   alpha-masked leaf quads (MASK: deterministic in AnyHit)
    ``alpha_blend``  stochastic alpha (BLEND, plain kernel only)
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    sc = Scene()
    extra = {}
    floor_p, floor_n, floor_t = _uv_quad((-4, 0, 4), (8, 0, 0), (0, 0, -8), tiles=4.0)
    wall_p, wall_n, wall_t = _uv_quad((-4, 0, -4), (8, 0, 0), (0, 5, 0), tiles=2.0)
    floor = sc.AddMesh(floor_p, floor_n, floor_t, name="floor")
    wall = sc.AddMesh(wall_p, wall_n, wall_t, name="wall")
    bpos, bnrm = _unit_box()
    box = sc.AddMesh(bpos, bnrm, name="box")
    sph_p, sph_n = displaced_icosphere(6, 0.08, seed)
    sph = sc.AddMesh(sph_p, sph_n, name="sphere")
    grey = sc.AddMaterial(baseColor=(0.7, 0.7, 0.7), roughness=0.6)
    m_floor, m_wall = grey, sc.AddMaterial(baseColor=(0.6, 0.5, 0.4), roughness=0.8)
    objs = [sc.AddMaterial(baseColor=rng.uniform(0.2, 0.9, 3), roughness=rng.uniform(0.1, 0.8),
                           metallic=float(k == 1), clearcoat=0.5 * (k == 2), ior=1.5) for k in range(3)]
    if kind == "textures":
        sc.SetTextureArraySize(64, 64)
        t_base = sc.AddTexture(_checker(64, (0.9, 0.8, 0.2), (0.2, 0.3, 0.8)), "base")
        mr = np.zeros((64, 64, 4), np.uint8)
        mr[..., 1] = (rng.random((64, 64)) * 200 + 40).astype(np.uint8)   # roughness (g)
        mr[..., 2] = np.where(rng.random((64, 64)) < 0.3, 255, 0)         # metallic (b)
        mr[..., 3] = 255
        t_mr = sc.AddTexture(mr, "metalrough")
        i, j = np.meshgrid(np.arange(64), np.arange(64), indexing="ij")
        nx, ny = 0.4 * np.sin(i / 64 * 6 * np.pi), 0.4 * np.cos(j / 64 * 4 * np.pi)
        nz = np.sqrt(1 - nx * nx - ny * ny)
        nm = np.stack([nx, ny, nz], -1) * 0.5 + 0.5
        nmap = np.concatenate([(nm * 255 + 0.5).astype(np.uint8), np.full((64, 64, 1), 255, np.uint8)], -1)
        t_n = sc.AddTexture(nmap, "normal")
        t_e = sc.AddTexture(_checker(64, (1, 0.5, 0.1), (0, 0, 0), cells=4), "emission")
        m_floor = sc.AddMaterial(baseColor=(1, 1, 1), roughness=0.5, baseColorTexId=t_base,
                                 metallicRoughnessTexID=t_mr, normalmapTexID=t_n)
        m_wall = sc.AddMaterial(baseColor=(0.8, 0.8, 0.8), roughness=0.4, normalmapTexID=t_n,
                                emissionmapTexID=t_e)
        extra["opengl_normal_map"] = 1
    sc.AddMeshInstance(floor, mat4(), m_floor)
    sc.AddMeshInstance(wall, mat4(), m_wall)
    sc.AddMeshInstance(box, mat4(translate=(-1.6, 0.6, -0.5), scale=(0.6, 0.6, 0.6), rot=random_rotation(rng)),
                       objs[0])
    sc.AddMeshInstance(sph, mat4(translate=(1.5, 0.9, 0.0), scale=(0.9, 0.9, 0.9)), objs[1])
    sc.AddMeshInstance(sph, mat4(translate=(0.0, 0.5, 1.5), scale=(0.5, 0.5, 0.5)), objs[2])
    if kind in ("media", "volmis"):
        # volume_cube.scene: refractive, fully blended-away surface around a medium
        for k, (mt, col, dens, pos) in enumerate([(1, (0.9, 0.4, 0.1), 2.0, (-0.2, 1.0, -1.6)),
                                                   (2, (0.2, 0.5, 0.9), 1.5, (1.8, 0.8, -2.0)),
                                                   (3, (0.5, 0.2, 0.1), 0.6, (-2.2, 1.2, 1.6))]):
            vm = sc.AddMaterial(baseColor=(1, 1, 1), specTrans=1.0, roughness=0.0, mediumType=mt,
                                mediumDensity=dens, mediumColor=col, alphaMode=1, opacity=0.0,
                                mediumAnisotropy=0.3 * k)
            sc.AddMeshInstance(box, mat4(translate=pos, scale=(0.7, 0.7, 0.7)), vm)
        if kind == "volmis":
            extra["enable_volume_mis"] = 1
    if kind in ("alpha_mask", "alpha_blend"):
        sc.SetTextureArraySize(32, 32)
        leaf = np.zeros((32, 32, 4), np.uint8)
        i, j = np.meshgrid(np.arange(32), np.arange(32), indexing="ij")
        leaf[..., 1] = 160
        leaf[..., 0] = 60
        leaf[..., 3] = np.where((i - 16) ** 2 + (j - 16) ** 2 * 2.5 < 180, 255, 0)
        t_leaf = sc.AddTexture(leaf, "leaf")
        q_p, q_n, q_t = _uv_quad((-0.5, 0, 0), (1, 0, 0), (0, 1, 0))
        quad = sc.AddMesh(q_p, q_n, q_t, name="leafquad")
        if kind == "alpha_mask":
            am = sc.AddMaterial(baseColor=(1, 1, 1), roughness=0.7, baseColorTexId=t_leaf, alphaMode=2,
                                alphaCutoff=0.5)
        else:
            am = sc.AddMaterial(baseColor=(0.4, 0.8, 0.3), roughness=0.7, alphaMode=1, opacity=0.45)
        for k in range(10):
            xf = mat4(translate=(rng.uniform(-3, 3), rng.uniform(0.2, 2.5), rng.uniform(-2.5, 2)),
                      scale=(rng.uniform(0.6, 1.4),) * 3, rot=random_rotation(rng))
            sc.AddMeshInstance(quad, xf, am)
    sc.AddLight(quad_light((-1, 4.5, -1), (1, 4.5, -1), (-1, 4.5, 1), (9, 8, 7)))
    if kind == "envmap":
        sc.AddEnvMap(sky_envmap())
        extra.update(enable_envmap=1, envmap_intensity=1.5, envmap_rot=30.0)
    sc.AddCamera((0.0, 2.6, 7.5), (0.0, 0.8, 0.0), 50.0)
    sc.ProcessScene()
    return sc, RenderSettings(width, height, spp, max_depth, enable_rr=True, rr_depth=2, extra=extra)


def leaf_quads(num_quads, seed, size=(0.12, 0.3)):
    """Triangle soup of `num_quads` randomly oriented square quads (UVs 0..1)
    with centres uniform in the unit ball: a bush of leaf cards."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.normal(size=(num_quads, 3))
    c *= (rng.random(num_quads) ** (1 / 3) / np.linalg.norm(c, axis=1))[:, None]
    a = rng.normal(size=(num_quads, 3))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    b = np.cross(a, rng.normal(size=(num_quads, 3)))
    b /= np.linalg.norm(b, axis=1, keepdims=True)
    h = rng.uniform(size[0], size[1], num_quads)[:, None] * 0.5
    u, v = a * h, b * h
    p0, p1, p2, p3 = c - u - v, c + u - v, c + u + v, c - u + v
    pos = np.stack([np.stack([p0, p1, p2], 1), np.stack([p0, p2, p3], 1)], 1).reshape(-1, 3, 3)
    n = np.cross(u, v)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    nrm = np.repeat(n[:, None, :], 6, axis=1).reshape(-1, 3, 3)
    t = np.array([[[0, 0], [1, 0], [1, 1]], [[0, 0], [1, 1], [0, 1]]], np.float32)
    uvs = np.tile(t, (num_quads, 1, 1))
    return pos.astype(np.float32), nrm.astype(np.float32), uvs


def leaf_texture(n=64, seed=0x1EAF):
    """RGBA8 leaf card: green with a binary alpha mask (an ellipse with a
    seeded ragged edge)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    i, j = np.meshgrid((np.arange(n) + 0.5) / n - 0.5, (np.arange(n) + 0.5) / n - 0.5, indexing="ij")
    ang = np.arctan2(i, j)
    edge = 0.42 + 0.05 * np.sin(7 * ang + rng.uniform(0, 6.3)) + 0.02 * rng.standard_normal((n, n))
    inside = (i / 1.0) ** 2 + (j / 0.55) ** 2 < edge ** 2
    img = np.zeros((n, n, 4), np.uint8)
    img[..., 0] = 70
    img[..., 1] = (150 + 60 * rng.random((n, n))).astype(np.uint8)
    img[..., 2] = 40
    img[..., 3] = np.where(inside, 255, 0)
    return img


HEAVY_SPHERE_FREQS = (32, 45, 64, 128, 316)    # 20,480 ... 1,997,120 triangles
HEAVY_LEAF_QUADS = (10000, 20000, 40000)         # 20k, 40k, 80k triangles
# instance counts per mesh (spheres then leaf clusters), 1000 in total
HEAVY_COUNTS = (611, 25, 6, 2, 1, 330, 20, 5)


def heavy_mixed(width=3840, height=2160, spp=16, max_depth=12, seed=0xC5EA0005, sphere_freqs=HEAVY_SPHERE_FREQS,
                leaf_quads_per_mesh=HEAVY_LEAF_QUADS, counts=HEAVY_COUNTS, extent=80.0, sky=(2048, 1024),
                num_materials=16, num_volumes=2):
    """Heavy mixed scene: instances of 8 seeded meshes (noise-displaced
    spheres and alpha-tested leaf-quad clusters) spread over a ground plane,
    two scattering volume cubes (density 0.1-2, albedo 0.8) and a procedural
    HDR sky that is importance-sampled; no analytic lights. The leaf material
    is stochastic (BLEND) alpha with opacity 1 times a binary texture alpha;
    with media present shadow rays ignore alpha (anyhit.glsl:118-141).
    `sky=None` lights the scene with a constant environment instead."""
    assert len(counts) == len(sphere_freqs) + len(leaf_quads_per_mesh)
    rng = np.random.Generator(np.random.PCG64(seed))
    sc = Scene()
    sc.SetTextureArraySize(64, 64)
    t_leaf = sc.AddTexture(leaf_texture(64, seed ^ 0x1EAF), "leaf")
    meshes, kinds = [], []
    for k, f in enumerate(sphere_freqs):
        p, n = displaced_icosphere(f, 0.1, seed + 17 * k)
        meshes.append(sc.AddMesh(p, n, name=f"sphere{f}"))
        kinds.append("sphere")
    for k, nq in enumerate(leaf_quads_per_mesh):
        p, n, t = leaf_quads(nq, seed + 101 * k)
        meshes.append(sc.AddMesh(p, n, t, name=f"leaves{nq}"))
        kinds.append("leaves")
    g_p, g_n, g_t = _uv_quad((-extent, 0, extent), (2 * extent, 0, 0), (0, 0, -2 * extent), tiles=8.0)
    ground = sc.AddMesh(g_p, g_n, g_t, name="ground")
    bpos, bnrm = _unit_box()
    box = sc.AddMesh(bpos, bnrm, name="box")
    mids = [sc.AddMaterial(m) for m in disney_materials(rng, num_materials)]
    m_leaf = sc.AddMaterial(baseColor=(1, 1, 1), roughness=0.7, baseColorTexId=t_leaf, alphaMode=1, opacity=1.0)
    m_ground = sc.AddMaterial(baseColor=(0.45, 0.42, 0.38), roughness=0.9)
    sc.AddMeshInstance(ground, mat4(), m_ground)
    order = []
    for mesh, kind, cnt in zip(meshes, kinds, counts):
        order += [(mesh, kind)] * cnt
    for idx in rng.permutation(len(order)):
        mesh, kind = order[idx]
        x, z = rng.uniform(-extent * 0.9, extent * 0.9, 2)
        if kind == "sphere":
            s = rng.uniform(0.5, 2.0)
            xf = mat4(translate=(x, 1.1 * s, z), scale=(s, s, s), rot=random_rotation(rng))
            sc.AddMeshInstance(mesh, xf, mids[int(rng.integers(num_materials))])
        else:
            s = rng.uniform(1.0, 3.0)
            xf = mat4(translate=(x, s, z), scale=(s, s, s), rot=random_rotation(rng))
            sc.AddMeshInstance(mesh, xf, m_leaf)
    for k in range(num_volumes):
        s = rng.uniform(4.0, 8.0)
        vm = sc.AddMaterial(baseColor=(1, 1, 1), specTrans=1.0, roughness=0.0, mediumType=2,
                            mediumDensity=float(rng.uniform(0.1, 2.0)), mediumColor=(0.8, 0.8, 0.8), alphaMode=1,
                            opacity=0.0, mediumAnisotropy=0.2)
        x, z = rng.uniform(-extent * 0.4, extent * 0.4, 2)
        # lifted off the ground: a bottom face coplanar with the ground would
        # make every ray below the cube an exact tie between two surfaces
        sc.AddMeshInstance(box, mat4(translate=(x, s + 0.05, z), scale=(s, s, s)), vm)
    if sky is not None:
        sc.AddEnvMap(sky_envmap(sky[0], sky[1], seed=seed ^ 0x5C1, sun=(0.3, 0.2), sun_power=2000.0))
    sc.AddCamera((0.0, extent * 0.16, extent * 1.2), (0.0, 2.0, 0.0), 40.0)
    sc.ProcessScene()
    if sky is None:
        return sc, RenderSettings(width, height, spp, max_depth, enable_rr=True, rr_depth=2,
                                  enable_uniform_light=True, uniform_light_col=(1.0, 1.0, 1.0))
    return sc, RenderSettings(width, height, spp, max_depth, enable_rr=True, rr_depth=2,
                              extra=dict(enable_envmap=1, envmap_intensity=1.0))


def tiny(width=32, height=32, spp=2, max_depth=3):
    """Two triangles under a quad light, seen from above."""
    sc = Scene()
    quad = np.array([[[-1, 0, -1], [1, 0, -1], [1, 0, 1]], [[-1, 0, -1], [1, 0, 1], [-1, 0, 1]]], np.float32)
    quad = quad[:, [0, 2, 1]]  # face up (+y)
    n = np.tile(np.array([0, 1, 0], np.float32), (2, 3, 1))
    mesh = sc.AddMesh(quad, n, name="floor")
    mid = sc.AddMaterial(baseColor=(0.8, 0.6, 0.4), roughness=0.4)
    sc.AddMeshInstance(mesh, mat4(), mid)
    sc.AddLight(quad_light((-0.3, 1.5, -0.3), (0.3, 1.5, -0.3), (-0.3, 1.5, 0.3), (10, 10, 10)))
    sc.AddCamera((0, 3, 2.5), (0, 0, 0), 45.0)
    sc.ProcessScene()
    return sc, RenderSettings(width, height, spp, max_depth, enable_rr=True, rr_depth=2)

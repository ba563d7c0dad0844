"""Instance scenes and transform sets for the TLAS rebuild tests.

Each case gives meshes (reference vertex layout), instances (mesh, material,
Mat4 data as 4x4 rows = columns) and a second set of transforms for the edit.
The cases aim at the branches of the reference's median build
(thirdparty/RadeonRays/bvh.cpp:68-243):
  random       general position, rotations, non-uniform scales
  coincident   every centroid equal: no partition, middle splits throughout
  clustered    a few distinct centroids with many duplicates
  chain        centroids spaced geometrically: one instance peeled per level
  signed_zero  flat instances whose world boxes hold -0.0 or +0.0 at the
               extreme, so a child box takes the sign of the first grown
  adjacent     centroids on two adjacent floats: the split plane rounds onto
               one of them and every element lands on one side (fallback split)
"""
import numpy as np

from glslpt.scene import Scene
from glslpt.scenes import random_rotation

CASES = ("random", "coincident", "clustered", "chain", "signed_zero", "adjacent")
QUAD_CASES = ("signed_zero", "adjacent")


def _soup(tris):
    p = np.asarray(tris, np.float32).reshape(-1, 3)
    n = np.zeros_like(p)
    n[:, 1] = 1.0
    uv = np.zeros((p.shape[0], 2), np.float32)
    v = np.concatenate([p, uv[:, :1]], 1).astype(np.float32)
    nn = np.concatenate([n, uv[:, 1:]], 1).astype(np.float32)
    return v, nn


def quad_mesh():
    """Two triangles in the y = 0 plane over x, z in [-1, 1]."""
    return _soup([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, -1], [1, 0, 1], [-1, 0, 1]])


def meshes(case, seed=0):
    if case in QUAD_CASES:
        return [quad_mesh()]
    rng = np.random.default_rng(seed)
    soup = (rng.normal(size=(12, 1, 3)) * 1.5 + rng.normal(size=(12, 3, 3)) * 0.4).reshape(-1, 3)
    tri = [[0, 0, 0], [1, 0.2, 0], [0, 1, 0.3]]
    return [_soup(soup), quad_mesh(), _soup(tri)]


def _xf(right, up, fwd, t):
    """Mat4 data (n, 4, 4): rows 0..2 = columns right, up, forward; row 3 = translation."""
    n = t.shape[0]
    m = np.zeros((n, 4, 4), np.float32)
    m[:, 0, :3] = right
    m[:, 1, :3] = up
    m[:, 2, :3] = fwd
    m[:, 3, :3] = t
    m[:, 3, 3] = 1.0
    return m


def transforms(case, n, seed):
    rng = np.random.default_rng(seed)
    ex, ey, ez = np.eye(3, dtype=np.float32)
    one = np.ones((n, 1), np.float32)
    if case == "random":
        m = np.zeros((n, 4, 4), np.float32)
        for i in range(n):
            r = random_rotation(rng)[:3, :3]
            s = np.float32(rng.uniform(0.3, 3.0)) * np.array([1.0, rng.uniform(0.5, 2.0), 1.0], np.float32)
            m[i, :3, :3] = r * s[:, None]
        m[:, 3, :3] = rng.uniform(-30, 30, (n, 3)).astype(np.float32)
        m[:, 3, 3] = 1.0
        return m
    if case == "coincident":
        r = random_rotation(rng)[:3, :3]
        t = np.repeat(rng.uniform(-5, 5, (1, 3)).astype(np.float32), n, 0)
        return _xf(r[0] * one, r[1] * one, r[2] * one, t)
    if case == "clustered":
        pts = rng.uniform(-10, 10, (4, 3)).astype(np.float32)
        t = pts[rng.integers(0, 4, n)]
        return _xf(ex * one, ey * one, ez * one, t)
    if case == "chain":
        t = np.zeros((n, 3), np.float32)
        t[:, 0] = 1.5 ** rng.permutation(n).astype(np.float64)
        return _xf(ex * one, ey * one, ez * one, t)
    if case == "signed_zero":
        flip = rng.random(n) < 0.7
        up = np.where(flip[:, None], -ey, ey).astype(np.float32)
        t = np.zeros((n, 3), np.float32)
        t[:, 0] = rng.uniform(-10, 10, n)
        t[:, 2] = rng.uniform(-10, 10, n)
        t[:, 1] = np.where(rng.random(n) < 0.5, np.float32(-0.0), np.float32(0.0))
        return _xf(ex * one, up, ez * one, t)
    if case == "adjacent":
        c0 = np.float32(1000.0)
        c1 = np.nextafter(c0, np.float32(2000.0))
        t = np.zeros((n, 3), np.float32)
        t[:, 0] = np.where(rng.random(n) < 0.5, c0, c1)
        return _xf(0.5 * ex * one, ey * one, ez * one, t)
    raise ValueError(case)


def instances(case, n, seed):
    """[(mesh, material, xf 4x4)] and the edited transforms (n, 4, 4)."""
    rng = np.random.default_rng(seed + 1000)
    nm = len(meshes(case, seed))
    mesh_ids = rng.integers(0, nm, n)
    mat_ids = rng.integers(0, 3, n)
    xf0 = transforms(case, n, seed)
    xf1 = transforms(case, n, seed + 1)
    return [(int(m), int(mt), xf0[i]) for i, (m, mt) in enumerate(zip(mesh_ids, mat_ids))], xf1


def build_scene(mesh_list, insts, num_materials=3):
    """Product scene (libglslpt_scene.so) from reference-layout meshes."""
    sc = Scene()
    for v, nn in mesh_list:
        p = v[:, :3].reshape(-1, 3, 3)
        nr = nn[:, :3].reshape(-1, 3, 3)
        uv = np.stack([v[:, 3], nn[:, 3]], 1).reshape(-1, 3, 2)
        sc.AddMesh(p, nr, uv)
    for _ in range(num_materials):
        sc.AddMaterial()
    for m, mt, xf in insts:
        sc.AddMeshInstance(m, xf, mt)
    sc.ProcessScene()
    return sc

"""Scene builder (product, libglslpt_scene.so) against the reference.

Golden fixtures (tests/golden/*.npz) hold the reference's flattened arrays for
fixed inputs, produced by the reference's own Scene / SplitBvh / Bvh /
BvhTranslator compiled from its sources (tests/golden/make_golden.py). The
product must reproduce them bit for bit. When the compiled reference harness
is present (oracle/_ref/libref_scene.so) randomised scenes are compared too.
Structural invariants (child boxes inside the parent, every triangle in
exactly one leaf, leaf sizes) hold independently of the reference.
"""
import os

import numpy as np
import pytest

import refscene
from glslpt.scene import Scene, mat4
from glslpt.scenes import random_rotation

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def build_product(meshes, insts, num_materials, camera=None):
    sc = Scene()
    for v, n in meshes:
        p = v[:, :3].reshape(-1, 3, 3)
        nn = n[:, :3].reshape(-1, 3, 3)
        uv = np.stack([v[:, 3], n[:, 3]], 1).reshape(-1, 3, 2)
        sc.AddMesh(p, nn, uv)
    for _ in range(num_materials):
        sc.AddMaterial()
    for m, mt, xf in insts:
        sc.AddMeshInstance(m, xf, mt)
    if camera is not None:
        sc.AddCamera(camera[:3], camera[3:6], camera[6])
    sc.ProcessScene()
    return sc


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32) if a.dtype == np.float32 else a


def load_fixture(name):
    d = np.load(os.path.join(GOLDEN, name))
    counts = d["mesh_tri_counts"]
    meshes, off = [], 0
    for c in counts:
        meshes.append((d["mesh_vertices_uvx"][off:off + 3 * c], d["mesh_normals_uvy"][off:off + 3 * c]))
        off += 3 * c
    insts = [(int(m), int(mt), xf.reshape(4, 4)) for m, mt, xf in
             zip(d["instance_mesh"], d["instance_material"], d["instance_transform"])]
    return d, meshes, insts


def camera_vec(sc):
    c = sc.camera()
    return np.array(list(c.position) + list(c.up) + list(c.right) + list(c.forward) + [c.fov, c.focalDist, c.aperture],
                    np.float32)


@pytest.mark.parametrize("name", ["scene_cornell.npz", "scene_spheres.npz"])
def test_golden_arrays_bit_exact(name):
    d, meshes, insts = load_fixture(name)
    sc = build_product(meshes, insts, int(d["num_materials"]), d["camera_input"])
    a = sc.arrays()
    assert a["top_level_index"] == int(d["top_level_index"])
    for k in ("nodes", "vert_indices", "vertices_uvx", "normals_uvy"):
        assert a[k].shape == d[k].shape, k
        assert np.array_equal(bits(a[k]), bits(d[k])), f"{name}: {k} differs from the reference"
    assert np.array_equal(bits(camera_vec(sc)), bits(d["camera"])), "camera basis differs from the reference"


def test_obj_loader_matches_reference(tmp_path):
    d = np.load(os.path.join(GOLDEN, "obj_sample.npz"))
    p = tmp_path / "sample.obj"
    p.write_text(str(d["obj_text"]))
    sc = Scene()
    mid = sc.AddMesh(filename=str(p))
    assert mid == 0
    sc.AddMaterial()
    sc.AddMeshInstance(0, mat4(), 0)
    sc.ProcessScene()
    a = sc.arrays()
    assert np.array_equal(bits(a["vertices_uvx"]), bits(d["vertices_uvx"]))
    assert np.array_equal(bits(a["normals_uvy"]), bits(d["normals_uvy"]))


def test_obj_missing_file_returns_minus_one():
    sc = Scene()
    assert sc.AddMesh(filename="/nonexistent/file.obj") == -1


def random_scene(seed, nmeshes=5, ninst=40):
    rng = np.random.Generator(np.random.PCG64(seed))
    meshes = []
    for k in range(nmeshes):
        n = int(rng.choice([1, 2, 3, 4, 7, 50, 600, 3000]))
        c = rng.normal(size=(n, 1, 3)) * rng.uniform(0.5, 5)
        p = (c + rng.normal(size=(n, 3, 3)) * rng.uniform(0.05, 1)).astype(np.float32)
        if k == 0:  # degenerate: many triangles with identical centroids
            p = np.repeat(p[:1], n, axis=0)
        nr = rng.normal(size=(n, 3, 3)).astype(np.float32)
        uv = rng.random((n, 3, 2)).astype(np.float32)
        v = np.concatenate([p.reshape(-1, 3), uv.reshape(-1, 2)[:, :1]], 1).astype(np.float32)
        nn = np.concatenate([nr.reshape(-1, 3), uv.reshape(-1, 2)[:, 1:]], 1).astype(np.float32)
        meshes.append((v, nn))
    insts = []
    for i in range(ninst):
        s = float(rng.uniform(0.3, 3))
        insts.append((int(rng.integers(0, nmeshes)), int(rng.integers(0, 3)),
                      mat4(translate=rng.uniform(-30, 30, 3), scale=(s, s * rng.uniform(0.5, 2), s),
                           rot=random_rotation(rng))))
    cam = np.concatenate([rng.normal(size=3) * 10, rng.normal(size=3), [rng.uniform(20, 90)]]).astype(np.float32)
    return meshes, insts, cam


@pytest.mark.skipif(not refscene.available(), reason="compiled reference harness not built here")
@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_random_scenes_match_reference(seed):
    meshes, insts, cam = random_scene(seed)
    r = refscene.build(meshes, insts, 3, camera=cam)
    sc = build_product(meshes, insts, 3, cam)
    a = sc.arrays()
    assert a["top_level_index"] == r["top_level_index"]
    for k in ("nodes", "vert_indices", "vertices_uvx", "normals_uvy"):
        assert np.array_equal(bits(a[k]), bits(r[k])), f"seed {seed}: {k}"
    assert np.array_equal(bits(camera_vec(sc)), bits(r["camera"]))


def test_bvh_invariants():
    meshes, insts, cam = random_scene(11)
    sc = build_product(meshes, insts, 3, cam)
    a = sc.arrays()
    nodes, tli, vi = a["nodes"], a["top_level_index"], a["vert_indices"]
    ntri = vi.shape[0]
    seen = np.zeros(ntri, np.int32)

    def walk(i, blas):
        lr = nodes[i, 6:9]
        kind = int(lr[2])
        if kind == 0:
            for c in (int(lr[0]), int(lr[1])):
                assert np.all(nodes[c, 0:3] >= nodes[i, 0:3]) and np.all(nodes[c, 3:6] <= nodes[i, 3:6])
                walk(c, blas)
        elif kind > 0:
            assert blas
            first, cnt = int(lr[0]), int(lr[1])
            assert 1 <= cnt <= 3
            seen[first:first + cnt] += 1
        else:
            assert not blas
            walk(int(lr[0]), True)

    walk(tli, False)
    # every BLAS row is referenced by exactly one leaf of its BLAS; instances
    # reuse BLASes, so count distinct BLAS roots once
    roots = {int(nodes[i, 6]) for i in range(tli, nodes.shape[0]) if nodes[i, 8] < 0}
    seen[:] = 0
    for rt in roots:
        walk(rt, True)
    assert np.all(seen == 1)
    # vertIndices rows reference unshared vertices (3 per triangle)
    assert np.all(vi[:, 1] == vi[:, 0] + 1) and np.all(vi[:, 2] == vi[:, 0] + 2)


def test_rebuild_instances_keeps_blas_and_updates_tlas():
    meshes, insts, cam = random_scene(5, ninst=10)
    sc = build_product(meshes, insts, 3, cam)
    before = sc.arrays()
    xf = mat4(translate=(100, 0, 0))
    sc.SetInstanceTransform(3, xf)
    sc.RebuildInstances()
    after = sc.arrays()
    tli = before["top_level_index"]
    assert np.array_equal(before["nodes"][:tli], after["nodes"][:tli])
    assert np.array_equal(after["transforms"][3], xf.reshape(16))
    # the edited TLAS equals a fresh build with the edited transform
    insts2 = list(insts)
    insts2[3] = (insts[3][0], insts[3][1], xf)
    fresh = build_product(meshes, insts2, 3, cam).arrays()
    assert np.array_equal(bits(fresh["nodes"]), bits(after["nodes"]))


def test_process_scene_rejects_bad_input():
    sc = Scene()
    with pytest.raises(RuntimeError):
        sc.ProcessScene()  # no meshes
    sc.AddMesh(np.zeros((1, 3, 3), np.float32) + np.eye(3, dtype=np.float32))
    sc.AddMeshInstance(0, mat4(), 5)  # material 5 missing
    with pytest.raises(RuntimeError):
        sc.ProcessScene()

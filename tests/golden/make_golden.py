"""Writes the golden fixtures of tests/golden/ from the reference's own host
C++ (Scene, Mesh, Camera, RadeonRays Bvh/SplitBvh/BvhTranslator), compiled
from the reference sources into oracle/_ref/libref_scene.so by oracle/Makefile.

Run in a container where /root/reference exists:
    make -C oracle ref && python tests/golden/make_golden.py

Fixtures (inputs and the reference's outputs, numpy .npz):
  scene_cornell.npz   Cornell box: the reference's cbox_*.obj meshes as loaded by
                      its OBJ loader, placed as assets/cornell_box_orig.scene does
                      (scale 0.01, position), camera from that file.
  scene_spheres.npz   3 meshes (two geodesic spheres, a random soup), 12 instances
                      with rotations and scales, camera from a seed.
  obj_sample.npz      a generated OBJ (quads, negative indices, texture coords)
                      and the triangle soup the reference loader makes of it;
                      the OBJ text itself is stored in the same file.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "glsl-pathtracer_amd"))

import refscene  # noqa: E402
from glslpt.scene import mat4  # noqa: E402
from glslpt.scenes import geodesic_sphere, random_rotation  # noqa: E402

REF = "/root/reference"

# assets/cornell_box_orig.scene:51-107 (mesh file, position) and :11-16 (camera)
CORNELL = [("cbox_ceiling.obj", (.278, .5488, .27955)), ("cbox_floor.obj", (.2756, 0, .2796)),
           ("cbox_back.obj", (.2764, .2744, .5592)), ("cbox_smallbox.obj", (.1855, .0835, .169)),
           ("cbox_largebox.obj", (.3685, .166, .35125)), ("cbox_greenwall.obj", (0, .2744, .2796)),
           ("cbox_redwall.obj", (.5536, .2744, .2796))]
CORNELL_CAMERA = (.276, .275, -.75, .276, .275, .10, 40.0)


def save(name, meshes, instances, num_materials, camera, out):
    np.savez_compressed(
        os.path.join(HERE, name),
        mesh_vertices_uvx=np.concatenate([m[0] for m in meshes]),
        mesh_normals_uvy=np.concatenate([m[1] for m in meshes]),
        mesh_tri_counts=np.array([m[0].shape[0] // 3 for m in meshes], np.int32),
        instance_mesh=np.array([i[0] for i in instances], np.int32),
        instance_material=np.array([i[1] for i in instances], np.int32),
        instance_transform=np.stack([np.asarray(i[2], np.float32).reshape(16) for i in instances]),
        num_materials=np.int32(num_materials),
        camera_input=np.asarray(camera, np.float32),
        nodes=out["nodes"], top_level_index=np.int32(out["top_level_index"]),
        vert_indices=out["vert_indices"], vertices_uvx=out["vertices_uvx"], normals_uvy=out["normals_uvy"],
        camera=out["camera"])


def cornell():
    meshes, insts = [], []
    for i, (f, pos) in enumerate(CORNELL):
        meshes.append(refscene.load_obj(os.path.join(REF, "assets", "cornell_box", f)))
        insts.append((i, i, mat4(translate=pos, scale=(0.01, 0.01, 0.01))))
    out = refscene.build(meshes, insts, len(CORNELL), camera=CORNELL_CAMERA)
    save("scene_cornell.npz", meshes, insts, len(CORNELL), CORNELL_CAMERA, out)


def soup(rng, n):
    c = rng.normal(size=(n, 1, 3)) * 2
    p = (c + rng.normal(size=(n, 3, 3)) * 0.4).astype(np.float32)
    nr = rng.normal(size=(n, 3, 3)).astype(np.float32)
    uv = rng.random((n, 3, 2)).astype(np.float32)
    v = np.concatenate([p.reshape(-1, 3), uv.reshape(-1, 2)[:, :1]], 1).astype(np.float32)
    nn = np.concatenate([nr.reshape(-1, 3), uv.reshape(-1, 2)[:, 1:]], 1).astype(np.float32)
    return v, nn


def sphere_mesh(freq, rng):
    v, f = geodesic_sphere(freq)
    p = (v[f] * (1 + 0.05 * rng.standard_normal((f.shape[0], 3, 1)))).astype(np.float32)
    nrm = v[f].astype(np.float32)
    tx = np.zeros(p.shape[:2] + (1,), np.float32)
    ty = np.ones(p.shape[:2] + (1,), np.float32)
    return (np.concatenate([p, tx], 2).reshape(-1, 4), np.concatenate([nrm, ty], 2).reshape(-1, 4))


def spheres():
    rng = np.random.Generator(np.random.PCG64(20261202))
    meshes = [sphere_mesh(3, rng), sphere_mesh(5, rng), soup(rng, 300)]
    insts = []
    for i in range(12):
        s = float(rng.uniform(0.5, 2.0))
        insts.append((i % 3, i % 4, mat4(translate=rng.uniform(-8, 8, 3), scale=(s, s, s), rot=random_rotation(rng))))
    cam = (1.5, 9.0, 14.0, 0.3, -0.2, 0.1, 50.0)
    out = refscene.build(meshes, insts, 4, camera=cam)
    save("scene_spheres.npz", meshes, insts, 4, cam, out)


def obj_text():
    """A small OBJ: a cube of quads with vt and vn, a pentagon, negative indices."""
    lines = ["# generated", "o cube"]
    vs = [(x, y, z) for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)]
    lines += [f"v {x:.6f} {y:.6f} {z:.6f}" for x, y, z in vs]
    lines += ["vt 0 0", "vt 1 0", "vt 1 1", "vt 0 1", "vt 0.5 0.25"]
    lines += ["vn 1 0 0", "vn -1 0 0", "vn 0 1 0", "vn 0 -1 0", "vn 0 0 1", "vn 0 0 -1"]
    faces = [((1, 2, 4, 3), 2), ((5, 7, 8, 6), 1), ((3, 4, 8, 7), 3), ((1, 5, 6, 2), 4), ((2, 6, 8, 4), 5),
             ((1, 3, 7, 5), 6)]
    for (a, b, c, d), n in faces:
        lines.append(f"f {a}/1/{n} {b}/2/{n} {c}/3/{n} {d}/4/{n}")
    lines += ["v 3 0 0", "v 4 0 0", "v 4.5 1 0", "v 3.5 1.8 0", "v 2.5 1 0"]
    lines.append("f -5/-5/-1 -4/-4/-1 -3/-3/-1 -2/-2/-1 -1/-1/-1")
    return "\n".join(lines) + "\n"


def obj_sample():
    txt = obj_text()
    path = os.path.join(HERE, "_sample_tmp.obj")
    with open(path, "w") as f:
        f.write(txt)
    try:
        v, n = refscene.load_obj(path)
    finally:
        os.remove(path)
    np.savez_compressed(os.path.join(HERE, "obj_sample.npz"), obj_text=np.array(txt), vertices_uvx=v, normals_uvy=n)


if __name__ == "__main__":
    if not refscene.available():
        sys.exit("oracle/_ref/libref_scene.so is missing: run `make -C oracle ref` where /root/reference exists")
    cornell()
    spheres()
    obj_sample()
    print("fixtures written to", HERE)

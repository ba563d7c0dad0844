"""Loader for oracle/_ref/libref_scene.so: the reference's own host C++
(Scene, Mesh, Camera, RadeonRays BVH + translator) compiled from its sources.
Test infrastructure only: it produces the expected flattened arrays.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "_ref", "libref_scene.so")
P = C.POINTER


def available():
    return os.path.exists(PATH)


_lib = None


def lib():
    global _lib
    if _lib is None:
        l = C.CDLL(PATH)
        l.ref_load_obj.argtypes = [C.c_char_p, P(C.c_int), P(C.c_float), P(C.c_float)]
        l.ref_build_scene.argtypes = [C.c_int, P(C.c_int), P(C.c_float), P(C.c_float), C.c_int, C.c_int,
                                      P(C.c_int), P(C.c_int), P(C.c_float), P(C.c_float)]
        l.ref_get_sizes.argtypes = [P(C.c_int)] * 4
        l.ref_get_arrays.argtypes = [P(C.c_float), P(C.c_int), P(C.c_float), P(C.c_float), P(C.c_float)]
        l.ref_rebuild_instances.argtypes = [P(C.c_float)]
        _lib = l
    return _lib


def fp(a):
    return a.ctypes.data_as(P(C.c_float))


def ip(a):
    return a.ctypes.data_as(P(C.c_int))


def load_obj(filename):
    """Reference OBJ loader: (verticesUVX[n,4], normalsUVY[n,4])."""
    n = C.c_int()
    if lib().ref_load_obj(filename.encode(), C.byref(n), None, None) != 0:
        raise IOError(filename)
    v = np.zeros((n.value, 4), np.float32)
    nn = np.zeros((n.value, 4), np.float32)
    lib().ref_load_obj(filename.encode(), C.byref(n), fp(v), fp(nn))
    return v, nn


def build(meshes, instances, num_materials, camera=None):
    """meshes: list of (verticesUVX[3n,4], normalsUVY[3n,4]); instances: list of
    (mesh_id, mat_id, transform[4,4] as Mat4 data). Returns the reference's arrays."""
    counts = np.array([m[0].shape[0] // 3 for m in meshes], np.int32)
    v = np.ascontiguousarray(np.concatenate([m[0] for m in meshes]), np.float32)
    n = np.ascontiguousarray(np.concatenate([m[1] for m in meshes]), np.float32)
    mids = np.array([i[0] for i in instances], np.int32)
    mats = np.array([i[1] for i in instances], np.int32)
    xf = np.ascontiguousarray(np.stack([np.asarray(i[2], np.float32).reshape(16) for i in instances]), np.float32)
    cam = None if camera is None else np.asarray(camera, np.float32)
    lib().ref_build_scene(len(meshes), ip(counts), fp(v), fp(n), num_materials, len(instances), ip(mids),
                          ip(mats), fp(xf), None if cam is None else fp(cam))
    s = [C.c_int() for _ in range(4)]
    lib().ref_get_sizes(*[C.byref(x) for x in s])
    nn, tli, nt, nv = [x.value for x in s]
    nodes = np.zeros((nn, 9), np.float32)
    vi = np.zeros((nt, 3), np.int32)
    vu = np.zeros((nv, 4), np.float32)
    nu = np.zeros((nv, 4), np.float32)
    camo = np.zeros(15, np.float32)
    lib().ref_get_arrays(fp(nodes), ip(vi), fp(vu), fp(nu), fp(camo))
    return dict(nodes=nodes, top_level_index=tli, vert_indices=vi, vertices_uvx=vu, normals_uvy=nu, camera=camo)


def rebuild_instances(transforms):
    """Reference Scene::RebuildInstances of the last build with new transforms
    (n, 4, 4) in Mat4 data order; returns the rebuilt flattened nodes."""
    xf = np.ascontiguousarray(transforms, np.float32).reshape(-1, 16)
    if lib().ref_rebuild_instances(fp(xf)) != 0:
        raise RuntimeError("no reference scene built")
    s = [C.c_int() for _ in range(4)]
    lib().ref_get_sizes(*[C.byref(x) for x in s])
    nn, tli, nt, nv = [x.value for x in s]
    nodes = np.zeros((nn, 9), np.float32)
    vi = np.zeros((nt, 3), np.int32)
    vu = np.zeros((nv, 4), np.float32)
    nu = np.zeros((nv, 4), np.float32)
    lib().ref_get_arrays(fp(nodes), ip(vi), fp(vu), fp(nu), None)
    return nodes


def load_scene_file(filename):
    """Reference .scene loader + ProcessScene: render options (35 values in
    pts_render_options order), flattened arrays, records and decoded textures."""
    l = lib()
    l.ref_load_scene_file.argtypes = [C.c_char_p, P(C.c_float)]
    l.ref_get_counts.argtypes = [P(C.c_int)]
    l.ref_get_records.argtypes = [P(C.c_float), P(C.c_float), P(C.c_float)]
    l.ref_get_texture.argtypes = [C.c_int, P(C.c_int), P(C.c_int), P(C.c_uint8)]
    opts = np.zeros(35, np.float32)
    if l.ref_load_scene_file(str(filename).encode(), fp(opts)) != 0:
        raise IOError(filename)
    s = [C.c_int() for _ in range(4)]
    l.ref_get_sizes(*[C.byref(x) for x in s])
    nn, tli, nt, nv = [x.value for x in s]
    nodes = np.zeros((nn, 9), np.float32)
    vi = np.zeros((nt, 3), np.int32)
    vu = np.zeros((nv, 4), np.float32)
    nu = np.zeros((nv, 4), np.float32)
    camo = np.zeros(15, np.float32)
    l.ref_get_arrays(fp(nodes), ip(vi), fp(vu), fp(nu), fp(camo))
    cnt = np.zeros(7, np.int32)
    l.ref_get_counts(ip(cnt))
    mats = np.zeros((cnt[0], 32), np.float32)
    lights = np.zeros((cnt[1], 15), np.float32)
    xf = np.zeros((cnt[2], 16), np.float32)
    l.ref_get_records(fp(mats), fp(lights), fp(xf))
    textures = []
    for i in range(cnt[3]):
        w, h = C.c_int(), C.c_int()
        l.ref_get_texture(i, C.byref(w), C.byref(h), None)
        t = np.zeros((h.value, w.value, 4), np.uint8)
        l.ref_get_texture(i, C.byref(w), C.byref(h), t.ctypes.data_as(P(C.c_uint8)))
        textures.append(t)
    env = None
    if cnt[6]:
        l.ref_get_env.argtypes = [P(C.c_int), P(C.c_int), P(C.c_float), P(C.c_float), P(C.c_float)]
        w, h, tot = C.c_int(), C.c_int(), C.c_float()
        l.ref_get_env(C.byref(w), C.byref(h), None, None, None)
        img = np.zeros((h.value, w.value, 3), np.float32)
        cdf = np.zeros(w.value * h.value, np.float32)
        l.ref_get_env(C.byref(w), C.byref(h), fp(img), fp(cdf), C.byref(tot))
        env = dict(img=img, cdf=cdf, total=tot.value)
    return dict(options=opts, nodes=nodes, top_level_index=tli, vert_indices=vi, vertices_uvx=vu, normals_uvy=nu,
                camera=camo, materials=mats, lights=lights, transforms=xf, textures=textures,
                tex_array=(int(cnt[4]), int(cnt[5])), has_env=bool(cnt[6]), env=env)

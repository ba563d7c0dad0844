"""ctypes face of tests/hostpath/_build/libhostpath.so: the renderer's
per-path device code (pt_path.h, pt_replay.h over the pt_layout.h scene
layout) compiled for the CPU. Test infrastructure, checked against the oracle;
the product never loads it."""
import ctypes as C
import os
import subprocess

import numpy as np

from glslpt.renderer import options_for

ROOT = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(ROOT, "hostpath", "_build", "libhostpath.so")
P = C.POINTER

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "hostpath")])


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = C.CDLL(PATH)
        _lib.host_render.restype = C.c_int
        _lib.host_adjoint.restype = C.c_int
    return _lib


def _view(sc, materials=None):
    from glslpt import _capi as api
    v = sc.view()
    keep = None
    if materials is not None:
        keep = np.ascontiguousarray(materials, np.float32)
        v.materials = C.cast(keep.ctypes.data, P(api.PtMaterial))
    return v, keep


def render(sc, rs, first, n, **opts):
    """Accumulation (height, width, 4) of samples [first, first + n) by the
    plain kernel's per-pixel loop on the CPU."""
    v, _ = _view(sc)
    o = options_for(rs, **opts)
    cam = sc.camera()
    acc = np.zeros((rs.height, rs.width, 4), np.float32)
    rc = lib().host_render(C.byref(v), C.byref(o), C.byref(cam), int(first), int(n), acc.ctypes.data_as(P(C.c_float)))
    if rc != 0:
        raise RuntimeError("host_render failed")
    return acc


def adjoint(sc, rs, dl_dimage, first, n, materials=None, medium=False, **opts):
    """(gradient (num_materials, 7) float64, dropped vertices, paths whose
    replayed radiance differs from the primal's) of the adjoint's primal +
    replay passes on the CPU; with medium, (gradient, medium gradient
    (num_materials, 4), dropped, mismatched)."""
    v, _ = _view(sc, materials)
    o = options_for(rs, **opts)
    cam = sc.camera()
    g = np.zeros((v.num_materials, 7), np.float64)
    mg = np.zeros((v.num_materials, 4), np.float64)
    w = np.ascontiguousarray(dl_dimage, np.float32)
    dropped, mismatch = C.c_int(0), C.c_int(0)
    rc = lib().host_adjoint(C.byref(v), C.byref(o), C.byref(cam), int(first), int(n), w.ctypes.data_as(P(C.c_float)),
                            g.ctypes.data_as(P(C.c_double)),
                            mg.ctypes.data_as(P(C.c_double)) if medium else None,
                            C.byref(dropped), C.byref(mismatch))
    if rc != 0:
        raise RuntimeError("host_adjoint failed")
    if medium:
        return g, mg, dropped.value, mismatch.value
    return g, dropped.value, mismatch.value


def closest_split(sc, rs, origins, dirs, budget, **opts):
    """(rays whose hit differs in any bit between the whole traversal and the
    traversal in pieces of `budget` iterations with the stack moved between
    pieces, pieces traced) on the CPU."""
    v, _ = _view(sc)
    o = options_for(rs, **opts)
    cam = sc.camera()
    ro = np.ascontiguousarray(origins, np.float32)
    rd = np.ascontiguousarray(dirs, np.float32)
    differ, pieces = C.c_int(0), C.c_longlong(0)
    rc = lib().host_closest_split(C.byref(v), C.byref(o), C.byref(cam), int(ro.shape[0]),
                                  ro.ctypes.data_as(P(C.c_float)), rd.ctypes.data_as(P(C.c_float)), int(budget),
                                  C.byref(differ), C.byref(pieces))
    if rc != 0:
        raise RuntimeError("host_closest_split failed")
    return differ.value, pieces.value


def any_orders(sc, rs, origins, dirs, maxd, **opts):
    """(rays whose any-hit answer differs between nearest-first and child-order
    visits, rays occluded) on the CPU."""
    v, _ = _view(sc)
    o = options_for(rs, **opts)
    cam = sc.camera()
    ro = np.ascontiguousarray(origins, np.float32)
    rd = np.ascontiguousarray(dirs, np.float32)
    md = np.ascontiguousarray(maxd, np.float32)
    differ, hits = C.c_int(0), C.c_int(0)
    rc = lib().host_any_orders(C.byref(v), C.byref(o), C.byref(cam), int(ro.shape[0]), ro.ctypes.data_as(P(C.c_float)),
                               rd.ctypes.data_as(P(C.c_float)), md.ctypes.data_as(P(C.c_float)), C.byref(differ),
                               C.byref(hits))
    if rc != 0:
        raise RuntimeError("host_any_orders failed")
    return differ.value, hits.value


def tlas4(sc):
    """(nodes (room, 32) float32, root) of the host re-layout's four-wide TLAS:
    breadth first, the root first, so entry i has array index root + i (root
    is -1 - instance when the TLAS is one instance)."""
    v, _ = _view(sc)
    room, root = C.c_int(0), C.c_int(0)
    if lib().host_tlas4(C.byref(v), None, 0, C.byref(room), C.byref(root)) != 0:
        raise RuntimeError("host_tlas4 failed")
    nodes = np.zeros((room.value, 32), np.float32)
    if lib().host_tlas4(C.byref(v), nodes.ctypes.data_as(P(C.c_float)), room.value, C.byref(room), C.byref(root)) != 0:
        raise RuntimeError("host_tlas4 failed")
    return nodes, root.value

"""The C-ABI libraries load and export every function include/glslpt.h
declares; error conventions hold without a GPU (no compute calls here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from glslpt import _capi as api
from glslpt.scene import Scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "glslpt.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(pts?_[a-z_0-9]+)\s*\(", txt))


def test_header_declares_both_surfaces():
    names = declared()
    assert {"pt_create", "pt_render", "pt_upload_scene", "pt_batch_render", "pt_adjoint"} <= names
    assert {"pts_create", "pts_add_mesh", "pts_process"} <= names


def test_scene_library_exports_declared_symbols():
    lib = C.CDLL(os.path.join(api.LIB_DIR, "libglslpt_scene.so"))
    for n in sorted(x for x in declared() if x.startswith("pts_")):
        assert hasattr(lib, n), n
    assert set(api.SCENE_SYMBOLS) == {x for x in declared() if x.startswith("pts_")}


def test_hip_library_exports_declared_symbols():
    path = os.path.join(api.LIB_DIR, "libglslpt_hip.so")
    assert os.path.exists(path), "HIP library not built"
    lib = C.CDLL(path)
    for n in sorted(x for x in declared() if x.startswith("pt_")):
        assert hasattr(lib, n), n
    assert set(api.HIP_SYMBOLS) == {x for x in declared() if x.startswith("pt_")}


def test_struct_sizes_match_header_layouts():
    assert C.sizeof(api.PtMaterial) == 128   # 8 x vec4 (src/core/Material.h)
    assert C.sizeof(api.PtLight) == 60       # 5 x vec3 (src/core/Scene.h:50-59)
    assert C.sizeof(api.PtBvhNode) == 36     # 3 x vec3 (bvh_translator.h:45-50)
    assert C.sizeof(api.PtCamera) == 60


def test_scene_error_codes():
    lib = api.scene_lib()
    h = C.c_void_p()
    assert lib.pts_create(C.byref(h)) == 0
    out = C.c_int32()
    # no positions
    assert lib.pts_add_mesh(h, None, None, None, 3, b"x", C.byref(out)) == api.PT_OK - 1
    assert b"positions" in lib.pts_last_error(h)
    # process before anything exists
    assert lib.pts_process(h) < 0
    v = api.PtSceneView()
    assert lib.pts_get_view(h, C.byref(v)) < 0
    assert lib.pts_destroy(h) == 0


def test_renderer_without_gpu_fails_loudly():
    """Without a GPU pt_create reports PT_ERR_NO_DEVICE; it first compares the
    record layouts the kernels' units and the host code were compiled with,
    which would return PT_ERR_STATE for a library linked from mismatched
    builds, so reaching PT_ERR_NO_DEVICE also shows that the layouts agree."""
    lib = api.hip_lib()
    n = C.c_int32(-1)
    assert lib.pt_device_count(C.byref(n)) == 0
    if n.value == 0:
        h = C.c_void_p()
        assert lib.pt_create(0, C.byref(h)) == -7, lib.pt_last_error(None)  # PT_ERR_NO_DEVICE
        from glslpt.renderer import Renderer
        with pytest.raises(api.PtError):
            Renderer(0)


def test_default_options_match_render_options():
    o = api.PtOptions()
    api.hip_lib().pt_default_options(C.byref(o))
    # RenderOptions() defaults (src/core/Renderer.h:41-72)
    assert (o.width, o.height, o.max_depth, o.enable_rr, o.rr_depth) == (1280, 720, 2, 1, 2)
    assert np.allclose(list(o.uniform_light_col), [0.3, 0.3, 0.3])
    assert o.envmap_intensity == 1.0 and o.opengl_normal_map == 1


def test_scene_view_counts():
    from glslpt import scenes
    sc, rs = scenes.cornell()
    v = sc.view()
    assert v.num_tris == 34 and v.num_instances == 7 and v.num_lights == 1 and v.num_materials == 7
    assert v.num_vertices == 102


HIPRENDERER_APP = r"""
#include <cstdio>
#include "glslpt_hip_renderer.hpp"
int main() {
    pts_scene s;
    if (pts_create(&s) != PT_OK) return 2;
    const float p[9] = {0, 0, 0, 1, 0, 0, 0, 1, 0};
    int32_t mesh, mat, inst;
    if (pts_add_mesh(s, p, nullptr, nullptr, 3, "tri", &mesh) != PT_OK) return 3;
    pt_material m{};
    m.baseColor[0] = m.baseColor[1] = m.baseColor[2] = 1.f; m.roughness = 0.5f; m.ior = 1.5f; m.opacity = 1.f;
    if (pts_add_material(s, &m, &mat) != PT_OK) return 4;
    const float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    if (pts_add_mesh_instance(s, mesh, I, mat, &inst) != PT_OK) return 5;
    const float eye[3] = {0.3f, 0.3f, 2.f}, at[3] = {0.3f, 0.3f, 0.f};
    if (pts_add_camera(s, eye, at, 45.f) != PT_OK) return 6;
    if (pts_process(s) != PT_OK) return 7;
    pt_options o;
    pt_default_options(&o);
    o.width = 16; o.height = 16;
    try {
        glslpt::HipRenderer r(s, o);
        for (int i = 0; i < 2; ++i) { r.Update(0.f); r.Render(); }
        unsigned char* img; int w, h;
        r.GetOutputBuffer(&img, w, h);
        std::printf("rendered %dx%d spp=%d first=%d\n", w, h, r.GetSampleCount(), img[0]);
        delete[] img;
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        pts_destroy(s);
        return 1;
    }
    pts_destroy(s);
    return 0;
}
"""


def test_hip_renderer_header_compiles(tmp_path):
    """include/glslpt_hip_renderer.hpp builds against both libraries with a
    plain host compiler; without a GPU the constructor throws the create
    error, with one it renders."""
    import subprocess
    src = tmp_path / "app.cpp"
    src.write_text(HIPRENDERER_APP)
    exe = tmp_path / "app"
    lib = api.LIB_DIR
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", str(src), "-I", os.path.join(ROOT, "include"),
                    "-L", lib, "-lglslpt_hip", "-lglslpt_scene", f"-Wl,-rpath,{lib}", "-o", str(exe)], check=True)
    n = C.c_int32(0)
    api.hip_lib().pt_device_count(C.byref(n))
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    if n.value == 0:
        assert out.returncode == 1 and "pt_create: no HIP device visible" in out.stdout, out
    else:
        assert out.returncode == 0 and "rendered 16x16 spp=2" in out.stdout, out

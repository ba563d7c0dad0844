"""ctypes declarations of include/glslpt.h and loaders for the two libraries.

The product libraries live in ``glslpt/lib/`` (built by the package Makefile).
Loading fails loudly when a library is missing: there is no Python fallback
for any part of the path.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")

PT_OK = 0
PT_ERR_INVALID_ARG = -1
PT_ERR_OVERFLOW = -4
PT_MODE_WAVEFRONT = 0
PT_MODE_PLAIN = 1

# pt_adjoint parameter order and the material-record float each one is
# (baseColor r, g, b, roughness, metallic, ior, clearcoat).
PT_ADJOINT_PARAMS = 7
ADJOINT_FIELDS = (0, 1, 2, 9, 8, 17, 14)
ADJOINT_NAMES = ("baseColor.r", "baseColor.g", "baseColor.b", "roughness", "metallic", "ior", "clearcoat")
# pt_adjoint_medium's medium parameters and their material-record floats
PT_ADJOINT_MEDIUM_PARAMS = 4
ADJOINT_MEDIUM_FIELDS = (19, 20, 21, 22)
ADJOINT_MEDIUM_NAMES = ("mediumDensity", "mediumColor.r", "mediumColor.g", "mediumColor.b")
STAGES = ("raygen", "extend", "shade", "shadow", "accum", "plain", "adjoint", "batch")

f3 = C.c_float * 3
i32 = C.c_int32
f32 = C.c_float
P = C.POINTER


class PtMaterial(C.Structure):
    _fields_ = [("baseColor", f3), ("anisotropic", f32), ("emission", f3), ("padding1", f32),
                ("metallic", f32), ("roughness", f32), ("subsurface", f32), ("specularTint", f32),
                ("sheen", f32), ("sheenTint", f32), ("clearcoat", f32), ("clearcoatGloss", f32),
                ("specTrans", f32), ("ior", f32), ("mediumType", f32), ("mediumDensity", f32),
                ("mediumColor", f3), ("mediumAnisotropy", f32),
                ("baseColorTexId", f32), ("metallicRoughnessTexID", f32), ("normalmapTexID", f32),
                ("emissionmapTexID", f32),
                ("opacity", f32), ("alphaMode", f32), ("alphaCutoff", f32), ("padding2", f32)]


class PtLight(C.Structure):
    _fields_ = [("position", f3), ("emission", f3), ("u", f3), ("v", f3),
                ("radius", f32), ("area", f32), ("type", f32)]


class PtCamera(C.Structure):
    _fields_ = [("position", f3), ("up", f3), ("right", f3), ("forward", f3),
                ("fov", f32), ("focalDist", f32), ("aperture", f32)]


class PtBvhNode(C.Structure):
    _fields_ = [("bboxmin", f3), ("bboxmax", f3), ("LRLeaf", f3)]


class PtSceneView(C.Structure):
    _fields_ = [("nodes", P(PtBvhNode)), ("num_nodes", i32), ("top_level_index", i32),
                ("vert_indices", P(i32)), ("num_tris", i32),
                ("vertices_uvx", P(f32)), ("normals_uvy", P(f32)), ("num_vertices", i32),
                ("materials", P(PtMaterial)), ("num_materials", i32),
                ("transforms", P(f32)), ("num_instances", i32),
                ("lights", P(PtLight)), ("num_lights", i32),
                ("texture_maps", P(C.c_uint8)), ("tex_width", i32), ("tex_height", i32),
                ("num_textures", i32),
                ("env_img", P(f32)), ("env_cdf", P(f32)), ("env_width", i32), ("env_height", i32),
                ("env_total_sum", f32)]


class PtOptions(C.Structure):
    _fields_ = [("width", i32), ("height", i32), ("max_depth", i32),
                ("enable_rr", i32), ("rr_depth", i32),
                ("enable_uniform_light", i32), ("uniform_light_col", f3),
                ("enable_envmap", i32), ("envmap_intensity", f32), ("envmap_rot", f32),
                ("hide_emitters", i32), ("enable_background", i32), ("transparent_background", i32),
                ("enable_roughness_mollification", i32), ("roughness_mollification_amt", f32),
                ("enable_volume_mis", i32), ("opengl_normal_map", i32),
                ("mode", i32), ("block_size", i32), ("stack_depth", i32), ("wave_spp", i32),
                ("use_graph", i32), ("count_work", i32), ("compact_every", i32), ("bvh_width", i32),
                ("nee_inline", i32), ("extend_cap", i32)]


class PtStats(C.Structure):
    _fields_ = [("ms_last_render", C.c_double), ("ms_stage", C.c_double * 8),
                ("launches_stage", C.c_int64 * 8),
                ("rays_closest", C.c_int64), ("nodes_visited", C.c_int64), ("tris_tested", C.c_int64),
                ("instances_entered", C.c_int64), ("lights_tested", C.c_int64),
                ("rays_shadow", C.c_int64), ("shadow_nodes", C.c_int64), ("shadow_tris", C.c_int64),
                ("shadow_instances", C.c_int64), ("shadow_lights", C.c_int64),
                ("stack_spills", C.c_int64), ("stack_overflow", i32), ("adjoint_dropped", i32),
                ("bounce_hits", C.c_int64 * 16)]


class PtJob(C.Structure):
    _fields_ = [("camera", PtCamera), ("width", i32), ("height", i32), ("spp", i32),
                ("max_depth", i32), ("accum_out", P(f32))]


class PtsRenderOptions(C.Structure):
    _fields_ = [("render_resolution", i32 * 2), ("window_resolution", i32 * 2),
                ("uniform_light_col", f32 * 3), ("background_col", f32 * 3),
                ("tile_width", i32), ("tile_height", i32), ("max_depth", i32), ("max_spp", i32), ("rr_depth", i32),
                ("tex_array_width", i32), ("tex_array_height", i32), ("denoiser_frame_cnt", i32),
                ("enable_rr", i32), ("enable_denoiser", i32), ("enable_tonemap", i32), ("enable_aces", i32),
                ("simple_aces_fit", i32), ("opengl_normal_map", i32), ("enable_envmap", i32),
                ("enable_uniform_light", i32), ("hide_emitters", i32), ("enable_background", i32),
                ("transparent_background", i32), ("independent_render_size", i32),
                ("enable_roughness_mollification", i32), ("enable_volume_mis", i32),
                ("envmap_intensity", f32), ("envmap_rot", f32), ("roughness_mollification_amt", f32)]


_SCENE_SIGS = {
    "pts_create": [P(C.c_void_p)],
    "pts_destroy": [C.c_void_p],
    "pts_add_mesh_obj": [C.c_void_p, C.c_char_p, P(i32)],
    "pts_add_mesh": [C.c_void_p, P(f32), P(f32), P(f32), i32, C.c_char_p, P(i32)],
    "pts_add_material": [C.c_void_p, P(PtMaterial), P(i32)],
    "pts_add_mesh_instance": [C.c_void_p, i32, P(f32), i32, P(i32)],
    "pts_add_light": [C.c_void_p, P(PtLight), P(i32)],
    "pts_add_texture": [C.c_void_p, P(C.c_uint8), i32, i32, C.c_char_p, P(i32)],
    "pts_add_camera": [C.c_void_p, P(f32), P(f32), f32],
    "pts_add_env_map": [C.c_void_p, P(f32), i32, i32],
    "pts_set_texture_array_size": [C.c_void_p, i32, i32],
    "pts_process": [C.c_void_p],
    "pts_rebuild_instances": [C.c_void_p],
    "pts_set_instance_transform": [C.c_void_p, i32, P(f32)],
    "pts_get_view": [C.c_void_p, P(PtSceneView)],
    "pts_get_camera": [C.c_void_p, P(PtCamera)],
    "pts_camera_offset_orientation": [C.c_void_p, f32, f32],
    "pts_get_blas_root": [C.c_void_p, i32, P(i32)],
    "pts_load_scene_file": [C.c_void_p, C.c_char_p, P(PtsRenderOptions)],
    "pts_get_texture": [C.c_void_p, i32, P(i32), P(i32), P(C.c_uint8)],
}

_HIP_SIGS = {
    "pt_device_count": [P(i32)],
    "pt_create": [i32, P(C.c_void_p)],
    "pt_destroy": [C.c_void_p],
    "pt_upload_scene": [C.c_void_p, P(PtSceneView)],
    "pt_update_instances": [C.c_void_p, P(f32), i32, P(PtMaterial), i32, P(PtBvhNode), i32, i32],
    "pt_set_options": [C.c_void_p, P(PtOptions)],
    "pt_set_camera": [C.c_void_p, P(PtCamera)],
    "pt_clear_accum": [C.c_void_p],
    "pt_render": [C.c_void_p, i32, i32],
    "pt_synchronize": [C.c_void_p],
    "pt_read_accum": [C.c_void_p, P(f32), C.c_size_t],
    "pt_read_ldr": [C.c_void_p, i32, i32, i32, i32, P(f32), P(C.c_uint8), C.c_size_t],
    "pt_get_stats": [C.c_void_p, P(PtStats)],
    "pt_reset_stats": [C.c_void_p],
    "pt_closest_hit": [C.c_void_p, P(f32), P(f32), i32, P(f32), P(i32), P(i32), P(i32), P(f32)],
    "pt_any_hit": [C.c_void_p, P(f32), P(f32), P(f32), i32, P(i32)],
    "pt_batch_render": [C.c_void_p, P(PtJob), i32],
    "pt_adjoint": [C.c_void_p, i32, i32, P(f32), P(f32)],
    "pt_adjoint_medium": [C.c_void_p, i32, i32, P(f32), P(f32), P(f32)],
    "pt_set_env_map": [C.c_void_p, P(f32), i32, i32],
    "pt_read_env_cdf": [C.c_void_p, P(f32), C.c_size_t, P(f32)],
    "pt_rebuild_instances": [C.c_void_p, P(f32), i32, P(i32)],
    "pt_read_tlas": [C.c_void_p, P(f32), i32, P(f32), i32, P(i32)],
    "pt_read_tlas4": [C.c_void_p, P(f32), i32, P(i32), P(i32)],
    "pt_env_cdf_noalloc": [P(f32), i32, i32, P(f32), P(f32), C.c_void_p],
}

# Every symbol include/glslpt.h declares, per library.
SCENE_SYMBOLS = tuple(_SCENE_SIGS) + ("pts_last_error", "pts_default_render_options")
HIP_SYMBOLS = tuple(_HIP_SIGS) + ("pt_last_error", "pt_default_options")

_scene_lib = None
_hip_lib = None


def _load(name):
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{name} is not built ({path} missing); run __graft_entry__.build() "
                           "or make -C glsl-pathtracer_amd")
    return C.CDLL(path)


def scene_lib():
    global _scene_lib
    if _scene_lib is None:
        lib = _load("libglslpt_scene.so")
        for fn, args in _SCENE_SIGS.items():
            getattr(lib, fn).argtypes = args
            getattr(lib, fn).restype = C.c_int
        lib.pts_last_error.argtypes = [C.c_void_p]
        lib.pts_last_error.restype = C.c_char_p
        lib.pts_default_render_options.argtypes = [P(PtsRenderOptions)]
        lib.pts_default_render_options.restype = None
        _scene_lib = lib
    return _scene_lib


def hip_lib():
    global _hip_lib
    if _hip_lib is None:
        lib = _load("libglslpt_hip.so")
        for fn, args in _HIP_SIGS.items():
            getattr(lib, fn).argtypes = args
            getattr(lib, fn).restype = C.c_int
        lib.pt_last_error.argtypes = [C.c_void_p]
        lib.pt_last_error.restype = C.c_char_p
        lib.pt_default_options.argtypes = [P(PtOptions)]
        lib.pt_default_options.restype = None
        _hip_lib = lib
    return _hip_lib


class PtError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed with {code}: {msg}")
        self.code = code


def fptr(a):
    """float32 numpy array -> POINTER(c_float) (None passes NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(P(f32))


def iptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(P(i32))

"""Python face of the host scene surface (libglslpt_scene.so).

Method names follow the reference's ``GLSLPT::Scene`` (src/core/Scene.h:74-84):
AddMesh, AddMaterial, AddMeshInstance, AddLight, AddCamera, AddEnvMap,
AddTexture, ProcessScene, RebuildInstances. Ids are returned like the
reference (AddMesh returns -1 when a file cannot be loaded, src/core/Scene.cpp:62-83).
"""
import ctypes as C

import numpy as np

from . import _capi as api

# Material defaults of src/core/Material.h:50-85.
MATERIAL_DEFAULTS = dict(
    baseColor=(1.0, 1.0, 1.0), anisotropic=0.0, emission=(0.0, 0.0, 0.0),
    metallic=0.0, roughness=0.5, subsurface=0.0, specularTint=0.0,
    sheen=0.0, sheenTint=0.0, clearcoat=0.0, clearcoatGloss=0.0,
    specTrans=0.0, ior=1.5, mediumType=0.0, mediumDensity=0.0,
    mediumColor=(1.0, 1.0, 1.0), mediumAnisotropy=0.0,
    baseColorTexId=-1.0, metallicRoughnessTexID=-1.0, normalmapTexID=-1.0, emissionmapTexID=-1.0,
    opacity=1.0, alphaMode=0.0, alphaCutoff=0.0)

QUAD_LIGHT, SPHERE_LIGHT, DISTANT_LIGHT = 0, 1, 2


def material(**kw):
    """pt_material with reference defaults, overridden by keyword."""
    m = api.PtMaterial()
    vals = dict(MATERIAL_DEFAULTS)
    for k in kw:
        if k not in vals:
            raise KeyError(f"unknown material field {k}")
    vals.update(kw)
    for k, v in vals.items():
        if isinstance(v, (tuple, list, np.ndarray)):
            getattr(m, k)[:] = [float(x) for x in v]
        else:
            setattr(m, k, float(v))
    return m


def material_array(m):
    return np.frombuffer(bytes(m), dtype=np.float32).copy()


def quad_light(position, v1, v2, emission):
    """Rect light from a corner and two other corners, as the .scene loader
    builds it (src/loaders/Loader.cpp:186-192): u = v1 - p, v = v2 - p,
    area = |u x v|."""
    p = np.asarray(position, np.float32)
    u = np.asarray(v1, np.float32) - p
    v = np.asarray(v2, np.float32) - p
    cr = np.array([u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]],
                  np.float32)
    area = np.float32(np.sqrt(np.float32(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2])))
    return _light(p, emission, u, v, 0.0, area, QUAD_LIGHT)


def sphere_light(position, radius, emission):
    """Sphere light, area 4 pi r^2 (src/loaders/Loader.cpp:193-197)."""
    r = np.float32(radius)
    area = np.float32(np.float32(4.0) * np.float32(np.pi) * r * r)
    return _light(position, emission, (0, 0, 0), (0, 0, 0), r, area, SPHERE_LIGHT)


def distant_light(direction, emission):
    """Distant light: position holds the direction towards the light, area 0."""
    return _light(direction, emission, (0, 0, 0), (0, 0, 0), 0.0, 0.0, DISTANT_LIGHT)


def _light(p, e, u, v, radius, area, typ):
    l = api.PtLight()
    l.position[:] = [float(x) for x in p]
    l.emission[:] = [float(x) for x in e]
    l.u[:] = [float(x) for x in u]
    l.v[:] = [float(x) for x in v]
    l.radius, l.area, l.type = float(radius), float(area), float(typ)
    return l


def mat4(translate=(0, 0, 0), scale=(1, 1, 1), rot=None):
    """Mat4 as the reference composes it, scale * rot * translate with
    data[i] = column i (src/loaders/Loader.cpp:435-440, src/math/Mat4.h)."""
    def ident():
        return np.eye(4, dtype=np.float32)

    def mul(a, b):  # reference Mat4::operator*, out[i][j] = sum_k a[i][k] * b[k][j], in float
        out = np.zeros((4, 4), np.float32)
        for i in range(4):
            for j in range(4):
                acc = np.float32(a[i, 0] * b[0, j])
                acc = np.float32(acc + np.float32(a[i, 1] * b[1, j]))
                acc = np.float32(acc + np.float32(a[i, 2] * b[2, j]))
                acc = np.float32(acc + np.float32(a[i, 3] * b[3, j]))
                out[i, j] = acc
        return out

    t = ident()
    t[3, :3] = np.asarray(translate, np.float32)
    s = ident()
    s[0, 0], s[1, 1], s[2, 2] = [np.float32(x) for x in scale]
    r = ident() if rot is None else np.asarray(rot, np.float32)
    return mul(mul(s, r), t)


class Scene:
    """Owns a native scene (pts_scene)."""

    def __init__(self):
        self._lib = api.scene_lib()
        h = C.c_void_p()
        self._check("pts_create", self._lib.pts_create(C.byref(h)))
        self._h = h
        self._keep = []

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pts_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, fn, rc):
        if rc != api.PT_OK:
            msg = self._lib.pts_last_error(getattr(self, "_h", None) or None)
            raise api.PtError(fn, rc, msg.decode() if msg else "")
        return rc

    # --- reference surface -------------------------------------------------
    def AddMesh(self, positions=None, normals=None, uvs=None, name="", filename=None):
        """AddMesh(filename=...) loads an OBJ; otherwise positions[n,3,3] (and
        optional normals[n,3,3], uvs[n,3,2]) give a triangle soup."""
        out = C.c_int32()
        if filename is not None:
            rc = self._lib.pts_add_mesh_obj(self._h, filename.encode(), C.byref(out))
            if rc != api.PT_OK:
                return -1
            return out.value
        p = np.ascontiguousarray(positions, np.float32).reshape(-1, 9)
        n = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 9)
        t = None if uvs is None else np.ascontiguousarray(uvs, np.float32).reshape(-1, 6)
        self._check("pts_add_mesh", self._lib.pts_add_mesh(
            self._h, api.fptr(p), api.fptr(n), api.fptr(t), p.shape[0], name.encode(), C.byref(out)))
        return out.value

    def AddMaterial(self, m=None, **kw):
        m = m if m is not None else material(**kw)
        out = C.c_int32()
        self._check("pts_add_material", self._lib.pts_add_material(self._h, C.byref(m), C.byref(out)))
        return out.value

    def AddMeshInstance(self, mesh_id, transform, material_id):
        t = np.ascontiguousarray(transform, np.float32).reshape(16)
        out = C.c_int32()
        self._check("pts_add_mesh_instance", self._lib.pts_add_mesh_instance(
            self._h, int(mesh_id), api.fptr(t), int(material_id), C.byref(out)))
        return out.value

    def AddLight(self, light):
        out = C.c_int32()
        self._check("pts_add_light", self._lib.pts_add_light(self._h, C.byref(light), C.byref(out)))
        return out.value

    def AddTexture(self, rgba, name=""):
        a = np.ascontiguousarray(rgba, np.uint8)
        h, w = a.shape[:2]
        out = C.c_int32()
        self._check("pts_add_texture", self._lib.pts_add_texture(
            self._h, a.ctypes.data_as(C.POINTER(C.c_uint8)), w, h, name.encode(), C.byref(out)))
        return out.value

    def SetTextureArraySize(self, width, height):
        self._check("pts_set_texture_array_size", self._lib.pts_set_texture_array_size(self._h, width, height))

    def AddCamera(self, eye, lookat, fov_degrees):
        e = np.asarray(eye, np.float32)
        l = np.asarray(lookat, np.float32)
        self._check("pts_add_camera", self._lib.pts_add_camera(self._h, api.fptr(e), api.fptr(l), float(fov_degrees)))

    def AddEnvMap(self, rgb):
        a = np.ascontiguousarray(rgb, np.float32)
        h, w = a.shape[:2]
        self._check("pts_add_env_map", self._lib.pts_add_env_map(self._h, api.fptr(a), w, h))

    def LoadSceneFile(self, filename, options=None):
        """LoadSceneFromFile (src/loaders/Loader.cpp:36-517) into this scene;
        returns the RenderOptions the file sets (pts_render_options, starting
        from `options` or the reference defaults). Messages about skipped
        elements are in self.load_log."""
        o = api.PtsRenderOptions()
        if options is None:
            self._lib.pts_default_render_options(C.byref(o))
        else:
            C.pointer(o)[0] = options
        self._check("pts_load_scene_file", self._lib.pts_load_scene_file(self._h, str(filename).encode(), C.byref(o)))
        msg = self._lib.pts_last_error(self._h)
        self.load_log = msg.decode() if msg else ""
        return o

    def texture(self, tex_id):
        """Decoded RGBA8 pixels of texture tex_id as added (height, width, 4)."""
        w, h = C.c_int32(), C.c_int32()
        self._check("pts_get_texture", self._lib.pts_get_texture(self._h, tex_id, C.byref(w), C.byref(h), None))
        out = np.zeros((h.value, w.value, 4), np.uint8)
        self._check("pts_get_texture", self._lib.pts_get_texture(self._h, tex_id, C.byref(w), C.byref(h),
                                                                 out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    def ProcessScene(self):
        self._check("pts_process", self._lib.pts_process(self._h))

    def RebuildInstances(self):
        self._check("pts_rebuild_instances", self._lib.pts_rebuild_instances(self._h))

    def SetInstanceTransform(self, instance, transform):
        t = np.ascontiguousarray(transform, np.float32).reshape(16)
        self._check("pts_set_instance_transform", self._lib.pts_set_instance_transform(self._h, instance, api.fptr(t)))

    def OffsetCameraOrientation(self, dx, dy):
        self._check("pts_camera_offset_orientation", self._lib.pts_camera_offset_orientation(self._h, dx, dy))

    # --- access to the flat arrays ----------------------------------------
    def view(self):
        """pt_scene_view pointing into the native arrays (valid until the next mutation)."""
        v = api.PtSceneView()
        self._check("pts_get_view", self._lib.pts_get_view(self._h, C.byref(v)))
        return v

    def camera(self):
        c = api.PtCamera()
        self._check("pts_get_camera", self._lib.pts_get_camera(self._h, C.byref(c)))
        return c

    def arrays(self):
        """Copies of the flattened arrays as numpy (reference layouts)."""
        v = self.view()

        def grab(ptr, n, dtype, width):
            if n == 0 or not ptr:
                return np.zeros((0, width), dtype)
            buf = C.cast(ptr, C.POINTER(C.c_byte * (n * width * 4))).contents
            return np.frombuffer(buf, dtype=dtype).reshape(n, width).copy()

        out = dict(
            nodes=grab(v.nodes, v.num_nodes, np.float32, 9),
            top_level_index=v.top_level_index,
            vert_indices=grab(v.vert_indices, v.num_tris, np.int32, 3),
            vertices_uvx=grab(v.vertices_uvx, v.num_vertices, np.float32, 4),
            normals_uvy=grab(v.normals_uvy, v.num_vertices, np.float32, 4),
            materials=grab(v.materials, v.num_materials, np.float32, 32),
            transforms=grab(v.transforms, v.num_instances, np.float32, 16),
            lights=grab(v.lights, v.num_lights, np.float32, 15),
        )
        if v.num_textures > 0 and v.texture_maps:
            n = v.num_textures * v.tex_width * v.tex_height * 4
            buf = C.cast(v.texture_maps, C.POINTER(C.c_uint8 * n)).contents
            out["textures"] = np.frombuffer(buf, np.uint8).reshape(v.num_textures, v.tex_height, v.tex_width, 4).copy()
            out["tex_size"] = (v.tex_width, v.tex_height)
        if v.env_width > 0:
            out["env_img"] = grab(v.env_img, v.env_width * v.env_height, np.float32, 3)
            out["env_cdf"] = grab(v.env_cdf, v.env_width * v.env_height, np.float32, 1)[:, 0]
            out["env_total_sum"] = v.env_total_sum
            out["env_size"] = (v.env_width, v.env_height)
        return out

    def blas_root(self, mesh_id):
        out = C.c_int32()
        self._check("pts_get_blas_root", self._lib.pts_get_blas_root(self._h, mesh_id, C.byref(out)))
        return out.value

"""Python face of the HIP renderer (libglslpt_hip.so), mirroring the
reference's ``GLSLPT::Renderer`` (src/core/Renderer.h:104-185): upload the
scene arrays, set RenderOptions and the camera, Render (add samples to the
RGBA32F accumulation buffer), read the output back.

There is no CPU fallback: constructing a Renderer without the HIP library or
without a GPU raises.
"""
import ctypes as C

import numpy as np

from . import _capi as api


def default_options(**kw):
    o = api.PtOptions()
    api.hip_lib().pt_default_options(C.byref(o))
    for k, v in kw.items():
        if k == "uniform_light_col":
            o.uniform_light_col[:] = [float(x) for x in v]
        else:
            setattr(o, k, v)
    return o


class Renderer:
    def __init__(self, device=0):
        self._lib = api.hip_lib()
        h = C.c_void_p()
        rc = self._lib.pt_create(int(device), C.byref(h))
        if rc != api.PT_OK:
            raise api.PtError("pt_create", rc, "no usable HIP device" if rc == -7 else "")
        self._h = h
        self.width = self.height = 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pt_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, fn, rc):
        if rc != api.PT_OK:
            msg = self._lib.pt_last_error(self._h)
            raise api.PtError(fn, rc, msg.decode() if msg else "")

    # --- GLSLPT::Renderer surface ---------------------------------------
    def upload_scene(self, scene, materials=None, device_env_cdf=False):
        """Upload the scene's arrays; `materials` ((num_materials, 32) float32)
        replaces the scene's material table in the upload; `device_env_cdf`
        builds the environment map's CDF on the device instead of uploading
        the host-built one."""
        v = scene.view()
        if device_env_cdf:
            v.env_cdf = None
        if materials is not None:
            m = np.ascontiguousarray(materials, np.float32)
            if m.shape != (v.num_materials, 32):
                raise ValueError(f"materials must be ({v.num_materials}, 32)")
            v.materials = C.cast(m.ctypes.data, C.POINTER(api.PtMaterial))
        self._check("pt_upload_scene", self._lib.pt_upload_scene(self._h, C.byref(v)))
        self.num_materials = v.num_materials

    def set_env_map(self, rgb):
        """Replace the environment map: (height, width, 3) float32, rows top-down;
        its CDF is built on the device."""
        img = np.ascontiguousarray(rgb, np.float32)
        if img.ndim != 3 or img.shape[2] != 3:
            raise ValueError("rgb must be (height, width, 3)")
        self._check("pt_set_env_map", self._lib.pt_set_env_map(self._h, api.fptr(img), img.shape[1], img.shape[0]))

    def read_env_cdf(self, count):
        """(cdf[count], total) of the environment map on the device."""
        cdf = np.zeros(count, np.float32)
        total = C.c_float(0.0)
        self._check("pt_read_env_cdf", self._lib.pt_read_env_cdf(self._h, api.fptr(cdf), count, C.byref(total)))
        return cdf, total.value

    def update_instances(self, scene, first_node=None):
        """Instance-edit refresh after Scene.RebuildInstances (Renderer.cpp:649-665)."""
        a = scene.arrays()
        v = scene.view()
        tli = v.top_level_index if first_node is None else first_node
        self._check("pt_update_instances", self._lib.pt_update_instances(
            self._h, v.transforms, v.num_instances, v.materials, v.num_materials,
            C.cast(C.byref(v.nodes.contents, tli * C.sizeof(api.PtBvhNode)), C.POINTER(api.PtBvhNode)),
            tli, v.num_nodes - tli))
        return a

    def rebuild_instances(self, transforms, material_ids=None):
        """Instance edit with the TLAS rebuilt on the device (Scene::RebuildInstances,
        Scene.cpp:200-214): transforms (num_instances, 4, 4) or (num_instances, 16)
        in the reference's Mat4 memory order; optional per-instance material ids."""
        t = np.ascontiguousarray(transforms, np.float32).reshape(-1, 16)
        m = None if material_ids is None else np.ascontiguousarray(material_ids, np.int32).reshape(-1)
        if m is not None and m.shape[0] != t.shape[0]:
            raise ValueError("one material id per instance")
        self._check("pt_rebuild_instances", self._lib.pt_rebuild_instances(
            self._h, api.fptr(t), t.shape[0], api.iptr(m)))

    def read_tlas(self, num_instances):
        """(nodes (n-1, 16) float32, instance records (n, 24) float32, root) of the
        device TLAS; node columns 12 and 13 hold the child references as int bits."""
        nodes = np.zeros((max(num_instances - 1, 0), 16), np.float32)
        insts = np.zeros((num_instances, 24), np.float32)
        root = C.c_int32(0)
        self._check("pt_read_tlas", self._lib.pt_read_tlas(
            self._h, api.fptr(nodes), nodes.shape[0], api.fptr(insts), num_instances, C.byref(root)))
        return nodes, insts, root.value

    def read_tlas4(self):
        """(nodes (room, 32) float32, root) of the four-wide TLAS; columns 24-27
        hold the child references as int bits, breadth first from the root."""
        room, root = C.c_int32(0), C.c_int32(0)
        self._check("pt_read_tlas4", self._lib.pt_read_tlas4(self._h, None, 0, C.byref(room), None))
        nodes = np.zeros((room.value, 32), np.float32)
        self._check("pt_read_tlas4", self._lib.pt_read_tlas4(
            self._h, api.fptr(nodes), room.value, None, C.byref(root)))
        return nodes, root.value

    def set_options(self, opts=None, **kw):
        o = opts if opts is not None else default_options(**kw)
        self._check("pt_set_options", self._lib.pt_set_options(self._h, C.byref(o)))
        self.width, self.height = o.width, o.height
        self.opts = o

    def set_camera(self, cam):
        self._check("pt_set_camera", self._lib.pt_set_camera(self._h, C.byref(cam)))

    def clear(self):
        self._check("pt_clear_accum", self._lib.pt_clear_accum(self._h))

    def render(self, first_sample, num_samples):
        self._check("pt_render", self._lib.pt_render(self._h, int(first_sample), int(num_samples)))

    def synchronize(self):
        self._check("pt_synchronize", self._lib.pt_synchronize(self._h))

    def read_accum(self):
        out = np.zeros((self.height, self.width, 4), np.float32)
        self._check("pt_read_accum", self._lib.pt_read_accum(self._h, api.fptr(out), out.size))
        return out

    def read_ldr(self, num_samples, tonemap=True, aces=False, simple_aces=False, background=(1, 1, 1)):
        out = np.zeros((self.height, self.width, 4), np.uint8)
        bg = np.asarray(background, np.float32)
        self._check("pt_read_ldr", self._lib.pt_read_ldr(
            self._h, int(num_samples), int(tonemap), int(aces), int(simple_aces), api.fptr(bg),
            out.ctypes.data_as(C.POINTER(C.c_uint8)), out.size))
        return out

    def stats(self):
        s = api.PtStats()
        self._check("pt_get_stats", self._lib.pt_get_stats(self._h, C.byref(s)))
        d = {f: getattr(s, f) for f, _ in s._fields_ if f not in ("ms_stage", "launches_stage", "bounce_hits")}
        d["bounce_hits"] = list(s.bounce_hits)
        d["ms_stage"] = {api.STAGES[i]: s.ms_stage[i] for i in range(8)}
        d["launches_stage"] = {api.STAGES[i]: s.launches_stage[i] for i in range(8)}
        return d

    def reset_stats(self):
        self._check("pt_reset_stats", self._lib.pt_reset_stats(self._h))

    # --- batched jobs ----------------------------------------------------
    def batch_render(self, jobs):
        """Render every job (dict: camera, width, height, spp, max_depth) with
        the uploaded scene and the current options; returns each job's
        accumulation (sum over its samples), (height, width, 4) float32."""
        arr = (api.PtJob * len(jobs))()
        outs = []
        for i, j in enumerate(jobs):
            out = np.zeros((j["height"], j["width"], 4), np.float32)
            outs.append(out)
            arr[i].camera = j["camera"]
            arr[i].width, arr[i].height = j["width"], j["height"]
            arr[i].spp, arr[i].max_depth = j["spp"], j["max_depth"]
            arr[i].accum_out = api.fptr(out)
        self._check("pt_batch_render", self._lib.pt_batch_render(self._h, arr, len(jobs)))
        return outs

    # --- adjoint ---------------------------------------------------------
    def adjoint(self, first_sample, num_samples, dl_dimage):
        """Gradient of a loss with respect to the material parameters
        (baseColor r, g, b, roughness, metallic, ior, clearcoat per material)
        for the samples [first_sample, first_sample + num_samples). dl_dimage is
        dLoss/d(mean image), (height, width, 3), rows bottom-up like read_accum."""
        g = np.ascontiguousarray(dl_dimage, np.float32)
        if g.shape != (self.height, self.width, 3):
            raise ValueError(f"dl_dimage must be ({self.height}, {self.width}, 3)")
        out = np.zeros((self.num_materials, api.PT_ADJOINT_PARAMS), np.float32)
        self._check("pt_adjoint", self._lib.pt_adjoint(self._h, int(first_sample), int(num_samples), api.fptr(g),
                                                       api.fptr(out)))
        return out

    def adjoint_medium(self, first_sample, num_samples, dl_dimage):
        """As adjoint, plus the gradient with respect to each material's
        medium (density, colour r, g, b): returns (material gradient
        (num_materials, 7), medium gradient (num_materials, 4))."""
        g = np.ascontiguousarray(dl_dimage, np.float32)
        if g.shape != (self.height, self.width, 3):
            raise ValueError(f"dl_dimage must be ({self.height}, {self.width}, 3)")
        out = np.zeros((self.num_materials, api.PT_ADJOINT_PARAMS), np.float32)
        med = np.zeros((self.num_materials, api.PT_ADJOINT_MEDIUM_PARAMS), np.float32)
        self._check("pt_adjoint_medium", self._lib.pt_adjoint_medium(self._h, int(first_sample), int(num_samples),
                                                                     api.fptr(g), api.fptr(out), api.fptr(med)))
        return out, med

    # --- ray queries ------------------------------------------------------
    def closest_hit(self, origins, dirs):
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        n = o.shape[0]
        t = np.zeros(n, np.float32)
        inst = np.zeros(n, np.int32)
        tri = np.zeros(n, np.int32)
        li = np.zeros(n, np.int32)
        uv = np.zeros((n, 2), np.float32)
        self._check("pt_closest_hit", self._lib.pt_closest_hit(
            self._h, api.fptr(o), api.fptr(d), n, api.fptr(t), api.iptr(inst), api.iptr(tri), api.iptr(li), api.fptr(uv)))
        return dict(t=t, instance=inst, triangle=tri, light=li, uv=uv)

    def any_hit(self, origins, dirs, max_dist):
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        m = np.ascontiguousarray(max_dist, np.float32)
        n = o.shape[0]
        occ = np.zeros(n, np.int32)
        self._check("pt_any_hit", self._lib.pt_any_hit(self._h, api.fptr(o), api.fptr(d), api.fptr(m), n, api.iptr(occ)))
        return occ


def options_for(settings, mode=api.PT_MODE_WAVEFRONT, **kw):
    """pt_options for a scenes.RenderSettings workload."""
    o = default_options(width=settings.width, height=settings.height, max_depth=settings.max_depth,
                        enable_rr=int(settings.enable_rr), rr_depth=settings.rr_depth,
                        enable_uniform_light=int(settings.enable_uniform_light),
                        uniform_light_col=settings.uniform_light_col, mode=mode)
    for k, v in settings.extra.items():
        setattr(o, k, v)
    for k, v in kw.items():
        setattr(o, k, v)
    return o

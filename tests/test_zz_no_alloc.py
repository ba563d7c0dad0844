"""pt_env_cdf_noalloc (the environment-map CDF build that allocates nothing,
its scratch in a fixed device array of the library) against pt_set_env_map's
CDF (pt_read_env_cdf), through the C ABI: bit-identical CDF and total on
calls in a row, each smaller than the one before, the middle one taking the
one-lane fallback (a negative luminance clears the scan's flag word, which
the next call must set again); a size beyond the maximum is refused before
any work."""
import ctypes as C

import numpy as np
import pytest

from glslpt import _capi as api
from glslpt import scenes
from glslpt.renderer import Renderer

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the renderer library itself loaded."""
    api.hip_lib()
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64.so" in ln}
    assert paths, "libamdhip64 not loaded"
    h = C.CDLL(sorted(paths)[0])
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


H2D, D2H = 1, 2


def _noalloc(hip, img):
    """Run pt_env_cdf_noalloc on device copies of img; returns (cdf, total)."""
    lib = api.hip_lib()
    h, w = img.shape[:2]
    n = w * h
    src = np.ascontiguousarray(img, np.float32)
    d_rgb, d_cdf, d_tot = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for p, nb in ((d_rgb, src.nbytes), (d_cdf, 4 * n), (d_tot, 4)):
        assert hip.hipMalloc(C.byref(p), nb) == 0
    try:
        assert hip.hipMemcpy(d_rgb, src.ctypes.data, src.nbytes, H2D) == 0
        rc = lib.pt_env_cdf_noalloc(C.cast(d_rgb, C.POINTER(C.c_float)), w, h, C.cast(d_cdf, C.POINTER(C.c_float)),
                                    C.cast(d_tot, C.POINTER(C.c_float)), None)
        assert rc == api.PT_OK, lib.pt_last_error(None)
        assert hip.hipDeviceSynchronize() == 0
        cdf = np.zeros(n, np.float32)
        tot = np.zeros(1, np.float32)
        assert hip.hipMemcpy(cdf.ctypes.data, d_cdf, 4 * n, D2H) == 0
        assert hip.hipMemcpy(tot.ctypes.data, d_tot, 4, D2H) == 0
    finally:
        for p in (d_rgb, d_cdf, d_tot):
            hip.hipFree(p)
    return cdf, float(tot[0])


def _existing(r, img):
    r.set_env_map(img)
    return r.read_env_cdf(img.shape[0] * img.shape[1])


def test_noalloc_cdf_equals_set_env_map():
    lib = api.hip_lib()
    hip = _hip()
    r = Renderer(0)
    sc, _ = scenes.feature_scene("envmap")
    r.upload_scene(sc)
    neg = scenes.sky_envmap(700, 300)
    neg[5, 7] = (-1.0, 0.5, 0.5)  # negative luminance: the one-lane fallback
    for label, img in (("2048x1024 sky", scenes.sky_envmap(2048, 1024)), ("700x300 negative texel", neg),
                       ("300x200 sky", scenes.sky_envmap(300, 200))):
        print(f"{label}", flush=True)
        got, gt = _noalloc(hip, img)
        want, wt = _existing(r, img)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), label
        assert np.float32(gt).view(np.uint32) == np.float32(wt).view(np.uint32), label
    # beyond the maximum: refused on the host, nothing launched
    big = C.cast(C.c_void_p(16), C.POINTER(C.c_float))
    rc = lib.pt_env_cdf_noalloc(big, 8192, 4097, big, big, None)
    assert rc == api.PT_ERR_OVERFLOW
    assert b"PT_ENV_NOALLOC_MAX_TEXELS" in lib.pt_last_error(None)
    r.close()

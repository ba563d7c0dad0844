"""Environment-map CDF built on the device (pt_set_env_map, or pt_upload_scene
without a host CDF) against the host build, which test_scene_loader pins to the
reference's EnvironmentMap::BuildCDF bit for bit: the CDF and total luminance
are identical, and renders that importance-sample the map are identical."""
import numpy as np
import pytest

from glslpt import scenes
from glslpt._capi import PT_MODE_PLAIN, PT_MODE_WAVEFRONT
from glslpt.renderer import Renderer, options_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def renderer():
    r = Renderer(0)
    yield r
    r.close()


def host_cdf(img):
    """Sequential float32 running sum of the luminance (the host build's order)."""
    f = np.ascontiguousarray(img, np.float32).reshape(-1, 3)
    w = np.float32(0.212671) * f[:, 0] + np.float32(0.715160) * f[:, 1]
    w = w + np.float32(0.072169) * f[:, 2]
    return np.cumsum(w, dtype=np.float32)


def ragged_image(w, h, seed):
    rng = np.random.default_rng(seed)
    img = (rng.random((h, w, 3)) * rng.choice([0.0, 1.0, 50.0], (h, w, 1))).astype(np.float32)
    img[0, 0] = 0.0
    return img


@pytest.mark.parametrize("size", [(1, 1), (7, 5), (64, 32), (4099, 3), (4099, 12), (512, 256), (2048, 1024)])
def test_device_cdf_equals_host(renderer, size):
    """Single texel, a tail shorter than one step, one chunk, several chunks with
    a short last one (ragged and sky), and the C5 sky size."""
    sc, rs = scenes.feature_scene("envmap")
    renderer.upload_scene(sc)
    w, h = size
    img = scenes.sky_envmap(w, h) if size in ((64, 32), (512, 256), (2048, 1024)) else ragged_image(w, h, 3)
    renderer.set_env_map(img)
    cdf, total = renderer.read_env_cdf(w * h)
    want = host_cdf(img)
    print(f"{w}x{h}: total {total:.7g}, max |diff| {np.abs(cdf - want).max()}")
    assert np.array_equal(cdf, want)
    assert np.float32(total) == want[-1]


def test_scene_upload_builds_cdf_on_device(renderer):
    sc, rs = scenes.feature_scene("envmap")
    a = sc.arrays()
    n = a["env_size"][0] * a["env_size"][1]
    renderer.upload_scene(sc, device_env_cdf=True)
    cdf, total = renderer.read_env_cdf(n)
    assert np.array_equal(cdf, a["env_cdf"])
    assert np.float32(total) == np.float32(a["env_total_sum"])


@pytest.mark.parametrize("mode", [PT_MODE_PLAIN, PT_MODE_WAVEFRONT])
def test_renders_identical_with_device_cdf(renderer, mode):
    sc, rs = scenes.feature_scene("envmap")
    out = []
    for dev in (False, True):
        renderer.upload_scene(sc, device_env_cdf=dev)
        renderer.set_options(options_for(rs, mode=mode, wave_spp=2))
        renderer.set_camera(sc.camera())
        renderer.clear()
        renderer.render(0, rs.spp)
        renderer.synchronize()
        out.append(renderer.read_accum())
    assert np.array_equal(out[0], out[1])


def test_set_env_map_equals_host_upload(renderer):
    """Swapping the map on the device gives the image of a scene uploaded
    with that map (and its host CDF) from the start."""
    sc, rs = scenes.feature_scene("envmap")
    sky2 = scenes.sky_envmap(96, 48, seed=0x5EC0, sun=(0.7, 0.3), sun_power=900.0)
    renderer.upload_scene(sc)
    renderer.set_env_map(sky2)
    renderer.set_options(options_for(rs, wave_spp=2))
    renderer.set_camera(sc.camera())
    renderer.clear()
    renderer.render(0, rs.spp)
    renderer.synchronize()
    g = renderer.read_accum()
    sc.AddEnvMap(sky2)
    renderer.upload_scene(sc)
    renderer.clear()
    renderer.render(0, rs.spp)
    renderer.synchronize()
    r = renderer.read_accum()
    assert np.array_equal(g, r)
    assert np.abs(g).max() > 0


def _green_for(target):
    """A green value whose float32 luminance (red = blue = 0) is exactly target."""
    def lum(g):
        return np.float32(0.715160) * np.float32(g)
    g = np.float32(target / 0.715160)
    for _ in range(4000):
        v = lum(g)
        if v == np.float32(target):
            return g
        g = np.nextafter(g, np.float32(np.inf) if v < target else np.float32(-np.inf))
    raise AssertionError(f"no green value for {target}")


def special_image(kind, w=4099, h=9):
    rng = np.random.default_rng(7)
    img = np.zeros((h, w, 3), np.float32)
    if kind == "constant":
        img[:] = 1.0
    elif kind == "bright_then_dim":
        img[:] = rng.random((h, w, 3)) * 1e-3
        img[0, 0] = 1e6
    elif kind == "subnormal":
        img[:] = rng.random((h, w, 3)) * 1e-40
    elif kind == "overflow":
        img[:] = 3e38
    elif kind == "ties":
        # running sum at 2^23 + 1 (grid spacing 1), then halves: every add is a tie
        img[..., 1] = _green_for(0.5)
        img[0, 0, 1] = _green_for(2.0 ** 23 + 1)
    elif kind == "negative":
        img[:] = rng.random((h, w, 3))
        img[1, 5, 0] = -3.0
    elif kind == "neg_zero":
        img[:] = np.float32(-0.0)
    elif kind == "inf":
        img[:] = rng.random((h, w, 3))
        img[2, 7, 2] = np.inf
    return img


@pytest.mark.parametrize("kind", ["constant", "bright_then_dim", "subnormal", "overflow", "ties", "negative",
                                  "neg_zero", "inf"])
def test_device_cdf_special_inputs(renderer, kind):
    """Ties to even, long runs of adds below half a grid step, subnormal sums,
    overflow to inf, and (negative, -0, inf luminance) the one-lane fallback:
    all bit-identical to the sequential float32 sum."""
    sc, rs = scenes.feature_scene("envmap")
    renderer.upload_scene(sc)
    img = special_image(kind)
    renderer.set_env_map(img)
    h, w = img.shape[:2]
    cdf, total = renderer.read_env_cdf(w * h)
    with np.errstate(over="ignore", invalid="ignore"):
        want = host_cdf(img)
    assert np.array_equal(cdf.view(np.uint32), want.view(np.uint32)), kind
    assert np.float32(total).view(np.uint32) == want[-1].view(np.uint32)

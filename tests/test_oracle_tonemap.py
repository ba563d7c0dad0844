"""Oracle tonemap (tonemap.glsl:59-133 + glGetTexImage byte conversion)
pinned by closed forms and invariants, independent of its code:
gamma-only and luminance-Reinhard curves, the simple ACES rational fit, the
fitted ACES transform's grey-axis invariance (each row of both ACES matrices
sums to 1, so only the correct row/column orientation keeps r = g = b), and
the background / transparent-checker compositing."""
import numpy as np

import oracle


def accum_of(rgb, alpha=1.0, spp=4):
    rgb = np.asarray(rgb, np.float32)
    a = np.concatenate([rgb, np.full(rgb.shape[:-1] + (1,), alpha, np.float32)], -1)
    return a * spp


def to_byte(x):
    return np.floor(np.clip(x, 0, 1) * 255 + 0.5).astype(np.int64)


def test_gamma_only():
    rng = np.random.default_rng(1)
    c = rng.uniform(0, 1.3, (16, 16, 3))
    out = oracle.tonemap(accum_of(c), 4, tonemap=False).astype(np.int64)
    want = to_byte(np.clip(c, 0, None) ** (1 / 2.2))
    assert np.abs(out[..., :3] - want).max() <= 1
    assert np.all(out[..., 3] == 255)


def test_luminance_tonemap():
    rng = np.random.default_rng(2)
    c = rng.uniform(0, 8, (16, 16, 3))
    lum = c @ np.array([0.212671, 0.715160, 0.072169])
    want = to_byte((c / (1 + lum / 1.5)[..., None]) ** (1 / 2.2))
    out = oracle.tonemap(accum_of(c), 4).astype(np.int64)
    assert np.abs(out[..., :3] - want).max() <= 1


def test_simple_aces_fit():
    c = np.linspace(0, 12, 256).reshape(16, 16, 1).repeat(3, -1)
    f = np.clip((c * (2.51 * c + 0.03)) / (c * (2.43 * c + 0.59) + 0.14), 0, 1)
    out = oracle.tonemap(accum_of(c), 4, aces=True, simple_aces=True).astype(np.int64)
    assert np.abs(out[..., :3] - to_byte(f ** (1 / 2.2))).max() <= 1


def test_fitted_aces_keeps_grey_and_is_monotone():
    g = np.linspace(0.0, 20.0, 256).reshape(16, 16, 1).repeat(3, -1)
    out = oracle.tonemap(accum_of(g), 4, aces=True).astype(np.int64)[..., :3]
    assert np.abs(out[..., 0] - out[..., 1]).max() <= 1 and np.abs(out[..., 1] - out[..., 2]).max() <= 1
    v = out[..., 1].ravel()
    assert np.all(np.diff(v) >= 0) and v[0] == 0 and v[-1] >= 250
    # a saturated colour keeps its hue ordering
    red = oracle.tonemap(accum_of(np.array([[[2.0, 0.2, 0.1]]])), 4, aces=True)[0, 0]
    assert red[0] > red[1] > red[2] - 1


def test_background_and_checker():
    c = np.full((20, 20, 3), 0.5)
    bg = (0.2, 0.4, 0.6)
    out = oracle.tonemap(accum_of(c, alpha=0.0), 4, tonemap=False, bg_mode=1, background=bg).astype(np.int64)
    assert np.all(np.abs(out[..., :3] - to_byte(np.array(bg))) <= 1) and np.all(out[..., 3] == 255)
    t = oracle.tonemap(accum_of(c, alpha=0.0), 4, tonemap=False, bg_mode=2).astype(np.int64)
    y, x = np.meshgrid(np.arange(20), np.arange(20), indexing="ij")
    cell = ((x // 10) + (y // 10)) % 2
    assert np.all(t[..., 0] == np.where(cell == 1, to_byte(0.2), to_byte(0.1)))
    assert np.all(t[..., 3] == 0)
    half = oracle.tonemap(accum_of(c, alpha=0.5), 4, tonemap=False, bg_mode=2).astype(np.int64)
    assert np.all(half[..., 3] == 128)

"""GPU adjoint with respect to homogeneous media (pt_adjoint_medium): the
gradient of an image loss with respect to each material's medium density and
colour (pathtrace.glsl:370-410), beside the Disney-parameter gradient.

* Against the CPU build of the same replay (tests/hostpath): equal up to the
  order of the gradient sums (1e-4 of the largest), on the media scene
  (absorbing, scattering and emissive boxes) and on the scaled-down heavy
  scene (two scattering volumes among alpha-tested foliage).
* Against central differences (eps 1e-3) of the oracle's frozen-sampling
  render, whose free-flight distances are drawn with the scene's density and
  weighted by the density ratio: 1% per parameter (floor 2% of the largest).
* Against ordinary central differences of the GPU renderer (sampling
  attached: a density change moves every scatter point): over 48 batches the
  mean difference of every parameter lies within 4 standard errors (plus 1%)
  and the squared z-scores sum below the chi-square(12) 0.999 quantile.
* The heavy scene at full size (3840x2160, 1 sample): finite, nothing
  dropped, a gradient on both volumes, and the Disney-parameter gradient equal
  to pt_adjoint's up to summation order.
"""
import numpy as np
import pytest

import hostpath
import oracle
from common import oracle_setup
from glslpt import scenes
from glslpt._capi import ADJOINT_MEDIUM_FIELDS, ADJOINT_MEDIUM_NAMES
from glslpt.renderer import Renderer, options_for
from test_gpu_heavy import small_heavy

pytestmark = pytest.mark.gpu

EPS = 1e-3


@pytest.fixture(scope="module")
def renderer():
    r = Renderer(0)
    yield r
    r.close()


def gpu_adjoint_medium(r, sc, rs, w, first=0, spp=None, materials=None, wave_spp=4):
    spp = spp or rs.spp
    r.upload_scene(sc, materials=materials)
    r.set_options(options_for(rs, wave_spp=wave_spp))
    r.set_camera(sc.camera())
    g, mg = r.adjoint_medium(first, spp, w)
    assert r.stats()["adjoint_dropped"] == 0
    return g, mg


def media_of(mats):
    return [m for m in range(mats.shape[0]) if mats[m, 18] != 0]


def _close(label, got, ref):
    scale = np.abs(ref).max()
    err = np.abs(got - ref).max()
    print(f"{label}: max |gpu - host| {err:.3g} of {scale:.3g}")
    assert scale > 0
    assert err <= 1e-4 * scale


def test_medium_adjoint_media_scene(renderer):
    sc, rs = scenes.feature_scene("media")
    spp = 8
    sa, cam, o = oracle_setup(sc, rs)
    mats = sa.a["materials"]
    w = np.random.default_rng(43).normal(size=(rs.height, rs.width, 3)).astype(np.float32)
    g, mg = gpu_adjoint_medium(renderer, sc, rs, w, spp=spp, wave_spp=spp)
    hg, hmg, dropped, mismatch = hostpath.adjoint(sc, rs, w, 0, spp, medium=True)
    assert dropped == 0 and mismatch == 0
    _close("media scene, medium gradient", mg, hmg)
    _close("media scene, material gradient", g, hg)
    media = media_of(mats)
    assert np.all(mg[[m for m in range(mats.shape[0]) if m not in media]] == 0)
    fd = np.zeros((len(media), 4))
    for i, m in enumerate(media):
        for k, f in enumerate(ADJOINT_MEDIUM_FIELDS):
            mp, mm = mats.copy(), mats.copy()
            mp[m, f] += EPS
            mm[m, f] -= EPS
            ip = oracle.render_frozen(sa, mp, cam, o, 0, spp)[..., :3].astype(np.float64)
            im = oracle.render_frozen(sa, mm, cam, o, 0, spp)[..., :3].astype(np.float64)
            fd[i, k] = np.sum(w * (ip - im)) / (2 * EPS) / spp
    got = mg[media]
    scale = np.abs(fd).max()
    rel = np.abs(got - fd) / np.maximum(np.abs(fd), 0.02 * scale)
    for i, m in enumerate(media):
        print(f"medium type {int(mats[m, 18])}: " + ", ".join(
            f"{ADJOINT_MEDIUM_NAMES[k]} {got[i, k]:.6g} / fd {fd[i, k]:.6g}" for k in range(4)))
    print(f"worst relative error {rel.max():.3g}")
    assert (np.abs(fd) > 0.02 * scale).sum() >= 10
    assert rel.max() <= 0.01


def test_medium_adjoint_small_heavy(renderer):
    sc, rs = small_heavy()
    mats = sc.arrays()["materials"]
    w = np.random.default_rng(44).normal(size=(rs.height, rs.width, 3)).astype(np.float32)
    g, mg = gpu_adjoint_medium(renderer, sc, rs, w)
    hg, hmg, dropped, mismatch = hostpath.adjoint(sc, rs, w, 0, rs.spp, medium=True)
    assert dropped == 0 and mismatch == 0
    _close("small heavy, medium gradient", mg, hmg)
    _close("small heavy, material gradient", g, hg)
    assert all(np.any(mg[m] != 0) for m in media_of(mats))


def test_medium_adjoint_unbiased_against_gpu_fd(renderer):
    """Ordinary central differences (eps 2e-2, common random numbers) of the
    GPU render vs the adjoint over 48 batches of 16 samples, the 12 medium
    parameters of the media scene. With sampling attached, a density change
    moves scatter points and flips scatter decisions, so the difference
    quotient is far noisier than the adjoint; both estimate the gradient of
    the expected image."""
    sc, rs = scenes.feature_scene("media", spp=16)
    mats = sc.arrays()["materials"]
    media = media_of(mats)
    rng = np.random.default_rng(47)
    w = rng.normal(size=(rs.height, rs.width, 3)).astype(np.float32)
    pick = [(m, k) for m in media for k in range(4)]
    B, eps = 48, 2e-2
    ga = np.zeros((B, len(pick)))
    gf = np.zeros((B, len(pick)))
    for b in range(B):
        first = 2000 + b * rs.spp
        _, mg = gpu_adjoint_medium(renderer, sc, rs, w, first=first, wave_spp=rs.spp)
        for j, (m, k) in enumerate(pick):
            mp, mm = mats.copy(), mats.copy()
            mp[m, ADJOINT_MEDIUM_FIELDS[k]] += eps
            mm[m, ADJOINT_MEDIUM_FIELDS[k]] -= eps
            img = []
            for mt in (mp, mm):
                renderer.upload_scene(sc, materials=mt)
                renderer.set_options(options_for(rs, wave_spp=rs.spp))
                renderer.set_camera(sc.camera())
                renderer.clear()
                renderer.render(first, rs.spp)
                renderer.synchronize()
                img.append(renderer.read_accum()[..., :3].astype(np.float64) / rs.spp)
            gf[b, j] = np.sum(w * (img[0] - img[1])) / (2 * eps)
            ga[b, j] = mg[m, k]
    d = ga - gf
    mean_fd = gf.mean(axis=0)
    # the emissive medium's factor involves no sampling, so its batches barely
    # differ: add the float32 rounding of the differenced images (1e-5 of the
    # gradient) to the standard error
    se = np.sqrt((d.std(axis=0, ddof=1) / np.sqrt(B)) ** 2 + (1e-5 * np.abs(mean_fd)) ** 2)
    z = d.mean(axis=0) / np.maximum(se, 1e-12)
    for j, (m, k) in enumerate(pick):
        print(f"medium type {int(mats[m, 18])} {ADJOINT_MEDIUM_NAMES[k]}: adjoint {ga[:, j].mean():.5g} "
              f"fd {mean_fd[j]:.5g} diff {d[:, j].mean():.3g} se {se[j]:.3g} z {z[j]:.2f}")
    print(f"sum z^2 = {np.sum(z * z):.2f} over {len(pick)} parameters")
    assert np.all(np.abs(d.mean(axis=0)) <= 4 * se + 0.01 * np.abs(mean_fd) + 1e-9)
    assert np.sum(z * z) <= 32.9


def test_medium_adjoint_heavy_full_size(renderer):
    sc, rs = scenes.heavy_mixed()
    mats = sc.arrays()["materials"]
    w = np.random.default_rng(48).normal(size=(rs.height, rs.width, 3)).astype(np.float32)
    g, mg = gpu_adjoint_medium(renderer, sc, rs, w, spp=1, wave_spp=1)
    g0 = renderer.adjoint(0, 1, w)
    assert renderer.stats()["adjoint_dropped"] == 0
    assert np.isfinite(g).all() and np.isfinite(mg).all()
    vols = media_of(mats)
    print("volume gradients:", {m: mg[m].tolist() for m in vols})
    assert len(vols) == 2 and all(np.any(mg[m] != 0) for m in vols)
    assert np.abs(g - g0).max() <= 1e-4 * np.abs(g0).max()

"""GPU adjoint (path-replay backprop, pt_adjoint) against finite differences.

* Against the CPU transcription: central differences (eps 1e-3) of the
  oracle's frozen-sampling render, which samples every path with the scene's
  materials and evaluates the BSDF values with the perturbed ones. That is the
  quantity the adjoint estimates sample by sample (sampling, MIS weights and
  Russian roulette detached), so the two agree to 1% per parameter (relative,
  with a floor of 2% of the largest gradient): on a small scene for every
  parameter, and on the full differentiable-materials workload (512x512,
  depth 4, 199,072 triangles, 64 materials) for 32 random parameters.
* Against ordinary finite differences of the GPU renderer (common random
  numbers, sampling attached): both estimate the same gradient of the
  expected image with different per-sample noise; over independent sample
  batches the mean difference of every parameter lies within 4 standard
  errors, and the sum of squared z-scores stays below the chi-square 0.999
  quantile.
* Size-independent properties at the full workload: linearity in
  dLoss/dimage, zero in -> zero out, nothing dropped as non-finite, and the
  time of the adjoint against the forward render.
"""
import numpy as np
import pytest

import oracle
from common import oracle_setup
from glslpt import scenes
from glslpt._capi import ADJOINT_FIELDS, ADJOINT_NAMES, PtError
from glslpt.renderer import Renderer, options_for

pytestmark = pytest.mark.gpu

EPS = 1e-3


@pytest.fixture(scope="module")
def renderer():
    r = Renderer(0)
    yield r
    r.close()


def small_boxes(**kw):
    args = dict(num_boxes=48, sphere_freq=8, num_spheres=3, num_materials=12, width=64, height=64, spp=4,
                max_depth=4, extent=6.0)
    args.update(kw)
    return scenes.material_boxes(**args)


def gpu_adjoint(r, sc, rs, w, first=0, spp=None, materials=None, wave_spp=4):
    spp = spp or rs.spp
    r.upload_scene(sc, materials=materials)
    r.set_options(options_for(rs, wave_spp=wave_spp))
    r.set_camera(sc.camera())
    g = r.adjoint(first, spp, w)
    assert r.stats()["adjoint_dropped"] == 0
    return g


def gpu_mean_image(r, sc, rs, materials, first, spp, wave_spp=4):
    r.upload_scene(sc, materials=materials)
    r.set_options(options_for(rs, wave_spp=wave_spp))
    r.set_camera(sc.camera())
    r.clear()
    r.render(first, spp)
    r.synchronize()
    return r.read_accum()[..., :3].astype(np.float64) / spp


def frozen_fd(sa, cam, o, w, spp, pick):
    """Central differences of the oracle's frozen-sampling render for the
    (material, parameter) pairs in pick."""
    mats = sa.a["materials"]
    fd = np.zeros(len(pick))
    for j, (m, k) in enumerate(pick):
        mp, mm = mats.copy(), mats.copy()
        mp[m, ADJOINT_FIELDS[k]] += EPS
        mm[m, ADJOINT_FIELDS[k]] -= EPS
        ip = oracle.render_frozen(sa, mp, cam, o, 0, spp)[..., :3].astype(np.float64)
        im = oracle.render_frozen(sa, mm, cam, o, 0, spp)[..., :3].astype(np.float64)
        fd[j] = np.sum(w * (ip - im)) / (2 * EPS) / spp
    return fd


def _compare_frozen(label, got, fd, pick):
    scale = np.abs(fd).max()
    floor = np.maximum(np.abs(fd), 0.02 * scale)
    rel = np.abs(got - fd) / floor
    for j in np.nonzero(rel > 0.01)[0]:
        m, k = pick[j]
        print(f"{label} mat {m} {ADJOINT_NAMES[k]}: adjoint {got[j]:.6g} fd {fd[j]:.6g}")
    big = np.abs(fd) > 0.02 * scale
    print(f"{label}: params {len(pick)}, above 2% of max {int(big.sum())}, worst relative error {rel.max():.3g}")
    return big, rel


def test_adjoint_matches_frozen_fd_of_oracle(renderer):
    sc, rs = small_boxes()
    sa, cam, o = oracle_setup(sc, rs)
    nm = sa.a["materials"].shape[0]
    w = np.random.default_rng(41).normal(size=(rs.height, rs.width, 3)).astype(np.float32)
    g = gpu_adjoint(renderer, sc, rs, w)
    pick = [(m, k) for m in range(nm) for k in range(7)]
    fd = frozen_fd(sa, cam, o, w, rs.spp, pick)
    big, rel = _compare_frozen("small boxes", np.array([g[m, k] for m, k in pick]), fd, pick)
    assert big.sum() >= 20
    assert rel.max() <= 0.01


def test_adjoint_full_workload_matches_frozen_fd(renderer):
    """The differentiable-materials workload at full size (512x512, depth 4,
    64 materials), 4 samples per pixel, against the oracle's frozen-sampling
    central differences (eps 1e-3) on 32 random parameters."""
    sc, rs = scenes.material_boxes()
    spp = 4
    sa, cam, o = oracle_setup(sc, rs)
    nm = sa.a["materials"].shape[0]
    rng = np.random.default_rng(45)
    w = rng.normal(size=(rs.height, rs.width, 3)).astype(np.float32)
    g = gpu_adjoint(renderer, sc, rs, w, spp=spp, wave_spp=spp)
    flat = rng.choice(nm * 7, size=32, replace=False)
    pick = [(int(i) // 7, int(i) % 7) for i in flat]
    fd = frozen_fd(sa, cam, o, w, spp, pick)
    big, rel = _compare_frozen("full workload", np.array([g[m, k] for m, k in pick]), fd, pick)
    assert big.sum() >= 16
    assert rel.max() <= 0.01


def test_adjoint_unbiased_against_gpu_fd(renderer):
    """Ordinary central differences with common random numbers vs the
    adjoint, over 64 independent batches of 32 samples: the mean difference
    of every parameter's gradient lies within 4 standard errors (plus 1% of
    the gradient) of zero, and the 12 squared z-scores sum to less than the
    chi-square(12) 0.999 quantile (32.9). (At 24 batches one metallic
    parameter sat at 2.3 standard errors; at 96 batches it is at 0.9 and the
    12 z-scores sum to 14.3, as for an unbiased estimate.) With sampling attached, a
    perturbation flips discrete choices (lobe, Russian roulette) on a
    fraction ~eps of the paths, so the difference quotient's variance grows
    like 1/eps: eps is 2e-2 here (the O(eps^2) bias is far below the noise)."""
    sc, rs = small_boxes(width=96, height=96, spp=32)
    mats = sc.arrays()["materials"]
    rng = np.random.default_rng(46)
    w = rng.normal(size=(rs.height, rs.width, 3)).astype(np.float32)
    pick = [(int(rng.integers(mats.shape[0])), int(rng.integers(7))) for _ in range(12)]
    B = 64
    eps = 2e-2
    ga = np.zeros((B, len(pick)))
    gf = np.zeros((B, len(pick)))
    for b in range(B):
        first = 1000 + b * rs.spp
        g = gpu_adjoint(renderer, sc, rs, w, first=first, wave_spp=rs.spp)
        for j, (m, k) in enumerate(pick):
            mp, mm = mats.copy(), mats.copy()
            mp[m, ADJOINT_FIELDS[k]] += eps
            mm[m, ADJOINT_FIELDS[k]] -= eps
            ip = gpu_mean_image(renderer, sc, rs, mp, first, rs.spp, wave_spp=rs.spp)
            im = gpu_mean_image(renderer, sc, rs, mm, first, rs.spp, wave_spp=rs.spp)
            gf[b, j] = np.sum(w * (ip - im)) / (2 * eps)
            ga[b, j] = g[m, k]
    d = ga - gf
    se = d.std(axis=0, ddof=1) / np.sqrt(B)
    mean_fd = gf.mean(axis=0)
    z = d.mean(axis=0) / np.maximum(se, 1e-12)
    for j, (m, k) in enumerate(pick):
        print(f"mat {m} {ADJOINT_NAMES[k]}: adjoint {ga[:, j].mean():.5g} fd {mean_fd[j]:.5g} "
              f"diff {d[:, j].mean():.3g} se {se[j]:.3g} z {z[j]:.2f}")
    print(f"sum z^2 = {np.sum(z * z):.2f} over {len(pick)} parameters")
    assert np.all(np.abs(d.mean(axis=0)) <= 4 * se + 0.01 * np.abs(mean_fd) + 1e-9)
    assert np.sum(z * z) <= 32.9


@pytest.fixture(scope="module")
def full_workload():
    return scenes.material_boxes()


def test_adjoint_full_size_linear_in_loss_gradient(renderer, full_workload):
    sc, rs = full_workload
    rng = np.random.default_rng(43)
    w1 = rng.normal(size=(rs.height, rs.width, 3)).astype(np.float32)
    w2 = rng.normal(size=(rs.height, rs.width, 3)).astype(np.float32)
    wave = 16
    g1 = gpu_adjoint(renderer, sc, rs, w1, wave_spp=wave)
    g2 = gpu_adjoint(renderer, sc, rs, w2, wave_spp=wave)
    g12 = gpu_adjoint(renderer, sc, rs, (2.0 * w1 - 0.5 * w2).astype(np.float32), wave_spp=wave)
    want = 2.0 * g1.astype(np.float64) - 0.5 * g2
    scale = np.abs(want).max()
    err = np.abs(g12 - want).max() / scale
    print(f"full size: |g| max {scale:.4g}, linearity error {err:.3g}, nonzero {int((g1 != 0).sum())}/{g1.size}")
    assert np.isfinite(g1).all() and np.isfinite(g2).all()
    assert err < 1e-4
    assert (g1 != 0).sum() > 0.9 * g1.size
    g0 = gpu_adjoint(renderer, sc, rs, np.zeros_like(w1), wave_spp=wave)
    assert np.all(g0 == 0)


def test_adjoint_full_size_time_ratio(renderer, full_workload):
    sc, rs = full_workload
    rng = np.random.default_rng(44)
    w = rng.normal(size=(rs.height, rs.width, 3)).astype(np.float32)
    renderer.upload_scene(sc)
    renderer.set_options(options_for(rs, wave_spp=16))
    renderer.set_camera(sc.camera())
    renderer.clear()
    renderer.render(0, rs.spp)
    renderer.synchronize()
    renderer.adjoint(0, rs.spp, w)
    renderer.clear()
    renderer.render(0, rs.spp)
    renderer.synchronize()
    fwd = renderer.stats()["ms_last_render"]
    renderer.adjoint(0, rs.spp, w)
    bwd = renderer.stats()["ms_last_render"]
    print(f"forward {fwd:.2f} ms, adjoint {bwd:.2f} ms, ratio {bwd / fwd:.2f}")
    assert bwd / fwd <= 4.0


def test_adjoint_refuses_unsupported_paths(renderer):
    """Stochastic alpha in shadow rays and volume MIS resolve visibility
    inside the shade stage, which the replay cannot reproduce from recorded
    bits: pt_adjoint returns PT_ERR_UNSUPPORTED before any device work."""
    for kind in ("alpha_blend", "volmis"):
        sc, rs = scenes.feature_scene(kind, width=32, height=32, spp=1)
        renderer.upload_scene(sc)
        renderer.set_options(options_for(rs))
        renderer.set_camera(sc.camera())
        with pytest.raises(PtError) as e:
            renderer.adjoint(0, 1, np.zeros((rs.height, rs.width, 3), np.float32))
        assert e.value.code == -6


def test_torch_function_descends_l2_loss(renderer):
    """The adjoint behind a torch.autograd.Function drives Adam on an L2 image
    loss against a target rendered from a second parameter set: the loss
    falls, and the Function's gradient equals a direct pt_adjoint call."""
    import torch
    from diffrender import RenderMaterials, clamp_params
    sc, rs = small_boxes(num_materials=6)
    base = sc.arrays()["materials"]
    n = base.shape[0]
    target_p = scenes.diff_material_params(0xD1FF0005, n)
    start_p = scenes.diff_material_params(0xD1FF0004, n)
    with torch.no_grad():
        target = RenderMaterials.apply(torch.from_numpy(target_p), renderer, sc, base, 0, rs.spp)
    p = torch.tensor(start_p, requires_grad=True)
    opt = torch.optim.Adam([p], lr=0.03)
    losses = []
    for it in range(25):
        opt.zero_grad()
        img = RenderMaterials.apply(p, renderer, sc, base, 0, rs.spp)
        loss = torch.mean((img - target) ** 2)
        loss.backward()
        if it == 0:
            direct = renderer.adjoint(0, rs.spp, (2.0 / img.numel() * (img - target)).detach().numpy())
            assert np.allclose(p.grad.numpy(), direct, rtol=1e-4, atol=1e-7 * np.abs(direct).max())
        losses.append(float(loss))
        opt.step()
        clamp_params(p)
    print("loss", " ".join(f"{x:.3g}" for x in losses))
    assert losses[-1] < 0.3 * losses[0]

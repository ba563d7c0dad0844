"""Converged images of the GPU renderer against the oracle's FP64 render.

The float GPU path and the oracle's double instantiation estimate the same
per-pixel expectation; with independent samples (the FP64 render draws
sample indices far from the GPU's) the squared difference of an N_g-sample
GPU mean and an N_r-sample FP64 mean has expectation
sigma^2 * (1/N_g + 1/N_r) per pixel channel, sigma^2 being the per-sample
variance, estimated from B independent GPU batches. Acceptance, stated per
workload below:

* RMSE(GPU, FP64) / predicted RMSE within [0.85, 1.15]: a bias of the float
  path adds to the measured RMSE and fails the upper bound;
* the image-mean difference within 4 of its standard errors (from the spread
  of the batches' image means);
* the plain and wavefront images are bit-identical, so their RMSE against the
  FP64 render is one number.

Workloads: the Cornell box (256x256, depth 5, RR from depth 2) with 1024 GPU
and 128 FP64 samples per pixel; a 1920x64 band (rows 508-571) of the 1080p
headline frame (depth 8, RR from depth 3) with 256 GPU and 64 FP64 samples per
pixel.
"""
import numpy as np
import pytest

import oracle
from common import oracle_setup
from glslpt import scenes
from glslpt._capi import PT_MODE_PLAIN, PT_MODE_WAVEFRONT
from glslpt.renderer import Renderer, options_for

pytestmark = pytest.mark.gpu

REF_FIRST = 1 << 20  # FP64 sample indices: disjoint from the GPU's


@pytest.fixture(scope="module")
def renderer():
    r = Renderer(0)
    yield r
    r.close()


def _batches(r, sc, rs, mode, n, B, wave):
    r.upload_scene(sc)
    r.set_options(options_for(rs, mode=mode, wave_spp=wave))
    r.set_camera(sc.camera())
    per = n // B
    out = []
    for b in range(B):
        r.clear()
        r.render(b * per, per)
        r.synchronize()
        out.append(r.read_accum()[..., :3].astype(np.float64) / per)
    return np.stack(out)


def _check(label, gb, ref, rows, n_g, n_r):
    g = gb.mean(axis=0)[rows]
    ref = ref[rows]
    B = gb.shape[0]
    var = gb[:, rows].var(axis=0, ddof=1) * (n_g // B)  # per-sample variance, per pixel channel
    pred = np.sqrt(np.mean(var) * (1.0 / n_g + 1.0 / n_r))
    rmse = np.sqrt(np.mean((g - ref) ** 2))
    dm = g.mean() - ref.mean()
    # standard error of the image-mean difference from the spread of the
    # batches' image means (robust to heavy-tailed pixels): a batch holds
    # n_g / B samples per pixel, the FP64 render n_r
    vb = gb[:, rows].mean(axis=(1, 2, 3)).var(ddof=1)
    se_m = np.sqrt(vb / B + vb * (n_g // B) / n_r)
    print(f"{label}: RMSE vs FP64 {rmse:.5g}, predicted {pred:.5g} (ratio {rmse / pred:.3f}); "
          f"mean {g.mean():.6g} vs {ref.mean():.6g} ({dm / se_m:+.2f} se)")
    assert 0.85 <= rmse / pred <= 1.15
    assert abs(dm) <= 4 * se_m
    return rmse


@pytest.mark.parametrize("case", ["cornell", "headline_band"])
def test_converged_against_fp64(renderer, case):
    if case == "cornell":
        sc, rs = scenes.cornell()
        n_g, n_r, B, wave, rows = 1024, 128, 16, 64, slice(0, 256)
    else:
        sc, rs = scenes.instanced_spheres()
        n_g, n_r, B, wave, rows = 256, 64, 16, 16, slice(508, 572)
    sa, cam, o = oracle_setup(sc, rs)
    r0, r1 = rows.start, rows.stop
    ref = oracle.render(sa, cam, o, REF_FIRST, n_r, precision=64, rows=(r0, r1, 1))[..., :3] / n_r
    res = {}
    imgs = {}
    for mode in (PT_MODE_PLAIN, PT_MODE_WAVEFRONT):
        gb = _batches(renderer, sc, rs, mode, n_g, B, wave)
        imgs[mode] = gb
        res[mode] = _check(f"{case} mode={mode}", gb, ref, rows, n_g, n_r)
    assert np.array_equal(imgs[PT_MODE_PLAIN], imgs[PT_MODE_WAVEFRONT])
    assert res[PT_MODE_PLAIN] == res[PT_MODE_WAVEFRONT]

"""The renderer's per-path device code compiled for the CPU (tests/hostpath)
against the oracle, on the CPU (BLAS as BVH2 and as four-wide nodes): the kernels' arithmetic, expression by
expression, is the oracle's, so with the same math library the images are
bit-identical on every option path (the GPU differs from the oracle only
where the device math library rounds a transcendental function differently).
The adjoint's replay, in the same CPU build: its radiance is bit-identical to
the primal pass's, and its gradients match central differences of the
oracle's frozen-sampling render within 1%."""
import numpy as np
import pytest

import hostpath
import oracle
from common import camera_rays, oracle_setup, random_dirs, small_instanced
from glslpt import scenes
from glslpt._capi import ADJOINT_FIELDS, ADJOINT_MEDIUM_FIELDS, ADJOINT_MEDIUM_NAMES, ADJOINT_NAMES


def _scene(kind):
    if kind == "cornell":
        return scenes.cornell(width=96, height=96, spp=3)
    if kind == "instanced":
        return small_instanced()
    if kind == "heavy":
        return scenes.heavy_mixed(width=64, height=48, spp=2, max_depth=8, sphere_freqs=(4, 5, 6, 7, 9),
                                  leaf_quads_per_mesh=(20, 30, 40), counts=(24, 8, 4, 2, 1, 20, 6, 3), extent=20.0,
                                  sky=(128, 64))
    if kind == "boxes":
        return scenes.material_boxes(num_boxes=48, sphere_freq=8, num_spheres=3, num_materials=12, width=64,
                                     height=64, spp=2, max_depth=4, extent=6.0)
    return scenes.feature_scene(kind)


KINDS = ["cornell", "instanced", "boxes", "heavy", "envmap", "textures", "media", "volmis", "alpha_mask",
         "alpha_blend"]


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("kind", KINDS)
def test_device_path_on_host_equals_oracle(kind, width):
    """width 4: the BLAS traversed as four-wide nodes collapsed from the
    reference's BVH2 (nearest child first), which reaches the same closest
    hits; on these scenes no ray ties, so the images stay bit-identical."""
    sc, rs = _scene(kind)
    h = hostpath.render(sc, rs, 0, rs.spp, bvh_width=width)
    sa, cam, o = oracle_setup(sc, rs)
    r = oracle.render(sa, cam, o, 0, rs.spp)
    same = (h == r).all(axis=2).mean()
    print(f"{kind} bvh{width}: {rs.width}x{rs.height} x{rs.spp} spp, bit-identical pixels {same:.6f}")
    assert np.array_equal(h, r)


def test_host_adjoint_matches_frozen_fd_of_oracle():
    sc, rs = _scene("boxes")
    spp = 4
    sa, cam, o = oracle_setup(sc, rs)
    mats = sa.a["materials"]
    w = np.random.default_rng(41).normal(size=(rs.height, rs.width, 3)).astype(np.float32)
    g, dropped, mismatch = hostpath.adjoint(sc, rs, w, 0, spp)
    assert dropped == 0 and mismatch == 0
    eps = 1e-3
    fd = np.zeros_like(g)
    for m in range(mats.shape[0]):
        for k, f in enumerate(ADJOINT_FIELDS):
            mp, mm = mats.copy(), mats.copy()
            mp[m, f] += eps
            mm[m, f] -= eps
            ip = oracle.render_frozen(sa, mp, cam, o, 0, spp)[..., :3].astype(np.float64)
            im = oracle.render_frozen(sa, mm, cam, o, 0, spp)[..., :3].astype(np.float64)
            fd[m, k] = np.sum(w * (ip - im)) / (2 * eps) / spp
    scale = np.abs(fd).max()
    rel = np.abs(g - fd) / np.maximum(np.abs(fd), 0.02 * scale)
    for m, k in zip(*np.nonzero(rel > 0.01)):
        print(f"mat {m} {ADJOINT_NAMES[k]}: adjoint {g[m, k]:.6g} fd {fd[m, k]:.6g}")
    print(f"params {fd.size}, above 2% of max {int((np.abs(fd) > 0.02 * scale).sum())}, worst {rel.max():.3g}")
    assert (np.abs(fd) > 0.02 * scale).sum() >= 20
    assert rel.max() <= 0.01


def test_host_medium_adjoint_matches_frozen_fd_of_oracle():
    """Medium gradients (density, colour) of an absorbing, a scattering and an
    emissive medium against central differences of the oracle's
    frozen-sampling render, whose scatter and crossing weights carry the
    free-flight density (distances drawn with the scene's density)."""
    sc, rs = _scene("media")
    spp = 8
    sa, cam, o = oracle_setup(sc, rs)
    mats = sa.a["materials"]
    w = np.random.default_rng(43).normal(size=(rs.height, rs.width, 3)).astype(np.float32)
    g, mg, dropped, mismatch = hostpath.adjoint(sc, rs, w, 0, spp, medium=True)
    assert dropped == 0 and mismatch == 0
    media = [m for m in range(mats.shape[0]) if mats[m, 18] != 0]
    assert sorted(int(mats[m, 18]) for m in media) == [1, 2, 3]
    assert np.all(mg[[m for m in range(mats.shape[0]) if m not in media]] == 0)
    eps = 1e-3
    fd = np.zeros((len(media), 4))
    for i, m in enumerate(media):
        for k, f in enumerate(ADJOINT_MEDIUM_FIELDS):
            mp, mm = mats.copy(), mats.copy()
            mp[m, f] += eps
            mm[m, f] -= eps
            ip = oracle.render_frozen(sa, mp, cam, o, 0, spp)[..., :3].astype(np.float64)
            im = oracle.render_frozen(sa, mm, cam, o, 0, spp)[..., :3].astype(np.float64)
            fd[i, k] = np.sum(w * (ip - im)) / (2 * eps) / spp
    got = mg[media]
    scale = np.abs(fd).max()
    rel = np.abs(got - fd) / np.maximum(np.abs(fd), 0.02 * scale)
    for i, m in enumerate(media):
        print(f"medium type {int(mats[m, 18])}: " + ", ".join(
            f"{ADJOINT_MEDIUM_NAMES[k]} {got[i, k]:.6g} / fd {fd[i, k]:.6g}" for k in range(4)))
    print(f"worst {rel.max():.3g}")
    assert (np.abs(fd) > 0.02 * scale).sum() >= 10
    assert rel.max() <= 0.01


@pytest.mark.parametrize("kind", ["envmap", "media", "alpha_mask", "textures"])
def test_host_replay_radiance_equals_primal(kind):
    """On every deferred-NEE option path the replay's radiance is
    bit-identical to the primal's, whatever the loss gradient."""
    sc, rs = _scene(kind)
    w = np.random.default_rng(7).normal(size=(rs.height, rs.width, 3)).astype(np.float32)
    g, dropped, mismatch = hostpath.adjoint(sc, rs, w, 0, 2)
    print(f"{kind}: dropped {dropped}, mismatched paths {mismatch}, |g| {np.abs(g).max():.3g}")
    assert mismatch == 0 and dropped == 0
    assert np.isfinite(g).all()


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("kind", ["cornell", "instanced", "heavy"])
def test_split_traversal_equals_whole(kind, width):
    """The capped extend stage's resumable traversal: stopped every `budget`
    iterations, its stack moved to another stack (16 entries in LDS, the rest
    in the backing store), and resumed, it gives every ray's hit bit for bit,
    for budgets from 1 to 32, on camera rays and rays from inside the scene."""
    sc, rs = _scene(kind)
    rng = np.random.default_rng(61)
    ro, rd = camera_rays(sc.camera(), rs.width, rs.height, rng, 4000)
    lo, hi = {"cornell": ((0.05,) * 3, (0.5,) * 3), "instanced": ((-6,) * 3, (6,) * 3),
              "heavy": ((-20, 0, -20), (20, 8, 20))}[kind]
    ri = rng.uniform(lo, hi, (4000, 3)).astype(np.float32)
    ro = np.concatenate([ro, ri])
    rd = np.concatenate([rd, random_dirs(rng, 4000)])
    for budget in (1, 2, 3, 7, 16, 32):
        differ, pieces = hostpath.closest_split(sc, rs, ro, rd, budget, bvh_width=width)
        print(f"{kind} bvh{width} budget {budget}: {len(ro)} rays in {pieces} pieces, differing {differ}")
        assert differ == 0
        if budget == 1:
            assert pieces > 5 * len(ro)


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("kind", ["instanced", "heavy", "alpha_mask"])
def test_any_hit_unordered_equals_ordered(kind, width):
    """The shadow stage's any-hit traversal visits a node's hit children in
    child order (no random draws in it); its answers equal the reference's
    nearest-first order on every ray, with and without alpha masks."""
    sc, rs = _scene(kind)
    rng = np.random.default_rng(62)
    lo, hi = {"instanced": ((-6,) * 3, (6,) * 3), "heavy": ((-20, 0, -20), (20, 8, 20)),
              "alpha_mask": ((-3, 0, -3), (3, 3, 3))}[kind]
    ro = rng.uniform(lo, hi, (20000, 3)).astype(np.float32)
    rd = random_dirs(rng, 20000)
    md = rng.uniform(0.1, 30, 20000).astype(np.float32)
    differ, hits = hostpath.any_orders(sc, rs, ro, rd, md, bvh_width=width)
    print(f"{kind} bvh{width}: {len(ro)} rays, occluded {hits}, differing {differ}")
    assert differ == 0 and 0 < hits < len(ro)

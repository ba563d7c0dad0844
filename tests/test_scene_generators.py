"""Procedural workload generators (host only): the vectorised geodesic
sphere equals the loop form it replaced, and the heavy mixed scene has the
structure its description states (instance and triangle counts, alpha-tested
share, media, sky, no coplanar volume floor)."""
import numpy as np
import pytest

from glslpt import scenes


@pytest.mark.parametrize("freq", [1, 2, 3, 5, 8, 17, 28])
def test_geodesic_vectorised_equals_loop(freq):
    v, f = scenes.geodesic_sphere(freq)
    v0, f0 = scenes._geodesic_sphere_loop(freq)
    assert f.shape == (20 * freq * freq, 3)
    assert np.array_equal(v, v0) and np.array_equal(f, f0)
    # closed surface: Euler characteristic V - E + F = 2, every edge shared by two faces
    edges = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert np.all(counts == 2)
    assert v.shape[0] - counts.shape[0] + f.shape[0] == 2


def test_heavy_counts_match_description():
    tris = [20 * f * f for f in scenes.HEAVY_SPHERE_FREQS] + [2 * q for q in scenes.HEAVY_LEAF_QUADS]
    assert len(tris) == 8 and min(tris) >= 20000 and max(tris) <= 2000000
    assert sum(scenes.HEAVY_COUNTS) == 1000
    inst = sum(c * t for c, t in zip(scenes.HEAVY_COUNTS, tris))
    leaves = sum(c * t for c, t in zip(scenes.HEAVY_COUNTS[5:], tris[5:]))
    assert 20e6 <= inst <= 26e6
    assert 0.25 <= leaves / inst <= 0.35


def test_heavy_small_structure():
    sc, rs = scenes.heavy_mixed(width=64, height=48, spp=2, max_depth=6, sphere_freqs=(3, 4, 5, 6, 7),
                                leaf_quads_per_mesh=(10, 15, 20), counts=(6, 3, 2, 1, 1, 5, 2, 1), extent=15.0,
                                sky=(64, 32))
    a = sc.arrays()
    assert a["transforms"].shape[0] == 21 + 1 + 2          # instances + ground + two volume cubes
    mats = a["materials"]
    vol = mats[mats[:, 18] == 2]                            # mediumType (float 18): scatter
    assert vol.shape[0] == 2 and np.all(vol[:, 20:23] == np.float32(0.8))
    assert np.all((vol[:, 19] >= 0.1) & (vol[:, 19] <= 2.0))  # density
    leaf = mats[(mats[:, 29] == 1) & (mats[:, 24] >= 0)]    # BLEND with a base-colour texture
    assert leaf.shape[0] == 1 and leaf[0, 28] == 1.0
    assert tuple(a["env_size"]) == (64, 32) and a["env_img"].shape == (64 * 32, 3) and a["lights"].shape[0] == 0
    assert rs.extra.get("enable_envmap") == 1 and (rs.width, rs.height, rs.max_depth) == (64, 48, 6)
    # volume cubes (unit box scaled by s, centre at s + 0.05) do not touch the ground plane
    for t in a["transforms"]:
        m = t.reshape(4, 4)
        s, y = m[0, 0], m[3, 1]
        if abs(m[1, 1] - s) < 1e-6 and abs(y - s - 0.05) < 1e-5:
            assert y - s > 0.0

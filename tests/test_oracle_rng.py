"""PCG4D and rand() of the oracle (globals.glsl:149-166) against an
independent numpy uint32 implementation and stored known answers."""
import json
import os

import numpy as np

import oracle

KAT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pcg4d_kat.json")


def pcg4d_np(v):
    """Jarzynski & Olano (JCGT 2020) pcg4d on a uint32[4] vector, in numpy."""
    v = np.array(v, dtype=np.uint32)
    with np.errstate(over="ignore"):
        v = v * np.uint32(1664525) + np.uint32(1013904223)
        v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2]
        v = v ^ (v >> np.uint32(16))
        v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2]
    return v


def rand_np(state):
    return np.float32(np.float32(state[0]) / np.float32(4294967296.0))


SEEDS = [(0, 0, 2, 0), (1, 2, 3, 3), (255, 255, 5, 510), (1919, 1079, 65, 2998), (7, 0, 2, 7),
         (0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFE), (12345, 678, 90, 13023), (3, 1, 1000002, 4)]


def test_pcg4d_matches_numpy():
    for seed in SEEDS:
        st, r = oracle.pcg4d(seed, 64)
        v = np.array(seed, np.uint32)
        for i in range(64):
            v = pcg4d_np(v)
            assert np.array_equal(st[i], v), (seed, i)
            assert r[i] == rand_np(v)


def test_pcg4d_known_answers():
    with open(KAT) as f:
        kat = json.load(f)
    for case in kat:
        st, r = oracle.pcg4d(case["seed"], len(case["states"]))
        assert st.tolist() == case["states"]
        assert [float(x) for x in r] == case["rand"]


def test_rand_edges():
    # rand() = float(x) / float(0xffffffff): the divisor rounds to 2^32, so the
    # range is [0, 1] inclusive; x >= 0xffffff80 rounds to 1.0.
    for x, want in ((0, 0.0), (0xFFFFFFFF, 1.0), (0xFFFFFF80, 1.0), (0xFFFFFF7F, float(np.float32(4294967040.0) / np.float32(2 ** 32)))):
        assert rand_np(np.array([x, 0, 0, 0], np.uint32)) == np.float32(want)
    assert float(np.float32(0xFFFFFFFF)) == 4294967296.0


def test_rand_uniformity():
    _, r = oracle.pcg4d((5, 6, 2, 11), 200000)
    h, _ = np.histogram(r, bins=20, range=(0, 1))
    e = len(r) / 20
    chi2 = ((h - e) ** 2 / e).sum()
    assert chi2 < 45.3  # 19 dof, p = 0.001


if __name__ == "__main__":  # regenerate the known-answer file from the numpy implementation
    out = []
    for seed in SEEDS:
        v = np.array(seed, np.uint32)
        states, rs = [], []
        for _ in range(16):
            v = pcg4d_np(v)
            states.append([int(x) for x in v])
            rs.append(float(rand_np(v)))
        out.append({"seed": list(seed), "states": states, "rand": rs})
    with open(KAT, "w") as f:
        json.dump(out, f)

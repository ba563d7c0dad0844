"""The binade form of the environment CDF's running sum (the parallel device
scan's arithmetic) against the sequential float32 running sum of the
reference's EnvironmentMap::BuildCDF (src/core/EnvironmentMap.cpp:51-53),
numpy's float32 cumsum here: between two crossings of a power of two every add
rounds on one grid u, so r = R*u and each add moves R by floor(w/u), plus one
above a half or, at exactly a half, when R + floor(w/u) is odd; an add whose
exact sum leaves the grid is a plain float add. The model walks that rule
element by element (the device scans it in parallel) and must give the same
bits on inputs that stress ties, grid crossings, stalls at 2^24, subnormals and
overflow."""
import math

import numpy as np
import pytest

TOP = 1 << 24


def binade_scan(w):
    n = len(w)
    out = np.empty(n, np.float32)
    if n == 0:
        return out
    out[0] = w[0]
    r = np.float32(w[0])
    i = 1
    while i < n:
        if np.isinf(r):
            out[i:] = np.float32(np.inf)
            break
        k = -149 if r < np.float32(2.0 ** -126) else math.frexp(float(r))[1] - 24
        u = 2.0 ** k
        R = int(float(r) / u)
        assert R * u == float(r) and R < TOP
        while i < n:
            q = float(w[i]) / u
            F = math.floor(q)
            if R + F >= TOP:
                with np.errstate(over="ignore"):
                    r = np.float32(np.float32(R * u) + w[i])
                out[i] = r
                i += 1
                break
            frac = q - F
            R += F + (1 if frac > 0.5 else (((R + F) & 1) if frac == 0.5 else 0))
            out[i] = np.float32(R * u)
            i += 1
            if R >= TOP:
                r = np.float32(R * u)
                break
    return out


def cases():
    rng = np.random.default_rng(1)
    yield "uniform", rng.random(20000)
    yield "lognormal", rng.lognormal(0, 3, 20000)
    yield "zeros_mixed", np.where(rng.random(20000) < 0.3, 0, rng.random(20000) * 50)
    yield "quarter_ties", rng.integers(0, 8, 50000) * 0.25
    yield "small_ties", rng.integers(1, 4, 50000) * 2.0 ** -20
    yield "subnormal", rng.random(5000) * 1e-40
    yield "scales", np.concatenate([rng.random(100) * 1e-42, rng.random(100), rng.random(100) * 1e36])
    yield "overflow", np.array([3e38, 3e38, 1.0])
    yield "all_zero", np.zeros(3)
    yield "single", np.array([5.0])
    yield "stall_2_24", np.concatenate([[16777211.0], np.ones(100)])
    yield "halves_at_2_24", np.concatenate([[16777215.0], np.full(50, 0.5), np.full(50, 1.5)])
    yield "halves_spacing_1", np.concatenate([[2.0 ** 23 + 1], np.full(200, 0.5), np.full(200, 1.5)])
    yield "round_to_inf", np.concatenate([[2.0 ** 104 * 16777215], np.full(5, 2.0 ** 103)])


@pytest.mark.parametrize("name,w", list(cases()), ids=[c[0] for c in cases()])
def test_binade_scan_equals_sequential_sum(name, w):
    w = np.asarray(w, np.float32)
    with np.errstate(over="ignore"):
        want = np.cumsum(w, dtype=np.float32)
    got = binade_scan(w)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), name


def test_plausible_mistake_is_caught():
    """Rounding ties upward instead of to even breaks the equality (halves added
    where the grid spacing is 1)."""
    w = np.concatenate([[2.0 ** 23 + 1], np.full(200, 0.5), np.full(200, 1.5)]).astype(np.float32)
    want = np.cumsum(w, dtype=np.float32)
    out = np.empty_like(w)
    out[0] = w[0]
    r = float(w[0])
    for i in range(1, len(w)):  # round-half-up restatement
        k = math.frexp(r)[1] - 24 if r > 0 else -149
        u = 2.0 ** k
        s = (r + float(w[i])) / u
        r = math.floor(s + 0.5) * u
        out[i] = r
    assert not np.array_equal(out.view(np.uint32), want.view(np.uint32))

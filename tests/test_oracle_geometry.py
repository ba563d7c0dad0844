"""Intersection and traversal of the oracle (intersection.glsl,
closest_hit.glsl, anyhit.glsl) against closed forms and an FP64 brute force
over every triangle of every instance and every light."""
import numpy as np
import pytest

import oracle
from common import camera_rays, oracle_setup, random_dirs, small_instanced
from glslpt import scenes


def test_aabb_closed_form():
    box = [0, 0, 0, 1, 1, 1]
    cases = [  # origin, dir, expected (t0 if t0 > 0 else t1 when hit, else -1)
        ((-1, 0.5, 0.5), (1, 0, 0), 1.0),
        ((0.5, 0.5, 0.5), (1, 0, 0), 0.5),      # inside: exit distance
        ((-1, 2.0, 0.5), (1, 0, 0), -1.0),      # miss
        ((2, 0.5, 0.5), (1, 0, 0), -1.0),       # behind
        ((-1, -1, -1), (1, 1, 1), 1.0),         # through the corner
    ]
    for o, d, want in cases:
        got = oracle.fn("AABBIntersect", np.array([box]), np.array([o]), np.array([d], np.float64))[0]
        assert got == pytest.approx(want, abs=1e-6)
    rng = np.random.default_rng(1)
    n = 20000
    lo = rng.uniform(-1, 0, (n, 3)); hi = lo + rng.uniform(0.1, 2, (n, 3))
    o = rng.uniform(-3, 3, (n, 3)); d = random_dirs(rng, n).astype(np.float64)
    got = oracle.fn("AABBIntersect", np.concatenate([lo, hi], 1), o, d)
    inv = 1 / d
    t1 = np.minimum((hi - o) * inv, (lo - o) * inv).max(1)
    t2 = np.maximum((hi - o) * inv, (lo - o) * inv).min(1)
    want = np.where(t2 >= t1, np.where(t1 > 0, t1, t2), -1.0)
    close = np.abs(t2 - t1) > 1e-5
    np.testing.assert_allclose(got[close], want[close], rtol=1e-5, atol=1e-5)


def test_sphere_and_rect_closed_form():
    # sphere radius 1 at the origin
    t = oracle.fn("SphereIntersect", np.array([1.0]), np.array([[0, 0, 0]]), np.array([[0, 0, -5.0]]), np.array([[0, 0, 1.0]]))
    assert t[0] == pytest.approx(4.0)
    t = oracle.fn("SphereIntersect", np.array([1.0]), np.array([[0, 0, 0]]), np.array([[0, 0, 0.0]]), np.array([[0, 0, 1.0]]))
    assert t[0] == pytest.approx(1.0)  # from inside: far root
    t = oracle.fn("SphereIntersect", np.array([1.0]), np.array([[0, 0, 0]]), np.array([[0, 2, -5.0]]), np.array([[0, 0, 1.0]]))
    assert t[0] == 1e6  # miss -> INF
    # rect at y = 1 spanning x, z in [0, 2]; the caller passes u/|u|^2, v/|v|^2
    pos, u, v = np.array([0, 1, 0.0]), np.array([2.0, 0, 0]), np.array([0, 0, 2.0])
    n = np.cross(u, v); n /= np.linalg.norm(n)
    plane = np.concatenate([n, [n @ pos]])
    rng = np.random.default_rng(2)
    o = np.column_stack([rng.uniform(-1, 3, 500), np.full(500, -1.0), rng.uniform(-1, 3, 500)])
    d = np.tile([0, 1.0, 0], (500, 1))
    args = np.concatenate([np.tile(pos, (500, 1)), np.tile(u / (u @ u), (500, 1)), np.tile(v / (v @ v), (500, 1)),
                           np.tile(plane, (500, 1)), o, d], 1)
    t = oracle.fn("RectIntersect", args)
    inside = (o[:, 0] >= 0) & (o[:, 0] <= 2) & (o[:, 2] >= 0) & (o[:, 2] <= 2)
    np.testing.assert_allclose(t[inside], 2.0, rtol=1e-6)
    assert np.all(t[~inside] == 1e6)


def world_triangles(arrays):
    """(instance, vertIndices row, 3x3 world vertices in double) of every triangle."""
    nodes, tli = arrays["nodes"], arrays["top_level_index"]
    vi, vu, xf = arrays["vert_indices"], arrays["vertices_uvx"], arrays["transforms"]
    out_inst, out_row, out_v = [], [], []

    def rows_of(root):
        st, rows = [root], []
        while st:
            i = st.pop()
            lr = nodes[i, 6:9]
            if lr[2] == 0:
                st += [int(lr[0]), int(lr[1])]
            else:
                rows += list(range(int(lr[0]), int(lr[0]) + int(lr[1])))
        return rows

    for i in range(tli, nodes.shape[0]):
        if nodes[i, 8] < 0:
            inst = -int(nodes[i, 8]) - 1
            M = xf[inst].reshape(4, 4).astype(np.float64)  # M[col][row]
            for r in rows_of(int(nodes[i, 6])):
                p = vu[vi[r], :3].astype(np.float64)
                w = p @ M[:3, :3] + M[3, :3]
                out_inst.append(inst); out_row.append(r); out_v.append(w)
    return np.array(out_inst), np.array(out_row), np.array(out_v)


def brute_closest(tris, lights, o, d):
    """FP64 Moller-Trumbore over all triangles plus the analytic lights; returns
    best t, second-best t, and the winner (kind, id)."""
    inst, rows, V = tris
    o = o.astype(np.float64); d = d.astype(np.float64)
    e0 = V[:, 1] - V[:, 0]; e1 = V[:, 2] - V[:, 0]
    n = o.shape[0]
    best = np.full(n, np.inf); second = np.full(n, np.inf); who = np.full(n, -1)
    for k in range(V.shape[0]):
        pv = np.cross(d, e1[k]); det = pv @ e0[k]
        tv = o - V[k, 0]; qv = np.cross(tv, e0[k])
        with np.errstate(divide="ignore", invalid="ignore"):
            u = np.sum(tv * pv, 1) / det; v = np.sum(d * qv, 1) / det; t = (qv @ e1[k]) / det
        ok = (u >= 0) & (v >= 0) & (u + v <= 1) & (t >= 0)
        t = np.where(ok, t, np.inf)
        better = t < best
        second = np.where(better, best, np.minimum(second, t))
        who = np.where(better, k, who)
        best = np.where(better, t, best)
    lt = np.full(n, np.inf); lwho = np.full(n, -1)
    for li, L in enumerate(lights):
        pos, u, v = L[0:3].astype(np.float64), L[6:9].astype(np.float64), L[9:12].astype(np.float64)
        if int(L[14]) == 0:
            nrm = np.cross(u, v); nrm /= np.linalg.norm(nrm)
            dt = d @ nrm
            with np.errstate(divide="ignore", invalid="ignore"):
                t = ((pos - o) @ nrm) / dt
            p = o + d * t[:, None] - pos
            a1 = p @ u / (u @ u); a2 = p @ v / (v @ v)
            ok = (dt <= 0) & (t > 3e-4) & (a1 >= 0) & (a1 <= 1) & (a2 >= 0) & (a2 <= 1)
            t = np.where(ok, t, np.inf)
        elif int(L[14]) == 1:
            op = pos - o
            b_ = np.sum(op * d, 1)
            det = b_ * b_ - np.sum(op * op, 1) + L[12] ** 2
            sq = np.sqrt(np.maximum(det, 0))
            t1, t2 = b_ - sq, b_ + sq
            t = np.where(det < 0, np.inf, np.where(t1 > 1e-3, t1, np.where(t2 > 1e-3, t2, np.inf)))
        else:
            continue
        b = t < lt
        lwho = np.where(b, li, lwho); lt = np.where(b, t, lt)
    return best, second, who, lt, lwho


@pytest.mark.parametrize("which", ["cornell", "instanced"])
def test_closest_hit_brute_force(which):
    sc, rs = scenes.cornell() if which == "cornell" else small_instanced()
    a = sc.arrays()
    sa, cam, o = oracle_setup(sc, rs)
    rng = np.random.default_rng(3)
    ro, rd = camera_rays(sc.camera(), rs.width, rs.height, rng, 1500)
    if which == "cornell":
        ri = rng.uniform(0.05, 0.5, (1500, 3)).astype(np.float32)
    else:
        ri = rng.uniform(-6, 6, (1500, 3)).astype(np.float32)
    ro = np.concatenate([ro, ri]); rd = np.concatenate([rd, random_dirs(rng, 1500)])
    res = oracle.closest_hit(sa, o, ro, rd)
    tris = world_triangles(a)
    with np.errstate(invalid="ignore"):
        best, second, who, lt, lwho = brute_closest(tris, a["lights"], ro, rd)
    light_wins = lt <= best
    bt = np.minimum(best, lt)
    hit = np.isfinite(bt)
    # unambiguous rays: the runner-up is clearly farther
    with np.errstate(invalid="ignore"):
        clear = hit & ((np.minimum(second, np.where(light_wins, best, lt)) - bt) > 1e-4 * bt)
    got_t = res["t"].astype(np.float64)
    assert np.all((res["t"] >= 1e6) == ~hit), "hit/miss disagreement"
    np.testing.assert_allclose(got_t[clear], bt[clear], rtol=1e-4, atol=1e-6)
    tri_rows = tris[1]
    tri_inst = tris[0]
    surf = clear & ~light_wins
    assert np.array_equal(res["triangle"][surf], tri_rows[who[surf]])
    assert np.array_equal(res["instance"][surf], tri_inst[who[surf]])
    lw = clear & light_wins
    assert np.array_equal(res["light"][lw], lwho[lw])
    assert clear.sum() > 0.5 * hit.sum()


def test_inverse_modes_differ_only_on_ties():
    """The per-visit fp32 inverse and the correctly rounded one change the
    winner only where two candidates are at (nearly) the same distance."""
    sc, rs = scenes.cornell()
    sa, cam, o = oracle_setup(sc, rs)
    rng = np.random.default_rng(4)
    ro = rng.uniform(0.05, 0.5, (20000, 3)).astype(np.float32)
    rd = random_dirs(rng, 20000)
    oracle.set_inverse_mode(0)
    try:
        a = oracle.closest_hit(sa, o, ro, rd)
    finally:
        oracle.set_inverse_mode(1)
    b = oracle.closest_hit(sa, o, ro, rd)
    diff = (a["triangle"] != b["triangle"]) | (a["instance"] != b["instance"])
    assert diff.mean() < 0.01
    # every disagreement is a near tie in distance
    assert np.all(np.abs(a["t"][diff] - b["t"][diff]) <= 1e-4 * np.maximum(a["t"][diff], 1e-3))


def test_any_hit_brute_force():
    sc, rs = small_instanced()
    a = sc.arrays()
    sa, cam, o = oracle_setup(sc, rs)
    rng = np.random.default_rng(5)
    n = 3000
    ro = rng.uniform(-6, 6, (n, 3)).astype(np.float32)
    rd = random_dirs(rng, n)
    md = rng.uniform(0.1, 10, n).astype(np.float32)
    occ = oracle.any_hit(sa, o, ro, rd, md)
    best, second, who, lt, lwho = brute_closest(world_triangles(a), a["lights"], ro, rd)
    want = (best < md) | (lt < md)
    margin = np.abs(np.minimum(best, lt) - md) > 1e-4 * md
    assert np.array_equal(occ[margin].astype(bool), want[margin])

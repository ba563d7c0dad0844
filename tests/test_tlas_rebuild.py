"""Instance TLAS rebuild on the CPU side.

* The product scene builder's RebuildInstances against the reference's
  Scene::RebuildInstances (src/core/Scene.cpp:200-214), compiled from its
  sources into oracle/_ref/libref_scene.so: the flattened nodes after an edit
  of every transform are identical, for the cases of tlas_cases (including
  signed-zero boxes and the one-sided fallback split).
* The closed form that the device rebuild uses for the reference's in-place
  two-pointer partition (thirdparty/RadeonRays/bvh.cpp:130-196): split index,
  the resulting order, and the order in which each child box is grown, checked
  against a direct simulation of that loop for every flag pattern up to
  length 12.
"""
import itertools

import numpy as np
import pytest

import refscene
import tlas_cases

SIZES = {"random": 300, "coincident": 50, "clustered": 200, "chain": 120, "signed_zero": 100, "adjacent": 40}


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.skipif(not refscene.available(), reason="compiled reference harness not built here")
@pytest.mark.parametrize("case", tlas_cases.CASES)
def test_rebuild_matches_reference(case):
    n = SIZES[case]
    ms = tlas_cases.meshes(case, 3)
    insts, xf1 = tlas_cases.instances(case, n, 3)
    refscene.build(ms, insts, 3)
    ref_nodes = refscene.rebuild_instances(xf1)
    sc = tlas_cases.build_scene(ms, insts)
    for i in range(n):
        sc.SetInstanceTransform(i, xf1[i])
    sc.RebuildInstances()
    a = sc.arrays()
    assert a["nodes"].shape == ref_nodes.shape
    assert np.array_equal(bits(a["nodes"]), bits(ref_nodes)), case
    if case == "signed_zero":
        tli = a["top_level_index"]
        ymin = a["nodes"][tli:, 1]
        assert (np.signbit(ymin) & (ymin == 0)).any() and (~np.signbit(ymin) & (ymin == 0)).any()


def _sequential(flags, s):
    """The reference loop over element ids s.. s+n-1 with the given
    goes-left flags: final order, split, left and right grow orders."""
    n = len(flags)
    order = list(range(s, s + n))
    fl = {s + i: flags[i] for i in range(n)}
    lg, rg = [], []
    first, last = s, s + n
    while True:
        while first != last and fl[order[first - s]]:
            lg.append(order[first - s])
            first += 1
        done = first == last
        last -= 1
        if done:
            break
        rg.append(order[first - s])
        while first != last and not fl[order[last - s]]:
            rg.append(order[last - s])
            last -= 1
        if first == last:
            break
        lg.append(order[last - s])
        order[first - s], order[last - s] = order[last - s], order[first - s]
        first += 1
    return order, first, lg, rg


def _closed_form(flags, s):
    """Split, order and grow ranks as the device rebuild computes them."""
    n = len(flags)
    e = s + n
    f = {s + i: flags[i] for i in range(n)}
    split = s + sum(flags)
    a = [p for p in range(s, split) if not f[p]]
    b = [p for p in range(e - 1, split - 1, -1) if f[p]]
    assert len(a) == len(b)
    m = len(a)
    order = list(range(s, e))
    for k in range(m):
        order[a[k] - s] = b[k]
        order[b[k] - s] = a[k]
    lrank = {p: p - s for p in range(s, split)}
    brank = {b[k]: k for k in range(m)}
    rrank = {}
    for p in range(split, e):
        if p in brank:
            k = brank[p]
            rrank[p] = e - (e if k == 0 else b[k - 1])
        elif p == split:
            rrank[p] = e - (e if m == 0 else b[m - 1])
        else:
            rrank[p] = e - p
    return order, split, lrank, rrank


def test_partition_closed_form_matches_reference_loop():
    cases = 0
    for n in range(1, 13):
        for s in (0, 1, 5):
            for flags in itertools.product((0, 1), repeat=n):
                o1, sp1, lg, rg = _sequential(list(flags), s)
                o2, sp2, lr, rr = _closed_form(list(flags), s)
                assert o1 == o2 and sp1 == sp2, (flags, s)
                pos = {el: s + i for i, el in enumerate(o1)}
                assert [lr[pos[el]] for el in lg] == list(range(len(lg))), (flags, s)
                assert [rr[pos[el]] for el in rg] == list(range(len(rg))), (flags, s)
                cases += 1
    assert cases == 3 * (2 ** 13 - 2)

"""Four-wide TLAS (pt_options.bvh_width = 4) of the host re-layout, on the
CPU (tests/hostpath). The reference's TLAS is binary
(thirdparty/RadeonRays/bvh.cpp:68-243, flattened by bvh_translator.cpp); the
four-wide form gathers each node's descendants until it holds four, so it
must keep exactly the reference's boxes and instances:

* every instance is a leaf exactly once; every internal entry is named once,
  by an entry before it (breadth first from the root), and the entries after
  the last one are zero;
* each child is a subtree of the reference's binary TLAS (the same set of
  instances under it) and carries that subtree's box, bit for bit;
* a node holds four children unless all of its children are leaves.

Traversal over this TLAS is compared with the oracle, image for image, by
test_hostpath (width 4) on the CPU and test_gpu_bvh4 on the GPU; the device
collapse after a TLAS rebuild is compared with this one by test_gpu_tlas.
"""
import numpy as np
import pytest

import hostpath
import tlas_cases
from common import small_instanced

EMPTY4 = -2 ** 31


def flat_subtrees(sc):
    """{frozenset of instances: (min xyz, max xyz)} of every node of the
    reference's flat TLAS, and {instance: box} of its leaves."""
    a = sc.arrays()
    nodes, top = a["nodes"], a["top_level_index"]
    boxes, leaves = {}, {}

    def walk(i):
        row = nodes[i]
        if row[8] < 0:
            inst = int(-row[8]) - 1
            leaves[inst] = (row[0:3].copy(), row[3:6].copy())
            return frozenset([inst])
        assert row[8] == 0
        s = walk(int(row[6])) | walk(int(row[7]))
        boxes[s] = (row[0:3].copy(), row[3:6].copy())
        return s

    walk(top)
    return boxes, leaves


def check(sc, n):
    t4, root = hostpath.tlas4(sc)
    assert t4.shape[0] >= max(n - 1, 1)
    if n == 1:
        assert root == -1 and not t4.any()
        return 0
    refs = t4[:, 24:28].view(np.int32)
    lo = np.stack([t4[:, 0:4], t4[:, 8:12], t4[:, 16:20]], axis=2)  # (entry, child, xyz)
    hi = np.stack([t4[:, 4:8], t4[:, 12:16], t4[:, 20:24]], axis=2)
    boxes, leaves = flat_subtrees(sc)
    named = np.zeros(t4.shape[0], np.int32)
    named[0] = 1
    count = 0
    while count < t4.shape[0] and named[count]:
        q = count
        r = refs[q]
        k = int((r != EMPTY4).sum())
        assert 2 <= k <= 4 and np.all(r[:k] != EMPTY4) and np.all(r[k:] == EMPTY4)
        if k < 4:
            assert np.all(r[:k] < 0), f"node {q}: {k} children with an internal one"
        for j in range(k):
            if r[j] >= 0:
                c = int(r[j]) - root
                assert q < c < t4.shape[0] and named[c] == 0
                named[c] = 1
        count += 1
    assert named.sum() == count
    assert not t4[count:].any()

    def under(q):
        out = set()
        for j in range(4):
            if refs[q, j] == EMPTY4:
                continue
            out |= under(int(refs[q, j]) - root) if refs[q, j] >= 0 else {-1 - int(refs[q, j])}
        return frozenset(out)

    every = under(0)
    assert sorted(every) == list(range(n))
    for q in range(count):
        for j in range(4):
            rj = int(refs[q, j])
            if rj == EMPTY4:
                continue
            if rj < 0:
                want = leaves[-1 - rj]
            else:
                s = under(rj - root)
                assert s in boxes, f"entry {q} child {j}: not a subtree of the reference's TLAS"
                want = boxes[s]
            assert np.array_equal(lo[q, j].view(np.uint32), want[0].view(np.uint32))
            assert np.array_equal(hi[q, j].view(np.uint32), want[1].view(np.uint32))
    return count


@pytest.mark.parametrize("case", tlas_cases.CASES)
def test_tlas4_cases(case):
    n = 300
    sc = tlas_cases.build_scene(tlas_cases.meshes(case, 5), tlas_cases.instances(case, n, 5)[0])
    count = check(sc, n)
    print(f"{case} n={n}: {count} four-wide nodes for {n - 1} binary ones")


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 17, 64, 1000])
def test_tlas4_sizes(n):
    sc = tlas_cases.build_scene(tlas_cases.meshes("random", 6), tlas_cases.instances("random", n, 6)[0])
    check(sc, n)


def test_tlas4_instanced_scene():
    sc, _ = small_instanced()
    check(sc, sc.view().num_instances)


def test_check_catches_a_wrong_box(monkeypatch):
    """The check fails when one child box is off by one ulp."""
    sc = tlas_cases.build_scene(tlas_cases.meshes("random", 5), tlas_cases.instances("random", 50, 5)[0])
    t4, root = hostpath.tlas4(sc)
    bad = t4.copy()
    bad[0, 4] = np.nextafter(bad[0, 4], np.float32(np.inf))
    monkeypatch.setattr(hostpath, "tlas4", lambda s: (bad, root))
    with pytest.raises(AssertionError):
        check(sc, 50)

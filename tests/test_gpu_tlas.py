"""Device TLAS rebuild (pt_rebuild_instances) against the host path: the scene
builder's RebuildInstances, which test_tlas_rebuild pins bit for bit to the
reference's Scene::RebuildInstances, uploaded with pt_update_instances. The
device nodes, instance records and root are identical, so is the four-wide
TLAS that the device collapses from them to the host's collapse (pinned by
test_tlas4), and so are the images rendered after either path, in both
forms."""
import numpy as np
import pytest

import tlas_cases
from glslpt._capi import PtError
from glslpt.renderer import Renderer, default_options

pytestmark = pytest.mark.gpu

SIZES = {"random": 1000, "coincident": 100, "clustered": 500, "chain": 150, "signed_zero": 300, "adjacent": 64}


@pytest.fixture(scope="module")
def renderer():
    r = Renderer(0)
    yield r
    r.close()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def same(x, y):
    return np.array_equal(bits(x[0]), bits(y[0])) and np.array_equal(bits(x[1]), bits(y[1])) and x[2] == y[2]


def same4(x, y):
    return x[0].shape == y[0].shape and np.array_equal(bits(x[0]), bits(y[0])) and x[1] == y[1]


def host_path(r, sc, xf):
    for i in range(xf.shape[0]):
        sc.SetInstanceTransform(i, xf[i])
    sc.RebuildInstances()
    r.update_instances(sc)
    return r.read_tlas(xf.shape[0])


def check_case(r, case, n, seed):
    ms = tlas_cases.meshes(case, seed)
    insts, xf1 = tlas_cases.instances(case, n, seed)
    xf0 = np.stack([x for _, _, x in insts])
    sc = tlas_cases.build_scene(ms, insts)
    r.upload_scene(sc)
    uploaded = r.read_tlas(n)
    uploaded4 = r.read_tlas4()
    # unchanged transforms: the rebuild reproduces the uploaded TLAS, and the
    # device collapse the host's four-wide one
    r.rebuild_instances(xf0)
    assert same(r.read_tlas(n), uploaded), f"{case} n={n}: rebuild of the uploaded transforms"
    assert same4(r.read_tlas4(), uploaded4), f"{case} n={n}: four-wide TLAS of the uploaded transforms"
    r.rebuild_instances(xf1)
    dev = r.read_tlas(n)
    dev4 = r.read_tlas4()
    host = host_path(r, sc, xf1)
    host4 = r.read_tlas4()
    assert np.array_equal(bits(dev[0]), bits(host[0])), f"{case} n={n}: nodes"
    assert np.array_equal(bits(dev[1]), bits(host[1])), f"{case} n={n}: instance records"
    assert dev[2] == host[2]
    assert same4(dev4, host4), f"{case} n={n}: four-wide TLAS"


@pytest.mark.parametrize("case", tlas_cases.CASES)
def test_rebuild_equals_host(renderer, case):
    print(f"case {case} n={SIZES[case]}", flush=True)
    check_case(renderer, case, SIZES[case], 7)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 1024, 1025, 1500, 4097])
def test_rebuild_sizes(renderer, n):
    print(f"random n={n}", flush=True)
    check_case(renderer, "random", n, 11)


@pytest.mark.parametrize("n", [1500, 4097])
def test_rebuild_signed_zero_sizes(renderer, n):
    """-0.0 box coordinates select the 64-bit rank keys; at 1500 instances the
    arrays sit in LDS and the keys in global memory, at 4097 both in global
    memory, so every key type and placement the launcher picks is compared."""
    print(f"signed_zero n={n}", flush=True)
    check_case(renderer, "signed_zero", n, 13)


def test_repeated_rebuilds(renderer):
    """Several edits in a row, each equal to the host path from the same state."""
    n = 200
    ms = tlas_cases.meshes("random", 5)
    insts, _ = tlas_cases.instances("random", n, 5)
    sc = tlas_cases.build_scene(ms, insts)
    renderer.upload_scene(sc)
    for k in range(3):
        xf = tlas_cases.transforms("clustered" if k == 1 else "random", n, 20 + k)
        renderer.rebuild_instances(xf)
        dev = renderer.read_tlas(n)
        host = host_path(renderer, sc, xf)
        assert same(dev, host), f"edit {k}"


@pytest.mark.parametrize("width", [2, 4])
def test_render_after_rebuild_equals_host(renderer, width):
    n = 40
    ms = tlas_cases.meshes("random", 9)
    insts, xf1 = tlas_cases.instances("random", n, 9)
    xf1[:, 3, :3] *= 0.3
    sc = tlas_cases.build_scene(ms, insts)
    renderer.upload_scene(sc)
    renderer.set_options(default_options(width=48, height=32, max_depth=3, enable_uniform_light=1,
                                          uniform_light_col=(1.0, 1.0, 1.0), bvh_width=width))
    renderer.set_camera(sc.camera())
    images = []
    for path in ("device", "host"):
        if path == "device":
            renderer.rebuild_instances(xf1)
        else:
            renderer.upload_scene(sc)
            host_path(renderer, sc, xf1)
        renderer.clear()
        renderer.render(0, 4)
        renderer.synchronize()
        images.append(renderer.read_accum())
    assert np.abs(images[0]).max() > 0
    assert np.array_equal(images[0], images[1])


def test_material_ids(renderer):
    """Per-instance material ids land in the instance records; the tree does not
    depend on them."""
    n = 50
    ms = tlas_cases.meshes("random", 4)
    insts, xf1 = tlas_cases.instances("random", n, 4)
    sc = tlas_cases.build_scene(ms, insts)
    renderer.upload_scene(sc)
    mats = (np.arange(n) % 3).astype(np.int32)[::-1].copy()
    renderer.rebuild_instances(xf1, mats)
    nodes, recs, root = renderer.read_tlas(n)
    assert np.array_equal(recs[:, 22].view(np.int32), mats)
    renderer.rebuild_instances(xf1)
    nodes2, recs2, _ = renderer.read_tlas(n)
    assert np.array_equal(bits(nodes), bits(nodes2))
    assert np.array_equal(recs2[:, 22].view(np.int32), mats)


def test_errors_leave_state_unchanged(renderer):
    n = 20
    ms = tlas_cases.meshes("random", 2)
    insts, xf1 = tlas_cases.instances("random", n, 2)
    sc = tlas_cases.build_scene(ms, insts)
    renderer.upload_scene(sc)
    before = renderer.read_tlas(n)
    with pytest.raises(PtError):
        renderer.rebuild_instances(xf1[:-1])
    bad = xf1.copy()
    bad[7, :3, :3] = 0.0
    with pytest.raises(PtError):
        renderer.rebuild_instances(bad)
    with pytest.raises(PtError):
        renderer.rebuild_instances(xf1, np.full(n, 3, np.int32))
    assert same(renderer.read_tlas(n), before)
    # pt_read_tlas4 copies exactly the four-wide room, nothing else
    import ctypes as C

    from glslpt import _capi as api
    room = renderer.read_tlas4()[0].shape[0]
    assert room >= n - 1
    buf = np.zeros((room + 1, 32), np.float32)
    rc = renderer._lib.pt_read_tlas4(renderer._h, api.fptr(buf), room + 1, None, None)
    assert rc == api.PT_ERR_INVALID_ARG and not buf.any()


def test_update_without_nodes_keeps_tlas(renderer):
    """pt_update_instances with no node range refreshes the instance records
    only; after a device rebuild it keeps the rebuilt TLAS."""
    import ctypes as C

    from glslpt import _capi as api
    n = 30
    ms = tlas_cases.meshes("random", 6)
    insts, xf1 = tlas_cases.instances("random", n, 6)
    sc = tlas_cases.build_scene(ms, insts)
    renderer.upload_scene(sc)
    renderer.rebuild_instances(xf1)
    dev = renderer.read_tlas(n)
    dev4 = renderer.read_tlas4()
    xf = np.ascontiguousarray(xf1, np.float32)
    rc = renderer._lib.pt_update_instances(renderer._h, api.fptr(xf), n, None, 0, None, 0, 0)
    assert rc == api.PT_OK
    assert same(renderer.read_tlas(n), dev)
    assert same4(renderer.read_tlas4(), dev4)
    # a partial node range after a device rebuild is refused
    v = sc.view()
    rc = renderer._lib.pt_update_instances(renderer._h, api.fptr(xf), n, None, 0,
                                           C.cast(C.byref(v.nodes.contents, 0), C.POINTER(api.PtBvhNode)), 0, 1)
    assert rc == -3  # PT_ERR_STATE


def test_tlas_range_update_equals_full(renderer):
    """pt_update_instances with exactly the TLAS node range re-lays out and
    uploads the TLAS only; the device nodes equal those of a whole-array update."""
    n = 300
    ms = tlas_cases.meshes("clustered", 8)
    insts, xf1 = tlas_cases.instances("clustered", n, 8)
    sc = tlas_cases.build_scene(ms, insts)
    renderer.upload_scene(sc)
    for i in range(n):
        sc.SetInstanceTransform(i, xf1[i])
    sc.RebuildInstances()
    renderer.update_instances(sc)
    part = renderer.read_tlas(n)
    part4 = renderer.read_tlas4()
    renderer.update_instances(sc, first_node=0)
    full = renderer.read_tlas(n)
    assert same(part, full)
    assert same4(part4, renderer.read_tlas4())

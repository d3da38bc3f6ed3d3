"""lbm_reduce_boxes_noalloc: the allocation-free variant of lbm_reduce_box
(block partials in the library's static device array) against lbm_reduce_box
itself, through the C ABI, on two calls in a row (the second with fewer and
smaller boxes), in every storage precision. FP64 sums in the same block
order: the bits must be equal."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OBST, FLUID, INJ, GHOST = 1, 2, 4, 8


def say(*a):
    print("[case]", *a, flush=True)


def _solvers(hip, prec, sizes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for size in sizes:
        s = hip.LbmSolverHIP(size, precision=prec, inv_tau=1.6, lid_velocity=0.05)
        full = ((0, 0, 0), s.size_h)
        n = int(np.prod(s.size_h))
        w = np.array([1 / 18] * 4 + [1 / 36] * 12 + [1 / 18] * 2 + [1 / 3])
        d = w[:, None] + rng.uniform(-3e-3, 3e-3, (19, n))
        s.setDensityDistributions(d.astype(s.dtype).ravel(), *full)
        s.setFlags(rng.choice([OBST, FLUID, FLUID, INJ, GHOST], n).astype(np.int32), *full)
        s.simulationStepAlpha()
        s.simulationStepBeta()
        out.append(s)
    return out


def _noalloc(hip, solvers, boxes, buf):
    from cavityflow_cuda_amd._lib import check, lib
    n = len(solvers)
    handles = (ctypes.c_void_p * n)(*[s.handle for s in solvers])
    if boxes is None:
        org = siz = None
    else:
        org = (ctypes.c_int * (3 * n))(*[v for b in boxes for v in b[0]])
        siz = (ctypes.c_int * (3 * n))(*[v for b in boxes for v in b[1]])
    check(lib().lbm_reduce_boxes_noalloc(handles, org, siz, n, buf.data_ptr(), None))
    solvers[0].sync()
    return buf.download(np.float64, 5 * n).reshape(n, 5)


@pytest.mark.parametrize("prec", ["f32", "f64", "f16s", "bf16s", "f16c"])
def test_noalloc_reduce_equals_reduce_box(hip, prec):
    from hiprt import DeviceBuffer
    say("noalloc reduce", prec)
    big = _solvers(hip, prec, [(40, 33, 21), (17, 9, 30), (64, 64, 64), (1, 1, 1)], 7)
    buf = DeviceBuffer(5 * 8 * 16)
    boxes = [((0, 0, 0), big[0].size_h), ((1, 1, 1), (17, 9, 30)), ((3, 5, 7), (50, 40, 30)),
             ((0, 1, 2), (3, 2, 1))]
    got = _noalloc(hip, big, boxes, buf)
    for m, (s, b) in enumerate(zip(big, boxes)):
        want = s.reduce_box(*b)
        assert got[m].tobytes() == np.asarray(want, np.float64).tobytes(), (prec, m)
        assert got[m][4] == np.prod(b[1])
    # second call: fewer, smaller boxes (whole domains); the array keeps the
    # first call's partials, which must not leak into these sums
    small = big[1::2]
    got2 = _noalloc(hip, small, None, buf)
    for m, s in enumerate(small):
        want = s.reduce_box((0, 0, 0), s.size_h)
        assert got2[m].tobytes() == np.asarray(want, np.float64).tobytes(), (prec, m)
    buf.free()
    for s in big:
        s.close()


def test_noalloc_reduce_rejects_bad_arguments(hip):
    from cavityflow_cuda_amd._lib import LbmError, NOALLOC_MAX_BOXES, check, lib
    from hiprt import DeviceBuffer
    say("noalloc reduce errors")
    a = hip.LbmSolverHIP((8, 8, 8), precision="f32")
    b = hip.LbmSolverHIP((8, 8, 8), precision="f64")
    buf = DeviceBuffer(5 * 8 * (NOALLOC_MAX_BOXES + 1))
    many = (ctypes.c_void_p * (NOALLOC_MAX_BOXES + 1))(*([a.handle] * (NOALLOC_MAX_BOXES + 1)))
    with pytest.raises(LbmError, match="count 17 outside 1..16"):
        check(lib().lbm_reduce_boxes_noalloc(many, None, None, NOALLOC_MAX_BOXES + 1,
                                             buf.data_ptr(), None))
    with pytest.raises(LbmError):
        check(lib().lbm_reduce_boxes_noalloc(many, None, None, 0, buf.data_ptr(), None))
    mixed = (ctypes.c_void_p * 2)(a.handle, b.handle)
    with pytest.raises(LbmError, match="precision"):
        check(lib().lbm_reduce_boxes_noalloc(mixed, None, None, 2, buf.data_ptr(), None))
    bad = (ctypes.c_int * 3)(0, 0, 5)
    one = (ctypes.c_void_p * 1)(a.handle)
    size = (ctypes.c_int * 3)(10, 10, 10)
    with pytest.raises(LbmError):
        check(lib().lbm_reduce_boxes_noalloc(one, bad, size, 1, buf.data_ptr(), None))
    # the full count is accepted
    full = (ctypes.c_void_p * NOALLOC_MAX_BOXES)(*([a.handle] * NOALLOC_MAX_BOXES))
    check(lib().lbm_reduce_boxes_noalloc(full, None, None, NOALLOC_MAX_BOXES, buf.data_ptr(),
                                         None))
    a.sync()
    got = buf.download(np.float64, 5 * NOALLOC_MAX_BOXES).reshape(NOALLOC_MAX_BOXES, 5)
    want = np.asarray(a.reduce_box((0, 0, 0), a.size_h), np.float64)
    assert all(r.tobytes() == want.tobytes() for r in got)
    buf.free()
    a.close()
    b.close()

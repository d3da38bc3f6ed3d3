"""Batched cavities: every member is bit-identical to the same cavity stepped
alone, for every launch shape, storage precision and with graph replay."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [((17, 9, 12), 0.05, 100.0), ((1, 1, 1), 0.1, 10.0), ((2, 1, 3), 0.1, 20.0), ((40, 33, 21), 0.08, 300.0),
         ((5, 64, 7), 0.03, 60.0), ((30, 30, 30), 0.1, 400.0)]


def specs(hip, precision):
    from cavityflow_cuda_amd.batch import CavitySpec
    out = []
    for size, u, re in SPECS:
        out.append(CavitySpec(size, u, re, hip.cavity_parameters(size, u, re, precision)))
    return out


def alone(hip, sp, precision, steps):
    from cavityflow_cuda_amd.batch import solve_alone
    m, _ = solve_alone(sp, steps, precision)
    return m


TUNINGS = [dict(block=256, cells_per_thread=1, waves_per_cu=0, options=(1, 2)),
           dict(block=512, cells_per_thread=2, waves_per_cu=4, options=(0, 0)),
           dict(block=256, cells_per_thread=4, waves_per_cu=0, options=(3, 3)),
           dict(block=512, cells_per_thread=1, waves_per_cu=16, options=(2, 1)),
           dict(block=256, cells_per_thread=1, waves_per_cu=0, options=(9, 8)),
           dict(block=512, cells_per_thread=2, waves_per_cu=0, options=(4, 5)),
           dict(block=256, cells_per_thread=2, waves_per_cu=0, options=(17, 16)),
           dict(block=256, cells_per_thread=1, waves_per_cu=0, options=(9, 16), group_mb=1),
           dict(block=512, cells_per_thread=2, waves_per_cu=2, options=(9, 16), group_mb=2),
           dict(block=256, cells_per_thread=1, waves_per_cu=0, options=(0, 0), pairs=1),
           dict(block=256, cells_per_thread=1, waves_per_cu=0, options=(2, 0), pairs=2),
           dict(block=256, cells_per_thread=1, waves_per_cu=0, options=(1, 0), pairs=3),
           dict(block=256, cells_per_thread=1, waves_per_cu=0, options=(3, 0), pairs=4),
           dict(block=1024, cells_per_thread=1, waves_per_cu=0, options=(1, 0)),
           dict(block=1024, cells_per_thread=2, waves_per_cu=2, options=(9, 16)),
           dict(block=1024, cells_per_thread=4, waves_per_cu=1, options=(16, 8), group_mb=1)]


@pytest.mark.parametrize("precision", ["f32", "f64", "f16s", "bf16s", "f16c"])
@pytest.mark.parametrize("tuning", range(len(TUNINGS)))
def test_batch_members_equal_solo_runs(hip, precision, tuning):
    from cavityflow_cuda_amd.batch import CavityBatch
    print("[case] batch", precision, TUNINGS[tuning], flush=True)
    steps = 40
    if precision == "f64" and TUNINGS[tuning].get("pairs"):
        pytest.skip("the pair pass has no FP64 variant (covered by the rejection test)")
    if precision == "f64" and TUNINGS[tuning]["block"] == 1024:
        pytest.skip("FP64 batches take 256 or 512 threads (covered by the rejection test)")
    sp = specs(hip, precision)
    b = CavityBatch(sp, precision)
    b.set_tuning(**TUNINGS[tuning])
    b.run(steps, chunk=8)
    b.sync()
    for i, s in enumerate(sp):
        want = alone(hip, s, precision, steps)
        assert b.moments(i).tobytes() == want.tobytes(), (precision, i)
    b.close()


@pytest.mark.parametrize("group_mb", [0, 1])
def test_graph_replay_equals_direct_launches(hip, group_mb):
    """Graph replays (several, with host work between) equal direct launches,
    on the default stream and on a stream (lbm_batch_launch), with and without
    member groups."""
    from hiprt import stream_create, stream_destroy
    from cavityflow_cuda_amd.batch import CavityBatch
    print("[case] batch graph replay, group_mb", group_mb, flush=True)
    st = stream_create()
    sp = specs(hip, "f32")
    a = CavityBatch(sp, "f32")
    b = CavityBatch(sp, "f32")
    c = CavityBatch(sp, "f32")
    for x in (a, b, c):
        x.set_tuning(group_mb=group_mb)
    for _ in range(4):                    # replay the same graph several times,
        a.run(20, stream=st, chunk=10)    # with host syncs and other work between
        a.sync(st)
        b.run(20, chunk=10)
        b.sync()
        c.run(20, stream=st, chunk=10, graph=False)
        c.sync(st)
        for i in range(len(sp)):
            assert a.moments(i).tobytes() == b.moments(i).tobytes()
            assert c.moments(i).tobytes() == b.moments(i).tobytes()
    a.close()
    b.close()
    c.close()
    stream_destroy(st)


@pytest.mark.parametrize("on_stream", [False, True])
def test_autotune_then_reset_restores_initial_state(hip, on_stream):
    """Autotune (its trials are graph-captured batch runs on a stream, direct
    launches on the default stream), then reset: the batch continues exactly
    like a fresh one."""
    from cavityflow_cuda_amd.batch import CavityBatch
    from hiprt import stream_create, stream_destroy
    print("[case] batch autotune, stream" if on_stream else "[case] batch autotune", flush=True)
    st = stream_create() if on_stream else None
    sp = specs(hip, "f32")
    fresh = CavityBatch(sp, "f32")
    b = CavityBatch(sp, "f32")
    t = b.autotune(trial_steps=2, stream=st)
    assert t["block"] in (256, 512, 1024) and t["cells_per_thread"] in (1, 2, 4)
    for i in range(len(sp)):
        assert b.moments(i).tobytes() == fresh.moments(i).tobytes()
    b.run(40, stream=st)
    fresh.run(40)
    b.sync(st)
    fresh.sync()
    for i in range(len(sp)):
        assert b.moments(i).tobytes() == fresh.moments(i).tobytes()
    b.close()
    fresh.close()
    if st is not None:
        stream_destroy(st)


@pytest.mark.parametrize("precision", ["f32", "f16s"])
def test_batch_residual_matches_solver_residual(hip, precision):
    """The batch residual of every member carries the same bits as that
    member's own lbm_residual (the same block decomposition and final order),
    so a served solve stops at the same step as it would alone."""
    from cavityflow_cuda_amd.batch import CavityBatch
    print("[case] batch residual", precision, flush=True)
    sp = specs(hip, precision)
    b = CavityBatch(sp, precision)
    c = CavityBatch(sp, precision)
    r0 = b.residuals()
    assert np.isinf(r0).all()
    for s in c.solvers:
        assert np.isinf(s.residual(0))
    for _ in range(3):
        b.run(100)
        c.run(100)
        r1 = b.residuals()
        r2 = np.array([s.residual(0) for s in c.solvers])
        assert r1.tobytes() == r2.tobytes()
    assert np.isfinite(r1).any()
    b.close()
    c.close()


def test_pair_pass_rejections(hip):
    from cavityflow_cuda_amd import LbmError, GHOST_LAYER, LbmSolverHIP
    from cavityflow_cuda_amd.batch import CavityBatch
    from cavityflow_cuda_amd import _lib
    import ctypes
    print("[case] batch pair rejections", flush=True)
    b = CavityBatch(specs(hip, "f64"), "f64")
    with pytest.raises(LbmError):
        b.set_tuning(pairs=1)
    b.close()
    # a GHOST_LAYER face (a subdomain of a split cavity) is not a whole domain
    s = LbmSolverHIP((8, 8, 8), (GHOST_LAYER,) + (_lib.OBSTACLE,) * 5, "f32", 1.5, 0.05)
    h = ctypes.c_void_p()
    arr = (ctypes.c_void_p * 1)(s.handle.value)
    from cavityflow_cuda_amd._lib import check, lib
    check(lib().lbm_batch_create(arr, 1, ctypes.byref(h)))
    t = _lib.lbm_batch_tuning(256, 1, 0, (ctypes.c_int * 2)(0, 0), 1)
    assert lib().lbm_batch_set_tuning(h, ctypes.byref(t)) == _lib.LBM_ERR_UNSUPPORTED
    t.pairs = 0
    check(lib().lbm_batch_set_tuning(h, ctypes.byref(t)))
    # member groups and the pair pass do not combine; group_mb >= 0
    t.pairs, t.group_mb = 1, 64
    assert lib().lbm_batch_set_tuning(h, ctypes.byref(t)) == _lib.LBM_ERR_ARG
    t.pairs, t.group_mb = 0, -1
    assert lib().lbm_batch_set_tuning(h, ctypes.byref(t)) == _lib.LBM_ERR_ARG
    # the pair pass needs multiples of 4 half-steps from phase 0
    b2 = CavityBatch(specs(hip, "f32"), "f32")
    b2.set_tuning(pairs=1)
    assert lib().lbm_batch_run(b2._h, 0, 6, None) == _lib.LBM_ERR_ARG
    assert lib().lbm_batch_run(b2._h, 1, 8, None) == _lib.LBM_ERR_ARG
    b2.close()
    lib().lbm_batch_destroy(h)
    s.close()


def test_workgroup_shape_rejections(hip):
    """1024-thread workgroups are refused for FP64 batches (the kernels would
    spill at 128 VGPRs per thread) and any other block size is an argument
    error; a refused tuning leaves the batch stepping as before."""
    from cavityflow_cuda_amd import LbmError, _lib
    from cavityflow_cuda_amd.batch import CavityBatch
    from cavityflow_cuda_amd._lib import lib
    import ctypes
    print("[case] batch workgroup shape rejections", flush=True)
    sp = specs(hip, "f64")
    b = CavityBatch(sp, "f64")
    b.set_tuning(block=512, cells_per_thread=2)
    before = b.tuning()
    t = _lib.lbm_batch_tuning(1024, 1, 0, (ctypes.c_int * 2)(0, 0), 0, 0)
    assert lib().lbm_batch_set_tuning(b._h, ctypes.byref(t)) == _lib.LBM_ERR_UNSUPPORTED
    for blk in (128, 768, 2048):
        t.block = blk
        assert lib().lbm_batch_set_tuning(b._h, ctypes.byref(t)) == _lib.LBM_ERR_ARG
    assert b.tuning() == before
    with pytest.raises(LbmError):
        b.set_tuning(block=1024)
    b.run(8, chunk=4)
    b.sync()
    want = alone(hip, sp[0], "f64", 8)
    assert b.moments(0).tobytes() == want.tobytes()
    b.close()

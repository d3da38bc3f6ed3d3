"""Launch variants (nontemporal accesses, XCD renumbering, occupancy caps,
64-bit indexing, 512-thread workgroups of two cells per thread, LDS-staged
beta) change speed only: alpha and beta of a random state with
every flag, gravity and a lid give the same bits under every option
combination. The 64-bit indexing variant is the code path of domains whose
planes exceed 4 GiB, checked here at a small size."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZE = (37, 21, 13)            # odd sizes: partial workgroups and wrap blocks


def state(hip, precision, rng, size=SIZE):
    from cavityflow_cuda_amd import _lib
    s = hip.LbmSolverHIP(size, (_lib.OBSTACLE, _lib.GHOST_LAYER, _lib.OBSTACLE,
                                _lib.VELOCITY_INJECTION, _lib.GHOST_LAYER, _lib.OBSTACLE),
                         precision, inv_tau=1.37, lid_velocity=0.04,
                         gravitation=(1e-4, -2e-4, 5e-5), store_densities=True,
                         store_velocities=True)
    nh = tuple(v + 2 for v in size)
    full = ((0, 0, 0), nh)
    cells = int(np.prod(nh))
    flags = rng.choice(np.array([1, 2, 4, 8], np.int32), cells, p=[0.15, 0.7, 0.1, 0.05])
    s.setFlags(flags, *full)
    w = np.array([1 / 18] * 6 + [1 / 36] * 12 + [1 / 3])
    dd = (np.repeat(w, cells) * (1 + 0.05 * rng.standard_normal(19 * cells)))
    s.setDensityDistributions(dd.astype(s.dtype), *full)
    return s, full


def run(s, full, alpha_opt, beta_opt):
    from cavityflow_cuda_amd._lib import check, lib
    check(lib().lbm_set_launch_options(s.handle, 0, alpha_opt))
    check(lib().lbm_set_launch_options(s.handle, 1, beta_opt))
    s.simulationStepAlpha()
    a = s.getDensityDistributions(*full).copy()
    s.simulationStepBeta()
    b = s.getDensityDistributions(*full).copy()
    return a, b, s.getVelocities().copy(), s.getDensities().copy()


@pytest.mark.parametrize("precision", ["f32", "f64", "f16s", "bf16s", "f16c"])
def test_every_launch_option_gives_the_same_bits(hip, precision):
    print("[case] launch options", precision, flush=True)
    want = None
    for opt in range(512):
        s, full = state(hip, precision, np.random.default_rng(11))
        got = run(s, full, opt, (opt * 5) % 512)
        s.close()
        if want is None:
            want = got
            continue
        for x, y in zip(got, want):
            assert x.tobytes() == y.tobytes(), opt


# per-precision defaults (alpha, beta) documented in include/lbm_abi.h
DEFAULTS = {"f32": (1 | 16, 16), "f64": (1 | 16, 128 | 16), "f16s": (0, 0), "bf16s": (0, 0),
            "f16c": (0, 0)}


@pytest.mark.parametrize("precision", sorted(DEFAULTS))
def test_default_options_are_valid_and_settable(hip, precision):
    from cavityflow_cuda_amd import LbmError
    from cavityflow_cuda_amd._lib import check, lib
    import ctypes
    s = hip.LbmSolverHIP((8, 8, 8), precision=precision)
    v = ctypes.c_int()
    for k in (0, 1):
        check(lib().lbm_get_launch_options(s.handle, k, ctypes.byref(v)))
        assert v.value == DEFAULTS[precision][k]
    with pytest.raises(LbmError):
        check(lib().lbm_set_launch_options(s.handle, 0, 512))
    # with a step pack FP32 alpha keeps 3 waves per SIMD; options the caller
    # set are kept; without the pack the first defaults return
    want_pack = {"f32": (1 | 8, 16)}.get(precision, DEFAULTS[precision])
    buf = ctypes.c_void_p()
    check(lib().lbm_device_alloc(0, s.face_elements(0) * 8, ctypes.byref(buf)))
    check(lib().lbm_set_step_pack(s.handle, buf, None))
    for k in (0, 1):
        check(lib().lbm_get_launch_options(s.handle, k, ctypes.byref(v)))
        assert v.value == want_pack[k]
    check(lib().lbm_set_launch_options(s.handle, 1, 2))
    check(lib().lbm_set_step_pack(s.handle, None, None))
    for k, want in ((0, DEFAULTS[precision][0]), (1, 2)):
        check(lib().lbm_get_launch_options(s.handle, k, ctypes.byref(v)))
        assert v.value == want
    s.close()
    check(lib().lbm_device_free(0, buf))


@pytest.mark.parametrize("precision,small,large", [
    ("f32", (1 | 16, 16), (64 | 1 | 16, 64 | 16)), ("f16s", (0, 0), (0, 256)),
    ("bf16s", (0, 0), (0, 256)), ("f16c", (0, 0), (0, 0))])
def test_large_domains_default_to_bigger_workgroups(hip, precision, small, large):
    """From 6 Mi cells (halo included) FP32 defaults to 512-thread
    workgroups and F16S / BF16S beta to 1024-thread workgroups; with a step
    pack FP32 beta there takes 1024-thread workgroups at 4 waves per SIMD,
    and with the fused unpack FP32 alpha 256-thread workgroups at 3 waves."""
    import ctypes
    from cavityflow_cuda_amd._lib import check, lib
    v = ctypes.c_int()
    for size, want in (((182, 182, 182), small), ((183, 183, 183), large)):
        s = hip.LbmSolverHIP(size, precision=precision, store_densities=False,
                             store_velocities=False)
        for k in (0, 1):
            check(lib().lbm_get_launch_options(s.handle, k, ctypes.byref(v)))
            assert v.value == want[k], (size, k)
        if precision in ("f16s", "bf16s") and size[0] == 183:
            # with the fused unpack, beta in 256-thread workgroups
            buf = ctypes.c_void_p()
            check(lib().lbm_device_alloc(0, s.face_elements(0) * 4, ctypes.byref(buf)))
            check(lib().lbm_set_step_unpack(s.handle, buf, buf))
            check(lib().lbm_get_launch_options(s.handle, 1, ctypes.byref(v)))
            assert v.value == 0
            check(lib().lbm_set_step_unpack(s.handle, None, None))
            check(lib().lbm_get_launch_options(s.handle, 1, ctypes.byref(v)))
            assert v.value == 256
            check(lib().lbm_device_free(0, buf))
        if precision == "f32" and size[0] == 183:
            buf = ctypes.c_void_p()
            check(lib().lbm_device_alloc(0, s.face_elements(0) * 4, ctypes.byref(buf)))
            check(lib().lbm_set_step_pack(s.handle, buf, buf))
            for k, w in ((0, 64 | 1 | 16), (1, 256 | 4)):
                check(lib().lbm_get_launch_options(s.handle, k, ctypes.byref(v)))
                assert v.value == w, ("pack", k)
            # with the fused unpack as well, alpha takes 256-thread workgroups
            # at 3 waves per SIMD; both off, the first defaults return
            check(lib().lbm_set_step_unpack(s.handle, buf, buf))
            for k, w in ((0, 1 | 8), (1, 256 | 4)):
                check(lib().lbm_get_launch_options(s.handle, k, ctypes.byref(v)))
                assert v.value == w, ("unpack", k)
            check(lib().lbm_set_step_unpack(s.handle, None, None))
            check(lib().lbm_set_step_pack(s.handle, None, None))
            for k in (0, 1):
                check(lib().lbm_get_launch_options(s.handle, k, ctypes.byref(v)))
                assert v.value == large[k], ("off", k)
            check(lib().lbm_device_free(0, buf))
        s.close()


@pytest.mark.parametrize("precision", ["f32", "f16s"])
def test_box_steps_same_bits_under_wide_workgroups(hip, precision):
    """Sub-box steps (the exchange's inner and boundary boxes) map workgroup
    cells through the box: 512-thread, two-cells-per-thread and 1024-thread
    workgroups give the bits of the 256-thread ones, partial last workgroup
    included."""
    from cavityflow_cuda_amd._lib import (LAUNCH_BIG, LAUNCH_NONTEMPORAL, LAUNCH_WIDE,
                                          LAUNCH_WAVES2, check, lib)
    print("[case] wide box steps", precision, flush=True)
    boxes = [((1, 1, 1), SIZE), ((2, 3, 1), (33, 15, 11)), ((1, 1, 1), (37, 21, 1)),
             ((36, 2, 2), (2, 19, 11))]
    want = None
    for opts in (0, LAUNCH_WIDE, LAUNCH_WIDE | LAUNCH_NONTEMPORAL | LAUNCH_WAVES2, LAUNCH_BIG,
                 LAUNCH_BIG | LAUNCH_NONTEMPORAL | LAUNCH_WAVES2):
        s, full = state(hip, precision, np.random.default_rng(5))
        check(lib().lbm_set_launch_options(s.handle, 0, opts))
        check(lib().lbm_set_launch_options(s.handle, 1, opts & ~LAUNCH_NONTEMPORAL))
        got = []
        for o, z in boxes:
            s.simulationStepAlpha(o, z)
            got.append(s.getDensityDistributions(*full).copy())
            s.simulationStepBeta(o, z)
            got.append(s.getDensityDistributions(*full).copy())
        s.close()
        if want is None:
            want = got
            continue
        for i, (x, y) in enumerate(zip(got, want)):
            assert x.tobytes() == y.tobytes(), (opts, i)


@pytest.mark.parametrize("prec", ["f32", "f16s", "f64"])
def test_pattern_probes_keep_or_reset_the_state(hip, prec):
    """lbm_probe_pattern (the bandwidth probes of bench.py --workload
    ceiling), on the solver's own elements (4, 2 or 8 bytes): the
    read-modify-write and read patterns leave every population as it was
    under every workgroup shape, the write pattern stores the rest state;
    bad arguments and option combinations are rejected."""
    import ctypes
    from cavityflow_cuda_amd import LbmError
    from cavityflow_cuda_amd._lib import check, lib
    print(f"[case] pattern probes {prec}", flush=True)
    s, full = state(hip, prec, np.random.default_rng(3))
    before = s.getDensityDistributions(*full).copy()
    nx, ny, _ = s.size_h
    pitch = (nx + 31) // 32 * 32
    span = {0: 0, 1: nx * ny + nx, 2: 0, 4: pitch * ny + pitch, 5: 1, 6: nx * ny + nx}
    vec = (256, 512, 512 | 17, 256 | 4) if prec == "f32" else ()
    shapes = (1024, 2048 | 4, 4096, 8192 | 1024 | 1, 16384 | 2048, 16384 | 17, 4096 | 1024 | 16)
    for mode in (0, 1, 2, 4, 5, 6):
        opts = (0, 1, 16, 17, 8, 4) + (() if mode == 2 else vec + shapes)
        for opt in opts:
            n = ctypes.c_uint64()
            check(lib().lbm_probe_pattern(s.handle, mode, opt, ctypes.byref(n), None))
            s.sync()
            per = 4 if opt & 512 else (2 if opt & 256 else 1)
            cells = int(np.prod(s.size_h)) - 2 * span[mode]
            assert n.value == cells // per * per, (mode, opt)
            assert s.getDensityDistributions(*full).tobytes() == before.tobytes(), (mode, opt)
    check(lib().lbm_probe_pattern(s.handle, 3, 1, None, None))
    s.sync()
    dt = np.float64 if prec == "f64" else np.float32
    w = np.array([1 / 18] * 4 + [1 / 36] * 12 + [1 / 18] * 2 + [1 / 3], dt)
    got = s.getDensityDistributions(*full).reshape(19, -1)
    assert np.array_equal(got, np.repeat(w[:, None], got.shape[1], axis=1))
    bad = [(8, 0), (-1, 0), (0, 2), (0, 32768), (0, 256 | 4096), (0, 1024 | 2048),
           (0, 4096 | 8192), (2, 4096), (3, 1024), (0, 2048 | 16), (0, 2048 | 8),
           (7, 256), (7, 512 | 1), (7, 16384 | 2048), (7, 2048 | 16)]
    if prec != "f32":
        bad += [(0, 256), (7, 0)]
    for mode, opt in bad:
        with pytest.raises(LbmError):
            check(lib().lbm_probe_pattern(s.handle, mode, opt, None, None))
    s.close()


def test_lane_shifted_probe_keeps_the_state(hip):
    """The lane-shifted FP32 probe (padded-row y / z shifts, x shifts moved
    across lanes with DPP wave shifts, one extra element at either end of a
    wave's run of chunks) stores every population back where it was loaded,
    for runs of 1-8 chunks per wave, including the short runs of the last
    workgroup (175 chunks here: 7 past a multiple of 8, 15 past 16 and 32) --
    a shift in the wrong direction or a lost end element changes the array."""
    import ctypes
    from cavityflow_cuda_amd._lib import check, lib
    print("[case] lane-shifted probe", flush=True)
    for size in ((37, 23, 13), SIZE):
        s, full = state(hip, "f32", np.random.default_rng(5), size)
        before = s.getDensityDistributions(*full).copy()
        nx, ny, nz = s.size_h
        pitch = (nx + 31) // 32 * 32
        span = pitch * ny + pitch + 1
        first = (span + 63) // 64 * 64
        cells = (nx * ny * nz - span - first) // 64 * 64
        for opt in (0, 1, 16, 4, 1024, 2048 | 4, 4096, 8192 | 1024 | 1, 16384 | 17, 16384 | 1024,
                    8192 | 2048, 4096 | 2048 | 1, 8192, 16384):
            n = ctypes.c_uint64()
            check(lib().lbm_probe_pattern(s.handle, 7, opt, ctypes.byref(n), None))
            s.sync()
            assert n.value == cells, (size, opt)
            assert s.getDensityDistributions(*full).tobytes() == before.tobytes(), (size, opt)
        s.close()

"""Fused x-face packing inside the step kernels equals the separate face pack
(the 5 populations each phase moves: after alpha those pointing back into the
solver, after beta those crossing the face), writes nothing else into the
buffers and leaves the step results unchanged."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OBST, FLUID, INJ, GHOST = 1, 2, 4, 8
# c_x = -1 / c_x = +1: the planes a LEFT / RIGHT face buffer carries after
# beta, and a RIGHT / LEFT one after alpha
MINUS_X, PLUS_X = [1, 5, 7, 9, 11], [0, 4, 6, 8, 10]
SENT = {(0, 0): PLUS_X, (1, 0): MINUS_X, (0, 1): MINUS_X, (1, 1): PLUS_X}   # (side, phase)


@pytest.mark.parametrize("prec,dtype", [("f32", np.float32), ("f64", np.float64),
                                        ("f16s", np.float32)])
@pytest.mark.parametrize("size", [(9, 6, 7), (1, 4, 5), (13, 3, 2), (70, 5, 4)])
def test_fused_x_pack_matches_face_pack(hip, oracle_cpu, prec, dtype, size):
    from hiprt import DeviceBuffer
    from cavityflow_cuda_amd import _lib
    print("[case] fused pack", prec, size, flush=True)
    L = _lib.lib()
    rng = np.random.default_rng(sum(size) + len(prec))
    s = hip.LbmSolverHIP(size, (OBST, GHOST, GHOST, OBST, GHOST, OBST), prec, 1.4, 0.06,
                         (0.001, -0.002, 0.0))
    o = oracle_cpu.OracleSolver(size, (OBST,) * 6, dtype=np.float64 if prec == "f64" else np.float32,
                                inv_tau=1.4, lid_velocity=0.06, gravitation=(0.001, -0.002, 0.0))
    full = ((0, 0, 0), s.size_h)
    if prec == "f16s":
        start = (o.dd.reshape(19, -1) + (rng.random((19, o.n)) - 0.5) * 2e-3).astype(np.float32).ravel()
    else:
        start = (rng.random(o.dd.size) * 0.08 + 0.01).astype(o.dtype)
    flags = rng.choice([OBST, FLUID, FLUID, FLUID, INJ, GHOST], size=o.n).astype(np.int32)
    s.setDensityDistributions(start, *full)
    s.setFlags(flags, *full)
    ne = s.face_elements(0)
    esz = 8 if prec == "f64" else 4
    fused = [DeviceBuffer(ne * esz), DeviceBuffer(ne * esz)]
    plain = DeviceBuffer(ne * esz)
    ref = hip.LbmSolverHIP(size, (OBST, GHOST, GHOST, OBST, GHOST, OBST), prec, 1.4, 0.06,
                           (0.001, -0.002, 0.0))
    ref.setDensityDistributions(start, *full)
    ref.setFlags(flags, *full)
    for phase in (0, 1):
        for b in fused:
            b.upload(np.full(ne, -7.0, dtype))
        _lib.check(L.lbm_set_step_pack(s.handle, ctypes.c_void_p(fused[0].data_ptr()),
                                       ctypes.c_void_p(fused[1].data_ptr())))
        (s.simulationStepBeta if phase else s.simulationStepAlpha)()
        (ref.simulationStepBeta if phase else ref.simulationStepAlpha)()
        _lib.check(L.lbm_set_step_pack(s.handle, None, None))
        s.sync()
        # the step result itself is unchanged by packing
        assert s.getDensityDistributions(*full).tobytes() == ref.getDensityDistributions(*full).tobytes()
        for side in (0, 1):
            s.pack_face(side, phase, plain)
            s.sync()
            want = plain.download(dtype, ne).reshape(19, -1)
            got = fused[side].download(dtype, ne).reshape(19, -1)
            for l in range(19):
                if l in SENT[(side, phase)]:
                    assert got[l].tobytes() == want[l].tobytes(), (phase, side, l)
                else:
                    assert (got[l] == -7.0).all(), (phase, side, l)
    for b in fused + [plain]:
        b.free()
    s.close()
    ref.close()

"""Fused x-face unpacking inside the step kernels (lbm_set_step_unpack): a
step that takes the x faces' received populations from the buffers equals
lbm_unpack_face on both x faces followed by the step, bit for bit (the whole
population array with its halo, rho, u and the F16C range flag), for every
storage format, both phases, the launch variants, box steps and cells of
every flag on the send planes; and a cavity whose x faces receive their own
fused-pack buffers (a wall reflected through the exchange path) runs exactly
as the plain cavity."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OBST, FLUID, INJ, GHOST = 1, 2, 4, 8
LEFT, RIGHT = 0, 1
W = np.array([1 / 18] * 4 + [1 / 36] * 12 + [1 / 18] * 2 + [1 / 3])
# alpha / beta launch options: default, base 256 threads, WIDE + nontemporal +
# 2 waves, BIG + 4 waves, XCD map, LDS-staged beta, 64-bit element indexing
OPTIONS = [None, (0, 0), (64 | 1 | 16, 64 | 16), (256 | 4, 256 | 4), (2 | 8, 2 | 4), (1, 128 | 4),
           (32, 32 | 64)]


def dtype_of(prec):
    return np.float64 if prec == "f64" else np.float32


def make(hip, size, prec, flags, start, options):
    from cavityflow_cuda_amd import _lib
    s = hip.LbmSolverHIP(size, (GHOST, GHOST, OBST, INJ, GHOST, OBST), prec, 1.3, 0.07,
                         (0.002, -0.001, 0.0005))
    full = ((0, 0, 0), s.size_h)
    s.setDensityDistributions(start, *full)
    s.setFlags(flags, *full)
    if options is not None:
        for k in (0, 1):
            _lib.check(_lib.lib().lbm_set_launch_options(s.handle, k, options[k]))
    return s, full


def state(s, full, prec):
    dd = s.getDensityDistributions(*full)
    u = s.getVelocities(*full)
    rho = s.getDensities(*full)
    flagged = bool(np.isnan(s.getMoments(0, (1, 1, 1), (1, 1, 1))).any())
    return dd.tobytes(), u.tobytes(), rho.tobytes(), flagged


@pytest.mark.parametrize("prec", ["f32", "f64", "f16s", "bf16s", "f16c"])
@pytest.mark.parametrize("size", [(9, 6, 7), (2, 4, 5), (13, 3, 2), (70, 5, 4)])
def test_fused_unpack_equals_unpack_then_step(hip, prec, size):
    from hiprt import DeviceBuffer
    print("[case] fused unpack", prec, size, flush=True)
    rng = np.random.default_rng(sum(size) * 7 + len(prec))
    dt = dtype_of(prec)
    nh = tuple(v + 2 for v in size)
    cells = int(np.prod(nh))
    start = (W[:, None] * (1 + 0.1 * rng.standard_normal((19, cells)))).astype(dt).ravel()
    flags = rng.choice([OBST, FLUID, FLUID, FLUID, INJ, GHOST], size=cells).astype(np.int32)
    for oi, options in enumerate(OPTIONS):
        a, full = make(hip, size, prec, flags, start, options)
        b, _ = make(hip, size, prec, flags, start, options)
        ne = a.face_elements(LEFT)
        recv = [DeviceBuffer(ne * np.dtype(dt).itemsize) for _ in range(2)]
        for step in range(4):
            kind = step & 1                # 0 alpha, 1 beta
            phase = 1 - kind               # the exchange phase each consumes
            for r in recv:
                vals = (np.repeat(W, ne // 19) * (1 + 0.1 * rng.standard_normal(ne))).astype(dt)
                if prec == "f16c" and step == 3 and oi == 0:
                    vals[7 * (ne // 19) + 1] = 0.9   # out of F16C range: both solvers flag it
                r.upload(vals)
            # one side only on the first steps, both afterwards
            sides = (recv[0], None) if step == 0 else ((None, recv[1]) if step == 1 else recv)
            a.set_step_unpack(*sides)
            for side, r in enumerate(sides):
                if r is not None:
                    b.unpack_face(side, phase, r)
            (a.simulationStepBeta if kind else a.simulationStepAlpha)()
            (b.simulationStepBeta if kind else b.simulationStepAlpha)()
            a.set_step_unpack(None, None)
            a.sync()
            b.sync()
            assert state(a, full, prec) == state(b, full, prec), (options, step)
        for r in recv:
            r.free()
        a.close()
        b.close()


@pytest.mark.parametrize("prec", ["f32", "f16s"])
def test_fused_unpack_box_steps_and_pack(hip, prec):
    """Box steps over slabs (the boundary-slabs-first order) with the fused
    pack and unpack on equal whole-domain unpack-then-step with the pack, and
    the pack buffers agree."""
    from hiprt import DeviceBuffer
    print("[case] fused unpack, box steps", prec, flush=True)
    size = (11, 7, 6)
    rng = np.random.default_rng(91)
    dt = dtype_of(prec)
    nh = tuple(v + 2 for v in size)
    cells = int(np.prod(nh))
    start = (W[:, None] * (1 + 0.1 * rng.standard_normal((19, cells)))).astype(dt).ravel()
    flags = rng.choice([OBST, FLUID, FLUID, FLUID, INJ, GHOST], size=cells).astype(np.int32)
    a, full = make(hip, size, prec, flags, start, None)
    b, _ = make(hip, size, prec, flags, start, None)
    ne = a.face_elements(LEFT)
    nb = ne * np.dtype(dt).itemsize
    recv = [DeviceBuffer(nb) for _ in range(2)]
    pack = [DeviceBuffer(nb) for _ in range(4)]
    for p in pack:
        p.upload(np.full(ne, -7.0, dt))
    a.set_step_pack(pack[0], pack[1])
    b.set_step_pack(pack[2], pack[3])
    a.set_step_unpack(recv[0], recv[1])
    # slabs in x (2 cells at each end), then the rest
    nx = nh[0]
    boxes = [((0, 0, 0), (2, nh[1], nh[2])), ((nx - 2, 0, 0), (2, nh[1], nh[2])),
             ((2, 0, 0), (nx - 4, nh[1], nh[2]))]
    for step in range(4):
        kind = step & 1                    # 0 alpha, 1 beta
        for r in recv:
            r.upload((np.repeat(W, ne // 19) * (1 + 0.1 * rng.standard_normal(ne))).astype(dt))
        for side in (LEFT, RIGHT):
            b.unpack_face(side, 1 - kind, recv[side])
        for o, z in boxes:
            (a.simulationStepBeta if kind else a.simulationStepAlpha)(o, z)
        (b.simulationStepBeta if kind else b.simulationStepAlpha)()
        a.sync()
        b.sync()
        assert state(a, full, prec) == state(b, full, prec), step
        got = [p.download(dt, ne).tobytes() for p in pack]
        assert got[0] == got[2] and got[1] == got[3], step
    for buf in recv + pack:
        buf.free()
    a.close()
    b.close()


@pytest.mark.parametrize("prec", ["f32", "f16s", "f64"])
@pytest.mark.parametrize("double", [False, True])
def test_cavity_receiving_its_own_x_faces_runs_as_the_plain_cavity(hip, prec, double):
    """The bench's x exchange: each x face packs inside the steps and the
    steps take its received populations from that face's own pack -- one
    buffer, or (double) two that swap roles every alpha + beta pair, packing
    into one while unpacking from the other. The fused pack stores exactly
    what the wall's implicit bounce-back returns, so velocities and the
    interior (FLUID) densities stay identical to the plain cavity."""
    import cavityflow_cuda_amd as cf
    from hiprt import DeviceBuffer
    print("[case] self-exchanged x faces", prec, flush=True)
    size = (20, 14, 12)
    dt = dtype_of(prec)
    solvers = []
    for fused in (False, True):
        s = cf.LbmSolverHIP(size, (OBST,) * 6, prec, inv_tau=1.6, lid_velocity=0.08)
        s.setDrivenCavitySzenario()
        solvers.append(s)
    plain, fused = solvers
    ne = fused.face_elements(LEFT)
    bufs = [DeviceBuffer(ne * np.dtype(dt).itemsize) for _ in range(4 if double else 2)]
    sets = [bufs[:2], bufs[2:]] if double else [bufs, bufs]
    for side in (LEFT, RIGHT):
        fused.pack_face(side, 0, sets[1][side])
        fused.pack_face(side, 1, sets[1][side])
    full = ((0, 0, 0), plain.size_h)
    nx, ny, nz = plain.size_h
    for step in range(60):
        j = (step >> 1) & 1
        fused.set_step_pack(*sets[j])
        fused.set_step_unpack(*sets[1 - j])
        for s in solvers:
            (s.simulationStepBeta if step & 1 else s.simulationStepAlpha)()
        if step % 20 == 19:
            for s in solvers:
                s.sync()
            assert plain.getVelocities(*full).tobytes() == fused.getVelocities(*full).tobytes(), step
            # the interior cells (all FLUID); wall cells' densities may differ
            rp = plain.getDensities(*full).reshape(nz, ny, nx)[1:-1, 1:-1, 1:-1]
            rf = fused.getDensities(*full).reshape(nz, ny, nx)[1:-1, 1:-1, 1:-1]
            assert rp.tobytes() == rf.tobytes(), step
    for b in bufs:
        b.free()
    for s in solvers:
        s.close()


def test_fused_unpack_rejects_narrow_x(hip):
    from cavityflow_cuda_amd import LbmError
    from hiprt import DeviceBuffer
    s = hip.LbmSolverHIP((1, 4, 4), (OBST,) * 6, "f32")
    buf = DeviceBuffer(4 * s.face_elements(LEFT))
    with pytest.raises(LbmError):
        s.set_step_unpack(buf, None)
    s.set_step_unpack(None, None)
    buf.free()
    s.close()

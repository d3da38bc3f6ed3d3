"""GPU parity of the HIP A-A kernels against the oracle and the golden vectors.

F32 and F64 are bit-exact by construction (same operation order, no FMA
contraction), so every comparison here is on bytes. Each test prints its case
before touching the GPU.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

OBST, FLUID, INJ, GHOST = 1, 2, 4, 8
PREC = {np.float32: "f32", np.float64: "f64"}


def say(*a):
    print("[case]", *a, flush=True)


def full_box(s):
    return (0, 0, 0), s.size_h


def load_state(g_solver, o_solver):
    """Copy the oracle solver's whole state into the GPU solver."""
    o, z = full_box(g_solver)
    g_solver.setDensityDistributions(o_solver.dd, o, z)
    g_solver.setFlags(o_solver.flags, o, z)
    g_solver.setVelocities(o_solver.u, o, z)
    g_solver.setDensities(o_solver.rho, o, z)


def same_state(g_solver, o_solver, what=""):
    o, z = full_box(g_solver)
    dd = g_solver.getDensityDistributions(o, z)
    u = g_solver.getVelocities(o, z)
    rho = g_solver.getDensities(o, z)
    assert dd.tobytes() == o_solver.dd.tobytes(), f"populations differ {what}"
    assert u.tobytes() == o_solver.u.tobytes(), f"velocities differ {what}"
    assert rho.tobytes() == o_solver.rho.tobytes(), f"densities differ {what}"


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("size,bc", [((6, 5, 7), (OBST,) * 6),
                                     ((9, 4, 6), (OBST, GHOST, OBST, OBST, GHOST, OBST)),
                                     ((3, 3, 3), (GHOST,) * 6)])
def test_init_matches_oracle(hip, oracle_cpu, dtype, size, bc):
    say("init", dtype.__name__, size, bc)
    g = hip.LbmSolverHIP(size, bc, PREC[dtype])
    o = oracle_cpu.OracleSolver(size, bc, dtype=dtype)
    same_state(g, o, "after init")
    assert (g.getFlags(*full_box(g)) == o.flags).all()
    g.close()


@pytest.mark.parametrize("dtype,name", [(np.float32, "step_f32.npz"), (np.float64, "step_f64.npz")])
def test_steps_match_reference_golden(hip, dtype, name):
    g = np.load(os.path.join(GOLDEN, name))
    size = tuple(int(v) for v in g["size"])
    say("golden step", dtype.__name__, size)
    s = hip.LbmSolverHIP(size, (OBST,) * 6, PREC[dtype], float(g["inv_tau"]), float(g["lid"]),
                         tuple(float(v) for v in g["g"]))
    o, z = full_box(s)
    s.setDensityDistributions(g["dd0"], o, z)
    s.setFlags(g["flags"], o, z)
    s.setVelocities(np.zeros(3 * int(np.prod(z)), dtype), o, z)
    s.setDensities(np.zeros(int(np.prod(z)), dtype), o, z)
    for phase, step in (("alpha", s.simulationStepAlpha), ("beta", s.simulationStepBeta)):
        step()
        assert s.getDensityDistributions(o, z).tobytes() == g[f"{phase}_dd"].tobytes(), phase
        assert s.getVelocities(o, z).tobytes() == g[f"{phase}_u"].tobytes(), phase
        assert s.getDensities(o, z).tobytes() == g[f"{phase}_rho"].tobytes(), phase
    bo, bz = tuple(int(v) for v in g["box_origin"]), tuple(int(v) for v in g["box_size"])
    s.simulationStepAlpha(bo, bz)
    s.simulationStepBeta(bo, bz)
    assert s.getDensityDistributions(o, z).tobytes() == g["box_dd"].tobytes()
    assert s.getVelocities(o, z).tobytes() == g["box_u"].tobytes()
    assert s.getDensities(o, z).tobytes() == g["box_rho"].tobytes()
    s.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("size", [(37, 19, 23), (64, 3, 5), (1, 1, 1), (130, 2, 9)])
def test_random_states_match_oracle(hip, oracle_cpu, dtype, size):
    say("random", dtype.__name__, size)
    rng = np.random.default_rng(hash(size) % 2**32)
    kw = dict(inv_tau=1.53, lid_velocity=0.07, gravitation=(0.0009, -0.0013, 0.0004))
    o = oracle_cpu.OracleSolver(size, (OBST,) * 6, dtype=dtype, **kw)
    o.dd[:] = (rng.random(o.dd.size) * 0.08 + 0.01).astype(dtype)
    o.flags[:] = rng.choice([OBST, FLUID, FLUID, FLUID, INJ, GHOST], size=o.n).astype(np.int32)
    s = hip.LbmSolverHIP(size, (OBST,) * 6, PREC[dtype], kw["inv_tau"], kw["lid_velocity"],
                         kw["gravitation"])
    load_state(s, o)
    nxh, nyh, nzh = o.size_h
    boxes = [None, ((1, 0, 1), (max(1, nxh - 2), nyh, max(1, nzh - 1))),
             ((0, nyh - 1, 0), (nxh, 1, nzh)), ((nxh - 1, 0, 0), (1, nyh, nzh))]
    for n in range(8):
        box = boxes[(n // 2) % len(boxes)]
        args = () if box is None else box
        if n & 1:
            o.beta(*args)
            s.simulationStepBeta(*args)
        else:
            o.alpha(*args)
            s.simulationStepAlpha(*args)
        same_state(s, o, f"step {n} box {box}")
    s.close()


def _cavity(hip, size, inv_tau, steps, split=(1, 1, 1)):
    c = hip.CavityController(size, u_lid=0.05, reynolds=100.0, split=split)
    assert np.float32(c.inv_tau) == np.float32(inv_tau)
    for _ in range(steps):
        c.computeNextStep()
    c.sync()
    return c


def test_cavity16_matches_reference_golden(hip):
    say("cavity16 200 steps")
    g = np.load(os.path.join(GOLDEN, "cavity16.npz"))
    c = _cavity(hip, (16, 16, 16), float(g["inv_tau"]), 200)
    assert c.getVelocities().tobytes() == g["u"].tobytes()
    assert c.getDensities().tobytes() == g["rho"].tobytes()
    c.close()


def test_config1_matches_reference_golden(hip):
    """64^3 cells including the wall layer, lid 0.05, Re 100, 1000 steps."""
    say("config1 62^3 interior 1000 steps")
    with open(os.path.join(GOLDEN, "config1.json")) as f:
        g = json.load(f)
    c = _cavity(hip, (62, 62, 62), g["inv_tau"], 1000)
    u = c.getVelocities()
    rho = c.getDensities()
    assert hashlib.sha256(u.tobytes()).hexdigest() == g["sha256_u"]
    assert hashlib.sha256(rho.tobytes()).hexdigest() == g["sha256_rho"]
    mid = np.load(os.path.join(GOLDEN, "config1_midplane.npz"))["u_mid_z31"]
    assert (u[:, 31] == mid).all()
    c.close()


@pytest.mark.parametrize("split", [(2, 1, 1), (1, 2, 1), (1, 1, 2), (2, 2, 2)])
def test_split_equals_monolithic(hip, split):
    """main.cpp:121-303 validation semantic, with on-device face exchange."""
    say("split", split)
    size = (24, 24, 24)
    mono = _cavity(hip, size, hip.cavity_parameters(size, 0.05, 100.0)["inv_tau"], 81)
    parts = _cavity(hip, size, mono.inv_tau, 81, split)
    assert np.abs(mono.getVelocities()).max() > 1e-3
    assert mono.getVelocities().tobytes() == parts.getVelocities().tobytes()
    assert mono.getDensities().tobytes() == parts.getDensities().tobytes()
    mono.close()
    parts.close()


@pytest.mark.parametrize("split,size", [((2, 1, 1), (24, 20, 18)), ((1, 2, 1), (16, 24, 20)),
                                        ((1, 1, 2), (18, 16, 24)), ((2, 2, 2), (24, 24, 24)),
                                        ((3, 1, 2), (30, 14, 16)), ((2, 2, 1), (4, 4, 9))])
@pytest.mark.parametrize("prec", ["f32", "f16s"])
def test_overlapped_exchange_equals_serial_order(hip, split, size, prec):
    """CController's boundary-first order (CController.cpp:209-309, 647-747):
    slabs next to the communicating faces, then the inner boxes while the
    faces are exchanged on a second stream. Bit-identical to whole-subdomain
    steps followed by the exchange, and to the undivided cavity; the smallest
    case has subdomains of 2 cells, so slabs fill them and no inner box is
    left."""
    say("overlapped exchange", split, size, prec)
    # Re scaled with the edge so tau stays inside the reference's limits
    p = hip.cavity_parameters(size, 0.08, 10.0 * max(size), prec)
    runs = {}
    for name, kw in (("mono", dict()), ("serial", dict(split=split)),
                     ("overlap", dict(split=split, overlap=True)),
                     ("overlap_x8", dict(split=split, overlap=True, xslab=8))):
        c = hip.CavityController(size, precision=prec, params=p, **kw)
        assert c.overlap == name.startswith("overlap")
        for _ in range(41):
            c.computeNextStep()
        runs[name] = (c.getVelocities().tobytes(), c.getDensities().tobytes())
        c.close()
    assert runs["overlap"] == runs["overlap_x8"] == runs["serial"] == runs["mono"]


@pytest.mark.parametrize("split,size", [((2, 1, 1), (24, 20, 18)), ((3, 1, 1), (30, 14, 16)),
                                        ((4, 1, 1), (8, 6, 7))])
@pytest.mark.parametrize("prec", ["f32", "f16s", "f64"])
def test_fused_x_exchange_equals_serial_order(hip, split, size, prec):
    """Subdomains split along x exchanging their x faces inside the step
    kernels (each reads its neighbours' fused-pack buffers, no copy kernel)
    equal the serial order with lbm_exchange_face and the undivided cavity,
    bit for bit; the smallest case has subdomains of 2 cells."""
    say("fused x exchange", split, size, prec)
    p = hip.cavity_parameters(size, 0.08, 10.0 * max(size), prec)
    runs = {}
    for name, kw in (("mono", dict()), ("serial", dict(split=split)),
                     ("fused", dict(split=split, fused_x=True))):
        c = hip.CavityController(size, precision=prec, params=p, **kw)
        for _ in range(41):
            c.computeNextStep()
        runs[name] = (c.getVelocities().tobytes(), c.getDensities().tobytes())
        c.close()
    assert runs["fused"] == runs["serial"] == runs["mono"]


def test_fused_x_exchange_rejects_other_splits(hip):
    for kw in (dict(split=(2, 2, 1)), dict(split=(1, 1, 2)), dict(split=(2, 1, 1), overlap=True)):
        with pytest.raises(ValueError):
            hip.CavityController((8, 8, 8), fused_x=True, **kw)


def _transfers(pairs):
    from cavityflow_cuda_amd import _lib
    arr = (_lib.lbm_face_transfer * len(pairs))()
    for k, (a, d, b) in enumerate(pairs):
        arr[k].from_, arr[k].direction, arr[k].to = a.handle.value, d, b.handle.value
    return arr


@pytest.mark.parametrize("prec", ["f32", "f64", "f16s", "bf16s", "f16c"])
def test_exchange_faces_equals_single_exchanges(hip, prec):
    """lbm_exchange_faces over all the transfers of one axis (3 x 3 subdomains
    in chains of 3 along it, of different extents along it: 12 transfers, two
    launches) leaves every solver exactly as lbm_exchange_face on each
    transfer in turn, for every axis and phase; so does a solver exchanging
    its two faces of an axis with itself (a wall reflected through the
    exchange)."""
    from cavityflow_cuda_amd._lib import check, lib
    say("exchange_faces", prec)
    rng = np.random.default_rng(53)
    dt = np.float64 if prec == "f64" else np.float32
    w = np.array([1 / 18] * 4 + [1 / 36] * 12 + [1 / 18] * 2 + [1 / 3])
    L = lib()
    for axis in range(3):
        # extents along the axis differ between the solvers of a chain (1, 3
        # and 6 interior cells), so one launch mixes row and plane pitches
        sizes = []
        for k in range(9):
            sz = [5, 4, 3]
            sz[axis] = (1, 3, 6)[k % 3]
            sizes.append(tuple(sz))
        for phase in (0, 1):
            grids = []
            starts = [(w[:, None] * (1 + 0.1 * rng.standard_normal(
                (19, int(np.prod([v + 2 for v in sz])))))).astype(dt).ravel() for sz in sizes]
            for _ in range(2):
                g = [hip.LbmSolverHIP(sz, (GHOST,) * 6, prec) for sz in sizes]
                for sv, st in zip(g, starts):
                    sv.setDensityDistributions(st, *full_box(sv))
                grids.append(g)
            pairs = []
            for c in range(3):              # chains of 3 along the axis
                for i in range(2):
                    a, b = c * 3 + i, c * 3 + i + 1
                    pairs.append((a, 2 * axis + 1, b))
                    pairs.append((b, 2 * axis, a))
            batched, single = grids
            hip.exchange_faces([(batched[a], d, batched[b]) for a, d, b in pairs], phase)
            for a, d, b in pairs:
                single[a].exchange_face(d, phase, single[b])
            # self exchange of solver 0's two faces on the axis, through the C ABI
            self_pairs = [(0, 2 * axis + 1, 0), (0, 2 * axis, 0)]
            check(L.lbm_exchange_faces(_transfers([(batched[a], d, batched[b])
                                                   for a, d, b in self_pairs]), 2, phase, None))
            for a, d, b in self_pairs:
                check(L.lbm_exchange_face(single[a].handle, d, phase, single[b].handle, None))
            for x, y in zip(batched, single):
                x.sync()
                y.sync()
                assert x.getDensityDistributions(*full_box(x)).tobytes() == \
                    y.getDensityDistributions(*full_box(y)).tobytes(), (axis, phase)
                fx = bool(np.isnan(x.getMoments(0, (1, 1, 1), (1, 1, 1))).any())
                fy = bool(np.isnan(y.getMoments(0, (1, 1, 1), (1, 1, 1))).any())
                assert fx == fy
            for g in grids:
                for sv in g:
                    sv.close()


def test_exchange_faces_rejections(hip):
    """Transfers on two axes, two into the same face, solvers of another
    precision, a negative count or a NULL list are argument errors and run
    nothing; an empty list is a no-op."""
    from cavityflow_cuda_amd import _lib
    from cavityflow_cuda_amd._lib import lib
    say("exchange_faces rejections")
    L = lib()
    a = hip.LbmSolverHIP((4, 4, 4), (GHOST,) * 6, "f32")
    b = hip.LbmSolverHIP((4, 4, 4), (GHOST,) * 6, "f32")
    c = hip.LbmSolverHIP((4, 4, 4), (GHOST,) * 6, "f16s")
    before = [s.getDensityDistributions(*full_box(s)).tobytes() for s in (a, b)]
    for pairs in ([(a, 1, b), (b, 2, a)],          # x and y
                  [(a, 1, b), (a, 1, b)],          # b's LEFT face twice
                  [(a, 1, b), (c, 0, a)]):         # another precision
        assert L.lbm_exchange_faces(_transfers(pairs), len(pairs), 0, None) == _lib.LBM_ERR_ARG
    ok = _transfers([(a, 1, b)])
    assert L.lbm_exchange_faces(ok, 1, 2, None) == _lib.LBM_ERR_ARG
    assert L.lbm_exchange_faces(ok, -1, 0, None) == _lib.LBM_ERR_ARG
    assert L.lbm_exchange_faces(None, 1, 0, None) == _lib.LBM_ERR_ARG
    assert L.lbm_exchange_faces(None, 0, 0, None) == _lib.LBM_OK
    for s, want in zip((a, b), before):
        s.sync()
        assert s.getDensityDistributions(*full_box(s)).tobytes() == want
    for s in (a, b, c):
        s.close()


def test_overlap_rejects_subdomains_too_thin(hip):
    with pytest.raises(ValueError):
        hip.CavityController((2, 8, 8), split=(2, 1, 1), overlap=True)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_box_copies_match_oracle(hip, oracle_cpu, dtype):
    say("box copies", dtype.__name__)
    rng = np.random.default_rng(11)
    size = (11, 7, 5)
    o = oracle_cpu.OracleSolver(size, (OBST,) * 6, dtype=dtype)
    s = hip.LbmSolverHIP(size, (OBST,) * 6, PREC[dtype])
    o.dd[:] = rng.random(o.dd.size).astype(dtype)
    load_state(s, o)
    for origin, box in [((0, 0, 0), (13, 9, 7)), ((2, 1, 3), (5, 6, 2)), ((12, 8, 6), (1, 1, 1))]:
        assert s.getDensityDistributions(origin, box).tobytes() == \
            o.getDensityDistributions(origin, box).tobytes()
        src = rng.random(19 * int(np.prod(box))).astype(dtype)
        for direction in (None, 0, 1, 2, 3, 4, 5):
            s.setDensityDistributions(src, origin, box, direction)
            o.setDensityDistributions(origin, box, src, direction)
            assert s.getDensityDistributions((0, 0, 0), (13, 9, 7)).tobytes() == o.dd.tobytes()
        uvals = rng.random(3 * int(np.prod(box))).astype(dtype)
        s.setVelocities(uvals, origin, box)
        assert s.getVelocities(origin, box).tobytes() == uvals.tobytes()
        rvals = rng.random(int(np.prod(box))).astype(dtype)
        s.setDensities(rvals, origin, box)
        assert s.getDensities(origin, box).tobytes() == rvals.tobytes()
        fl = rng.choice([1, 2, 4, 8], size=int(np.prod(box))).astype(np.int32)
        s.setFlags(fl, origin, box)
        assert (s.getFlags(origin, box) == fl).all()
    # interior default boxes
    assert s.getVelocities().size == 3 * 11 * 7 * 5
    s.close()


def test_box_errors(hip):
    say("box errors")
    s = hip.LbmSolverHIP((4, 4, 4))
    for origin, box in [((-1, 0, 0), (1, 1, 1)), ((0, 0, 0), (7, 1, 1)), ((6, 5, 5), (1, 1, 1)),
                        ((0, 0, 0), (0, 1, 1))]:
        with pytest.raises(hip.LbmError) as e:
            s.simulationStepAlpha(origin, box)
        assert e.value.status == -2
        with pytest.raises(hip.LbmError):
            s.getDensityDistributions(origin, box)
    with pytest.raises(hip.LbmError) as e:
        s.setFlags(np.full(8, 3, np.int32), (1, 1, 1), (2, 2, 2))
    assert e.value.status == -1
    t = hip.LbmSolverHIP((4, 4, 4), store_densities=False, store_velocities=False)
    with pytest.raises(hip.LbmError) as e:
        t.getVelocities()
    assert e.value.status == -7
    s.close()
    t.close()


def test_pack_unpack_faces_match_oracle(hip, oracle_cpu):
    """lbm_pack_face writes exactly the 5 planes of its face and phase
    (phase 0: the vectors pointing back into the sender; phase 1: those
    pointing out of it) as the oracle's box copy reads them, and leaves the
    other 14 planes of the buffer alone; lbm_unpack_face writes those planes
    of the receiving face as the oracle's direction-filtered box copy does."""
    from hiprt import DeviceBuffer
    say("pack/unpack faces")
    rng = np.random.default_rng(12)
    size = (9, 6, 7)
    o = oracle_cpu.OracleSolver(size, (OBST,) * 6, dtype=np.float32)
    s = hip.LbmSolverHIP(size, (OBST,) * 6, "f32")
    o.dd[:] = rng.random(o.dd.size).astype(np.float32)
    load_state(s, o)
    for direction in range(6):
        n = s.face_elements(direction)
        for phase in (0, 1):
            send = oracle_cpu.direction_mask(direction if phase == 0 else direction ^ 1)
            buf = DeviceBuffer(4 * n)
            buf.upload(np.full(n, -7.0, np.float32))
            s.pack_face(direction, phase, buf)
            s.sync()
            origin, box = s.face_box(direction, 0 if phase == 0 else 1)
            packed = buf.download(np.float32, n).reshape(19, -1)
            want = o.getDensityDistributions(origin, box).reshape(19, -1)
            for l in range(19):
                if (send >> l) & 1:
                    assert packed[l].tobytes() == want[l].tobytes(), (direction, phase, l)
                else:
                    assert (packed[l] == -7.0).all(), (direction, phase, l)
            src = rng.random(n).astype(np.float32)
            buf.upload(src)
            # the receiving face `direction` takes what its neighbour's opposite face sent
            s.unpack_face(direction, phase, buf)
            s.sync()
            buf.free()
            origin, box = s.face_box(direction, 1 if phase == 0 else 0)
            o.setDensityDistributions(origin, box, src, direction ^ 1 if phase == 0 else direction)
            assert s.getDensityDistributions((0, 0, 0), s.size_h).tobytes() == o.dd.tobytes()
    s.close()


@pytest.mark.parametrize("prec", ["f32", "f64", "f16s", "bf16s", "f16c"])
@pytest.mark.parametrize("size", [(9, 6, 7), (1, 4, 5), (5, 1, 1)])
def test_pack_face_pair_equals_two_packs(hip, prec, size):
    """lbm_pack_face_pair(axis) fills both buffers exactly as the two
    single-face packs do, planes outside each face's 5 untouched, for every
    axis and phase."""
    from hiprt import DeviceBuffer
    say("pack_face_pair", prec, size)
    rng = np.random.default_rng(37)
    dt = np.float64 if prec == "f64" else np.float32
    s = hip.LbmSolverHIP(size, (OBST,) * 6, prec)
    o, z = full_box(s)
    cells = z[0] * z[1] * z[2]
    w = np.array([1 / 18] * 4 + [1 / 36] * 12 + [1 / 18] * 2 + [1 / 3])
    s.setDensityDistributions((w[:, None] * (1 + 0.1 * rng.standard_normal((19, cells))))
                              .astype(dt).ravel(), o, z)
    for axis in range(3):
        n = s.face_elements(2 * axis)
        for phase in (0, 1):
            bufs = [DeviceBuffer(n * np.dtype(dt).itemsize) for _ in range(4)]
            for b in bufs:
                b.upload(np.full(n, -7.0, dt))
            s.pack_face(2 * axis, phase, bufs[0])
            s.pack_face(2 * axis + 1, phase, bufs[1])
            s.pack_face_pair(axis, phase, bufs[2], bufs[3])
            s.sync()
            got = [b.download(dt, n).tobytes() for b in bufs]
            for b in bufs:
                b.free()
            assert got[2] == got[0] and got[3] == got[1], (axis, phase)
    from cavityflow_cuda_amd import LbmError
    buf = DeviceBuffer(64)
    with pytest.raises(LbmError):
        s.pack_face_pair(3, 0, buf, buf)
    with pytest.raises(LbmError):
        s.pack_face_pair(1, 2, buf, buf)
    buf.free()
    s.close()


@pytest.mark.parametrize("prec", ["f32", "f64", "f16s", "bf16s", "f16c"])
@pytest.mark.parametrize("size", [(9, 6, 7), (1, 4, 5), (5, 1, 1)])
def test_unpack_face_pair_equals_two_unpacks(hip, prec, size):
    """lbm_unpack_face_pair(axis) leaves exactly the state of the two
    single-face unpacks, for every axis and phase (one launch per axis; a
    one-cell axis puts both phase-1 planes in one plane of the array, with
    disjoint lattice vectors)."""
    from hiprt import DeviceBuffer
    say("unpack_face_pair", prec, size)
    rng = np.random.default_rng(31)
    dt = np.float64 if prec == "f64" else np.float32
    a = hip.LbmSolverHIP(size, (OBST,) * 6, prec)
    b = hip.LbmSolverHIP(size, (OBST,) * 6, prec)
    o, z = full_box(a)
    cells = z[0] * z[1] * z[2]
    w = np.array([1 / 18] * 4 + [1 / 36] * 12 + [1 / 18] * 2 + [1 / 3])
    state = (w[:, None] * (1 + 0.1 * rng.standard_normal((19, cells)))).astype(dt)
    for s in (a, b):
        s.setDensityDistributions(state.ravel(), o, z)
    for axis in range(3):
        for phase in (0, 1):
            n = a.face_elements(2 * axis)
            src = [(w[:, None] * (1 + 0.1 * rng.standard_normal((19, n // 19)))).astype(dt)
                   for _ in range(2)]
            bufs = [DeviceBuffer(src[0].nbytes) for _ in range(2)]
            for buf, v in zip(bufs, src):
                buf.upload(v)
            a.unpack_face(2 * axis, phase, bufs[0])
            a.unpack_face(2 * axis + 1, phase, bufs[1])
            b.unpack_face_pair(axis, phase, bufs[0], bufs[1])
            a.sync()
            b.sync()
            for buf in bufs:
                buf.free()
            assert (a.getDensityDistributions(o, z).tobytes()
                    == b.getDensityDistributions(o, z).tobytes()), (axis, phase)
    from cavityflow_cuda_amd import LbmError
    buf = DeviceBuffer(64)
    with pytest.raises(LbmError):
        b.unpack_face_pair(3, 0, buf, buf)
    with pytest.raises(LbmError):
        b.unpack_face_pair(0, 2, buf, buf)
    buf.free()
    a.close()
    b.close()


@pytest.mark.parametrize("prec,tol", [("f32", 2e-6), ("f64", 1e-13)])
def test_mass_is_conserved_in_closed_box(hip, prec, tol):
    """No lid: the total of all populations (walls included) is invariant."""
    say("mass", prec)
    s = hip.LbmSolverHIP((20, 16, 12), precision=prec, inv_tau=1.7)
    dtype = np.float64 if prec == "f64" else np.float32
    rng = np.random.default_rng(13)
    d = s.getDensityDistributions((1, 1, 1), (20, 16, 12))
    d = (d * (1 + 0.02 * (rng.random(d.size) - 0.5))).astype(dtype)
    s.setDensityDistributions(d, (1, 1, 1), (20, 16, 12))
    m0 = s.reduce_box()[0]
    for n in range(200):
        (s.simulationStepBeta if n & 1 else s.simulationStepAlpha)()
    m1 = s.reduce_box()[0]
    assert abs(m1 - m0) / m0 < tol * 200
    s.close()


def test_residual_decreases_toward_steady_state(hip):
    say("residual")
    c = hip.CavityController((16, 16, 16), u_lid=0.05, reynolds=100.0)
    s = next(iter(c.solvers.values()))
    res = []
    for n in range(2000):
        c.computeNextStep()
        if n % 200 == 199:
            res.append(s.residual(layout=0))   # after a beta step
    assert res[0] == float("inf")
    assert all(np.isfinite(res[1:]))
    assert res[-1] < res[1] * 0.5
    c.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("seed", range(16))
def test_seeded_random_cases_match_oracle(hip, oracle_cpu, dtype, seed):
    """Seeded sweep: random shape (thin axes and rows that straddle workgroup
    boundaries included), random flags, tau, lid and gravity, random launch
    options for each kernel, and random boxes for the steps; every cell of
    every population equals the oracle after each half-step."""
    from cavityflow_cuda_amd._lib import check, lib
    rng = np.random.default_rng(1000 + seed)
    size = tuple(int(v) for v in rng.choice([1, 2, 3, 7, 16, 31, 33, 47, 63, 65], 3))
    say("seeded", dtype.__name__, seed, size)
    kw = dict(inv_tau=float(rng.uniform(0.6, 1.9)), lid_velocity=float(rng.uniform(0, 0.1)),
              gravitation=tuple(float(g) for g in rng.uniform(-2e-3, 2e-3, 3)))
    o = oracle_cpu.OracleSolver(size, (OBST,) * 6, dtype=dtype, **kw)
    o.dd[:] = (rng.random(o.dd.size) * 0.08 + 0.01).astype(dtype)
    o.flags[:] = rng.choice([OBST, FLUID, FLUID, FLUID, INJ, GHOST], size=o.n).astype(np.int32)
    s = hip.LbmSolverHIP(size, (OBST,) * 6, PREC[dtype], kw["inv_tau"], kw["lid_velocity"],
                         kw["gravitation"])
    for k in (0, 1):
        check(lib().lbm_set_launch_options(s.handle, k, int(rng.integers(0, 256))))
    load_state(s, o)
    nh = o.size_h
    for n in range(6):
        if rng.random() < 0.5:
            box = ()
        else:
            org = tuple(int(rng.integers(0, nh[a])) for a in range(3))
            ext = tuple(int(rng.integers(1, nh[a] - org[a] + 1)) for a in range(3))
            box = (org, ext)
        if n & 1:
            o.beta(*box)
            s.simulationStepBeta(*box)
        else:
            o.alpha(*box)
            s.simulationStepAlpha(*box)
        same_state(s, o, f"seed {seed} step {n} box {box}")
    s.close()

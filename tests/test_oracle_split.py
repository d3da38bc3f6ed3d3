"""The reference's validation semantic (main.cpp:121-303): a run split into
subdomains with halo exchange equals the monolithic run. Checked here on the
CPU oracle with the exchange sequence of CController::stepAlpha/stepBeta
(CController.cpp:369-570, 807-1008) and the CComm face boxes
(CManager.cpp:135-202). The GPU controller is checked against the same
semantic in tests/test_gpu_parity.py."""
import itertools

import numpy as np
import pytest

OBST, FLUID, INJ, GHOST = 1, 2, 4, 8


def face_box(size_h, direction, which):
    axis, high = direction // 2, direction & 1
    origin = [0, 0, 0]
    size = list(size_h)
    size[axis] = 1
    if which == 0:   # send plane
        origin[axis] = size_h[axis] - 2 if high else 1
    else:            # halo plane
        origin[axis] = size_h[axis] - 1 if high else 0
    return tuple(origin), tuple(size)


def exchange(cpu, a, direction, phase, b):
    """a's face `direction` -> b across the opposite face."""
    o, z = face_box(a.size_h, direction, 0 if phase == 0 else 1)
    buf = a.getDensityDistributions(o, z)
    o2, z2 = face_box(b.size_h, direction ^ 1, 1 if phase == 0 else 0)
    b.setDensityDistributions(o2, z2, buf, None if phase == 0 else direction ^ 1)


def build(cpu, size, split, inv_tau, lid):
    sub = tuple(size[a] // split[a] for a in range(3))
    solvers = {}
    for k, j, i in itertools.product(range(split[2]), range(split[1]), range(split[0])):
        bc = [GHOST] * 6
        for axis, idx in enumerate((i, j, k)):
            if idx == 0:
                bc[2 * axis] = OBST
            if idx == split[axis] - 1:
                bc[2 * axis + 1] = OBST
        s = cpu.OracleSolver(sub, bc, dtype=np.float32, inv_tau=inv_tau, lid_velocity=lid)
        if j == split[1] - 1:
            s.set_driven_cavity()
        solvers[(i, j, k)] = s
    return sub, solvers


def step(cpu, solvers, n):
    phase = n & 1
    for s in solvers.values():
        s.beta() if phase else s.alpha()
    for axis in range(3):
        for key, a in solvers.items():
            nb = list(key)
            nb[axis] += 1
            b = solvers.get(tuple(nb))
            if b is None:
                continue
            exchange(cpu, a, 2 * axis + 1, phase, b)
            exchange(cpu, b, 2 * axis, phase, a)


def gather(solvers, size, sub, field):
    nc = 3 if field == "u" else 1
    out = np.empty((nc,) + size[::-1], np.float32)
    for (i, j, k), s in solvers.items():
        part = (s.getVelocities() if field == "u" else s.getDensities()).reshape(
            nc, sub[2], sub[1], sub[0])
        out[:, k * sub[2]:(k + 1) * sub[2], j * sub[1]:(j + 1) * sub[1],
            i * sub[0]:(i + 1) * sub[0]] = part
    return out


@pytest.mark.parametrize("split", [(2, 1, 1), (1, 2, 1), (1, 1, 2), (2, 2, 2), (3, 1, 2)])
def test_split_equals_monolithic_on_oracle(oracle_cpu, split):
    size = (12, 12, 12) if split != (3, 1, 2) else (12, 10, 8)
    inv_tau, lid = 1.6863406896591187, 0.05
    _, mono = build(oracle_cpu, size, (1, 1, 1), inv_tau, lid)
    sub, parts = build(oracle_cpu, size, split, inv_tau, lid)
    for n in range(41):
        step(oracle_cpu, mono, n)
        step(oracle_cpu, parts, n)
        if n in (39, 40):   # after a beta and after an alpha step
            u_m = gather(mono, size, size, "u")
            u_s = gather(parts, size, sub, "u")
            r_m = gather(mono, size, size, "rho")
            r_s = gather(parts, size, sub, "rho")
            assert np.abs(u_m).max() > 1e-3
            assert u_m.tobytes() == u_s.tobytes()
            assert r_m.tobytes() == r_s.tobytes()


def exchange_five(cpu, a, direction, phase, b):
    """The exchange of lbm_pack_face / lbm_unpack_face: of the face buffer only
    the 5 planes the receiver applies travel (phase 0: the vectors pointing
    back into the sender, direction_mask(direction); phase 1: those pointing
    out of it, direction_mask(direction ^ 1)); the other 14 planes arrive as
    NaN, so any use of them would show."""
    o, z = face_box(a.size_h, direction, 0 if phase == 0 else 1)
    buf = a.getDensityDistributions(o, z).copy().reshape(19, -1)
    mask = cpu.direction_mask(direction if phase == 0 else direction ^ 1)
    assert bin(mask).count("1") == 5
    for l in range(19):
        if not (mask >> l) & 1:
            buf[l] = np.nan
    o2, z2 = face_box(b.size_h, direction ^ 1, 1 if phase == 0 else 0)
    b.setDensityDistributions(o2, z2, buf.ravel(), direction if phase == 0 else direction ^ 1)


def interior_populations(solvers, size, sub):
    out = np.empty((19,) + size[::-1], np.float32)
    for (i, j, k), s in solvers.items():
        nh = s.size_h
        dd = s.dd.reshape(19, nh[2], nh[1], nh[0])[:, 1:-1, 1:-1, 1:-1]
        out[:, k * sub[2]:(k + 1) * sub[2], j * sub[1]:(j + 1) * sub[1],
            i * sub[0]:(i + 1) * sub[0]] = dd
    return out


@pytest.mark.parametrize("split,size", [((2, 1, 1), (12, 12, 12)), ((1, 2, 1), (12, 12, 12)),
                                        ((1, 1, 2), (12, 12, 12)), ((2, 2, 2), (12, 12, 12)),
                                        ((3, 1, 2), (12, 10, 8)), ((2, 3, 2), (10, 12, 8)),
                                        ((3, 3, 3), (9, 9, 9)), ((2, 2, 1), (6, 4, 5))])
def test_five_plane_exchange_keeps_every_interior_population(oracle_cpu, split, size):
    """The device exchange moves 5 of the 19 planes per face and phase. In the
    beta step plane l at cell p is read and written only by cell p - c_l, so
    the other 14 planes of a halo plane stay in the halo or are carried (by
    the linear wrap and the halo cells' pass-through) into interior slots the
    phase-1 unpack overwrites before they are read. Checked here: every
    interior population of every subdomain equals the monolithic run's after
    every exchange, for 60 half-steps, with the untransmitted planes NaN."""
    inv_tau, lid = 1.6863406896591187, 0.05
    _, mono = build(oracle_cpu, size, (1, 1, 1), inv_tau, lid)
    sub, parts = build(oracle_cpu, size, split, inv_tau, lid)
    for n in range(60):
        phase = n & 1
        for s in mono.values():
            s.beta() if phase else s.alpha()
        for s in parts.values():
            s.beta() if phase else s.alpha()
        for axis in range(3):
            for key, a in parts.items():
                nb = list(key)
                nb[axis] += 1
                b = parts.get(tuple(nb))
                if b is None:
                    continue
                exchange_five(oracle_cpu, a, 2 * axis + 1, phase, b)
                exchange_five(oracle_cpu, b, 2 * axis, phase, a)
        want = interior_populations(mono, size, size)
        got = interior_populations(parts, size, sub)
        assert got.tobytes() == want.tobytes(), n
    assert np.abs(gather(mono, size, size, "u")).max() > 1e-3

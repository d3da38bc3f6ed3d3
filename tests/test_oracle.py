"""Pins the CPU oracle (oracle/lbm_oracle.c) before it is trusted as the checker.

1. Golden vectors written by tests/golden/make_golden.py from the reference's
   own compiled CPU kernels: bit-for-bit equality.
2. Closed forms and invariants that a dropped term, a wrong sign, a wrong
   index or a swapped slot would break.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

OBST, FLUID, INJ, GHOST = 1, 2, 4, 8

# D3Q19 velocities E0..E18 (common.h:29-47), written out independently of the oracle
C = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [1, 1, 0], [-1, -1, 0],
              [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1],
              [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1], [0, 0, 1], [0, 0, -1],
              [0, 0, 0]])
OPP = [1, 0, 3, 2, 5, 4, 7, 6, 9, 8, 11, 10, 13, 12, 15, 14, 17, 16, 18]


def load(name):
    return np.load(os.path.join(GOLDEN, name))


@pytest.mark.parametrize("dtype,name", [(np.float32, "step_f32.npz"), (np.float64, "step_f64.npz")])
def test_step_matches_reference_golden(oracle_cpu, dtype, name):
    g = load(name)
    size = tuple(int(v) for v in g["size"])
    # flag_order=1: the reference CPU init enumerates z fastest (non-cubic domain)
    s = oracle_cpu.OracleSolver(size, (OBST,) * 6, dtype=dtype, inv_tau=g["inv_tau"],
                                lid_velocity=g["lid"], gravitation=g["g"], flag_order=1)
    assert s.dd.tobytes() == g["init_dd"].tobytes()
    assert s.u.tobytes() == g["init_u"].tobytes()
    assert s.rho.tobytes() == g["init_rho"].tobytes()
    assert (s.flags == g["init_flags"]).all()
    s.dd[:] = g["dd0"]
    s.flags[:] = g["flags"]
    s.u[:] = 0
    s.rho[:] = 0
    s.alpha()
    assert s.dd.tobytes() == g["alpha_dd"].tobytes()
    assert s.u.tobytes() == g["alpha_u"].tobytes()
    assert s.rho.tobytes() == g["alpha_rho"].tobytes()
    s.beta()
    assert s.dd.tobytes() == g["beta_dd"].tobytes()
    assert s.u.tobytes() == g["beta_u"].tobytes()
    assert s.rho.tobytes() == g["beta_rho"].tobytes()
    o, b = tuple(g["box_origin"]), tuple(g["box_size"])
    s.alpha(o, b)
    s.beta(o, b)
    assert s.dd.tobytes() == g["box_dd"].tobytes()
    assert s.u.tobytes() == g["box_u"].tobytes()
    assert s.rho.tobytes() == g["box_rho"].tobytes()


def test_flag_init_orders(oracle_cpu):
    """order=1 reproduces the reference CPU enumeration (z fastest); order=0 is
    the x-fastest layout every kernel reads, i.e. its transpose."""
    g = load("flags_init.npz")
    for i in range(3):
        size = tuple(int(v) for v in g[f"size_{i}"])
        size_h = tuple(v + 2 for v in size)
        bc = tuple(int(v) for v in g[f"bc_{i}"])
        ref_flags = g[f"flags_{i}"].astype(np.int32)
        zfast = oracle_cpu.set_flags(size_h, bc, order=1)
        assert (zfast == ref_flags).all()
        xfast = oracle_cpu.set_flags(size_h, bc, order=0)
        as_xyz = ref_flags.reshape(size_h[0], size_h[1], size_h[2])   # [x][y][z]
        assert (xfast.reshape(size_h[2], size_h[1], size_h[0]) == as_xyz.transpose(2, 1, 0)).all()


def _oracle_cavity(cpu, size, params, steps, nthreads=8):
    s = cpu.OracleSolver(size, (OBST,) * 6, dtype=np.float32, inv_tau=params["inv_tau"],
                         lid_velocity=params["lid"], nthreads=nthreads)
    s.set_driven_cavity()
    for i in range(steps):
        if i & 1:
            s.beta()
        else:
            s.alpha()
    return s


def test_cavity16_matches_reference_golden(oracle_cpu):
    g = load("cavity16.npz")
    s = _oracle_cavity(oracle_cpu, (16, 16, 16), dict(inv_tau=float(g["inv_tau"]), lid=0.05), 200)
    assert s.getVelocities().tobytes() == g["u"].tobytes()
    assert s.getDensities().tobytes() == g["rho"].tobytes()


@pytest.mark.slow
def test_config1_matches_reference_golden(oracle_cpu):
    import hashlib
    with open(os.path.join(GOLDEN, "config1.json")) as f:
        g = json.load(f)
    s = _oracle_cavity(oracle_cpu, (62, 62, 62), dict(inv_tau=g["inv_tau"], lid=0.05), 1000)
    u = s.getVelocities()
    rho = s.getDensities()
    assert hashlib.sha256(u.tobytes()).hexdigest() == g["sha256_u"]
    assert hashlib.sha256(rho.tobytes()).hexdigest() == g["sha256_rho"]
    mid = load("config1_midplane.npz")["u_mid_z31"]
    assert (u.reshape(3, 62, 62, 62)[:, 31] == mid).all()


def test_solver_params_match_reference_golden(oracle_cpu):
    with open(os.path.join(GOLDEN, "params.json")) as f:
        cases = json.load(f)
    for entry in cases:
        c = entry["case"]
        length = c.get("length") or c["size"]
        for name, dtype in (("float32", np.float32), ("float64", np.float64)):
            ref = entry[name]
            st, out = oracle_cpu.solver_params(length[0], c["size"][0], max(length), c["velocity"],
                                               c["acceleration"], c["viscosity"], c["timestep"],
                                               c["max_vel"], c["max_acc"], dtype=dtype)
            if ref["exit_code"] != 0:
                assert st == -1, c
                continue
            v = np.array(ref["values"], dtype=dtype)
            assert st == 0
            assert (out["velocity_dimless"] == v[0:3]).all()
            assert (out["acceleration_dimless"] == v[3:6]).all()
            assert out["viscosity_dimless"] == v[6]
            assert out["tau"] == v[7]
            assert out["inv_tau"] == v[8]
            assert out["timestep"] == v[9]


# ---------------------------------------------------------------- closed forms

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rest_equilibrium_is_lattice_weights(oracle_cpu, dtype):
    s = oracle_cpu.OracleSolver((3, 4, 5), dtype=dtype)
    T = dtype
    w = [T(1) / T(18)] * 4 + [T(1) / T(36)] * 12 + [T(1) / T(18)] * 2 + [T(1) / T(3)]
    dd = s.dd.reshape(19, -1)
    for l in range(19):
        assert (dd[l] == w[l]).all()
    assert abs(sum(float(x) for x in w) - 1.0) < 1e-6


def _random_fluid(cpu, rng, size, dtype, inv_tau):
    s = cpu.OracleSolver(size, (OBST,) * 6, dtype=dtype, inv_tau=inv_tau, lid_velocity=0.0)
    s.dd[:] = (rng.random(s.dd.size) * 0.08 + 0.01).astype(dtype)
    s.flags[:] = FLUID
    return s


def _eq(rho, j):
    """Incompressible D3Q19 equilibrium from the lattice, in float64."""
    out = []
    usq = (j ** 2).sum(axis=0)
    for l in range(19):
        w = 1 / 3 if l == 18 else (1 / 18 if (l < 4 or l >= 16) else 1 / 36)
        cu = (C[l][:, None] * j).sum(axis=0)
        out.append(w * (rho - 1.5 * usq + 3 * cu + 4.5 * cu ** 2))
    return np.array(out)


def test_tau_one_gives_equilibrium(oracle_cpu):
    """inv_tau = 1: post-collision populations equal the equilibrium of the
    pre-collision moments, stored in the opposite slot (alpha)."""
    rng = np.random.default_rng(3)
    s = _random_fluid(oracle_cpu, rng, (4, 4, 4), np.float64, 1.0)
    d0 = s.dd.reshape(19, -1).copy()
    rho = d0.sum(axis=0)
    j = (C.T[:, :, None] * d0[None]).sum(axis=1)
    s.alpha()
    d1 = s.dd.reshape(19, -1)
    eq = _eq(rho, j)
    for l in range(19):
        assert np.allclose(d1[OPP[l]], eq[l], rtol=0, atol=1e-15)


@pytest.mark.parametrize("dtype,tol", [(np.float32, 2e-6), (np.float64, 1e-14)])
def test_fluid_collision_conserves_mass_and_momentum(oracle_cpu, dtype, tol):
    rng = np.random.default_rng(4)
    s = _random_fluid(oracle_cpu, rng, (5, 4, 3), dtype, 1.7)
    d0 = s.dd.reshape(19, -1).astype(np.float64).copy()
    s.alpha()
    d1 = s.dd.reshape(19, -1).astype(np.float64)[OPP]   # direction l from slot opp(l)
    assert np.allclose(d1.sum(0), d0.sum(0), rtol=0, atol=tol * 2)
    for a in range(3):
        j0 = (C[:, a][:, None] * d0).sum(0)
        j1 = (C[:, a][:, None] * d1).sum(0)
        assert np.allclose(j0, j1, rtol=0, atol=tol * 4)


@pytest.mark.parametrize("size", [(3, 4, 5), (4, 4, 4), (6, 2, 3)])
def test_all_ghost_beta_is_neighbour_shift(oracle_cpu, size):
    """All-GHOST beta moves slot opp(l) at id - c_l to slot l at id + c_l, with
    the reference's linear-index wrap: new_l = roll(old_opp(l), 2 * delta_l)."""
    rng = np.random.default_rng(5)
    s = oracle_cpu.OracleSolver(size, (GHOST,) * 6, dtype=np.float64)
    s.flags[:] = GHOST
    s.dd[:] = rng.random(s.dd.size)
    old = s.dd.reshape(19, -1).copy()
    s.beta()
    new = s.dd.reshape(19, -1)
    nx, ny = s.size_h[0], s.size_h[1]
    for l in range(19):
        delta = C[l][0] + C[l][1] * nx + C[l][2] * nx * ny
        assert (new[l] == np.roll(old[OPP[l]], 2 * delta)).all(), l


def test_obstacle_beta_is_identity_on_populations(oracle_cpu):
    rng = np.random.default_rng(6)
    s = oracle_cpu.OracleSolver((4, 3, 5), dtype=np.float32)
    s.flags[:] = OBST
    s.dd[:] = rng.random(s.dd.size).astype(np.float32)
    before = s.dd.copy()
    s.beta()
    assert (s.dd == before).all()
    assert (s.u == 0).all()


def test_box_steps_equal_masked_full_steps(oracle_cpu):
    """Alpha is cell-local: a box step equals the full step on the box cells and
    leaves every other cell untouched."""
    rng = np.random.default_rng(7)
    size = (6, 5, 4)
    a = _random_fluid(oracle_cpu, rng, size, np.float32, 1.3)
    a.flags[:] = rng.choice([OBST, FLUID, INJ, GHOST], size=a.flags.size).astype(np.int32)
    b = oracle_cpu.OracleSolver(size, dtype=np.float32, inv_tau=1.3)
    b.dd[:] = a.dd
    b.flags[:] = a.flags
    origin, box = (1, 2, 0), (4, 3, 5)
    a.alpha(origin, box)
    b.alpha()
    nxh, nyh, nzh = a.size_h
    m = np.zeros((nzh, nyh, nxh), bool)
    m[origin[2]:origin[2] + box[2], origin[1]:origin[1] + box[1], origin[0]:origin[0] + box[0]] = True
    da, db = a.dd.reshape(19, -1), b.dd.reshape(19, -1)
    assert (da[:, m.ravel()] == db[:, m.ravel()]).all()


def test_direction_masks(oracle_cpu):
    expected = {0: {0, 4, 6, 8, 10}, 1: {1, 5, 7, 9, 11}, 2: {2, 4, 7, 12, 14},
                3: {3, 5, 6, 13, 15}, 4: {8, 11, 12, 15, 16}, 5: {9, 10, 13, 14, 17}}
    for d, ls in expected.items():
        m = oracle_cpu.direction_mask(d)
        assert {l for l in range(19) if (m >> l) & 1} == ls


# accumulation sequences of CLbmAlphaCPU.cpp:87-182, written out from the
# reference's statement order: (slot, sign) per step after the first term
_JX = [(0, 1), (1, -1), (4, 1), (5, -1), (6, 1), (7, -1), (8, 1), (9, -1), (10, 1), (11, -1)]
_JY = [(2, 1), (3, -1), (4, 1), (5, -1), (6, -1), (7, 1), (12, 1), (13, -1), (14, 1), (15, -1)]
_JZ = [(8, 1), (9, -1), (10, -1), (11, 1), (12, 1), (13, -1), (14, -1), (15, 1), (16, 1),
       (17, -1)]


def _sequential(d, seq, dtype):
    acc = d[seq[0][0]].astype(dtype).copy()
    for l, sign in seq[1:]:
        acc = (acc + d[l]).astype(dtype) if sign > 0 else (acc - d[l]).astype(dtype)
    return acc


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_moments_follow_the_reference_order(oracle_cpu, dtype):
    """oracle_moments (the checker of lbm_get_moments): rho is the index-order
    sum, j the reference's statement order, equal bit for bit to the momentum
    alpha stores for FLUID cells (alpha is pinned by the golden vectors); the
    swapped layout of an opp-permuted state gives the same bits."""
    size = (7, 5, 6)
    rng = np.random.default_rng(21)
    o = oracle_cpu.OracleSolver(size, (OBST,) * 6, dtype=dtype, inv_tau=1.3)
    o.dd[:] = (rng.random(o.dd.size) * 0.08 + 0.01).astype(dtype)
    o.flags[:] = rng.choice([OBST, FLUID, INJ, GHOST], size=o.n).astype(np.int32)
    full = ((0, 0, 0), o.size_h)
    m = o.moments(0, *full).reshape(4, -1)
    d = o.dd.reshape(19, -1)
    assert m[0].tobytes() == _sequential(d, [(l, 1) for l in range(19)], dtype).tobytes()
    for k, seq in ((1, _JX), (2, _JY), (3, _JZ)):
        assert m[k].tobytes() == _sequential(d, seq, dtype).tobytes()
    swapped = o.dd.reshape(19, -1)[OPP].copy()
    keep = o.dd.copy()
    o.dd[:] = swapped.reshape(-1)
    assert o.moments(1, *full).tobytes() == m.tobytes()
    o.dd[:] = keep
    o.alpha()
    u = o.u.reshape(3, -1)
    fluid = o.flags == FLUID
    for k in range(3):
        assert u[k][fluid].tobytes() == m[k + 1][fluid].tobytes()
    # a box is the same cells' values
    box = ((1, 2, 0), (4, 2, 5))
    mb = oracle_cpu.OracleSolver(size, (OBST,) * 6, dtype=dtype)
    mb.dd[:] = keep
    want = m.reshape(4, o.size_h[2], o.size_h[1], o.size_h[0])[:, 0:5, 2:4, 1:5]
    assert mb.moments(0, *box).tobytes() == np.ascontiguousarray(want).tobytes()

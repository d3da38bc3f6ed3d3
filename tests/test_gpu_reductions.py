"""The FP64 diagnostics against independent computations.

* lbm_get_moments: byte-equal to the oracle's moments (reference accumulation
  order, pinned in tests/test_oracle.py) in both slot layouts, on random
  states with every flag, whole interior and boxes; for the 16-bit formats on
  the decoded populations (oracle/storage.py).
* lbm_reduce_box: numpy FP64 sums of the populations the solver returns
  (getDensityDistributions), sum f and sum c_l f, and the cell count.
* lbm_residual: numpy FP64 ||j - j_prev|| / ||j|| over FLUID cells, with j the
  oracle's moments; +inf on the first call, NaN for a non-finite state (also on
  a first call).
The device sums run in another order than numpy's, so the FP64 comparisons
allow 1e-12 relative; a wrong mask, slot or layout is off by far more.
"""
import numpy as np
import pytest

from oracle import storage

pytestmark = pytest.mark.gpu

OBST, FLUID, INJ, GHOST = 1, 2, 4, 8
PRECS = ["f32", "f64", "f16s", "bf16s", "f16c"]


def say(*a):
    print("[case]", *a, flush=True)


def random_solver(hip, oracle_cpu, prec, size, seed):
    """A GPU solver and an oracle solver (FP32 / FP64) holding the same
    populations (decoded values for the 16-bit formats) and flags."""
    rng = np.random.default_rng(seed)
    dtype = np.float64 if prec == "f64" else np.float32
    g = hip.LbmSolverHIP(size, (OBST,) * 6, prec, inv_tau=1.4, lid_velocity=0.05)
    o = oracle_cpu.OracleSolver(size, (OBST,) * 6, dtype=dtype, inv_tau=1.4, lid_velocity=0.05)
    n = o.n
    x = rng.standard_normal((19, n)) * 5e-3
    f = (storage.W[:, None].astype(np.float64) + x).astype(dtype)
    o.flags[:] = rng.choice([OBST, FLUID, FLUID, INJ, GHOST], size=n).astype(np.int32)
    full = ((0, 0, 0), g.size_h)
    g.setFlags(o.flags, *full)
    g.setDensityDistributions(f.reshape(-1), *full)
    o.dd[:] = g.getDensityDistributions(*full)      # what the solver stores
    return g, o, full


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("layout", [0, 1])
def test_moments_equal_oracle(hip, oracle_cpu, prec, layout):
    say("moments", prec, layout)
    g, o, full = random_solver(hip, oracle_cpu, prec, (13, 6, 9), 31 + layout)
    for box in (None, full, ((2, 0, 3), (9, 8, 1)), ((0, 5, 0), (15, 3, 11))):
        args = () if box is None else box
        got = g.getMoments(layout, *args)
        want = o.moments(layout, *args)
        assert got.tobytes() == want.tobytes(), (prec, layout, box)
    g.close()


@pytest.mark.parametrize("prec", PRECS)
def test_reduce_box_equals_numpy_fp64(hip, oracle_cpu, prec):
    say("reduce_box", prec)
    g, o, full = random_solver(hip, oracle_cpu, prec, (21, 10, 17), 41)
    rng = np.random.default_rng(42)
    boxes = [full, ((1, 1, 1), g.size), ((3, 0, 2), (5, 12, 1))]
    for _ in range(4):
        org = tuple(int(rng.integers(0, g.size_h[a])) for a in range(3))
        ext = tuple(int(rng.integers(1, g.size_h[a] - org[a] + 1)) for a in range(3))
        boxes.append((org, ext))
    for box in boxes:
        got = g.reduce_box(*box)
        d = g.getDensityDistributions(*box).reshape(19, -1).astype(np.float64)
        want = np.array([d.sum(), (storage.C[:, 0, None] * d).sum(),
                         (storage.C[:, 1, None] * d).sum(), (storage.C[:, 2, None] * d).sum(),
                         d.shape[1]])
        scale = np.abs(d).sum()
        assert got[4] == want[4]
        assert np.all(np.abs(got[:4] - want[:4]) <= 1e-12 * scale), (prec, box, got, want)
    g.close()


def _residual_numpy(o, layout, prev):
    m = o.moments(layout, (0, 0, 0), o.size_h).reshape(4, -1).astype(np.float64)
    fluid = o.flags == FLUID
    j = m[1:, fluid]
    num = ((j - prev[:, fluid]) ** 2).sum() if prev is not None else None
    den = (j ** 2).sum()
    full_j = np.zeros((3, o.n))
    full_j[:, fluid] = j
    return (np.sqrt(num / den) if num is not None else None), full_j


@pytest.mark.parametrize("prec", PRECS)
def test_residual_equals_numpy_fp64(hip, oracle_cpu, prec):
    say("residual", prec)
    g, o, full = random_solver(hip, oracle_cpu, prec, (12, 11, 10), 51)
    assert g.residual(0) == float("inf")             # first call: no previous field
    _, prev = _residual_numpy(o, 0, None)
    for step in range(4):
        phase = step & 1
        (g.simulationStepBeta if phase else g.simulationStepAlpha)()
        layout = 0 if phase else 1                    # after alpha the slots are swapped
        o.dd[:] = g.getDensityDistributions(*full)
        got = g.residual(layout)
        want, prev = _residual_numpy(o, layout, prev)
        assert np.isfinite(got) and want > 0
        assert abs(got - want) <= 1e-12 * want, (prec, step, got, want)
    g.close()


def test_residual_and_sums_of_a_non_finite_state_are_nan(hip):
    say("residual NaN")
    for prec in ("f32", "f64"):
        g = hip.LbmSolverHIP((6, 6, 6), precision=prec)
        full = ((0, 0, 0), g.size_h)
        d = g.getDensityDistributions(*full).reshape(19, -1)
        d[5, 100] = np.nan
        g.setDensityDistributions(d.reshape(-1), *full)
        assert np.isnan(g.residual(0))                # NaN even on the first call
        assert np.isnan(g.reduce_box()[0])
        g.reset()
        assert g.residual(0) == float("inf")
        assert g.residual(0) == 0.0                   # rest state: nothing moves
        g.close()

"""Bit-exact parity of the shifted 16-bit step arithmetic (F16S, BF16S, F16C).

The kernels decode the 16-bit codes of a cell, run the FP32 collision (the
arithmetic that is bit-exact with the reference's float kernels) and encode
what they store. The oracle does the same in three separate pieces: the
codecs of oracle/storage.py (pinned on the CPU by tests/test_storage_formats.py),
the oracle's FP32 alpha / beta (pinned to the compiled reference by the golden
vectors) and the mask of the (slot, cell) pairs a step writes. Codes the step
does not write keep their bits; OBSTACLE cells, whose swap writes back what it
read, keep theirs too. Every code and every rho / u output must match, after
each of several whole-domain and box steps on random states with every flag.
"""
import numpy as np
import pytest

from oracle import storage

pytestmark = pytest.mark.gpu

OBST, FLUID, INJ, GHOST = 1, 2, 4, 8


def say(*a):
    print("[case]", *a, flush=True)


def _codes(g):
    n = int(np.prod(g.size_h))
    raw = g.save_state()
    return raw[: 19 * n * 2].view(np.uint16).reshape(19, n).copy(), raw[19 * n * 2:]


@pytest.mark.parametrize("prec", storage.PRECISIONS)
@pytest.mark.parametrize("size,seed", [((11, 7, 9), 1), ((37, 5, 6), 2), ((1, 1, 1), 3),
                                       ((64, 3, 4), 4)])
def test_shifted_steps_match_oracle_bitwise(hip, oracle_cpu, prec, size, seed):
    say("shifted parity", prec, size, seed)
    rng = np.random.default_rng(seed)
    kw = dict(inv_tau=float(rng.uniform(0.6, 1.9)), lid_velocity=float(rng.uniform(0.01, 0.1)),
              gravitation=tuple(float(v) for v in rng.uniform(-1e-3, 1e-3, 3)))
    o = oracle_cpu.OracleSolver(size, (OBST,) * 6, dtype=np.float32, **kw)
    n = o.n
    # deviations of realistic size, some tiny (subnormal range of F16C)
    x = rng.standard_normal((19, n)) * 4e-3
    x[:, ::3] *= 1e-3
    codes = storage.encode((storage.W[:, None] + x).astype(np.float32), prec)
    o.flags[:] = rng.choice([OBST, FLUID, FLUID, FLUID, INJ, GHOST], size=n).astype(np.int32)
    g = hip.LbmSolverHIP(size, (OBST,) * 6, prec, kw["inv_tau"], kw["lid_velocity"],
                         kw["gravitation"])
    g.load_state(np.concatenate([codes.reshape(-1).view(np.uint8),
                                 o.flags.astype(np.uint8)]))
    full = ((0, 0, 0), g.size_h)
    nh = o.size_h
    boxes = [None, ((1, 0, 1), (max(1, nh[0] - 2), nh[1], max(1, nh[2] - 1))),
             ((0, nh[1] - 1, 0), (nh[0], 1, nh[2]))]
    for step in range(8):
        box = boxes[(step // 2) % len(boxes)]
        phase = step & 1
        args = () if box is None else box
        flag = [False]
        codes = storage.step(o, codes, prec, phase, *(box or (None, None)), overflow=flag)
        (g.simulationStepBeta if phase else g.simulationStepAlpha)(*args)
        got, got_flags = _codes(g)
        assert np.array_equal(got_flags, o.flags.astype(np.uint8))
        bad = np.argwhere(got != codes)
        assert bad.size == 0, (prec, step, box, bad[:5], got[tuple(bad[0])], codes[tuple(bad[0])])
        assert g.getVelocities(*full).tobytes() == o.u.tobytes(), (prec, step, "u")
        assert g.getDensities(*full).tobytes() == o.rho.tobytes(), (prec, step, "rho")
        assert not flag[0]
    g.close()


def test_f16c_overflow_poisons_the_reductions(hip, oracle_cpu):
    """A store out of F16C's range (|f - w| >= 1/4) marks the solver: the
    residual, the box sums and the moments become NaN until reset."""
    say("f16c overflow")
    g = hip.LbmSolverHIP((6, 5, 4), precision="f16c", inv_tau=1.2)
    full = ((0, 0, 0), g.size_h)
    n = int(np.prod(g.size_h))
    assert np.isfinite(g.reduce_box()).all()
    assert g.residual(0) == float("inf")
    f = np.repeat(storage.W[:, None], n, axis=1).astype(np.float32)
    f[3, 17] += np.float32(0.6)                            # cannot be stored
    g.setDensityDistributions(f.ravel(), *full)
    assert np.isnan(g.reduce_box()[:4]).all()
    assert np.isnan(g.residual(0))
    assert np.isnan(g.getMoments(0)).all()
    g.reset()
    assert np.isfinite(g.reduce_box()).all()
    # a step whose collision leaves the range raises it as well: every
    # population storable, the oracle's step then overflows
    f = np.repeat(storage.W[:, None], n, axis=1).astype(np.float32)
    f += np.float32(0.2) * (np.arange(19)[:, None] % 3 == 0)
    g.setDensityDistributions(f.ravel(), *full)
    assert np.isfinite(g.reduce_box()).all()
    o = oracle_cpu.OracleSolver((6, 5, 4), dtype=np.float32, inv_tau=1.2)
    flag = [False]
    storage.step(o, storage.encode(f, "f16c"), "f16c", 0, overflow=flag)
    assert flag[0]
    g.simulationStepAlpha()
    assert np.isnan(g.reduce_box()[0])
    g.close()

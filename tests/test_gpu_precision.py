"""Shifted 16-bit population storage (f_l - w_l in 16 bits, FP32 collision)
against an FP64 run of the same kernels on the same cavity, in the relative L2
norm of the interior velocity field, ||u - u64|| / ||u64||.

The bar for shifted 16-bit storage is 1e-3 on a cavity run to its steady
state: 24^3, Re 100, lid 0.1, stepped until the FP64 run's residual
||u^n - u^(n-100)|| / ||u^n|| falls to 1e-6. F16S and F16C must meet it; F16C
(one more significand bit) must beat F16S. During a transient (32^3 at Re 400
and 48x24x40 at Re 1000, far from steady) rounding differences are carried by
the flow and 16-bit storage does not meet 1e-3; those runs are bounded at
about 1.3x the errors measured on MI355X, to catch regressions, and printed.
The FP64 run itself is bit-exact with the reference <double> kernels
(tests/test_gpu_parity.py), which pins it as the oracle here.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BAR = 1e-3


def run(hip, prec, size, u_lid, re, steps):
    c = hip.CavityController(size, u_lid=u_lid, reynolds=re, precision=prec)
    for _ in range(steps):
        c.computeNextStep()
    c.sync()
    u = c.getVelocities().astype(np.float64)
    c.close()
    return u


def rel_l2(u, u64):
    return float(np.sqrt(((u - u64) ** 2).sum() / (u64 ** 2).sum()))


def test_converged_cavity_error_vs_fp64(hip):
    size, u_lid, re = (24, 24, 24), 0.1, 100.0
    print("[case] converged precision", size, re, flush=True)
    c = hip.CavityController(size, u_lid=u_lid, reynolds=re, precision="f64")
    s = next(iter(c.solvers.values()))
    steps, res = 0, float("inf")
    while res > 1e-6 and steps < 20000:
        for _ in range(100):
            c.computeNextStep()
        steps += 100
        res = s.residual(layout=0)            # after beta: natural slots
    c.sync()
    u64 = c.getVelocities().astype(np.float64)
    c.close()
    print(f"  f64 residual {res:.2e} after {steps} steps", flush=True)
    assert res <= 1e-6
    errs = {}
    for prec in ("f32", "f16s", "f16c", "bf16s"):
        u = run(hip, prec, size, u_lid, re, steps)
        assert np.isfinite(u).all()
        errs[prec] = rel_l2(u, u64)
        print(f"  {prec}: rel L2 {errs[prec]:.3e}  Linf {np.abs(u - u64).max():.3e}", flush=True)
    # measured on MI355X: 2.6e-5, 5.4e-4, 3.8e-4, 1.6e-3 after 4000 steps
    assert errs["f32"] <= 1e-4
    assert errs["f16s"] <= BAR
    assert errs["f16c"] <= BAR
    assert errs["f16c"] < errs["f16s"]
    assert errs["bf16s"] <= 2.5e-3


# measured on MI355X: 9.2e-5 / 1.56e-3 / 1.18e-3 / 3.8e-3 and
# 8.0e-5 / 1.31e-3 / 9.6e-4 / 3.3e-3
TRANSIENT = {((32, 32, 32), 400.0, 2000): {"f32": 1.2e-4, "f16s": 2e-3, "f16c": 1.5e-3,
                                           "bf16s": 5e-3},
             ((48, 24, 40), 1000.0, 1200): {"f32": 1.1e-4, "f16s": 1.7e-3, "f16c": 1.25e-3,
                                            "bf16s": 4.3e-3}}


@pytest.mark.parametrize("case", list(TRANSIENT))
def test_transient_cavity_error_vs_fp64(hip, case):
    size, re, steps = case
    print("[case] transient precision", size, re, steps, flush=True)
    u64 = run(hip, "f64", size, 0.1, re, steps)
    assert np.sqrt((u64 ** 2).sum()) > 0
    errs = {}
    for prec, bound in TRANSIENT[case].items():
        u = run(hip, prec, size, 0.1, re, steps)
        assert np.isfinite(u).all()
        errs[prec] = rel_l2(u, u64)
        print(f"  {prec}: rel L2 {errs[prec]:.3e}  Linf {np.abs(u - u64).max():.3e}", flush=True)
        assert errs[prec] <= bound, (prec, errs[prec])
    assert errs["f32"] < errs["f16c"] < errs["f16s"] < errs["bf16s"]


W = np.array([1 / 18] * 4 + [1 / 36] * 12 + [1 / 18] * 2 + [1 / 3], np.float32)


def _bf16_rne(x):
    """float32 -> bfloat16 bits, round to nearest even (finite inputs)."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


@pytest.mark.parametrize("prec", ["f16s", "bf16s", "f16c"])
def test_shifted_codes_are_exact(hip, prec):
    """The device's 16-bit codes and their decode, bit for bit against numpy:
    F16S code = binary16 RNE of (f - w) * 2^12 (float32 arithmetic), decode =
    float32(code) * 2^-12 + w; BF16S code = bfloat16 RNE of f - w, decode =
    float32(code << 16) + w; F16C as restated in oracle/storage.py."""
    from oracle import storage
    print("[case] shifted codes", prec, flush=True)
    s = hip.LbmSolverHIP((6, 5, 4), precision=prec)
    full = ((0, 0, 0), s.size_h)
    cells = int(np.prod(s.size_h))
    rng = np.random.default_rng(5)
    dev = rng.standard_normal((19, cells)).astype(np.float32) * np.float32(3e-3)
    dev[:, ::7] *= np.float32(1e-3)                      # small deviations too
    f = (W[:, None] + dev).astype(np.float32)
    s.setDensityDistributions(f.ravel(), *full)
    raw = s.save_state()[: 19 * cells * 2].view(np.uint16).reshape(19, cells)
    d = (f - W[:, None]).astype(np.float32)
    if prec == "f16s":
        want = (d * np.float32(4096)).astype(np.float16).view(np.uint16)
        back = (raw.view(np.float16).astype(np.float32) * np.float32(2.0 ** -12)
                + W[:, None]).astype(np.float32)
    elif prec == "f16c":
        want = storage.encode(f, "f16c")
        back = storage.decode(raw, "f16c")
    else:
        want = _bf16_rne(d)
        back = ((raw.astype(np.uint32) << 16).view(np.float32) + W[:, None]).astype(np.float32)
    assert np.array_equal(raw, want)
    got = s.getDensityDistributions(*full).reshape(19, cells)
    assert got.tobytes() == back.tobytes()
    s.close()


def test_shifted_storage_round_trip(hip):
    """Rest equilibrium stores exactly zero deviations; set/get of f values
    round to the 16-bit grid of f - w."""
    print("[case] shifted round trip", flush=True)
    for prec, eps in (("f16s", 2.0 ** -11), ("bf16s", 2.0 ** -8), ("f16c", 2.0 ** -12)):
        s = hip.LbmSolverHIP((6, 5, 4), precision=prec)
        full = ((0, 0, 0), s.size_h)
        d = s.getDensityDistributions(*full).reshape(19, -1)
        w = np.array([1 / 18] * 4 + [1 / 36] * 12 + [1 / 18] * 2 + [1 / 3], np.float32)
        assert np.array_equal(d, np.repeat(w[:, None], d.shape[1], axis=1))
        rng = np.random.default_rng(1)
        dev = (rng.random(d.shape) - 0.5).astype(np.float32) * 0.01
        s.setDensityDistributions((w[:, None] + dev).astype(np.float32).ravel(), *full)
        back = s.getDensityDistributions(*full).reshape(19, -1) - w[:, None]
        # 16-bit rounding of the deviation plus two FP32 roundings near w
        assert np.all(np.abs(back - dev) <= np.abs(dev) * eps + 1e-7)
        s.close()

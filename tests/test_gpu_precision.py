"""Shifted 16-bit population storage (f_l - w_l in FP16 / BF16, FP32 collision)
against an FP64 run of the same kernels on the same cavity.

Tolerances (relative L2 of the interior velocity field vs FP64). They are
set to catch defects (a wrong shift, weight or slot gives errors of order 1),
not to certify accuracy; the measured values are printed:
  F32   <= 1e-3   (unshifted FP32, bit-exact with the reference; ~1e-4 measured)
  F16S  <= 5e-3   (11-bit significand of f - w; ~1.3e-3 to 1.6e-3 measured)
  BF16S <= 5e-2   (8-bit significand; reported)
The FP64 run itself is bit-exact with the reference <double> kernels
(tests/test_gpu_parity.py), which pins it as the oracle here.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {"f32": 1e-3, "f16s": 5e-3, "bf16s": 5e-2}


def run(hip, prec, size, u_lid, re, steps):
    c = hip.CavityController(size, u_lid=u_lid, reynolds=re, precision=prec)
    for _ in range(steps):
        c.computeNextStep()
    c.sync()
    u = c.getVelocities().astype(np.float64)
    rho = c.getDensities().astype(np.float64)
    c.close()
    return u, rho


@pytest.mark.parametrize("size,re,steps", [((32, 32, 32), 400.0, 2000), ((48, 24, 40), 1000.0, 1200)])
def test_shifted_16bit_storage_error_vs_fp64(hip, size, re, steps):
    print("[case] precision", size, re, steps, flush=True)
    u64, r64 = run(hip, "f64", size, 0.1, re, steps)
    norm = np.sqrt((u64 ** 2).sum())
    assert norm > 0
    errs = {}
    for prec in ("f32", "f16s", "bf16s"):
        u, _ = run(hip, prec, size, 0.1, re, steps)
        errs[prec] = np.sqrt(((u - u64) ** 2).sum()) / norm
        linf = np.abs(u - u64).max()
        print(f"  {prec}: rel L2 {errs[prec]:.3e}  Linf {linf:.3e}", flush=True)
        assert np.isfinite(u).all()
        assert errs[prec] <= TOL[prec], (prec, errs[prec])
    assert errs["f32"] < errs["f16s"]


W = np.array([1 / 18] * 4 + [1 / 36] * 12 + [1 / 18] * 2 + [1 / 3], np.float32)


def _bf16_rne(x):
    """float32 -> bfloat16 bits, round to nearest even (finite inputs)."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


@pytest.mark.parametrize("prec", ["f16s", "bf16s"])
def test_shifted_codes_are_exact(hip, prec):
    """The device's 16-bit codes and their decode, bit for bit against numpy:
    F16S code = binary16 RNE of (f - w) * 2^12 (float32 arithmetic), decode =
    float32(code) * 2^-12 + w; BF16S code = bfloat16 RNE of f - w, decode =
    float32(code << 16) + w."""
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
    for prec, eps in (("f16s", 2.0 ** -11), ("bf16s", 2.0 ** -8)):
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

"""The shifted 16-bit storage codecs of the oracle (oracle/storage.py) pinned
against independent definitions, on the CPU: F16S against numpy's binary16
conversion and a float64 model of its rounding, BF16S against float64 rounding
to 8 significant bits, F16C exhaustively over its 65536 codes against the
format's definition from (sign, exponent, significand). The GPU tests use
these codecs as the checker of the kernels' 16-bit arithmetic."""
import numpy as np
import pytest

from oracle import storage

W = storage.W


def _rne_bits(v: np.ndarray, bits: int) -> np.ndarray:
    """Round float64 values to `bits` significant bits, ties to even (normal
    range only)."""
    m, e = np.frexp(v)
    scaled = m * 2.0 ** bits
    return np.ldexp(np.round(scaled), e - bits)   # np.round: half to even


def test_f16c_decode_matches_definition_for_every_code():
    codes = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    x = storage.f16c_x(codes).astype(np.float64)
    want = storage.f16c_value(codes)
    assert np.array_equal(x, want)
    # sign symmetry, monotone magnitudes, the extremes of the format
    assert np.array_equal(x[:32768], -x[32768:])
    assert np.all(np.diff(x[:32768]) > 0)
    assert x[1] == 2.0 ** -28                      # smallest subnormal
    assert x[0x800] == 2.0 ** -17                  # smallest normal
    assert x[0x7FF] == 2047 * 2.0 ** -28           # largest subnormal
    assert x[0x7FFF] == (2 - 2.0 ** -11) * 2.0 ** -3


def test_f16c_encode_round_trips_every_code():
    codes = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    x = storage.f16c_x(codes)
    back, r = storage.f16c_encode_x(x)
    assert np.array_equal(back, codes)
    assert (r < storage.F16C_OVERFLOW).all()


def test_f16c_encode_rounds_to_nearest_even():
    """Every midpoint between two consecutive positive codes goes to the even
    code; the nearest value on either side goes to the nearer code. Below
    2^-17 the encode's scaled value is a float32 subnormal (step 2^-149, i.e.
    2^-40 in x), so the neighbours there are taken 2^-40 away: finer x would
    be rounded twice, but x = f - w of float32 populations is a multiple of
    2^-29 in that range."""
    codes = np.arange(0, 0x7FFF, dtype=np.uint32).astype(np.uint16)
    lo = storage.f16c_value(codes)
    hi = storage.f16c_value(codes + np.uint16(1))
    mid = ((lo + hi) / 2).astype(np.float32)
    assert np.array_equal(mid.astype(np.float64), (lo + hi) / 2)   # exact in float32
    even = np.where(codes % 2 == 0, codes, codes + 1).astype(np.uint16)
    got, _ = storage.f16c_encode_x(mid)
    assert np.array_equal(got, even)
    got_neg, _ = storage.f16c_encode_x(-mid)
    assert np.array_equal(got_neg, even | np.uint16(0x8000))
    sub = codes < 0x800
    below = np.where(sub, mid - np.float32(2.0 ** -40), np.nextafter(mid, np.float32(0)))
    above = np.where(sub, mid + np.float32(2.0 ** -40), np.nextafter(mid, np.float32(1)))
    assert np.array_equal(storage.f16c_encode_x(below)[0], codes)
    assert np.array_equal(storage.f16c_encode_x(above)[0], codes + np.uint16(1))


def test_f16c_overflow_boundary_and_non_finite():
    top = storage.f16c_value(np.array([0x7FFF], np.uint16))[0]
    limit = 0.25                                    # the next binade's first value
    mid = np.float32((top + limit) / 2)
    _, r = storage.f16c_encode_x(np.array([np.nextafter(mid, np.float32(0))], np.float32))
    assert r[0] < storage.F16C_OVERFLOW             # rounds down to the largest code
    for v in (mid, np.float32(0.25), np.float32(3.0), np.float32(np.inf), np.float32(np.nan),
              np.float32(-np.inf)):
        _, r = storage.f16c_encode_x(np.array([v, -v], np.float32))
        assert (r >= storage.F16C_OVERFLOW).all(), v
    flag = [False]
    storage.encode((W[:, None] + np.full((19, 1), 0.3, np.float32)).astype(np.float32), "f16c",
                   flag)
    assert flag[0]


def test_f16c_resolves_the_fp32_step_near_every_weight():
    """Populations f near w_l survive encode -> decode exactly when f - w is a
    multiple of the format's quantum: the subnormal step 2^-28 equals the FP32
    spacing of f near 1/18, so no representable f between w - 2^-17 and
    w + 2^-17 is lost for the axis directions."""
    w = np.float32(1 / 18)
    f = (w + np.arange(-512, 513, dtype=np.float32) * np.float32(2.0 ** -28)).astype(np.float32)
    block = np.repeat(f[None, :], 19, axis=0)
    block[4:] = W[4:, None]
    back = storage.decode(storage.encode(block, "f16c"), "f16c")
    assert np.array_equal(back[0], f)


def test_f16s_matches_numpy_binary16():
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((19, 4000)) * 3e-3).astype(np.float32)
    x[:, ::5] *= np.float32(1e-4)
    f = (W[:, None] + x).astype(np.float32)
    codes = storage.encode(f, "f16s")
    d = (f - W[:, None]).astype(np.float32)
    # float64 model: x * 2^12 rounded to 11 significant bits in binary16's normal range
    scaled = d.astype(np.float64) * 4096.0
    normal = np.abs(scaled) >= 2.0 ** -14
    model = _rne_bits(scaled, 11)
    got = codes.view(np.float16).astype(np.float64)
    assert np.array_equal(got[normal], model[normal])
    back = storage.decode(codes, "f16s")
    assert np.array_equal(back, (got.astype(np.float32) * np.float32(2.0 ** -12)
                                 + W[:, None]).astype(np.float32))


def test_bf16s_matches_float64_rounding():
    rng = np.random.default_rng(4)
    x = (rng.standard_normal((19, 4000)) * 3e-3).astype(np.float32)
    f = (W[:, None] + x).astype(np.float32)
    codes = storage.encode(f, "bf16s")
    d = (f - W[:, None]).astype(np.float32).astype(np.float64)
    model = _rne_bits(d, 8)
    got = (codes.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    assert np.array_equal(got, model)


@pytest.mark.parametrize("prec", storage.PRECISIONS)
def test_written_mask_matches_an_all_ghost_shift(oracle_cpu, prec):
    """An all-GHOST beta is a pure neighbour move: the mask covers every slot
    exactly once and step() equals moving the codes themselves whenever
    decode -> encode is the identity on them."""
    size = (4, 3, 5)
    o = oracle_cpu.OracleSolver(size, (oracle_cpu.GHOST_LAYER,) * 6, dtype=np.float32)
    o.flags[:] = oracle_cpu.GHOST_LAYER
    n = o.n
    rng = np.random.default_rng(9)
    x = (rng.standard_normal((19, n)) * 2e-3).astype(np.float32)
    codes = storage.encode((W[:, None] + x).astype(np.float32), prec)
    codes = storage.encode(storage.decode(codes, prec), prec)   # on the fixed points
    mask = storage.written_mask(o.flags, o.size_h, 1)
    assert mask.all()
    new = storage.step(o, codes, prec, 1)
    nx, ny, _ = o.size_h
    want = np.empty_like(codes)
    for l in range(19):
        opp = l ^ 1 if l < 18 else 18
        delta = storage.C[l, 0] + storage.C[l, 1] * nx + storage.C[l, 2] * nx * ny
        # slot l at id + c_l receives plane opp(l) read at id - c_l
        src = np.roll(codes[opp], delta)
        want[l] = np.roll(src, delta)
    assert np.array_equal(new, want)

"""Numpy restatement of the HIP solver's shifted 16-bit population storage
(TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker, never by the
product).

The reference stores populations as T (float/double) only; these formats are
this project's (storage policies in cavityflow_cuda_amd/csrc/lbm_storage.hpp).
Each stores x = f_l - w_l of plane l, w_l the D3Q19 lattice weight
(common.h:132-160: 1/18 for the six axis directions, 1/36 for the twelve
diagonals, 1/3 at rest), and collides in FP32:

  F16S : IEEE binary16 of x * 2^12 (round to nearest even);
         decode float32(code) * 2^-12 + w with one rounding.
  BF16S: bfloat16 of x (round to nearest even); decode float32(code << 16) + w.
  F16C : 1 sign, 4 exponent, 11 significand bits of x (FP16C, Lehmann et al.,
         Phys. Rev. E 106, 015308, 2022) with e = 1..15 -> 2^(e-18)(1 + m/2^11)
         and e = 0 -> m 2^-28; no infinities or NaN. Encode: |y| = |x| * 2^-109
         in float32 (subnormal results kept), its 12 low bits rounded to nearest
         even in integer arithmetic, the sign of x on top; decode (code << 12
         with the sign kept) * 2^109 + w with one rounding.

Pinning (tests/test_storage_formats.py): F16S and BF16S against numpy's own
float16 and an independent float64 rounding; F16C exhaustively over all 65536
codes against its definition from (sign, e, m), round trips, ties-to-even at
every midpoint, and the overflow boundary.

step(): one alpha or beta step of a state held as 16-bit codes: decode, the
oracle's FP32 step (oracle/lbm_oracle.c, the restatement of CLbmAlphaCPU /
CLbmBetaCPU), encode the slots that step writes. That is what the HIP
kernels compute: the per-cell arithmetic is the FP32 path's, which is
bit-exact with the reference.
"""
from __future__ import annotations

import numpy as np

from . import cpu

W = np.array([1 / 18] * 4 + [1 / 36] * 12 + [1 / 18] * 2 + [1 / 3], np.float32)

# D3Q19 lattice vectors E0..E18 (common.h:29-47)
C = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [1, 1, 0], [-1, -1, 0], [1, -1, 0],
              [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1], [0, 1, 1],
              [0, -1, -1], [0, 1, -1], [0, -1, 1], [0, 0, 1], [0, 0, -1], [0, 0, 0]])

PRECISIONS = ("f16s", "bf16s", "f16c")
F16C_OVERFLOW = 0x08000000


def decode(codes: np.ndarray, prec: str) -> np.ndarray:
    """[19, n] uint16 codes -> float32 populations (plane l of row l)."""
    codes = np.asarray(codes, np.uint16)
    w = W[:, None]
    if prec == "f16s":
        return (codes.view(np.float16).astype(np.float32) * np.float32(2.0 ** -12) + w)
    if prec == "bf16s":
        return (codes.astype(np.uint32) << 16).view(np.float32) + w
    if prec == "f16c":
        return (f16c_x(codes).astype(np.float64) + w.astype(np.float64)).astype(np.float32)
    raise ValueError(prec)


def encode(f: np.ndarray, prec: str, overflow: list | None = None) -> np.ndarray:
    """[19, n] float32 populations -> uint16 codes (overflow[0] set for F16C
    values out of range)."""
    x = (np.asarray(f, np.float32) - W[:, None]).astype(np.float32)
    if prec == "f16s":
        return (x * np.float32(4096)).astype(np.float16).view(np.uint16)
    if prec == "bf16s":
        u = x.view(np.uint32).astype(np.uint64)
        return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    if prec == "f16c":
        code, r = f16c_encode_x(x)
        if overflow is not None and (r >= F16C_OVERFLOW).any():
            overflow[0] = True
        return code
    raise ValueError(prec)


def f16c_x(codes: np.ndarray) -> np.ndarray:
    """F16C codes -> x as float32 (exact), by the decode's bit arithmetic."""
    c = np.asarray(codes, np.uint16)
    y = ((c.view(np.int16).astype(np.int32).astype(np.uint32) << 12) & np.uint32(0x87FFFFFF))
    return (y.view(np.float32).astype(np.float64) * 2.0 ** 109).astype(np.float32)


def f16c_encode_x(x: np.ndarray):
    """float32 x -> (F16C codes, rounded |y| words); word >= F16C_OVERFLOW
    means out of range."""
    x = np.asarray(x, np.float32)
    a = (np.abs(x).astype(np.float64) * 2.0 ** -109).astype(np.float32).view(np.uint32)
    r = a + np.uint32(0x7FF) + ((a >> np.uint32(12)) & np.uint32(1))
    # the sign in bit 31 and the rounded magnitude in bits 16..30: the code is
    # the high half
    hi = (x.view(np.uint32) & np.uint32(0x80000000)) | (r << np.uint32(4))
    return (hi >> np.uint32(16)).astype(np.uint16), r


def f16c_value(codes: np.ndarray) -> np.ndarray:
    """F16C codes -> x in float64 from the definition (sign, e, m)."""
    c = np.asarray(codes, np.uint16).astype(np.int64)
    s = np.where(c >> 15, -1.0, 1.0)
    e = (c >> 11) & 0xF
    m = c & 0x7FF
    mag = np.where(e > 0, np.ldexp(1.0 + m / 2048.0, (e - 18).astype(np.int32)),
                   m * 2.0 ** -28)
    return s * mag


def written_mask(flags: np.ndarray, size_h, phase: int, origin=None, box=None) -> np.ndarray:
    """[19, n] bool: the (slot, cell) pairs a step stores. Alpha: all 19 slots
    of FLUID / VELOCITY_INJECTION cells (CLbmAlphaCPU.cpp:75-494; OBSTACLE and
    GHOST_LAYER store nothing in place). Beta: slot l of cell id + c_l (linear
    wrap, CLbmBetaCPU.cpp:44-50) for every cell id that is not OBSTACLE (whose
    swap writes back what it read, CLbmBetaCPU.cpp:302-311)."""
    nx, ny, nz = size_h
    n = nx * ny * nz
    flags = np.asarray(flags).reshape(-1)
    inbox = np.zeros(n, bool)
    if origin is None:
        inbox[:] = True
    else:
        v = inbox.reshape(nz, ny, nx)
        (ox, oy, oz), (bx, by, bz) = origin, box
        v[oz:oz + bz, oy:oy + by, ox:ox + bx] = True
    mask = np.zeros((19, n), bool)
    if phase == 0:
        cells = inbox & ((flags == cpu.FLUID) | (flags == cpu.VELOCITY_INJECTION))
        mask[:, cells] = True
    else:
        ids = np.nonzero(inbox & (flags != cpu.OBSTACLE))[0]
        for l in range(19):
            delta = C[l, 0] + C[l, 1] * nx + C[l, 2] * nx * ny
            mask[l, (ids + delta) % n] = True
    return mask


def step(solver: "cpu.OracleSolver", codes: np.ndarray, prec: str, phase: int, origin=None,
         box=None, overflow: list | None = None) -> np.ndarray:
    """One step of a shifted-storage state: the oracle FP32 solver holds the
    flags, parameters and rho/u outputs; `codes` [19, n] are the stored
    populations. Returns the new codes (solver.dd holds the FP32 result)."""
    assert solver.dtype == np.float32
    solver.dd[:] = decode(codes, prec).reshape(-1)
    args = () if origin is None else (origin, box)
    (solver.beta if phase else solver.alpha)(*args)
    mask = written_mask(solver.flags, solver.size_h, phase, origin, box)
    new = encode(solver.dd.reshape(19, -1), prec, overflow)
    return np.where(mask, new, codes).astype(np.uint16)

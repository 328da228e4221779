"""Seeded synthetic inputs and the benchmark workloads.

Inputs come from splitmix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014): output k of seed s mixes s + (k+1)*0x9E3779B97F4A7C15.  The
stream is position-addressable, so any slice of a large batch can be regenerated on its
own (the parity tests check slices of full-size batches that way).

    uniform_pm1   k*2^-23 - 1 for k = top 24 bits: uniform on [-1, 1), exact in float32
    q15_uniform   top 15 bits - 16384: integers uniform on [-16384, 16383]
    normal        Box-Muller in FP64 from two outputs, rounded once to float32
"""
from __future__ import annotations

import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * _G
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform_pm1(seed: int, count: int, offset: int = 0) -> np.ndarray:
    z = splitmix64(seed, count, offset) >> np.uint64(40)
    return (z.astype(np.float64) * (1.0 / 8388608.0) - 1.0).astype(np.float32)


def q15_uniform(seed: int, count: int, offset: int = 0) -> np.ndarray:
    z = splitmix64(seed, count, offset) >> np.uint64(49)
    return (z.astype(np.int64) - 16384).astype(np.int16)


def normal(seed: int, count: int, offset: int = 0) -> np.ndarray:
    z = splitmix64(seed, 2 * count, 2 * offset)
    u1 = ((z[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = (z[1::2] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return (np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)).astype(np.float32)


def complex_pm1(seed: int, n: int, howmany: int, first: int = 0) -> np.ndarray:
    """(howmany, n, 2) float32: re, im i.i.d. uniform on [-1, 1); items first..first+howmany."""
    v = uniform_pm1(seed, 2 * n * howmany, 2 * n * first)
    return v.reshape(howmany, n, 2)


def q15_complex(seed: int, n: int, howmany: int, first: int = 0) -> np.ndarray:
    v = q15_uniform(seed, 2 * n * howmany, 2 * n * first)
    return v.reshape(howmany, n, 2)


def two_tone(seed: int, n: int, howmany: int, first: int = 0) -> np.ndarray:
    """(howmany, n) float32: N(0,1) noise plus cos tones at bins 100 and 1500.5 (amplitude 1)."""
    noise = normal(seed, n * howmany, n * first).reshape(howmany, n).astype(np.float64)
    t = np.arange(n, dtype=np.float64)
    tones = np.cos(2 * np.pi * 100.0 * t / n) + np.cos(2 * np.pi * 1500.5 * t / n)
    return (noise + tones[None, :]).astype(np.float32)


# The reference's benchmark workloads, one GPU each.
CONFIGS = {
    "c2c_1200": dict(kind="c2c", n=1200, batch=256, seed=1, dtype="float"),
    "r2c_4096": dict(kind="r2c", n=4096, batch=65536, seed=2, dtype="float"),
    "r2c_4000": dict(kind="r2c", n=4000, batch=65536, seed=2, dtype="float"),
    "q15_512": dict(kind="c2c", n=512, batch=1 << 20, seed=3, dtype="int16_t"),
    "nd_64x96x120": dict(kind="nd", dims=(64, 96, 120), batch=16, seed=4, dtype="float"),
    "ndr_1024x1000": dict(kind="ndr", dims=(1024, 1000), batch=32, seed=5, dtype="float"),
}


def algorithmic_bytes(kind: str, n_or_dims, scalar_bytes: int = 4) -> int:
    """One read of the input plus one write of the output, per transform."""
    if kind in ("c2c", "nd"):
        tot = int(np.prod(n_or_dims))
        return 2 * 2 * scalar_bytes * tot
    dims = (n_or_dims,) if np.isscalar(n_or_dims) else tuple(n_or_dims)
    tot = int(np.prod(dims))
    bins = tot // dims[-1] * (dims[-1] // 2 + 1)
    return scalar_bytes * tot + 2 * scalar_bytes * bins

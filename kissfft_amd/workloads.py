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


def generate(fn, seed: int, n: int, howmany: int, workers: int = 8, chunk: int = 8192) -> np.ndarray:
    """fn(seed, n, howmany) for large batches: slices of `chunk` items generated side by side
    (the stream is position-addressable, so every slice is regenerated on its own and the
    result equals one call of fn)."""
    import concurrent.futures

    first = fn(seed, n, min(chunk, howmany), 0)
    out = np.empty((howmany,) + first.shape[1:], first.dtype)
    out[:first.shape[0]] = first

    def one(b0: int) -> None:
        m = min(chunk, howmany - b0)
        out[b0:b0 + m] = fn(seed, n, m, b0)

    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, workers)) as pool:
        list(pool.map(one, range(first.shape[0], howmany, chunk)))
    return out


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


# Serving stream: Poisson arrivals of mixed requests.
SERVE_SIZES = (256, 1000, 1009, 3072, 65536, 1048576)
SERVE_KINDS = ("c2c", "r2c", "inverse")
SERVE_MAX_BATCH = 4096


def serving_stream(count: int, rate_per_s: float, seed: int = 6) -> list[dict]:
    """`count` requests: arrival times (Poisson process of `rate_per_s`), kind, n, batch.

    kind is uniform over c2c / r2c / inverse; "inverse" is the real inverse (c2r), except at
    odd n (1009), where neither real kind exists (real transforms need even n) and r2c maps
    to the complex forward, inverse to the complex inverse.  n is uniform over SERVE_SIZES;
    batch = floor(exp(U(0, ln 4096))), log-uniform in [1, 4096].  Uniforms come from
    splitmix64 (four per request)."""
    z = splitmix64(seed, 4 * count)
    u = ((z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53).reshape(count, 4)
    out, t = [], 0.0
    for i in range(count):
        t += -np.log1p(-u[i, 0]) / rate_per_s
        kind = SERVE_KINDS[min(2, int(u[i, 1] * 3))]
        n = SERVE_SIZES[min(5, int(u[i, 2] * 6))]
        if n % 2:
            kind = {"c2c": "c2c", "r2c": "c2c", "inverse": "c2c_inv"}[kind]
        elif kind == "inverse":
            kind = "c2r"
        batch = max(1, min(SERVE_MAX_BATCH, int(np.exp(u[i, 3] * np.log(SERVE_MAX_BATCH)))))
        out.append(dict(t=t, kind=kind, n=n, batch=batch))
    return out


def request_bytes(kind: str, n: int, batch: int, scalar_bytes: int = 4) -> tuple[int, int]:
    """(input bytes, output bytes) of one dense request."""
    cpx, real, half = 2 * scalar_bytes * n, scalar_bytes * n, 2 * scalar_bytes * (n // 2 + 1)
    i, o = {"c2c": (cpx, cpx), "c2c_inv": (cpx, cpx), "r2c": (real, half), "c2r": (half, real)}[kind]
    return i * batch, o * batch

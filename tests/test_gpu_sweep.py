"""GPU parity over a sweep of lengths, fixed point and float, complex and real.

Every length from 1 to 72 (complex) and every even length from 2 to 144 (real) plus lengths
with mixed and generic radices, each through one batched call of 3 transforms in both
directions: Q15 and Q31 bit-exact against the oracle (itself pinned bit-exact to the
reference build), float within the float bar of FP64.  The sweep reaches every kernel
family a length can fall to: the general line kernel (any radix, kf_bfly_generic for 7,
11, 13, ...), the compiled registry, and runtime-compiled kernels from 128 points up (144,
160, 400, 1331, 2310; KISSFFT_AMD_RTC_MIN=128 keeps the shorter lines on the general
kernel here).  test_short_lengths_compiled runs short lengths through the runtime-compiled
kernels they take by default (from 16 points).
"""
import sys

import numpy as np
import pytest

import kissfft_amd
from helpers import ABS_FLOAT_BOUND, cplx, rel_l2
from oracle import Oracle

pytestmark = pytest.mark.gpu

C2C_SIZES = list(range(1, 73)) + [77, 91, 96, 100, 121, 125, 143, 144, 160, 169, 243, 400, 1331, 2310]
REAL_SIZES = list(range(2, 145, 2)) + [154, 182, 200, 250, 338, 400, 486, 2310]
B = 3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rnd(rng, dtype, shape):
    if dtype == "float":
        return rng.uniform(-1, 1, shape).astype(np.float32)
    if dtype == "int16_t":
        return rng.integers(-32768, 32768, shape).astype(np.int16)
    return rng.integers(-(1 << 31), 1 << 31, shape, dtype=np.int64).astype(np.int32)


@pytest.fixture
def general_short(monkeypatch):
    monkeypatch.setenv("KISSFFT_AMD_RTC_MIN", "128")


@pytest.mark.parametrize("dtype", ["int16_t", "int32_t", "float"])
def test_c2c_length_sweep(dtype, general_short):
    lib = kissfft_amd.load(dtype)
    O = Oracle(dtype) if dtype != "float" else None
    rng = np.random.default_rng(71)
    worst = 0.0
    for n in C2C_SIZES:
        x = rnd(rng, dtype, (B, n, 2))
        for inv in (False, True):
            say("case sweep c2c", dtype, n, "inverse", inv)
            y = np.zeros_like(x)
            lib.fft_batch(lib.fft_alloc(n, inv), x, y, B, n, n)
            if O is not None:
                want = O.fft(x, inv)
                assert np.array_equal(y, want), f"{dtype} n={n} inverse={inv}: differs from the oracle"
            else:
                want = np.fft.ifft(cplx(x), axis=-1) * n if inv else np.fft.fft(cplx(x), axis=-1)
                err = rel_l2(cplx(y), want)
                worst = max(worst, err)
                assert err <= ABS_FLOAT_BOUND, (n, inv, err)
    say("worst float rel-L2", worst)


@pytest.mark.parametrize("dtype", ["int16_t", "int32_t", "float"])
def test_real_length_sweep(dtype, general_short):
    lib = kissfft_amd.load(dtype)
    O = Oracle(dtype) if dtype != "float" else None
    rng = np.random.default_rng(72)
    worst = 0.0
    for n in REAL_SIZES:
        say("case sweep real", dtype, n)
        x = rnd(rng, dtype, (B, n))
        X = np.zeros((B, n // 2 + 1, 2), lib.scalar)
        lib.fftr_batch(lib.fftr_alloc(n, False), x, X, B, n, n // 2 + 1)
        Z = rnd(rng, dtype, (B, n // 2 + 1, 2))
        z = np.zeros((B, n), lib.scalar)
        lib.fftri_batch(lib.fftr_alloc(n, True), Z, z, B, n // 2 + 1, n)
        if O is not None:
            assert np.array_equal(X, O.fftr(x)), f"{dtype} n={n}: forward differs from the oracle"
            assert np.array_equal(z, O.fftri(Z, n)), f"{dtype} n={n}: inverse differs from the oracle"
        else:
            err = rel_l2(cplx(X), np.fft.rfft(x.astype(np.float64), axis=-1))
            # kiss_fftri ignores the imaginary parts of the DC and Nyquist bins (kiss_fftr.c:131-134)
            zc = cplx(Z)
            zc[..., 0] = zc[..., 0].real
            zc[..., -1] = zc[..., -1].real
            erri = rel_l2(z.astype(np.float64), np.fft.irfft(zc, n, axis=-1) * n)
            worst = max(worst, err, erri)
            assert err <= ABS_FLOAT_BOUND and erri <= ABS_FLOAT_BOUND, (n, err, erri)
    say("worst float rel-L2", worst)


SHORT_C2C = [16, 20, 36, 49, 60, 77, 100, 121, 126]
SHORT_REAL = [32, 40, 98, 120, 154]


@pytest.mark.parametrize("dtype", ["int16_t", "int32_t", "float"])
def test_short_lengths_compiled(dtype, monkeypatch, capfd):
    """Short lines (16 to 127 points) on the kernels compiled at run time by default:
    bit-exact against the oracle in fixed point, within the float bar of FP64."""
    monkeypatch.setenv("KISSFFT_AMD_VERBOSE", "1")
    lib = kissfft_amd.load(dtype)
    lib.L.kiss_fft_cleanup()  # kernel choices made afresh (and reported) for these lengths
    O = Oracle(dtype) if dtype != "float" else None
    rng = np.random.default_rng(73)
    for n in SHORT_C2C:
        x = rnd(rng, dtype, (17, n, 2))
        for inv in (False, True):
            say("case short c2c", dtype, n, "inverse", inv)
            y = np.zeros_like(x)
            lib.fft_batch(lib.fft_alloc(n, inv), x, y, 17, n, n)
            if O is not None:
                assert np.array_equal(y, O.fft(x, inv)), f"{dtype} n={n} inverse={inv}"
            else:
                want = np.fft.ifft(cplx(x), axis=-1) * n if inv else np.fft.fft(cplx(x), axis=-1)
                assert rel_l2(cplx(y), want) <= ABS_FLOAT_BOUND, (n, inv)
    for n in SHORT_REAL:
        say("case short real", dtype, n)
        x = rnd(rng, dtype, (17, n))
        X = np.zeros((17, n // 2 + 1, 2), lib.scalar)
        lib.fftr_batch(lib.fftr_alloc(n, False), x, X, 17, n, n // 2 + 1)
        if O is not None:
            assert np.array_equal(X, O.fftr(x)), f"{dtype} n={n}"
        else:
            assert rel_l2(cplx(X), np.fft.rfft(x.astype(np.float64), axis=-1)) <= ABS_FLOAT_BOUND, n
    err = capfd.readouterr().err
    assert "fast variant rtc_c2c60" in err and "fast variant rtc_r2c49" in err, err[-2000:]

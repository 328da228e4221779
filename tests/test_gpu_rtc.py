"""GPU tests of runtime-compiled row kernels (kfa_rtc.hip) for lengths outside the
built-in registry, and of the persisted table of tuned choices.

Lengths whose radices are 2, 3, 4, 5, 7, 11, 13 and whose line fits one LDS line get the
compile-time Stockham kernel k_fast compiled on first use (fixed point with the reference
schedule of kf_factor, kiss_fft.c:306-328, and kf_bfly_generic's arithmetic for 7, 11
and 13, kiss_fft.c:191-233).  Results: bit-exact against the oracle for Q15 / Q31,
within the float bar of FP64 for float; the engine's own report (KISSFFT_AMD_VERBOSE=1)
shows that the runtime-compiled variant ran.
"""
import os
import sys

import numpy as np
import pytest

import kissfft_amd
from helpers import ABS_FLOAT_BOUND, cplx, rel_l2
from oracle import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture
def rtc_env(monkeypatch, tmp_path):
    monkeypatch.setenv("KISSFFT_AMD_CACHE", str(tmp_path))
    monkeypatch.setenv("KISSFFT_AMD_TUNE_TABLE", str(tmp_path / "tuned.txt"))
    monkeypatch.setenv("KISSFFT_AMD_VERBOSE", "1")
    return tmp_path


def rnd(rng, dtype, shape):
    if dtype == "float":
        return rng.uniform(-1, 1, shape).astype(np.float32)
    if dtype == "int16_t":
        return rng.integers(-16384, 16384, shape).astype(np.int16)
    return rng.integers(-(1 << 30), 1 << 30, shape).astype(np.int32)


# (lengths no earlier test file compiles: the test checks that this test's compile
# reached the disk cache, and a kernel another test loaded is not compiled again; and
# lengths whose kernels fit the registers: Q31 3003 = 13*11*7*3 spills at every thread
# shape, the widest included, and stays on the general kernel)
@pytest.mark.parametrize("dtype,n", [("float", 1536), ("float", 2187), ("float", 3000), ("float", 6561),
                                     ("float", 7680), ("float", 2401), ("float", 3003), ("float", 1331),
                                     ("float", 2197), ("float", 7 * 11 * 13 * 5), ("float", 784),
                                     ("int16_t", 1536), ("int16_t", 3000), ("int16_t", 6144),
                                     ("int16_t", 3003), ("int16_t", 2401), ("int16_t", 1331),
                                     ("int32_t", 1536), ("int32_t", 2187), ("int32_t", 1331), ("int32_t", 2401)])
def test_rtc_c2c(rtc_env, capfd, dtype, n):
    lib = kissfft_amd.load(dtype)
    lib.L.kiss_fft_cleanup()
    rng = np.random.default_rng(n)
    b = 37
    x = rnd(rng, dtype, (b, n, 2))
    for inv in (False, True):
        say("case rtc", dtype, n, "inverse", inv)
        y = np.zeros_like(x)
        lib.fft_batch(lib.fft_alloc(n, inv), x, y, b, n, n)
        if dtype == "float":
            want = np.fft.ifft(cplx(x), axis=-1) * n if inv else np.fft.fft(cplx(x), axis=-1)
            err = rel_l2(cplx(y), want)
            say("rel-L2", err)
            assert err <= ABS_FLOAT_BOUND
        else:
            np.testing.assert_array_equal(y, Oracle(dtype).fft(x, inv))
    # every other input element (in_stride 2) and output rows n + 3 apart: the kernel's
    # general-stride body (fixed point otherwise takes the unit-stride one)
    xs = rnd(rng, dtype, (5, 2 * n, 2))
    ys = np.zeros((5, n + 3, 2), lib.scalar)
    lib.fft_batch(lib.fft_alloc(n), xs, ys, 5, 2 * n, n + 3, in_stride=2)
    xd = np.ascontiguousarray(xs[:, ::2])
    if dtype == "float":
        assert rel_l2(cplx(ys[:, :n]), np.fft.fft(cplx(xd), axis=-1)) <= ABS_FLOAT_BOUND
    else:
        np.testing.assert_array_equal(ys[:, :n], Oracle(dtype).fft(xd))
    assert (ys[:, n:] == 0).all()
    out = capfd.readouterr().err
    assert f"n={n} kind=0 inverse=0: fast variant rtc_c2c{n}" in out, out[-2000:]
    assert f"n={n} kind=0 inverse=1: fast variant rtc_c2c{n}" in out, out[-2000:]
    assert any(f.name.startswith("rtc-") for f in rtc_env.iterdir())  # code objects cached


@pytest.mark.parametrize("dtype,n", [("float", 3072 * 2), ("float", 6000), ("float", 2 * 2401), ("float", 2002),
                                     ("int16_t", 3000), ("int16_t", 2002), ("int32_t", 2 * 2401)])
def test_rtc_real(rtc_env, capfd, dtype, n):
    lib = kissfft_amd.load(dtype)
    lib.L.kiss_fft_cleanup()
    rng = np.random.default_rng(n + 1)
    b = 9
    x = rnd(rng, dtype, (b, n))
    X = np.zeros((b, n // 2 + 1, 2), lib.scalar)
    lib.fftr_batch(lib.fftr_alloc(n), x, X, b, n, n // 2 + 1)
    z = np.zeros_like(x)
    lib.fftri_batch(lib.fftr_alloc(n, True), X, z, b, n // 2 + 1, n)
    say("case rtc real", dtype, n)
    if dtype == "float":
        assert rel_l2(cplx(X), np.fft.rfft(x.astype(np.float64), axis=-1)) <= ABS_FLOAT_BOUND
        assert rel_l2(z, x.astype(np.float64) * n) <= ABS_FLOAT_BOUND
    else:
        O = Oracle(dtype)
        np.testing.assert_array_equal(X, O.fftr(x))
        np.testing.assert_array_equal(z, O.fftri(X, n))
    out = capfd.readouterr().err
    assert f"n={n // 2} kind=1 inverse=0: fast variant rtc_r2c{n // 2}" in out, out[-2000:]
    assert f"n={n // 2} kind=2 inverse=1: fast variant rtc_c2r{n // 2}" in out, out[-2000:]


def test_tuned_table_persists(rtc_env, capfd):
    """The first use times the candidates and records the choice; after a cleanup the same
    decision comes from the table without timing."""
    lib = kissfft_amd.load("float")
    lib.L.kiss_fft_cleanup()
    n, b = 3600, 64  # a length no other test of this process has decided
    x = np.random.default_rng(5).uniform(-1, 1, (b, n, 2)).astype(np.float32)
    y = np.zeros_like(x)
    lib.fft_batch(lib.fft_alloc(n), x, y, b, n, n)
    first = capfd.readouterr().err
    assert "candidate" in first  # timed
    table = (rtc_env / "tuned.txt").read_text()
    assert f"n={n} inverse=0" in table
    lib.L.kiss_fft_cleanup()
    y2 = np.zeros_like(x)
    lib.fft_batch(lib.fft_alloc(n), x, y2, b, n, n)
    second = capfd.readouterr().err
    assert "candidate" not in second  # taken from the table
    np.testing.assert_array_equal(y, y2)


def test_rtc_off_keeps_general_kernel(rtc_env, capfd, monkeypatch):
    monkeypatch.setenv("KISSFFT_AMD_RTC", "0")
    lib = kissfft_amd.load("float")
    lib.L.kiss_fft_cleanup()
    n = 1536
    x = np.random.default_rng(6).uniform(-1, 1, (5, n, 2)).astype(np.float32)
    y = np.zeros_like(x)
    lib.fft_batch(lib.fft_alloc(n), x, y, 5, n, n)
    assert rel_l2(cplx(y), np.fft.fft(cplx(x), axis=-1)) <= ABS_FLOAT_BOUND
    assert "rtc_" not in capfd.readouterr().err
    lib.L.kiss_fft_cleanup()

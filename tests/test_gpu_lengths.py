"""GPU parity of lengths whose radix factors do not fit one LDS line.

kf_factor (kiss_fft.c:306-328) takes a cofactor above sqrt(n) whole, so a length with a
large prime factor p has a p-point stage (kf_bfly_generic, kiss_fft.c:191-233).  Above one
LDS line (8192 points for float and Q31, 16384 for Q15) that stage runs as a global-memory
pass and the remaining stages as a second pass.  Fixed point is bit-exact against the
oracle (itself pinned to the reference build); float is within the float bar of FP64.
N-D leading axes longer than one LDS line run their strided lines through the multi-pass
path.
"""
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


def rnd(rng, dtype, shape):
    if dtype == "float":
        return rng.uniform(-1, 1, shape).astype(np.float32)
    if dtype == "int16_t":
        return rng.integers(-16384, 16384, shape).astype(np.int16)
    return rng.integers(-(1 << 30), 1 << 30, shape).astype(np.int32)


@pytest.mark.parametrize("dtype,n,b", [("float", 10007, 3), ("float", 20014, 2), ("float", 8209 * 3, 2),
                                       ("int32_t", 10007, 2), ("int32_t", 20014, 2),
                                       ("int16_t", 16411, 2), ("int16_t", 32822, 2), ("int16_t", 16411 * 3, 1)])
def test_large_prime_factor_c2c(dtype, n, b):
    lib = kissfft_amd.load(dtype)
    rng = np.random.default_rng(n)
    x = rnd(rng, dtype, (b, n, 2))
    for inv in (False, True):
        say("case", dtype, n, "inverse", inv)
        y = np.zeros_like(x)
        lib.fft_batch(lib.fft_alloc(n, inv), x, y, b, n, n)
        if dtype == "float":
            want = np.fft.ifft(cplx(x), axis=-1) * n if inv else np.fft.fft(cplx(x), axis=-1)
            err = rel_l2(cplx(y), want)
            say("rel-L2", err)
            assert err <= ABS_FLOAT_BOUND
        else:
            np.testing.assert_array_equal(y, Oracle(dtype).fft(x, inv))
        # the reference-shaped call on one transform, in place on device memory
        import torch
        d = torch.from_numpy(np.ascontiguousarray(x[0])).cuda()
        lib.kiss_fft(lib.fft_alloc(n, inv), d, d)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d.cpu().numpy(), y[0])


@pytest.mark.parametrize("dtype,n", [("int16_t", 2 * 16411), ("float", 2 * 10007)])
def test_large_prime_factor_real(dtype, n):
    lib = kissfft_amd.load(dtype)
    rng = np.random.default_rng(n + 1)
    x = rnd(rng, dtype, (2, n))
    X = np.zeros((2, n // 2 + 1, 2), lib.scalar)
    lib.fftr_batch(lib.fftr_alloc(n), x, X, 2, n, n // 2 + 1)
    z = np.zeros_like(x)
    lib.fftri_batch(lib.fftr_alloc(n, True), X, z, 2, n // 2 + 1, n)
    say("case real", dtype, n)
    if dtype == "float":
        assert rel_l2(cplx(X), np.fft.rfft(x.astype(np.float64), axis=-1)) <= ABS_FLOAT_BOUND
        assert rel_l2(z, x.astype(np.float64) * n) <= ABS_FLOAT_BOUND
    else:
        O = Oracle(dtype)
        np.testing.assert_array_equal(X, O.fftr(x))
        np.testing.assert_array_equal(z, O.fftri(X, n))


@pytest.mark.parametrize("dtype,dims,b", [("float", (16384, 4), 2), ("float", (3, 20000, 5), 2),
                                          ("int16_t", (32768, 4), 2), ("int32_t", (16384, 3), 2),
                                          ("int16_t", (2, 16411, 2), 1)])
def test_nd_leading_axis_longer_than_a_line(dtype, dims, b):
    lib = kissfft_amd.load(dtype)
    rng = np.random.default_rng(sum(dims))
    x = rnd(rng, dtype, (b,) + dims + (2,))
    for inv in (False, True):
        say("case nd", dtype, dims, "inverse", inv)
        y = np.zeros_like(x)
        lib.fftnd_batch(lib.fftnd_alloc(dims, inv), x, y, b)
        if dtype == "float":
            axes = tuple(range(1, len(dims) + 1))
            want = np.fft.ifftn(cplx(x), axes=axes) * np.prod(dims) if inv else np.fft.fftn(cplx(x), axes=axes)
            assert rel_l2(cplx(y), want) <= ABS_FLOAT_BOUND
        else:
            np.testing.assert_array_equal(y, Oracle(dtype).fftnd(x, dims, inv))


@pytest.mark.parametrize("dims,b", [((16384, 512), 1), ((96, 600), 3), ((5, 1000), 2)])
def test_ndr_inverse_leading_axis_lengths(dims, b):
    """The inverse real 2-D transform with a leading axis longer than one LDS line (the
    multi-pass path into the inverse's temporary), one within a line and a short one;
    against numpy, forward then inverse returns the input times the size, and the inverse
    leaves its input untouched."""
    lib = kissfft_amd.load("float")
    rng = np.random.default_rng(sum(dims) + b)
    x = rng.standard_normal((b,) + dims).astype(np.float32)
    X = np.zeros((b, dims[0], dims[1] // 2 + 1, 2), np.float32)
    lib.fftndr_batch(lib.fftndr_alloc(dims), x, X, b)
    Z = X.copy()
    z = np.zeros_like(x)
    lib.fftndri_batch(lib.fftndr_alloc(dims, True), Z, z, b)
    say("case ndri", dims, b)
    want = np.fft.irfftn(cplx(X), dims, axes=(1, 2)) * np.prod(dims)
    assert rel_l2(z.astype(np.float64), want) <= ABS_FLOAT_BOUND
    assert rel_l2(z.astype(np.float64), x.astype(np.float64) * np.prod(dims)) <= 2 * ABS_FLOAT_BOUND
    np.testing.assert_array_equal(Z, X)  # the input is left untouched

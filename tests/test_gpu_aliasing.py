"""GPU tests of overlapping input and output buffers.

The reference accepts in-place calls (kiss_fft.c:377-395 copies through a temporary; the
real and N-D transforms go through cfg->tmpbuf, kiss_fftndr.c:45-46).  Device buffers that
overlap must give the same result as separate buffers: identical spans of one-pass lines
run in place, every other overlap runs through a temporary.
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
    return rng.integers(-16384, 16384, shape).astype(np.int16)


@pytest.mark.parametrize("dtype", ["float", "int16_t"])
def test_ndr_forward_in_place(dtype):
    import torch
    lib = kissfft_amd.load(dtype)
    rng = np.random.default_rng(81)
    dims = (64, 64)
    x = rnd(rng, dtype, (3,) + dims)
    want = np.zeros((3, 64, 33, 2), lib.scalar)
    lib.fftndr_batch(lib.fftndr_alloc(dims), x, want, 3)
    if dtype == "int16_t":
        np.testing.assert_array_equal(want, Oracle(dtype).fftndr(x, dims))
    else:
        assert rel_l2(cplx(want), np.fft.rfftn(x.astype(np.float64), axes=(1, 2))) <= ABS_FLOAT_BOUND
    tdt = torch.float32 if dtype == "float" else torch.int16
    buf = torch.zeros(want.size, dtype=tdt, device="cuda")
    buf[:x.size] = torch.from_numpy(x.ravel()).cuda()
    say("case ndr in place", dtype)
    lib.fftndr_batch(lib.fftndr_alloc(dims), buf, buf, 3)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(buf.cpu().numpy().reshape(want.shape), want)
    # the reference-shaped single call, in place
    one = torch.zeros(64 * 33 * 2, dtype=tdt, device="cuda")
    one[:64 * 64] = torch.from_numpy(x[1].ravel()).cuda()
    lib.kiss_fftndr(lib.fftndr_alloc(dims), one, one)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(one.cpu().numpy().reshape(64, 33, 2), want[1])


@pytest.mark.parametrize("dtype,n", [("float", 4096), ("float", 40000), ("int16_t", 512)])
def test_real_batch_in_place(dtype, n):
    import torch
    lib = kissfft_amd.load(dtype)
    rng = np.random.default_rng(n)
    b, nb = 5, n // 2 + 1
    x = rnd(rng, dtype, (b, n))
    want = np.zeros((b, nb, 2), lib.scalar)
    lib.fftr_batch(lib.fftr_alloc(n), x, want, b, n, nb)
    tdt = torch.float32 if dtype == "float" else torch.int16
    # padded rows (2 * nb scalars) and dense rows (n scalars) in the output's buffer
    for idist in (2 * nb, n):
        say("case real in place", dtype, n, "idist", idist)
        buf = torch.zeros(b * nb * 2, dtype=tdt, device="cuda")
        for r in range(b):
            buf[r * idist:r * idist + n] = torch.from_numpy(x[r]).cuda()
        lib.fftr_batch(lib.fftr_alloc(n), buf, buf, b, idist, nb)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(buf.cpu().numpy().reshape(b, nb, 2), want)
        # and back: the inverse in place into dense real rows
        lib.fftri_batch(lib.fftr_alloc(n, True), buf, buf, b, nb, n)
        torch.cuda.synchronize()
        back = np.zeros((b, n), lib.scalar)
        lib.fftri_batch(lib.fftr_alloc(n, True), want, back, b, nb, n)
        np.testing.assert_array_equal(buf.cpu().numpy()[:b * n].reshape(b, n), back)


@pytest.mark.parametrize("dtype,n", [("float", 1200), ("int16_t", 512), ("float", 65536), ("int16_t", 16411)])
def test_c2c_partial_overlap(dtype, n):
    import torch
    lib = kissfft_amd.load(dtype)
    rng = np.random.default_rng(n + 3)
    b = 3 if n < 10000 else 2
    x = rnd(rng, dtype, (b, n, 2))
    want = np.zeros_like(x)
    lib.fft_batch(lib.fft_alloc(n), x, want, b, n, n)
    if dtype == "int16_t":
        np.testing.assert_array_equal(want, Oracle(dtype).fft(x))
    tdt = torch.float32 if dtype == "float" else torch.int16
    for shift in (7, -5, 0):
        say("case c2c overlap", dtype, n, "shift", shift)
        buf = torch.zeros((b * n + 16) * 2, dtype=tdt, device="cuda")
        i0 = 8
        o0 = i0 + shift
        buf[2 * i0:2 * (i0 + b * n)] = torch.from_numpy(x.ravel()).cuda()
        lib.fft_batch(lib.fft_alloc(n), buf[2 * i0:], buf[2 * o0:], b, n, n)
        torch.cuda.synchronize()
        got = buf.cpu().numpy()[2 * o0:2 * (o0 + b * n)].reshape(b, n, 2)
        np.testing.assert_array_equal(got, want)
    # strided input in place, output rows further apart than their length (gaps kept)
    buf = torch.zeros(b * (n + 5) * 2 + 2 * n, dtype=tdt, device="cuda")
    flat = np.zeros((b * (n + 5) + n, 2), lib.scalar)
    for r in range(b):
        flat[r * (n + 5):r * (n + 5) + n] = x[r]
    flat[b * (n + 5) - 5:b * (n + 5)] = 3  # the gap after the last row keeps its values
    buf.copy_(torch.from_numpy(flat.ravel()))
    lib.fft_batch(lib.fft_alloc(n), buf, buf, b, n + 5, n + 5)
    torch.cuda.synchronize()
    got = buf.cpu().numpy().reshape(-1, 2)
    for r in range(b):
        np.testing.assert_array_equal(got[r * (n + 5):r * (n + 5) + n], want[r])
    np.testing.assert_array_equal(got[b * (n + 5) - 5:b * (n + 5)], flat[b * (n + 5) - 5:b * (n + 5)])

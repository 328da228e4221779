"""GPU tests of fast FIR filtering by overlap-scrap (include/kiss_fastfir.h).

Streaming a signal through kiss_fastfir in pieces and flushing yields exactly the 'valid'
part of the linear convolution: with the rotated impulse response of
tools/kiss_fastfir.c:137-141, good sample t of block k is sum_i h[i] x[k*ngood+t+nh-1-i].
The reference is numpy's FP64 convolution.  The tolerance is the float bar of the other
transforms: rel-L2 <= min(1e-6, 2x the reference float build's own error on the same
input), that error measured by restating the reference's float filter (rotated impulse
response, its spectrum times a float 1/nfft, per block FFT, C_MUL, inverse FFT,
tools/kiss_fastfir.c:137-176) with the oracle's float transforms.
"""
import sys

import numpy as np
import pytest

import kissfft_amd
from helpers import rel_l2

from oracle import Oracle

pytestmark = pytest.mark.gpu


def reference_float_filter(h, x, nfft, cpx):
    """The reference's float overlap-scrap filter (tools/kiss_fastfir.c:137-176) on every
    complete block of x, with the oracle's float transforms: the 'valid' convolution."""
    O = Oracle("float")
    nh = h.shape[0]
    ngood = nfft - nh + 1
    shape = (nfft, 2) if cpx else (nfft,)
    tmp = np.zeros(shape, np.float32)
    tmp[0] = h[nh - 1]
    tmp[nfft - nh + 1:] = h[:nh - 1]
    H = O.fft(tmp) if cpx else O.fftr(tmp)
    H = (H * np.float32(1.0 / nfft)).astype(np.float32)  # float scale (kiss_fastfir.c:70,153-161)
    nblocks = (x.shape[0] - nfft) // ngood + 1
    idx = np.arange(nblocks)[:, None] * ngood + np.arange(nfft)[None, :]
    blocks = np.ascontiguousarray(x[idx])
    F = O.fft(blocks) if cpx else O.fftr(blocks)
    fr, fi, hr, hi = F[..., 0], F[..., 1], H[..., 0], H[..., 1]
    P = np.stack([fr * hr - fi * hi, fr * hi + fi * hr], axis=-1).astype(np.float32)  # C_MUL
    y = O.fft(P, inverse=True) if cpx else O.fftri(P, nfft)
    return y[:, :ngood].reshape((-1, 2) if cpx else (-1,))


def bound(h, x, nfft, cpx, want):
    """min(1e-6, 2x the reference float filter's rel-L2 against FP64)"""
    ref = reference_float_filter(h, x, nfft, cpx)
    n = ref.shape[0]
    got = ref[:, 0] + 1j * ref[:, 1] if cpx else ref
    err = rel_l2(got, want[:n])
    say("reference float filter rel-L2", err)
    return min(1e-6, 2 * err)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def stream(ff, x, chunk, cpx):
    """The reference's do_file_filter loop (kiss_fastfir.c:260-310) over an array."""
    n = x.shape[0]
    width = (n + ff.nfft) * (2 if cpx else 1)
    inbuf = np.zeros(width, np.float32)
    outbuf = np.zeros(width, np.float32)
    res, pos = [], 0
    k = 2 if cpx else 1
    while pos < n:
        take = min(chunk, n - pos)
        inbuf[ff.offset * k:(ff.offset + take) * k] = x[pos:pos + take].reshape(-1)
        w = ff.filter(inbuf, outbuf, take)
        res.append(outbuf[:w * k].copy())
        pos += take
    w = ff.filter(inbuf, outbuf, 0)
    res.append(outbuf[:w * k].copy())
    y = np.concatenate(res)
    return y.reshape(-1, 2) if cpx else y


@pytest.mark.parametrize("nh,n,chunk,nfft", [(100, 20000, 3000, 0), (17, 5000, 5000, 256), (1000, 70000, 9999, 0),
                                             (2049, 40000, 40000, 8192)])
def test_complex_filter_stream(nh, n, chunk, nfft):
    lib = kissfft_amd.load("float")
    rng = np.random.default_rng(nh)
    h = rng.uniform(-1, 1, (nh, 2)).astype(np.float32)
    x = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    ff = lib.fastfir_alloc(h, real=False, nfft=nfft)
    say("complex fir", nh, n, "nfft", ff.nfft)
    y = stream(ff, x, chunk, True)
    want = np.convolve(x[:, 0] + 1j * x[:, 1], h[:, 0].astype(np.float64) + 1j * h[:, 1], mode="valid")
    assert y.shape[0] == want.shape[0]
    err = rel_l2(y[:, 0] + 1j * y[:, 1], want)
    tol = bound(h, x, ff.nfft, True, want)
    say("rel-L2", err, "bound", tol)
    assert err <= tol


@pytest.mark.parametrize("nh,n,chunk", [(64, 30001, 4096), (513, 100000, 100000), (2000, 9000, 700)])
def test_real_filter_stream(nh, n, chunk):
    # two real blocks per complex transform, including an odd block count
    lib = kissfft_amd.load("float")
    rng = np.random.default_rng(nh + 1)
    h = rng.standard_normal(nh).astype(np.float32)
    x = rng.standard_normal(n).astype(np.float32)
    ff = lib.fastfir_alloc(h, real=True)
    say("real fir", nh, n, "nfft", ff.nfft)
    y = stream(ff, x, chunk, False)
    want = np.convolve(x.astype(np.float64), h.astype(np.float64), mode="valid")
    assert y.shape[0] == want.shape[0]
    err = rel_l2(y, want)
    tol = bound(h, x, ff.nfft, False, want)
    say("rel-L2", err, "bound", tol)
    assert err <= tol


def test_device_buffers_one_call():
    import torch
    lib = kissfft_amd.load("float")
    rng = np.random.default_rng(9)
    h = rng.standard_normal(300).astype(np.float32)
    x = rng.standard_normal(1 << 20).astype(np.float32)
    ff = lib.fastfir_alloc(h, real=True)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros_like(xd)
    w = ff.filter(xd, yd, x.shape[0])
    torch.cuda.synchronize()
    want = np.convolve(x.astype(np.float64), h.astype(np.float64), mode="valid")
    assert w == (x.shape[0] - ff.nfft) // ff.ngood * ff.ngood + ff.ngood
    assert rel_l2(yd[:w].cpu().numpy(), want[:w]) <= bound(h, x, ff.nfft, False, want)

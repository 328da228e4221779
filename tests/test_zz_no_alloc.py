"""GPU test of kiss_fft_batch_noalloc, the entry point that allocates nothing.

Its results must equal kiss_fft_batch's (bit for bit: the same kernels run), on two calls in
a row, the second with a smaller input, for lengths whose kernels use scratch (multi-pass
lengths: the four-step temporary, carved from the library's static device array) and for
single-pass ones.  A length whose one-transform temporary exceeds the array, a plan not yet
prepared, and host buffers are refused with an error.
"""
import sys

import numpy as np
import pytest

import kissfft_amd
from helpers import ABS_FLOAT_BOUND, cplx, rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [("float", 1048576, 6, 3), ("float", 65536, 300, 17), ("float", 1200, 64, 5), ("int16_t", 65536, 40, 9),
         ("int16_t", 512, 1000, 3), ("int32_t", 32768, 20, 2), ("float", 2 * 10007, 5, 2),
         ("int16_t", 16411, 3, 1)]


@pytest.mark.parametrize("dtype,n,b1,b2", CASES)
def test_noalloc_equals_batched_call(dtype, n, b1, b2):
    import torch
    lib = kissfft_amd.load(dtype)
    tdt = {"float": torch.float32, "int16_t": torch.int16, "int32_t": torch.int32}[dtype]
    g = torch.Generator(device="cuda").manual_seed(n)
    if dtype == "float":
        x = torch.rand(b1, n, 2, generator=g, device="cuda") * 2 - 1
    else:
        hi = 16384 if dtype == "int16_t" else 1 << 30
        x = torch.randint(-hi, hi, (b1, n, 2), generator=g, device="cuda", dtype=tdt)
    for inv in (False, True):
        cfg = lib.fft_alloc(n, inv)
        want = torch.empty_like(x)
        lib.fft_batch(cfg, x, want, b1, n, n)  # prepares the device plan and kernel choice
        for b in (b1, b2):  # two calls in a row, the second smaller
            say("case noalloc", dtype, n, "inverse", inv, "batch", b)
            got = torch.full_like(x, 3)
            st = torch.cuda.Stream()
            lib.fft_batch_noalloc(cfg, x, got, b, n, n, stream=st)
            st.synchronize()
            assert torch.equal(got[:b], want[:b]), (dtype, n, inv, b)
            assert torch.equal(got[b:], torch.full_like(got[b:], 3))
        if dtype == "float":
            assert rel_l2(cplx(want[:2].cpu().numpy()),
                          (np.fft.ifft(cplx(x[:2].cpu().numpy()), axis=-1) * n) if inv
                          else np.fft.fft(cplx(x[:2].cpu().numpy()), axis=-1)) <= ABS_FLOAT_BOUND


def test_noalloc_refusals():
    import torch
    lib = kissfft_amd.load("float")
    x = torch.zeros(1, 4096, 2, device="cuda")
    y = torch.empty_like(x)
    lib.L.kiss_fft_cleanup()
    with pytest.raises(kissfft_amd.KissFFTError, match="no-allocation"):
        lib.fft_batch_noalloc(lib.fft_alloc(4096), x, y, 1, 4096, 4096)  # not prepared
    lib.fft_batch(lib.fft_alloc(4096), x, y, 1, 4096, 4096)
    lib.fft_batch_noalloc(lib.fft_alloc(4096), x, y, 1, 4096, 4096)
    xh = np.zeros((1, 4096, 2), np.float32)
    with pytest.raises(kissfft_amd.KissFFTError, match="device memory"):
        lib.fft_batch_noalloc(lib.fft_alloc(4096), xh, y, 1, 4096, 4096)
    n = 1 << 23  # a 64 MiB temporary per transform: more than the static array holds
    big = torch.zeros(1, n, 2, device="cuda")
    bout = torch.empty_like(big)
    cfg = lib.fft_alloc(n)
    lib.fft_batch(cfg, big, bout, 1, n, n)
    with pytest.raises(kissfft_amd.KissFFTError, match="static"):
        lib.fft_batch_noalloc(cfg, big, bout, 1, n, n)

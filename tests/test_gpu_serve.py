"""GPU tests of the request server (include/kiss_fft_serve.h): continuous batching.

Each request's result must be identical whether it runs alone (a batched call) or inside
a mixed stream served concurrently; host-buffer requests are staged through the server's
HBM budget and admitted only when they fit; requests larger than the budget fail cleanly.
"""
import sys

import numpy as np
import pytest

import kissfft_amd
from helpers import ABS_FLOAT_BOUND, cplx, rel_l2
from oracle import Oracle

pytestmark = pytest.mark.gpu

C2C_FWD, C2C_INV, R2C, C2R = 0, 1, 2, 3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def shapes(kind, n, b):
    if kind in (C2C_FWD, C2C_INV):
        return (b, n, 2), (b, n, 2)
    if kind == R2C:
        return (b, n), (b, n // 2 + 1, 2)
    return (b, n // 2 + 1, 2), (b, n)


def alone(lib, kind, n, x):
    """The same request through the plain batched call."""
    b = x.shape[0]
    if kind in (C2C_FWD, C2C_INV):
        y = np.zeros_like(x)
        lib.fft_batch(lib.fft_alloc(n, kind == C2C_INV), x, y, b, n, n)
    elif kind == R2C:
        y = np.zeros((b, n // 2 + 1, 2), lib.scalar)
        lib.fftr_batch(lib.fftr_alloc(n), x, y, b, n, n // 2 + 1)
    else:
        y = np.zeros((b, n), lib.scalar)
        lib.fftri_batch(lib.fftr_alloc(n, True), x, y, b, n // 2 + 1, n)
    return y


def test_mixed_stream_equals_alone_float():
    import torch
    lib = kissfft_amd.load("float")
    srv = lib.server(0, 4, 64)
    rng = np.random.default_rng(61)
    reqs = []
    sizes = [256, 1000, 1009, 3072, 1200, 4096, 4000, 65536, 1001, 2002]  # (1001 = 7*11*13)
    for i in range(60):
        n = int(rng.choice(sizes))
        kind = int(rng.integers(0, 4))
        if n % 2 and kind in (R2C, C2R):
            kind = C2C_FWD
        b = int(np.exp(rng.uniform(0, np.log(64 if n >= 65536 else 512))))
        ish, osh = shapes(kind, n, b)
        x = rng.uniform(-1, 1, ish).astype(np.float32)
        xd = torch.from_numpy(x).cuda()
        yd = torch.empty(osh, dtype=torch.float32, device="cuda")
        reqs.append((kind, n, b, x, xd, yd))
    torch.cuda.synchronize()
    tickets = [srv.submit(k, n, xd, yd, b) for k, n, b, x, xd, yd in reqs]
    for t in tickets:
        assert srv.wait(t) > 0
    st = srv.stats()
    say("launches", st["launches"], "requests", st["completed"], "p50", st["p50_s"], "p99", st["p99_s"])
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["launches"] < len(reqs)  # same-size requests were grouped
    for (kind, n, b, x, xd, yd) in reqs:
        got = yd.cpu().numpy()
        want = alone(lib, kind, n, x)
        np.testing.assert_array_equal(got, want, err_msg=f"kind {kind} n {n} b {b}")
    # and the results are transforms (not just self-consistent)
    kind, n, b, x, xd, yd = next(r for r in reqs if r[0] == C2C_FWD)
    assert rel_l2(cplx(yd.cpu().numpy()), np.fft.fft(cplx(x), axis=-1)) <= ABS_FLOAT_BOUND
    srv.close()


def test_host_requests_staged_within_budget():
    lib = kissfft_amd.load("float")
    n, b = 2048, 512
    per_req = 2 * n * 8 * b  # staged input + output
    srv = lib.server(int(2.5 * per_req), 2, 8)  # two requests fit at a time
    rng = np.random.default_rng(62)
    xs = [rng.uniform(-1, 1, (b, n, 2)).astype(np.float32) for _ in range(6)]
    ys = [np.zeros_like(x) for x in xs]
    tickets = [srv.submit(C2C_FWD, n, x, y, b) for x, y in zip(xs, ys)]
    for t in tickets:
        srv.wait(t)
    st = srv.stats()
    assert st["completed"] == 6 and st["in_use_bytes"] == 0
    for x, y in zip(xs, ys):
        np.testing.assert_array_equal(y, alone(lib, C2C_FWD, n, x))
    # a request needing more than the whole budget fails instead of waiting forever
    big = np.zeros((4 * b, n, 2), np.float32)
    out = np.zeros_like(big)
    t = srv.submit(C2C_FWD, n, big, out, 4 * b)
    with pytest.raises(kissfft_amd.KissFFTError):
        srv.wait(t)
    assert srv.stats()["failed"] == 1
    srv.close()


def test_q15_served_bit_exact():
    lib = kissfft_amd.load("int16_t")
    srv = lib.server(0, 3, 16)
    O = Oracle("int16_t")
    rng = np.random.default_rng(63)
    jobs = []
    for n in (512, 1200, 512, 60, 512, 4096, 1001, 2401):
        x = rng.integers(-32768, 32768, (int(rng.integers(1, 40)), n, 2)).astype(np.int16)
        y = np.zeros_like(x)
        jobs.append((n, x, y, srv.submit(C2C_FWD, n, x, y, x.shape[0])))
    for n, x, y, t in jobs:
        srv.wait(t)
        np.testing.assert_array_equal(y, O.fft(x))
    srv.close()


def test_nd_requests_in_the_mix():
    import torch
    lib = kissfft_amd.load("float")
    srv = lib.server(0, 3, 16)
    rng = np.random.default_rng(64)
    jobs = []
    # N-D requests (complex 3-D, real 2-D forward and inverse) between 1-D ones
    for i in range(4):
        x = rng.uniform(-1, 1, (5, 16, 12, 10, 2)).astype(np.float32)
        y = np.zeros_like(x)
        jobs.append(("nd", (16, 12, 10), x, y, srv.submit_nd(C2C_FWD, (16, 12, 10), x, y, 5)))
        xr = torch.from_numpy(rng.standard_normal((3, 64, 50)).astype(np.float32)).cuda()
        Xr = torch.empty(3, 64, 26, 2, device="cuda")
        jobs.append(("ndr", (64, 50), xr, Xr, srv.submit_nd(R2C, (64, 50), xr, Xr, 3)))
        X = rng.uniform(-1, 1, (2, 32, 9, 2)).astype(np.float32)
        yr = np.zeros((2, 32, 16), np.float32)
        jobs.append(("ndri", (32, 16), X, yr, srv.submit_nd(C2R, (32, 16), X, yr, 2)))
        x1 = rng.uniform(-1, 1, (40, 1200, 2)).astype(np.float32)
        y1 = np.zeros_like(x1)
        jobs.append(("c2c", 1200, x1, y1, srv.submit(C2C_FWD, 1200, x1, y1, 40)))
    for kind, dims, x, y, t in jobs:
        srv.wait(t)
    for kind, dims, x, y, t in jobs:
        say("served", kind, dims)
        if kind == "nd":
            want = np.zeros_like(x)
            lib.fftnd_batch(lib.fftnd_alloc(dims), x, want, x.shape[0])
            np.testing.assert_array_equal(y, want)
            assert rel_l2(cplx(y), np.fft.fftn(cplx(x), axes=(1, 2, 3))) <= ABS_FLOAT_BOUND
        elif kind == "ndr":
            want = torch.empty_like(y)
            lib.fftndr_batch(lib.fftndr_alloc(dims), x, want, x.shape[0])
            torch.cuda.synchronize()
            assert torch.equal(y, want)
        elif kind == "ndri":
            want = np.zeros_like(y)
            lib.fftndri_batch(lib.fftndr_alloc(dims, True), x, want, x.shape[0])
            np.testing.assert_array_equal(y, want)
        else:
            np.testing.assert_array_equal(y, alone(lib, C2C_FWD, dims, x))
    st = srv.stats()
    assert st["completed"] == len(jobs) and st["failed"] == 0 and st["in_use_bytes"] == 0
    srv.close()


def test_q15_nd_served_bit_exact():
    lib = kissfft_amd.load("int16_t")
    O = Oracle("int16_t")
    srv = lib.server(0, 2, 8)
    rng = np.random.default_rng(65)
    x = rng.integers(-32768, 32768, (3, 12, 60, 2)).astype(np.int16)
    y = np.zeros_like(x)
    srv.wait(srv.submit_nd(C2C_FWD, (12, 60), x, y, 3))
    np.testing.assert_array_equal(y, O.fftnd(x, (12, 60)))
    xr = rng.integers(-32768, 32768, (2, 8, 40)).astype(np.int16)
    Y = np.zeros((2, 8, 21, 2), np.int16)
    srv.wait(srv.submit_nd(R2C, (8, 40), xr, Y, 2))
    np.testing.assert_array_equal(Y, O.fftndr(xr, (8, 40)))
    srv.close()

"""GPU tests of the library's thread-safety and of kiss_fft_cleanup's safety.

The reference's kiss_fft is reentrant and its cfgs may be shared (README.md:217); the
engine keeps a locked device-plan cache and per-thread streams.  Several host threads call
kiss_fft / kiss_fftr / kiss_fftnd on shared cfgs -- right after a cleanup, so first-use
kernel tuning races between them -- and submit to one server; every result must be
bit-identical to the single-threaded result.  kiss_fft_cleanup while requests are queued
in a server, or while other threads transform, must not disturb them.
"""
import sys
import threading

import numpy as np
import pytest

import kissfft_amd
from oracle import Oracle

pytestmark = pytest.mark.gpu

NTHREADS = 6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def jobs_for(lib, rng, scalar):
    """(name, callable producing a result) tuples on shared cfgs."""
    def rnd(shape):
        if scalar == np.float32:
            return rng.uniform(-1, 1, shape).astype(np.float32)
        return rng.integers(-16384, 16384, shape).astype(np.int16)
    cfg_c = lib.fft_alloc(1200)
    cfg_ci = lib.fft_alloc(3072, True)
    cfg_r = lib.fftr_alloc(4096)
    cfg_ri = lib.fftr_alloc(4000, True)
    cfg_nd = lib.fftnd_alloc((16, 12, 10))
    cfg_long = lib.fft_alloc(65536)
    out = []
    for i in range(4):
        x = rnd((1200, 2))
        out.append((f"kiss_fft 1200 #{i}", lambda x=x: _call(lib.kiss_fft, cfg_c, x, (1200, 2), scalar)))
        xb = rnd((17, 3072, 2))
        out.append((f"batch inverse 3072 #{i}",
                    lambda xb=xb: _batch(lib.fft_batch, cfg_ci, xb, xb.shape, 17, 3072, 3072)))
        xr = rnd((4096,))
        out.append((f"kiss_fftr 4096 #{i}", lambda xr=xr: _call(lib.kiss_fftr, cfg_r, xr, (2049, 2), scalar)))
        Xr = rnd((2001, 2))
        out.append((f"kiss_fftri 4000 #{i}", lambda Xr=Xr: _call(lib.kiss_fftri, cfg_ri, Xr, (4000,), scalar)))
        xn = rnd((16, 12, 10, 2))
        out.append((f"kiss_fftnd #{i}", lambda xn=xn: _call(lib.kiss_fftnd, cfg_nd, xn, xn.shape, scalar)))
        xl = rnd((65536, 2))
        out.append((f"kiss_fft 65536 #{i}", lambda xl=xl: _call(lib.kiss_fft, cfg_long, xl, (65536, 2), scalar)))
    return out


def _call(fn, cfg, x, oshape, scalar):
    y = np.zeros(oshape, scalar)
    fn(cfg, x, y)
    return y


def _batch(fn, cfg, x, oshape, b, idist, odist):
    y = np.zeros(oshape, x.dtype)
    fn(cfg, x, y, b, idist, odist)
    return y


def run_threads(jobs, nthreads):
    results = [None] * len(jobs)
    errors = []

    def worker(t):
        try:
            for j in range(t, len(jobs), nthreads):
                results[j] = jobs[j][1]()
        except Exception as e:  # noqa: BLE001
            errors.append(repr(e))

    th = [threading.Thread(target=worker, args=(t,)) for t in range(nthreads)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    return results


@pytest.mark.parametrize("dtype", ["float", "int16_t"])
def test_threads_share_cfgs_bit_identical(dtype):
    lib = kissfft_amd.load(dtype)
    rng = np.random.default_rng(91)
    jobs = jobs_for(lib, rng, lib.scalar)
    for rep in range(2):
        lib.L.kiss_fft_cleanup()  # every plan and kernel choice is built again, concurrently
        say("threads", dtype, "round", rep, len(jobs), "jobs")
        got = run_threads(jobs, NTHREADS)
        want = [fn() for _, fn in jobs]  # single-threaded, after the plans exist
        for (name, _), g, w in zip(jobs, got, want):
            np.testing.assert_array_equal(g, w, err_msg=name)
    if dtype == "int16_t":
        x = jobs[0][1].__defaults__[0]
        np.testing.assert_array_equal(want[0], Oracle(dtype).fft(x))


def test_threads_submit_to_one_server():
    import torch
    lib = kissfft_amd.load("float")
    lib.L.kiss_fft_cleanup()
    srv = lib.server(0, 4, 32)
    rng = np.random.default_rng(92)
    menu = [(0, 256), (0, 1009), (1, 3072), (2, 4096), (3, 4000), (0, 65536), (2, 40000)]
    reqs = []
    for i in range(48):
        kind, n = menu[i % len(menu)]
        b = int(rng.integers(1, 24))
        ish = {0: (b, n, 2), 1: (b, n, 2), 2: (b, n), 3: (b, n // 2 + 1, 2)}[kind]
        osh = {0: (b, n, 2), 1: (b, n, 2), 2: (b, n // 2 + 1, 2), 3: (b, n)}[kind]
        x = torch.from_numpy(rng.uniform(-1, 1, ish).astype(np.float32)).cuda()
        reqs.append((kind, n, b, x, torch.empty(osh, device="cuda")))
    torch.cuda.synchronize()
    jobs = [(f"req {i}", (lambda r=r: srv.wait(srv.submit(r[0], r[1], r[3], r[4], r[2])))) for i, r in enumerate(reqs)]
    run_threads(jobs, NTHREADS)
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    srv.close()
    for kind, n, b, x, y in reqs:
        want = torch.empty_like(y)
        if kind in (0, 1):
            lib.fft_batch(lib.fft_alloc(n, kind == 1), x, want, b, n, n)
        elif kind == 2:
            lib.fftr_batch(lib.fftr_alloc(n), x, want, b, n, n // 2 + 1)
        else:
            lib.fftri_batch(lib.fftr_alloc(n, True), x, want, b, n // 2 + 1, n)
        torch.cuda.synchronize()
        assert torch.equal(y, want), (kind, n, b)


def test_cleanup_while_requests_are_queued():
    """Cleanup drops the caches while a server holds requests; their plans and tables stay
    alive with the requests, and every result equals the same transform run afterwards."""
    import torch
    lib = kissfft_amd.load("float")
    srv = lib.server(0, 2, 8)
    rng = np.random.default_rng(93)
    reqs = []
    for kind, n, b in [(0, 1048576, 4), (2, 4096, 3000), (0, 1200, 500), (3, 65536, 40), (0, 1009, 300),
                       (2, 1048576, 3), (1, 3072, 800), (0, 256, 4000)] * 2:
        ish = {0: (b, n, 2), 1: (b, n, 2), 2: (b, n), 3: (b, n // 2 + 1, 2)}[kind]
        osh = {0: (b, n, 2), 1: (b, n, 2), 2: (b, n // 2 + 1, 2), 3: (b, n)}[kind]
        x = torch.from_numpy(rng.uniform(-1, 1, ish).astype(np.float32)).cuda()
        reqs.append((kind, n, b, x, torch.empty(osh, device="cuda")))
    torch.cuda.synchronize()
    tickets = []
    for i, (kind, n, b, x, y) in enumerate(reqs):
        tickets.append(srv.submit(kind, n, x, y, b))
        if i % 3 == 2:
            lib.L.kiss_fft_cleanup()
    lib.L.kiss_fft_cleanup()
    for t in tickets:
        srv.wait(t)
    srv.close()
    lib.L.kiss_fft_cleanup()
    for kind, n, b, x, y in reqs:
        say("cleanup case", kind, n, b)
        want = torch.empty_like(y)
        if kind in (0, 1):
            lib.fft_batch(lib.fft_alloc(n, kind == 1), x, want, b, n, n)
        elif kind == 2:
            lib.fftr_batch(lib.fftr_alloc(n), x, want, b, n, n // 2 + 1)
        else:
            lib.fftri_batch(lib.fftr_alloc(n, True), x, want, b, n // 2 + 1, n)
        torch.cuda.synchronize()
        assert torch.equal(y, want), (kind, n, b)


def test_cleanup_while_other_threads_transform():
    lib = kissfft_amd.load("int16_t")
    O = Oracle("int16_t")
    rng = np.random.default_rng(94)
    xs = [rng.integers(-16384, 16384, (64, n, 2)).astype(np.int16) for n in (512, 1200, 60, 4096, 32768)]
    wants = [O.fft(x) for x in xs]
    stop = threading.Event()
    errors = []

    def worker(k):
        try:
            i = 0
            while not stop.is_set() and i < 40:
                x = xs[(k + i) % len(xs)]
                y = np.zeros_like(x)
                n = x.shape[1]
                lib.fft_batch(lib.fft_alloc(n), x, y, x.shape[0], n, n)
                if not np.array_equal(y, wants[(k + i) % len(xs)]):
                    errors.append((k, i, n))
                i += 1
        except Exception as e:  # noqa: BLE001
            errors.append(repr(e))

    th = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for _ in range(20):
        lib.L.kiss_fft_cleanup()
    stop.set()
    for t in th:
        t.join()
    assert not errors, errors[:5]

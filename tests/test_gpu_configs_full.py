"""GPU parity of two workloads at their full stated sizes.

* Q15 c2c, nfft=512, 2^20 transforms of the seed-3 stream (int16 re/im uniform in
  [-16384, 16383]), one kiss_fft_batch on device buffers: every transform bit-exact
  against the oracle (itself bit-exact against the reference build's fixtures, which hold
  the first 16 transforms of this stream).  The engine picks its compiled variant by
  timing at the call's batch size, so the check runs for the tuned choice and then for
  every registered Q15 512-point variant forced in turn, each on the whole batch.
* The serving stream: requests of c2c / r2c / inverse kinds with n in {256, 1000, 1009,
  3072, 65536, 1048576} and batch counts up to 4096, submitted at once to one
  kiss_fft_server.  Every request's result is bit-identical to the same request run alone
  (one batched call), and sampled transforms of every request are within the float bar of
  an FP64 FFT.
"""
import sys

import numpy as np
import pytest

import kissfft_amd
from helpers import ABS_FLOAT_BOUND, cplx, rel_l2
from kissfft_amd import workloads as W
from oracle import Oracle

pytestmark = pytest.mark.gpu

C2C_FWD, C2C_INV, R2C, C2R = 0, 1, 2, 3
Q15_512_VARIANTS = ["q15x_c2c512_t16x8", "q15x_c2c512_t16x16", "q15x_c2c512_t16x8_o6", "q15x_c2c512_t16x4_o6",
                    "q15w_c2c512_t16x16", "q15w_c2c512_t32x8", "q15p_c2c512_t16x16_sb", "q15p_c2c512_t32x8_sb",
                    "q15c2c512_t128x2_sb"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def q15_full():
    """(input, oracle output) of the 2^20-transform workload, host arrays."""
    x = W.generate(W.q15_complex, 3, 512, 1 << 20, workers=16)
    say("q15 512 x 2^20: input generated")
    want = Oracle("int16_t").fft(x)
    say("q15 512 x 2^20: oracle done")
    return x, want


def _first_mismatch(got, want):
    bad = (got != want).reshape(got.shape[0], -1).any(dim=1)
    idx = bad.nonzero()
    return int(bad.sum()), (int(idx[0, 0]) if idx.numel() else -1)


def test_q15_512_full_batch_bit_exact(q15_full, golden_q15, monkeypatch, capfd):
    import torch

    x, want = q15_full
    # the oracle agrees with the reference build on the fixture's slice of this stream
    np.testing.assert_array_equal(want[:16], golden_q15["c2c_512_0_out"][:16])
    np.testing.assert_array_equal(x[:16], golden_q15["c2c_512_0_in"][:16])
    lib = kissfft_amd.load("int16_t")
    B = x.shape[0]
    xd = torch.from_numpy(x).cuda()
    wd = torch.from_numpy(want).cuda()
    yd = torch.empty_like(xd)
    monkeypatch.setenv("KISSFFT_AMD_VERBOSE", "1")
    for name in [None] + Q15_512_VARIANTS:
        lib.L.kiss_fft_cleanup()
        if name is None:
            monkeypatch.delenv("KISSFFT_AMD_VARIANT", raising=False)
        else:
            monkeypatch.setenv("KISSFFT_AMD_VARIANT", name)
        say("case q15 512 x 2^20 variant", name or "tuned")
        yd.fill_(0x5A5A)
        capfd.readouterr()
        lib.fft_batch(lib.fft_alloc(512), xd, yd, B, 512, 512)
        torch.cuda.synchronize()
        err = capfd.readouterr().err
        if name is not None:
            assert f"fast variant {name}\n" in err, f"variant {name} was not the one launched"
        nbad, first = _first_mismatch(yd, wd)
        assert nbad == 0, f"variant {name or 'tuned'}: {nbad} of {B} transforms differ, first {first}"
    lib.L.kiss_fft_cleanup()


def _shapes(kind, n, b):
    if kind in (C2C_FWD, C2C_INV):
        return (b, n, 2), (b, n, 2)
    if kind == R2C:
        return (b, n), (b, n // 2 + 1, 2)
    return (b, n // 2 + 1, 2), (b, n)


def _alone(lib, kind, n, xd, b):
    import torch

    _, osh = _shapes(kind, n, b)
    y = torch.empty(osh, dtype=torch.float32, device="cuda")
    if kind in (C2C_FWD, C2C_INV):
        lib.fft_batch(lib.fft_alloc(n, kind == C2C_INV), xd, y, b, n, n)
    elif kind == R2C:
        lib.fftr_batch(lib.fftr_alloc(n), xd, y, b, n, n // 2 + 1)
    else:
        lib.fftri_batch(lib.fftr_alloc(n, True), xd, y, b, n // 2 + 1, n)
    return y


def _fp64(kind, n, x):
    """FP64 reference of host rows x (the library's unnormalised conventions)."""
    if kind == C2C_FWD:
        return np.fft.fft(cplx(x), axis=-1)
    if kind == C2C_INV:
        return np.fft.ifft(cplx(x), axis=-1) * n
    if kind == R2C:
        return np.fft.rfft(x.astype(np.float64), axis=-1)
    # kiss_fftri ignores the imaginary parts of the DC and Nyquist bins (kiss_fftr.c:131-134)
    z = cplx(x)
    z[..., 0] = z[..., 0].real
    z[..., -1] = z[..., -1].real
    return np.fft.irfft(z, n, axis=-1) * n


def _serving_requests():
    """The seeded serving stream (first 40 requests, batches capped at 1 GiB of input),
    plus requests that make sure every kind meets n = 1048576 and that batches of 2048
    and 4096 are served."""
    kinds = {"c2c": C2C_FWD, "c2c_inv": C2C_INV, "r2c": R2C, "c2r": C2R}
    reqs = []
    for r in W.serving_stream(40, 2000.0, seed=6):
        n, b = r["n"], r["batch"]
        ib, _ = W.request_bytes(r["kind"], n, 1)
        b = max(1, min(b, (1 << 30) // ib))
        reqs.append((kinds[r["kind"]], n, b))
    reqs += [(C2C_FWD, 1048576, 3), (C2C_INV, 1048576, 2), (R2C, 1048576, 3), (C2R, 1048576, 2),
             (C2C_FWD, 256, 4096), (R2C, 3072, 2048), (C2C_INV, 1009, 2048), (C2R, 1000, 4096),
             (C2C_FWD, 65536, 64), (R2C, 65536, 96)]
    return reqs


def test_serving_stream_at_stated_sizes():
    import torch

    lib = kissfft_amd.load("float")
    reqs = _serving_requests()
    assert any(n == 1048576 and k == kd for k, n, b in reqs for kd in (C2C_FWD, R2C, C2R))
    assert any(b >= 2048 for k, n, b in reqs)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(66)
    jobs = []
    for kind, n, b in reqs:
        ish, osh = _shapes(kind, n, b)
        xd = torch.rand(ish, generator=gen, device="cuda", dtype=torch.float32) * 2 - 1
        yd = torch.full(osh, float("nan"), device="cuda", dtype=torch.float32)
        jobs.append((kind, n, b, xd, yd))
    torch.cuda.synchronize()
    srv = lib.server(0, 4, 64)
    say("serving", len(jobs), "requests, largest",
        max(W.request_bytes({0: "c2c", 1: "c2c_inv", 2: "r2c", 3: "c2r"}[k], n, b)[0] for k, n, b in reqs), "B in")
    tickets = [srv.submit(k, n, xd, yd, b) for k, n, b, xd, yd in jobs]
    lat = [srv.wait(t) for t in tickets]
    st = srv.stats()
    say("served: launches", st["launches"], "p50", st["p50_s"], "p99", st["p99_s"], "max", max(lat))
    assert st["completed"] == len(jobs) and st["failed"] == 0
    srv.close()
    worst = 0.0
    for i, (kind, n, b, xd, yd) in enumerate(jobs):
        say("check request", i, "kind", kind, "n", n, "batch", b)
        want = _alone(lib, kind, n, xd, b)
        torch.cuda.synchronize()
        assert torch.equal(yd, want), f"request {i} (kind {kind}, n {n}, batch {b}) differs from its alone call"
        rows = sorted({0, b // 2, b - 1})
        xs = xd[rows].cpu().numpy()
        ys = yd[rows].cpu().numpy()
        got = cplx(ys) if kind != C2R else ys.astype(np.float64)
        err = rel_l2(got, _fp64(kind, n, xs))
        worst = max(worst, err)
        assert err <= ABS_FLOAT_BOUND, f"request {i} (kind {kind}, n {n}): rel-L2 {err:.3e}"
        del want
    say("worst sampled rel-L2", worst)

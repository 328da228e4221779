"""GPU parity of the complex 1-D path (kiss_fft / kiss_fft_stride / kiss_fft_batch / kfc).

Fixed point: bit-exact against the golden fixtures (reference outputs) and against the
oracle.  Float: rel-L2 against FP64 <= min(1e-6, 2x the reference float build's error on
the same inputs), the tolerance written in helpers.float_bound.
"""
import sys

import numpy as np
import pytest

import kissfft_amd
from helpers import ABS_FLOAT_BOUND, cplx, fixture_cases, float_bound, pairs, rel_l2
from kissfft_amd import workloads as W
from oracle import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lf():
    L = kissfft_amd.load("float")
    assert L.device_available(), "no GPU visible"
    return L


@pytest.fixture(scope="module")
def lq():
    return kissfft_amd.load("int16_t")


def run_batch(lib, n, inverse, x):
    """x: (howmany, n, 2) host array -> same shape, one batched call."""
    cfg = lib.fft_alloc(n, inverse)
    out = np.zeros_like(x)
    lib.fft_batch(cfg, x, out, x.shape[0], n, n)
    return out


@pytest.mark.parametrize("datatype,gname", [("int16_t", "golden_q15"), ("int32_t", "golden_q31")])
def test_fixed_c2c_matches_reference_fixtures(datatype, gname, request):
    lib = kissfft_amd.load(datatype)
    g = request.getfixturevalue(gname)
    cases = 0
    for stem, x, want in fixture_cases(g, "c2c"):
        _, n, inv = stem.split("_")
        n, inv = int(n), int(inv)
        say("case", datatype, stem)
        got = run_batch(lib, n, bool(inv), np.ascontiguousarray(x))
        np.testing.assert_array_equal(got, want, err_msg=stem)
        # and the reference-shaped single call on the first item
        cfg = lib.fft_alloc(n, bool(inv))
        one = np.zeros((n, 2), lib.scalar)
        lib.kiss_fft(cfg, np.ascontiguousarray(x[0]), one)
        np.testing.assert_array_equal(one, want[0], err_msg=stem + " single")
        cases += 1
    assert cases >= 8


def test_q15_stride_matches_reference_fixtures(lq, golden_q15):
    for stem, x, want in fixture_cases(golden_q15, "stride"):
        _, n, s = stem.split("_")
        say("case stride", stem)
        cfg = lq.fft_alloc(int(n))
        out = np.zeros((int(n), 2), np.int16)
        lq.kiss_fft_stride(cfg, np.ascontiguousarray(x), out, int(s))
        np.testing.assert_array_equal(out, want, err_msg=stem)


def test_q15_batch_equals_loop_of_single_calls(lq):
    O = Oracle("int16_t")
    for n in (512, 1200, 60, 1009):
        x = W.q15_complex(21, n, 8)
        say("case q15 batch-vs-loop", n)
        got = run_batch(lq, n, False, x)
        cfg = lq.fft_alloc(n)
        for b in range(8):
            one = np.zeros((n, 2), np.int16)
            lq.kiss_fft(cfg, x[b], one)
            np.testing.assert_array_equal(one, got[b])
        np.testing.assert_array_equal(got, O.fft(x))


@pytest.mark.parametrize("n", [2048, 4096, 16384, 32768, 65536])
def test_q15_long_lines_bit_exact_vs_oracle(lq, n):
    # above the 16384-point LDS line the transform runs as two passes; the split keeps the
    # reference's stage arithmetic, so results stay bit-identical
    O = Oracle("int16_t")
    x = W.q15_complex(22, n, 2)
    for inv in (False, True):
        say("case q15 long", n, inv)
        np.testing.assert_array_equal(run_batch(lq, n, inv, x), O.fft(x, inv))


def test_float_c2c_1200_workload(lf, golden_float):
    # 256 transforms of 1200 points, re/im uniform on [-1,1), seed 1
    x = W.complex_pm1(1, 1200, 256)
    say("case float c2c 1200 x256")
    got = cplx(run_batch(lf, 1200, False, x))
    err = rel_l2(got, np.fft.fft(cplx(x), axis=-1))
    say("rel-L2", err, "bound", float_bound(golden_float, "c2c_1200"))
    assert err <= float_bound(golden_float, "c2c_1200")


SIZES = list(range(1, 41)) + [49, 64, 77, 96, 120, 121, 128, 243, 256, 500, 512, 1000, 1001, 1024, 1200, 2000,
                              2048, 3072, 4096, 8192]


def test_float_c2c_many_sizes(lf):
    rng = np.random.default_rng(9)
    worst = 0.0
    for n in SIZES:
        x = rng.uniform(-1, 1, (3, n, 2)).astype(np.float32)
        for inv in (False, True):
            got = cplx(run_batch(lf, n, inv, x))
            want = np.fft.ifft(cplx(x), axis=-1) * n if inv else np.fft.fft(cplx(x), axis=-1)
            err = rel_l2(got, want)
            worst = max(worst, err)
            assert err <= ABS_FLOAT_BOUND, (n, inv, err)
    say("worst rel-L2 over sizes", worst)


def test_float_c2c_prime_1009(lf, golden_float):
    x = W.complex_pm1(5, 1009, 64)
    got = cplx(run_batch(lf, 1009, False, x))
    err = rel_l2(got, np.fft.fft(cplx(x), axis=-1))
    say("1009 rel-L2", err)
    assert err <= float_bound(golden_float, "c2c_1009")


@pytest.mark.parametrize("n,howmany", [(16384, 3), (32768, 3), (65536, 4), (131072, 2), (262144, 2),
                                       (1048576, 2), (107520, 2)])
@pytest.mark.parametrize("inverse", [False, True])
def test_float_c2c_long_lines(lf, n, howmany, inverse):
    # four-step split: N1*N2 tile passes where both halves are registered, else the
    # general two-pass split in the reference's stage order (107520 = 3*5*7*1024)
    x = W.complex_pm1(6, n, howmany)
    say("case float long", n, "inverse", inverse)
    got = cplx(run_batch(lf, n, inverse, x))
    want = np.fft.ifft(cplx(x), axis=-1) * n if inverse else np.fft.fft(cplx(x), axis=-1)
    err = rel_l2(got, want)
    say("rel-L2", err)
    assert err <= ABS_FLOAT_BOUND


def test_float_c2c_long_lines_strided(lf):
    # istride / idist / odist through both four-step passes
    n, howmany, s = 65536, 3, 2
    rng = np.random.default_rng(17)
    idist, odist = n * s + 5, n + 3
    buf = rng.uniform(-1, 1, (howmany * idist, 2)).astype(np.float32)
    out = np.full((howmany * odist, 2), 9.0, np.float32)
    lf.fft_batch(lf.fft_alloc(n), buf, out, howmany, idist, odist, in_stride=s)
    for b in range(howmany):
        xb = cplx(buf[b * idist: b * idist + n * s: s])
        assert rel_l2(cplx(out[b * odist: b * odist + n]), np.fft.fft(xb)) <= ABS_FLOAT_BOUND, b
        assert (out[b * odist + n: (b + 1) * odist] == 9.0).all(), b  # the gaps keep their values


def test_float_stride_and_in_place(lf):
    rng = np.random.default_rng(10)
    n, s = 1200, 3
    buf = rng.uniform(-1, 1, ((n - 1) * s + 1, 2)).astype(np.float32)
    cfg = lf.fft_alloc(n)
    out = np.zeros((n, 2), np.float32)
    lf.kiss_fft_stride(cfg, buf, out, s)
    assert rel_l2(cplx(out), np.fft.fft(cplx(buf[::s]))) <= ABS_FLOAT_BOUND
    x = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    want = np.fft.fft(cplx(x))
    lf.kiss_fft(cfg, x, x)  # fin == fout
    assert rel_l2(cplx(x), want) <= ABS_FLOAT_BOUND


def test_inverse_of_forward_is_n_times_identity(lf):
    x = W.complex_pm1(12, 1000, 4)
    X = run_batch(lf, 1000, False, x)
    y = run_batch(lf, 1000, True, X)
    assert rel_l2(cplx(y), 1000 * cplx(x)) <= ABS_FLOAT_BOUND


def test_device_pointers_and_async_stream(lf, lq):
    import torch
    dev = torch.device("cuda:0")
    n, howmany = 1200, 64
    x = W.complex_pm1(13, n, howmany)
    xd = torch.from_numpy(x).to(dev)
    yd = torch.empty_like(xd)
    cfg = lf.fft_alloc(n)
    before = lf.stats()
    lf.fft_batch(cfg, xd, yd, howmany, n, n)  # device pointers: used in place, no staging
    after = lf.stats()
    assert after["h2d_bytes"] == before["h2d_bytes"]
    assert rel_l2(cplx(yd.cpu().numpy()), np.fft.fft(cplx(x), axis=-1)) <= ABS_FLOAT_BOUND
    s = torch.cuda.Stream()
    y2 = torch.empty_like(xd)
    lf.fft_batch(cfg, xd, y2, howmany, n, n, stream=s)
    lf.stream_sync(s)
    assert torch.equal(y2, yd)
    # Q15 on device, in place
    xq = W.q15_complex(14, 512, 32)
    tq = torch.from_numpy(xq).to(dev)
    cq = lq.fft_alloc(512)
    lq.fft_batch(cq, tq, tq, 32, 512, 512)
    np.testing.assert_array_equal(tq.cpu().numpy(), Oracle("int16_t").fft(xq))


def test_kfc_cached_calls(lf):
    lf.L.kfc_cleanup()
    assert lf.L.kfc_cached_count() == 0
    x = W.complex_pm1(15, 512, 1)[0]
    y = np.zeros_like(x)
    lf.L.kfc_fft(512, x.ctypes.data, y.ctypes.data)
    assert lf.L.kfc_cached_count() == 1
    lf.L.kfc_fft(512, x.ctypes.data, y.ctypes.data)
    assert lf.L.kfc_cached_count() == 1
    z = np.zeros_like(x)
    lf.L.kfc_ifft(512, y.ctypes.data, z.ctypes.data)
    assert lf.L.kfc_cached_count() == 2
    assert rel_l2(cplx(z), 512 * cplx(x)) <= ABS_FLOAT_BOUND
    lf.L.kfc_cleanup()
    assert lf.L.kfc_cached_count() == 0


def test_float_long_lines_chunked_batches(lf):
    # batches whose four-step temporary exceeds the 2 GiB scratch bound run in chunks of
    # equal size (2^20 points, 260 transforms: two chunks of 130, complex and real): sampled
    # transforms against FP64 on both sides of the chunk edge, every transform through the
    # round trip
    import torch
    dev = torch.device("cuda:0")
    n, B = 1 << 20, 260
    g = torch.Generator(device=dev).manual_seed(77)
    x = torch.rand(B, n, 2, device=dev, generator=g) * 2 - 1
    y = torch.empty_like(x)
    lf.fft_batch(lf.fft_alloc(n), x, y, B, n, n)
    torch.cuda.synchronize()
    for b in (0, 129, 130, 259):
        say("chunked c2c sample", b)
        assert rel_l2(cplx(y[b].cpu().numpy()), np.fft.fft(cplx(x[b].cpu().numpy()))) <= ABS_FLOAT_BOUND, b
    z = torch.empty_like(x)
    lf.fft_batch(lf.fft_alloc(n, True), y, z, B, n, n)
    torch.cuda.synchronize()
    rt = float(torch.linalg.vector_norm((z.double() - n * x.double()).flatten()) /
               torch.linalg.vector_norm((n * x.double()).flatten()))
    say("chunked c2c round trip", rt)
    assert rt <= 2 * ABS_FLOAT_BOUND
    del x, y, z
    B = 260
    xr = torch.rand(B, n, device=dev, generator=g) * 2 - 1
    X = torch.empty(B, n // 2 + 1, 2, device=dev)
    lf.fftr_batch(lf.fftr_alloc(n), xr, X, B, n, n // 2 + 1)
    torch.cuda.synchronize()
    for b in (0, 129, 130, 259):
        say("chunked r2c sample", b)
        want = np.fft.rfft(xr[b].cpu().numpy().astype(np.float64))
        assert rel_l2(cplx(X[b].cpu().numpy()), want) <= ABS_FLOAT_BOUND, b
    yr = torch.empty_like(xr)
    lf.fftri_batch(lf.fftr_alloc(n, True), X, yr, B, n // 2 + 1, n)
    torch.cuda.synchronize()
    rt = float(torch.linalg.vector_norm((yr.double() - n * xr.double()).flatten()) /
               torch.linalg.vector_norm((n * xr.double()).flatten()))
    say("chunked real round trip", rt)
    assert rt <= 2 * ABS_FLOAT_BOUND


@pytest.mark.parametrize("n", [17, 34, 1009, 1011, 2018, 2039, 2053])
def test_float_large_prime_lengths(lf, n, monkeypatch):
    # lengths with a prime factor above 16: Bluestein convolution (M = 64 .. 4096) for
    # n <= 2048, the direct generic butterfly above; both directions, strided input, and
    # the generic path forced off Bluestein for comparison
    x = W.complex_pm1(90 + n, n, 8)
    for inv in (False, True):
        say("case prime", n, inv)
        got = cplx(run_batch(lf, n, inv, x))
        want = np.fft.ifft(cplx(x), axis=-1) * n if inv else np.fft.fft(cplx(x), axis=-1)
        err = rel_l2(got, want)
        say("rel-L2", err)
        assert err <= ABS_FLOAT_BOUND, (n, inv)
    s = 3
    buf = W.complex_pm1(91, n * s, 1)[0]
    out = np.zeros((n, 2), np.float32)
    lf.kiss_fft_stride(lf.fft_alloc(n), buf, out, s)
    assert rel_l2(cplx(out), np.fft.fft(cplx(buf[::s]))) <= ABS_FLOAT_BOUND
    monkeypatch.setenv("KISSFFT_AMD_BLUESTEIN", "0")
    lf.L.kiss_fft_cleanup()
    try:
        got = cplx(run_batch(lf, n, False, x))
        assert rel_l2(got, np.fft.fft(cplx(x), axis=-1)) <= ABS_FLOAT_BOUND
    finally:
        lf.L.kiss_fft_cleanup()

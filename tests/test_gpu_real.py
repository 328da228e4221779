"""GPU parity of the real 1-D path (kiss_fftr / kiss_fftri / batched forms)."""
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


def r2c(lib, n, x):
    cfg = lib.fftr_alloc(n, False)
    out = np.zeros((x.shape[0], n // 2 + 1, 2), lib.scalar)
    lib.fftr_batch(cfg, np.ascontiguousarray(x), out, x.shape[0], n, n // 2 + 1)
    return out


def c2r(lib, n, X):
    cfg = lib.fftr_alloc(n, True)
    out = np.zeros((X.shape[0], n), lib.scalar)
    lib.fftri_batch(cfg, np.ascontiguousarray(X), out, X.shape[0], n // 2 + 1, n)
    return out


@pytest.mark.parametrize("datatype,gname", [("int16_t", "golden_q15"), ("int32_t", "golden_q31")])
def test_fixed_real_matches_reference_fixtures(datatype, gname, request):
    lib = kissfft_amd.load(datatype)
    g = request.getfixturevalue(gname)
    for stem, x, want in fixture_cases(g, "r2c"):
        n = int(stem.split("_")[1])
        say("case", datatype, stem)
        np.testing.assert_array_equal(r2c(lib, n, x), want, err_msg=stem)
        cfg = lib.fftr_alloc(n, False)
        one = np.zeros((n // 2 + 1, 2), lib.scalar)
        lib.kiss_fftr(cfg, np.ascontiguousarray(x[0]), one)
        np.testing.assert_array_equal(one, want[0], err_msg=stem + " single")
    for stem, X, want in fixture_cases(g, "c2r"):
        n = int(stem.split("_")[1])
        say("case", datatype, stem)
        np.testing.assert_array_equal(c2r(lib, n, X), want, err_msg=stem)
        cfg = lib.fftr_alloc(n, True)
        one = np.zeros((n,), lib.scalar)
        lib.kiss_fftri(cfg, np.ascontiguousarray(X[0]), one)
        np.testing.assert_array_equal(one, want[0], err_msg=stem + " single")


def test_q15_real_bit_exact_vs_oracle():
    lib = kissfft_amd.load("int16_t")
    O = Oracle("int16_t")
    rng = np.random.default_rng(31)
    for n in (4096, 4000, 1024, 2, 6, 250, 8192, 32768):
        x = rng.integers(-32768, 32768, (4, n)).astype(np.int16)
        say("case q15 real", n)
        np.testing.assert_array_equal(r2c(lib, n, x), O.fftr(x))
        X = rng.integers(-32768, 32768, (4, n // 2 + 1, 2)).astype(np.int16)
        np.testing.assert_array_equal(c2r(lib, n, X), O.fftri(X, n))


@pytest.mark.parametrize("datatype,n", [("int16_t", 65536), ("int16_t", 40000), ("int32_t", 32768),
                                         ("int32_t", 20000)])
def test_fixed_real_multipass_bit_exact_vs_oracle(datatype, n):
    # half length beyond one LDS line: four-step complex pass + separate split/merge pass
    lib = kissfft_amd.load(datatype)
    O = Oracle(datatype)
    rng = np.random.default_rng(32)
    info = np.iinfo(lib.scalar)
    x = rng.integers(info.min, info.max, (3, n), endpoint=True).astype(lib.scalar)
    say("case real multipass", datatype, n)
    np.testing.assert_array_equal(r2c(lib, n, x), O.fftr(x))
    X = rng.integers(info.min, info.max, (3, n // 2 + 1, 2), endpoint=True).astype(lib.scalar)
    np.testing.assert_array_equal(c2r(lib, n, X), O.fftri(X, n))


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "unfused"])
@pytest.mark.parametrize("n,howmany", [(32768, 8), (65536, 4), (1 << 20, 2), (3 * 5 * 7 * 1024, 2)])
def test_float_real_multipass(n, howmany, fused, monkeypatch):
    """Long real lines: the split / merge fused into the four-step passes (when a split of
    n/2 is registered; 107520 has none and always takes the three-pass form) and the
    three-pass form (complex transform, then split / merge as a pass of its own)."""
    monkeypatch.setenv("KISSFFT_AMD_REAL_FUSED", "1" if fused else "0")
    lib = kissfft_amd.load("float")
    x = W.two_tone(5, n, howmany)
    say("case float real multipass", n, "fused" if fused else "unfused")
    X = r2c(lib, n, x)
    want = np.fft.rfft(x.astype(np.float64), axis=-1)
    err = rel_l2(cplx(X), want)
    say("r2c rel-L2", err)
    assert err <= ABS_FLOAT_BOUND
    Xin = pairs(want.astype(np.complex64))
    y = c2r(lib, n, Xin)
    err = rel_l2(y, np.fft.irfft(cplx(Xin), n=n, axis=-1) * n)
    say("c2r rel-L2", err)
    assert err <= ABS_FLOAT_BOUND


@pytest.mark.parametrize("n", [65536, 1 << 20])
def test_float_real_multipass_padded_rows(n):
    """Rows further apart than one transform on both sides (idist, odist padded)."""
    lib = kissfft_amd.load("float")
    howmany, ipad, opad = 3, 6, 5
    x = np.zeros((howmany, n + ipad), np.float32)
    x[:, :n] = W.two_tone(9, n, howmany)
    say("case float real padded rows", n)
    X = np.full((howmany, n // 2 + 1 + opad, 2), 7.0, np.float32)
    lib.fftr_batch(lib.fftr_alloc(n, False), x, X, howmany, n + ipad, n // 2 + 1 + opad)
    want = np.fft.rfft(x[:, :n].astype(np.float64), axis=-1)
    assert rel_l2(cplx(X[:, : n // 2 + 1]), want) <= ABS_FLOAT_BOUND
    assert (X[:, n // 2 + 1 :] == 7.0).all()
    Xin = np.zeros((howmany, n // 2 + 1 + opad, 2), np.float32)
    Xin[:, : n // 2 + 1] = pairs(want.astype(np.complex64))
    y = np.full((howmany, n + ipad), 5.0, np.float32)
    lib.fftri_batch(lib.fftr_alloc(n, True), Xin, y, howmany, n // 2 + 1 + opad, n + ipad)
    assert rel_l2(y[:, :n], np.fft.irfft(cplx(Xin[:, : n // 2 + 1]), n=n, axis=-1) * n) <= ABS_FLOAT_BOUND
    assert (y[:, n:] == 5.0).all()


@pytest.mark.parametrize("n", [4096, 4000])
def test_float_real_workload_sample(n, golden_float):
    lib = kissfft_amd.load("float")
    x = W.two_tone(2, n, 256)
    say("case float r2c", n)
    X = r2c(lib, n, x)
    want = np.fft.rfft(x.astype(np.float64), axis=-1)
    err = rel_l2(cplx(X), want)
    say("r2c rel-L2", err, "bound", float_bound(golden_float, f"r2c_{n}"))
    assert err <= float_bound(golden_float, f"r2c_{n}")
    Xin = pairs(want.astype(np.complex64))
    y = c2r(lib, n, Xin)
    want_t = np.fft.irfft(cplx(Xin), n=n, axis=-1) * n
    err = rel_l2(y, want_t)
    say("c2r rel-L2", err, "bound", float_bound(golden_float, f"c2r_{n}"))
    assert err <= float_bound(golden_float, f"c2r_{n}")


@pytest.mark.parametrize("n", [4096, 4000])
def test_float_real_full_batch_on_device(n, golden_float):
    # the full 65536-signal batch, resident in HBM: sampled signals against FP64, and the
    # whole batch through the unnormalised round trip c2r(r2c(x)) = n*x
    import torch
    lib = kissfft_amd.load("float")
    dev = torch.device("cuda:0")
    B = 65536
    g = torch.Generator(device=dev).manual_seed(2)
    t = torch.arange(n, device=dev, dtype=torch.float64)
    tones = (torch.cos(2 * np.pi * 100.0 * t / n) + torch.cos(2 * np.pi * 1500.5 * t / n)).float()
    x = torch.randn(B, n, generator=g, device=dev) + tones
    X = torch.empty(B, n // 2 + 1, 2, device=dev)
    say("case float full batch", n)
    lib.fftr_batch(lib.fftr_alloc(n), x, X, B, n, n // 2 + 1)
    torch.cuda.synchronize()
    idx = np.array([0, 1, 777, 12345, 40000, B - 1])
    xs = x[idx].cpu().numpy()
    err = rel_l2(cplx(X[idx].cpu().numpy()), np.fft.rfft(xs.astype(np.float64), axis=-1))
    say("sampled r2c rel-L2", err)
    assert err <= float_bound(golden_float, f"r2c_{n}")
    y = torch.empty(B, n, device=dev)
    lib.fftri_batch(lib.fftr_alloc(n, True), X, y, B, n // 2 + 1, n)
    torch.cuda.synchronize()
    num = torch.linalg.vector_norm((y.double() - n * x.double()).flatten())
    den = torch.linalg.vector_norm((n * x.double()).flatten())
    rt = float(num / den)
    say("round-trip rel-L2", rt)
    assert rt <= 2 * ABS_FLOAT_BOUND


def test_real_direction_misuse_writes_nothing():
    lib = kissfft_amd.load("float")
    inv_cfg = lib.fftr_alloc(64, True)
    x = np.ones(64, np.float32)
    out = np.full((33, 2), 5.0, np.float32)
    lib.kiss_fftr(inv_cfg, x, out)  # wrong direction: the reference logs and returns
    assert np.all(out == 5.0)
    with pytest.raises(kissfft_amd.KissFFTError):
        lib.fftr_batch(inv_cfg, x, out, 1, 64, 33)

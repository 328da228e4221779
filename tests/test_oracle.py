"""CPU tests of the oracle (the CPU restatement in oracle/kfo.c).

The oracle is pinned three ways before any GPU result is compared with it:
  * against the golden fixtures written from the reference compiled from its own sources
    (bit-exact for Q15, Q31 and float -- the float restatement keeps the reference's
    operation order);
  * against oracle/_ref itself on further random sizes, where that build is present;
  * against what the mathematics fixes: brute-force DFTs, closed-form radix schedules and
    twiddles, impulse/tone spectra, Hermitian symmetry, round trips, Parseval, the N-D DC
    term.  Each of these fails for a dropped term, a wrong sign or a wrong index.
"""
import numpy as np
import pytest

from helpers import cplx, fixture_cases, pairs, parse_dims, rel_l2
from oracle import Oracle

import refkiss


@pytest.fixture(scope="module")
def oq15():
    return Oracle("int16_t")


@pytest.fixture(scope="module")
def oq31():
    return Oracle("int32_t")


@pytest.fixture(scope="module")
def of32():
    return Oracle("float")


# ---------------------------------------------------------------------------------------
# golden fixtures (reference outputs), bit for bit
# ---------------------------------------------------------------------------------------
def _check_fixed_fixture(O, g):
    n_checked = 0
    for stem, x, want in fixture_cases(g, "c2c"):
        _, n, inv = stem.split("_")
        np.testing.assert_array_equal(O.fft(x, int(inv)), want, err_msg=stem)
        n_checked += 1
    for stem, x, want in fixture_cases(g, "stride"):
        _, n, s = stem.split("_")
        np.testing.assert_array_equal(O.fft_stride(x, int(n), int(s)), want, err_msg=stem)
        n_checked += 1
    for stem, x, want in fixture_cases(g, "r2c"):
        np.testing.assert_array_equal(O.fftr(x), want, err_msg=stem)
        n_checked += 1
    for stem, X, want in fixture_cases(g, "c2r"):
        n = int(stem.split("_")[1])
        np.testing.assert_array_equal(O.fftri(X, n), want, err_msg=stem)
        n_checked += 1
    for stem, x, want in fixture_cases(g, "nd"):
        _, dims, inv = stem.split("_")
        np.testing.assert_array_equal(O.fftnd(x, parse_dims(dims), int(inv)), want, err_msg=stem)
        n_checked += 1
    for stem, x, want in fixture_cases(g, "ndr"):
        np.testing.assert_array_equal(O.fftndr(x, parse_dims(stem.split("_")[1])), want, err_msg=stem)
        n_checked += 1
    for stem, X, want in fixture_cases(g, "ndri"):
        np.testing.assert_array_equal(O.fftndri(X, parse_dims(stem.split("_")[1])), want, err_msg=stem)
        n_checked += 1
    for k in g.files:
        if k.startswith("tw_"):
            _, n, inv = k.split("_")
            np.testing.assert_array_equal(O.twiddles(int(n), int(inv)), g[k], err_msg=k)
        elif k.startswith("stw_"):
            _, n, inv = k.split("_")
            np.testing.assert_array_equal(O.super_twiddles(int(n), int(inv)), g[k], err_msg=k)
    return n_checked


def test_q15_oracle_matches_reference_fixtures(oq15, golden_q15):
    assert _check_fixed_fixture(oq15, golden_q15) > 40


def test_q31_oracle_matches_reference_fixtures(oq31, golden_q31):
    assert _check_fixed_fixture(oq31, golden_q31) > 15


def test_float_oracle_matches_reference_fixtures(of32, golden_float):
    for stem, x, want in fixture_cases(golden_float, "c2c"):
        np.testing.assert_array_equal(of32.fft(x, int(stem.split("_")[2])), want, err_msg=stem)
    for stem, x, want in fixture_cases(golden_float, "r2c"):
        np.testing.assert_array_equal(of32.fftr(x), want, err_msg=stem)
    for stem, x, want in fixture_cases(golden_float, "nd"):
        _, dims, inv = stem.split("_")
        np.testing.assert_array_equal(of32.fftnd(x, parse_dims(dims), int(inv)), want, err_msg=stem)


def test_q15_tie_prone_twiddles_differ_from_naive_rounding(oq15, golden_q15):
    # N=1200 hits x.5 ties: the reference expression and np.round(32767*cos) disagree
    # somewhere, and the oracle must follow the reference (fixture), not the naive form.
    tw = golden_q15["tw_1200_0"].astype(np.int64)
    k = np.arange(1200)
    naive_r = np.round(32767 * np.cos(-2 * np.pi * k / 1200)).astype(np.int64)
    naive_i = np.round(32767 * np.sin(-2 * np.pi * k / 1200)).astype(np.int64)
    assert np.abs(tw[:, 0] - naive_r).max() <= 1 and np.abs(tw[:, 1] - naive_i).max() <= 1
    np.testing.assert_array_equal(oq15.twiddles(1200, 0), golden_q15["tw_1200_0"])


# ---------------------------------------------------------------------------------------
# against the reference build itself (present in the container that has the reference)
# ---------------------------------------------------------------------------------------
@pytest.mark.skipif(not refkiss.available("int16_t"), reason="oracle/_ref not built here")
@pytest.mark.parametrize("datatype", ["int16_t", "int32_t", "float"])
def test_oracle_matches_reference_build(datatype):
    O, R = Oracle(datatype), refkiss.RefKiss(datatype)
    rng = np.random.default_rng(7)
    for n in [6, 10, 14, 18, 20, 24, 27, 36, 45, 81, 100, 125, 143, 169, 200, 243, 360, 625, 2187]:
        for inv in (0, 1):
            if datatype == "float":
                x = rng.uniform(-1, 1, (2, n, 2)).astype(np.float32)
            elif datatype == "int16_t":
                x = rng.integers(-32768, 32768, (2, n, 2)).astype(np.int16)  # full range: wraps
            else:
                x = rng.integers(-(2 ** 31), 2 ** 31, (2, n, 2)).astype(np.int32)
            np.testing.assert_array_equal(O.fft(x, inv), R.fft(x, inv), err_msg=f"{datatype} n={n} inv={inv}")


# ---------------------------------------------------------------------------------------
# what the mathematics fixes
# ---------------------------------------------------------------------------------------
SCHEDULES = {
    1200: [(4, 300), (4, 75), (3, 25), (5, 5), (5, 1)],
    512: [(4, 128), (4, 32), (4, 8), (4, 2), (2, 1)],
    2048: [(4, 512), (4, 128), (4, 32), (4, 8), (4, 2), (2, 1)],
    2000: [(4, 500), (4, 125), (5, 25), (5, 5), (5, 1)],
    96: [(4, 24), (4, 6), (2, 3), (3, 1)],
    120: [(4, 30), (2, 15), (3, 5), (5, 1)],
    1000: [(4, 250), (2, 125), (5, 25), (5, 5), (5, 1)],
    1009: [(1009, 1)],
    3072: [(4, 768), (4, 192), (4, 48), (4, 12), (4, 3), (3, 1)],
    1: [(1, 1)],
    7: [(7, 1)],
    1001: [(7, 143), (11, 13), (13, 1)],
}


@pytest.mark.parametrize("n", sorted(SCHEDULES))
def test_radix_schedule_closed_form(of32, n):
    assert of32.factor(n) == SCHEDULES[n]


def test_radix_schedule_invariants(of32):
    for n in range(1, 3000):
        f = of32.factor(n)
        prod, rem = 1, n
        for p, m in f:
            assert rem % p == 0 and m == rem // p
            rem = m
            prod *= p
        assert prod == n and f[-1][1] == 1


def test_next_fast_size_brute_force(of32):
    def smooth(k):
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        return k == 1
    for n in range(1, 2000):
        want = next(k for k in range(n, 4 * n + 10) if smooth(k))
        assert of32.next_fast_size(n) == want


def test_float_twiddles_are_rounded_fp64(of32):
    for n in (7, 120, 1200, 4096):
        for inv in (0, 1):
            k = np.arange(n)
            w = np.exp((1j if inv else -1j) * 2 * np.pi * k / n)
            tw = of32.twiddles(n, inv)
            np.testing.assert_array_equal(tw[:, 0], w.real.astype(np.float32))
            np.testing.assert_allclose(tw[:, 1], w.imag.astype(np.float32), rtol=0, atol=1.2e-7)


def _dft(x, inverse=False):
    n = x.shape[-1]
    k = np.arange(n)
    M = np.exp((1j if inverse else -1j) * 2 * np.pi * np.outer(k, k) / n)
    return x @ M.T


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 16, 25, 30, 49, 60, 77, 121])
def test_float_oracle_vs_brute_force_dft(of32, n):
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, (3, n, 2)).astype(np.float32)
    for inv in (0, 1):
        got = cplx(of32.fft(x, inv))
        want = _dft(cplx(x), inv)
        assert rel_l2(got, want) < 1e-6, (n, inv)


def test_float_oracle_real_transforms(of32):
    rng = np.random.default_rng(3)
    for n in (2, 4, 12, 30, 96, 1000, 4096):
        x = rng.standard_normal((2, n)).astype(np.float32)
        X = of32.fftr(x)
        want = np.fft.rfft(x.astype(np.float64), axis=-1)
        assert rel_l2(cplx(X), want) < 1e-6
        assert np.all(X[:, 0, 1] == 0) and np.all(X[:, -1, 1] == 0)
        y = of32.fftri(X, n)
        assert rel_l2(y, n * x.astype(np.float64)) < 1e-6  # unnormalised round trip


def test_float_oracle_nd_dc_and_parseval(of32):
    rng = np.random.default_rng(4)
    dims = (6, 10, 12)
    x = rng.uniform(-1, 1, (1,) + dims + (2,)).astype(np.float32)
    X = cplx(of32.fftnd(x, dims))
    assert abs(X[0, 0, 0, 0] - cplx(x).sum()) < 1e-3
    assert rel_l2(X, np.fft.fftn(cplx(x), axes=(1, 2, 3))) < 1e-6
    e_t = (np.abs(cplx(x)) ** 2).sum()
    e_f = (np.abs(X) ** 2).sum() / np.prod(dims)
    assert abs(e_t - e_f) / e_t < 1e-5


def test_float_oracle_ndr(of32):
    rng = np.random.default_rng(5)
    for dims in [(8, 12), (3, 4, 10), (16,)]:
        x = rng.standard_normal((2,) + dims).astype(np.float32)
        X = of32.fftndr(x, dims)
        want = np.fft.rfftn(x.astype(np.float64), axes=tuple(range(1, len(dims) + 1)))
        assert rel_l2(cplx(X), want) < 1e-6, dims
        y = of32.fftndri(X, dims)
        assert rel_l2(y, np.prod(dims) * x.astype(np.float64)) < 1e-6, dims


def test_q15_impulse_and_tone(oq15):
    # A DC input of amplitude a gives X[0] ~= a (each radix-p stage divides by ~p), others ~0;
    # a tone at bin k0 lands in bin k0 (forward) with the right sign convention.
    n = 512
    x = np.zeros((1, n, 2), np.int16)
    x[0, :, 0] = 8000
    X = oq15.fft(x).astype(np.int64)
    assert abs(X[0, 0, 0] - 8000) < 16 and np.abs(X[0, 1:]).max() <= 2
    k0 = 37
    t = np.arange(n)
    x = pairs(16000 * np.exp(2j * np.pi * k0 * t / n), np.int16)[None]
    X = cplx(oq15.fft(x))[0]
    assert np.argmax(np.abs(X)) == k0
    assert abs(X[k0] - 16000) < 40


def test_q15_snr_like_reference_harness(oq15):
    # test/testkiss.py:95-101: input x32767, output xN/32767, SNR >= 10 dB vs NumPy.
    rng = np.random.default_rng(6)
    for dims in [(4,), (3, 4), (2, 3, 4)]:
        z = rng.uniform(-1, 1, dims) + 1j * rng.uniform(-1, 1, dims)
        xq = pairs(32767 * z, np.int16)[None]
        if len(dims) == 1:
            X = cplx(oq15.fft(xq))[0]
        else:
            X = cplx(oq15.fftnd(xq, dims))[0]
        X = X * np.prod(dims) / 32767.0
        want = np.fft.fftn(z)
        snr = 10 * np.log10(np.vdot(want, want).real / np.vdot(X - want, X - want).real)
        assert snr >= 10, (dims, snr)


def test_oracle_detects_a_planted_mistake(of32):
    # A plausible mistake (conjugated twiddles = wrong sign) must fail the brute-force pin.
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, (1, 60, 2)).astype(np.float32)
    wrong = cplx(of32.fft(x, 1))  # inverse direction: exactly the sign slip
    assert rel_l2(wrong, _dft(cplx(x), False)) > 0.5

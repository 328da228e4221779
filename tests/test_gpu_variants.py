"""Every registered fast-kernel variant, forced one at a time, against the oracle / FP64.

The engine times the variants of a size on first use and keeps the fastest; this test
does not rely on that choice: it forces each variant by name (KISSFFT_AMD_VARIANT),
drops the plan cache between variants (kiss_fft_cleanup), and checks the result --
bit-exact for Q15 and Q31, rel-L2 for float.  The variant list is read from the registry source.
"""
import os
import re
import sys

import numpy as np
import pytest

import kissfft_amd
from conftest import ROOT
from helpers import ABS_FLOAT_BOUND, cplx, pairs, rel_l2
from kissfft_amd import workloads as W
from oracle import Oracle

pytestmark = pytest.mark.gpu

REG = os.path.join(ROOT, "kissfft_amd", "csrc", "engine", "kfa_fast_reg.hip")


def registry():
    src = open(REG).read()
    out = []
    fk = {"FK_C2C": "c2c", "FK_R2C": "r2c", "FK_C2R": "c2r"}
    for m in re.finditer(r"add_(real|variant|fused|wave|xpose)<([^>]*)>\(\*v,\s*(?:KFA_\w+,\s*)?\"(\w+)\"\)", src):
        fn, targs, name = m.groups()
        a = [x.strip() for x in targs.split(",")]
        if fn in ("wave", "xpose"):  # <T, LPB, radices...> / <LPB, WPE, radices...>: Q15 complex
            n = int(np.prod([int(r) for r in a[2:]]))
            out.append(("Q15P", "c2c", n, name))
            continue
        if fn == "real":  # <A, T, LPB, DB, radices...>
            kinds, radices = ["r2c", "c2r"], a[4:]
        elif fn == "variant":  # <A, FK, T, LPB, DB, FUSE, radices...>
            kinds, radices = [fk[a[1]]], a[6:]
        else:  # fused: <A, FK, T, LPB, radices...>
            kinds, radices = [fk[a[1]]], a[4:]
        n = int(np.prod([int(r) for r in radices]))
        for k in kinds:
            out.append((a[0], k, n, name))
    return out


CASES = registry()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_registry_parsed():
    assert len(CASES) >= 20


@pytest.mark.parametrize("typ,kind,n,name", CASES, ids=[f"{c[3]}-{c[1]}" for c in CASES])
def test_forced_variant(typ, kind, n, name, monkeypatch, capfd):
    datatype = {"F32": "float", "Q15": "int16_t", "Q15P": "int16_t", "Q31": "int32_t"}[typ]
    lib = kissfft_amd.load(datatype)
    lib.L.kiss_fft_cleanup()
    monkeypatch.setenv("KISSFFT_AMD_VARIANT", name)
    monkeypatch.setenv("KISSFFT_AMD_VERBOSE", "1")
    say("variant", name, kind, n)
    rng = np.random.default_rng(hash(name) & 0xFFFF)
    B = 37  # not a multiple of any lines-per-workgroup: exercises the partial last workgroup
    try:
        if datatype in ("int16_t", "int32_t"):
            O = Oracle(datatype)
            dt = np.int16 if datatype == "int16_t" else np.int32
            lo, hi = (-32768, 32768) if datatype == "int16_t" else (-(1 << 31), 1 << 31)
            if kind == "c2c":
                x = rng.integers(lo, hi, (B, n, 2)).astype(dt)
                for inv in (False, True):
                    cfg = lib.fft_alloc(n, inv)
                    y = np.zeros_like(x)
                    lib.fft_batch(cfg, x, y, B, n, n)
                    np.testing.assert_array_equal(y, O.fft(x, inv))
                # the strided instance: every other input element (in_stride 2), output rows
                # n + 3 apart; the gaps between output rows keep the caller's bytes
                xs = rng.integers(lo, hi, (B, 2 * n, 2)).astype(dt)
                ys = np.full((B, n + 3, 2), 7, dt)
                lib.fft_batch(lib.fft_alloc(n), xs, ys, B, 2 * n, n + 3, in_stride=2)
                np.testing.assert_array_equal(ys[:, :n], O.fft(np.ascontiguousarray(xs[:, ::2])))
                assert (ys[:, n:] == 7).all()
            elif kind == "r2c":
                x = rng.integers(lo, hi, (B, 2 * n)).astype(dt)
                y = np.zeros((B, n + 1, 2), dt)
                lib.fftr_batch(lib.fftr_alloc(2 * n), x, y, B, 2 * n, n + 1)
                np.testing.assert_array_equal(y, O.fftr(x))
            else:
                X = rng.integers(lo, hi, (B, n + 1, 2)).astype(dt)
                y = np.zeros((B, 2 * n), dt)
                lib.fftri_batch(lib.fftr_alloc(2 * n, True), X, y, B, n + 1, 2 * n)
                np.testing.assert_array_equal(y, O.fftri(X, 2 * n))
        else:
            if kind == "c2c":
                x = rng.uniform(-1, 1, (B, n, 2)).astype(np.float32)
                for inv in (False, True):
                    cfg = lib.fft_alloc(n, inv)
                    y = np.zeros_like(x)
                    lib.fft_batch(cfg, x, y, B, n, n)
                    want = np.fft.ifft(cplx(x), axis=-1) * n if inv else np.fft.fft(cplx(x), axis=-1)
                    assert rel_l2(cplx(y), want) <= ABS_FLOAT_BOUND
            elif kind == "r2c":
                x = W.two_tone(7, 2 * n, B)
                y = np.zeros((B, n + 1, 2), np.float32)
                lib.fftr_batch(lib.fftr_alloc(2 * n), x, y, B, 2 * n, n + 1)
                assert rel_l2(cplx(y), np.fft.rfft(x.astype(np.float64), axis=-1)) <= 3e-7
            else:
                Xc = np.fft.rfft(W.two_tone(8, 2 * n, B).astype(np.float64), axis=-1).astype(np.complex64)
                X = np.stack([Xc.real, Xc.imag], -1).astype(np.float32)
                y = np.zeros((B, 2 * n), np.float32)
                lib.fftri_batch(lib.fftr_alloc(2 * n, True), X, y, B, n + 1, 2 * n)
                want = np.fft.irfft(Xc.astype(np.complex128), n=2 * n, axis=-1) * 2 * n
                assert rel_l2(y, want) <= 4e-7
        # the forced variant is the one that ran (a name the plan cannot take is ignored)
        assert f"fast variant {name}\n" in capfd.readouterr().err, name
    finally:
        lib.L.kiss_fft_cleanup()


TREG = os.path.join(ROOT, "kissfft_amd", "csrc", "engine", "kfa_tile_reg.hip")


def tile_registry():
    out = []
    for m in re.finditer(r"add_tile<([^>]*)>\(\*v,\s*\"(\w+)\"\)", open(TREG).read()):
        a = [int(x) for x in m.group(1).split(",")]
        out.append((int(np.prod(a[2:])), m.group(2)))
    return out


TILES = tile_registry()


def test_tile_registry_parsed():
    assert len(TILES) >= 6


@pytest.mark.parametrize("n,name", TILES, ids=[t[1] for t in TILES])
def test_forced_tile(n, name, monkeypatch):
    """Each tile variant forced: as the column pass of a 2-D transform (n x 40), and as both
    four-step passes of an n*n line where that line is longer than one LDS line."""
    lib = kissfft_amd.load("float")
    monkeypatch.setenv("KISSFFT_AMD_TILE", name)
    monkeypatch.setenv("KISSFFT_AMD_VERBOSE", "1")
    lib.L.kiss_fft_cleanup()
    try:
        dims = (n, 40)
        x = W.complex_pm1(70, n * 40, 2).reshape((2,) + dims + (2,))
        for inv in (False, True):
            say("tile", name, "nd", dims, inv)
            cfg = lib.fftnd_alloc(dims, inv)
            y = np.zeros_like(x)
            lib.fftnd_batch(cfg, x, y, 2)
            want = np.fft.ifftn(cplx(x), axes=(1, 2)) * n * 40 if inv else np.fft.fftn(cplx(x), axes=(1, 2))
            assert rel_l2(cplx(y), want) <= ABS_FLOAT_BOUND, (name, inv)
        N = n * n
        if N > 8192:
            xs = W.complex_pm1(71, N, 2)
            for inv in (False, True):
                say("tile", name, "four-step", N, inv)
                ys = np.zeros_like(xs)
                lib.fft_batch(lib.fft_alloc(N, inv), xs, ys, 2, N, N)
                want = np.fft.ifft(cplx(xs), axis=-1) * N if inv else np.fft.fft(cplx(xs), axis=-1)
                assert rel_l2(cplx(ys), want) <= ABS_FLOAT_BOUND, (name, inv)
    finally:
        lib.L.kiss_fft_cleanup()


def rtile_registry():
    out = []
    for m in re.finditer(r"add_rtile<([^>]*)>\(\*v,\s*\"(\w+)\"\)", open(TREG).read()):
        a = [int(x) for x in m.group(1).split(",")]
        out.append((int(np.prod(a[2:])), m.group(2)))
    return out


RTILES = rtile_registry()


def test_rtile_registry_parsed():
    assert len(RTILES) >= 4


@pytest.mark.parametrize("n,name", RTILES, ids=[t[1] for t in RTILES])
def test_forced_rtile(n, name, monkeypatch):
    """Each fused real split / merge variant forced, on real lines of 2*n*n points (half
    length n*n, split n x n)."""
    lib = kissfft_amd.load("float")
    monkeypatch.setenv("KISSFFT_AMD_TILE", name)
    monkeypatch.setenv("KISSFFT_AMD_VERBOSE", "1")
    lib.L.kiss_fft_cleanup()
    try:
        N = 2 * n * n
        x = W.two_tone(72, N, 3)
        say("rtile", name, "real", N)
        X = np.zeros((3, N // 2 + 1, 2), np.float32)
        lib.fftr_batch(lib.fftr_alloc(N, False), x, X, 3, N, N // 2 + 1)
        want = np.fft.rfft(x.astype(np.float64), axis=-1)
        assert rel_l2(cplx(X), want) <= ABS_FLOAT_BOUND, name
        Xin = pairs(want.astype(np.complex64))
        y = np.zeros((3, N), np.float32)
        lib.fftri_batch(lib.fftr_alloc(N, True), Xin, y, 3, N // 2 + 1, N)
        assert rel_l2(y, np.fft.irfft(cplx(Xin), n=N, axis=-1) * N) <= ABS_FLOAT_BOUND, name
    finally:
        lib.L.kiss_fft_cleanup()

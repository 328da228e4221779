"""GPU parity of the N-D paths (kiss_fftnd / kiss_fftndr / kiss_fftndri / batched forms)."""
import os
import re
import sys

import numpy as np
import pytest

import kissfft_amd
from conftest import ROOT
from helpers import ABS_FLOAT_BOUND, cplx, fixture_cases, float_bound, pairs, parse_dims, rel_l2
from kissfft_amd import workloads as W
from oracle import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def nd(lib, dims, x, inverse=False):
    cfg = lib.fftnd_alloc(dims, inverse)
    out = np.zeros_like(x)
    lib.fftnd_batch(cfg, np.ascontiguousarray(x), out, x.shape[0])
    return out


def ndr(lib, dims, x):
    cfg = lib.fftndr_alloc(dims, False)
    out = np.zeros((x.shape[0],) + tuple(dims[:-1]) + (dims[-1] // 2 + 1, 2), lib.scalar)
    lib.fftndr_batch(cfg, np.ascontiguousarray(x), out, x.shape[0])
    return out


def ndri(lib, dims, X):
    cfg = lib.fftndr_alloc(dims, True)
    out = np.zeros((X.shape[0],) + tuple(dims), lib.scalar)
    lib.fftndri_batch(cfg, np.ascontiguousarray(X), out, X.shape[0])
    return out


@pytest.mark.parametrize("datatype,gname", [("int16_t", "golden_q15"), ("int32_t", "golden_q31")])
def test_fixed_nd_matches_reference_fixtures(datatype, gname, request):
    lib = kissfft_amd.load(datatype)
    g = request.getfixturevalue(gname)
    for stem, x, want in fixture_cases(g, "nd"):
        _, dims, inv = stem.split("_")
        dims = parse_dims(dims)
        say("case", datatype, stem)
        np.testing.assert_array_equal(nd(lib, dims, x, bool(int(inv))), want, err_msg=stem)
        cfg = lib.fftnd_alloc(dims, bool(int(inv)))
        one = np.zeros_like(x[0])
        lib.kiss_fftnd(cfg, np.ascontiguousarray(x[0]), one)
        np.testing.assert_array_equal(one, want[0], err_msg=stem + " single")
    for stem, x, want in fixture_cases(g, "ndr"):
        dims = parse_dims(stem.split("_")[1])
        say("case", datatype, stem)
        np.testing.assert_array_equal(ndr(lib, dims, x), want, err_msg=stem)
    for stem, X, want in fixture_cases(g, "ndri"):
        dims = parse_dims(stem.split("_")[1])
        say("case", datatype, stem)
        np.testing.assert_array_equal(ndri(lib, dims, X), want, err_msg=stem)


def test_q15_nd_bit_exact_vs_oracle_larger_shapes():
    lib = kissfft_amd.load("int16_t")
    O = Oracle("int16_t")
    rng = np.random.default_rng(41)
    for dims in [(64, 96, 120), (32, 1024), (3, 5, 7, 9), (1009,)]:
        x = rng.integers(-32768, 32768, (1,) + dims + (2,)).astype(np.int16)
        say("case q15 nd", dims)
        np.testing.assert_array_equal(nd(lib, dims, x), O.fftnd(x, dims))
    for dims in [(64, 100), (16, 12, 10), (1024, 1000)]:
        x = rng.integers(-32768, 32768, (1,) + dims).astype(np.int16)
        say("case q15 ndr", dims)
        np.testing.assert_array_equal(ndr(lib, dims, x), O.fftndr(x, dims))
        X = rng.integers(-32768, 32768, (1,) + dims[:-1] + (dims[-1] // 2 + 1, 2)).astype(np.int16)
        np.testing.assert_array_equal(ndri(lib, dims, X), O.fftndri(X, dims))


@pytest.mark.parametrize("dtype", ["int16_t", "int32_t", "float"])
def test_nd_generic_radix_axes(dtype):
    """N-D arrays whose axes have the generic radices 7, 11, 13: a last axis of 1001 (its
    contiguous lines on the kernel compiled at run time), an outer axis of 143 and a real
    last axis of 2002, against the oracle (fixed point, bit-exact) or FP64 (float)."""
    lib = kissfft_amd.load(dtype)
    rng = np.random.default_rng(44)
    lo, hi, dt = {"int16_t": (-32768, 32768, np.int16), "int32_t": (-(1 << 31), 1 << 31, np.int32),
                  "float": (None, None, np.float32)}[dtype]
    def rnd(shape):
        if dtype == "float":
            return rng.uniform(-1, 1, shape).astype(np.float32)
        return rng.integers(lo, hi, shape, dtype=np.int64).astype(dt)
    O = Oracle(dtype) if dtype != "float" else None
    for dims in [(6, 1001), (143, 12), (3, 77, 13)]:
        x = rnd((2,) + dims + (2,))
        say("case nd generic", dtype, dims)
        got = nd(lib, dims, x)
        if O is not None:
            np.testing.assert_array_equal(got, O.fftnd(x, dims))
        else:
            axes = tuple(range(1, len(dims) + 1))
            assert rel_l2(cplx(got), np.fft.fftn(cplx(x), axes=axes)) <= ABS_FLOAT_BOUND, dims
    dims = (5, 2002)
    x = rnd((2,) + dims)
    say("case ndr generic", dtype, dims)
    got = ndr(lib, dims, x)
    if O is not None:
        np.testing.assert_array_equal(got, O.fftndr(x, dims))
    else:
        assert rel_l2(cplx(got), np.fft.rfftn(x.astype(np.float64), axes=(1, 2))) <= ABS_FLOAT_BOUND


def test_float_nd_workload(golden_float):
    # 16 transforms of 64x96x120, re/im uniform on [-1,1), seed 4
    lib = kissfft_amd.load("float")
    dims = (64, 96, 120)
    x = W.complex_pm1(4, int(np.prod(dims)), 16).reshape((16,) + dims + (2,))
    say("case float nd 64x96x120 x16")
    got = cplx(nd(lib, dims, x))
    err = rel_l2(got, np.fft.fftn(cplx(x), axes=(1, 2, 3)))
    say("rel-L2", err, "bound", float_bound(golden_float, "nd_64x96x120"))
    assert err <= float_bound(golden_float, "nd_64x96x120")


def test_float_ndr_workload(golden_float):
    # 32 transforms of 1024x1000 real N(0,1), seed 5 (the reference's alloc rejects this shape)
    lib = kissfft_amd.load("float")
    dims = (1024, 1000)
    x = W.normal(5, int(np.prod(dims)) * 32).reshape((32,) + dims)
    say("case float ndr 1024x1000 x32")
    X = ndr(lib, dims, x)
    want = np.fft.rfftn(x.astype(np.float64), axes=(1, 2))
    err = rel_l2(cplx(X), want)
    say("rel-L2", err, "bound", float_bound(golden_float, "ndr_1024x1000"))
    assert err <= float_bound(golden_float, "ndr_1024x1000")
    y = ndri(lib, dims, X)
    rt = rel_l2(y, np.prod(dims) * x.astype(np.float64))
    say("round trip", rt)
    assert rt <= 2 * ABS_FLOAT_BOUND


def test_float_nd_shapes():
    lib = kissfft_amd.load("float")
    rng = np.random.default_rng(42)
    for dims in [(2, 3), (4, 4, 4), (7, 11), (2, 3, 4, 5), (1, 8), (5,), (30, 1, 6)]:
        x = rng.uniform(-1, 1, (2,) + dims + (2,)).astype(np.float32)
        for inv in (False, True):
            got = cplx(nd(lib, dims, x, inv))
            axes = tuple(range(1, len(dims) + 1))
            want = np.fft.ifftn(cplx(x), axes=axes) * np.prod(dims) if inv else np.fft.fftn(cplx(x), axes=axes)
            assert rel_l2(got, want) <= ABS_FLOAT_BOUND, (dims, inv)
    for dims in [(4, 6), (3, 5, 8), (12,), (2, 2, 2, 2)]:
        x = rng.standard_normal((2,) + dims).astype(np.float32)
        X = ndr(lib, dims, x)
        axes = tuple(range(1, len(dims) + 1))
        assert rel_l2(cplx(X), np.fft.rfftn(x.astype(np.float64), axes=axes)) <= ABS_FLOAT_BOUND, dims
        y = ndri(lib, dims, X)
        assert rel_l2(y, np.prod(dims) * x.astype(np.float64)) <= 2 * ABS_FLOAT_BOUND, dims


@pytest.mark.parametrize("dims", [(96, 64), (1024, 24), (256, 10), (512, 3), (128, 5, 7), (64, 64, 64),
                                  (3, 1024, 9)])
def test_float_nd_tile_axes(dims):
    # non-last axes of registered lengths run as column tiles
    lib = kissfft_amd.load("float")
    x = W.complex_pm1(44, int(np.prod(dims)), 3).reshape((3,) + dims + (2,))
    axes = tuple(range(1, len(dims) + 1))
    for inv in (False, True):
        say("case nd tiles", dims, inv)
        got = cplx(nd(lib, dims, x, inv))
        want = np.fft.ifftn(cplx(x), axes=axes) * np.prod(dims) if inv else np.fft.fftn(cplx(x), axes=axes)
        assert rel_l2(got, want) <= ABS_FLOAT_BOUND, (dims, inv)


def test_float_ndr_long_real_axis():
    # real last axis longer than one LDS line (multi-pass real rows) under a column axis
    lib = kissfft_amd.load("float")
    dims = (64, 65536)
    x = W.normal(45, int(np.prod(dims)) * 2).reshape((2,) + dims)
    X = ndr(lib, dims, x)
    assert rel_l2(cplx(X), np.fft.rfftn(x.astype(np.float64), axes=(1, 2))) <= ABS_FLOAT_BOUND
    y = ndri(lib, dims, X)
    assert rel_l2(y, np.prod(dims) * x.astype(np.float64)) <= 2 * ABS_FLOAT_BOUND


def test_nd_in_place_single_call():
    lib = kissfft_amd.load("float")
    dims = (8, 12, 10)
    x = W.complex_pm1(43, int(np.prod(dims)), 1).reshape(dims + (2,))
    want = np.fft.fftn(cplx(x))
    cfg = lib.fftnd_alloc(dims)
    lib.kiss_fftnd(cfg, x, x)
    assert rel_l2(cplx(x), want) <= ABS_FLOAT_BOUND


def plane_registry():
    src = open(os.path.join(ROOT, "kissfft_amd", "csrc", "engine", "kfa_tile_reg.hip")).read()
    out = []
    # add_plane<TA, TB, NTH, Radices<...>, Radices<...>[, SLAB[, PF]]>(*v, "name")
    pat = r"add_plane<\d+, \d+, \d+, Radices<2, ([\d, ]+)>, Radices<2, ([\d, ]+)>((?:, (?:true|false))*)>\(\*v, \"(\w+)\"\)"
    for m in re.finditer(pat, src):
        A = int(np.prod([int(v) for v in m.group(1).split(",")]))
        B = int(np.prod([int(v) for v in m.group(2).split(",")]))
        flags = [f.strip() for f in m.group(3).split(",") if f.strip()]
        out.append((A, B, m.group(4), not flags or flags[0] == "true"))
    return out


PLANES = [p[:3] for p in plane_registry()]
SLABS = [p[:3] for p in plane_registry() if p[3]]


def test_plane_registry_parsed():
    assert len(PLANES) >= 3 and (64, 64, "plane64x64_t8x8") in PLANES


@pytest.mark.parametrize("A,B,name", PLANES, ids=[p[2] for p in PLANES])
@pytest.mark.parametrize("lead", [(), (3,), (2, 5)], ids=["2d", "3d", "4d"])
def test_float_nd_plane(A, B, name, lead, monkeypatch):
    """The last two axes in one pass (plane kernel forced by name), under 0-2 leading axes,
    both directions, against numpy; the same call with the plane pass off agrees too."""
    lib = kissfft_amd.load("float")
    dims = lead + (A, B)
    x = W.complex_pm1(46, int(np.prod(dims)), 2).reshape((2,) + dims + (2,))
    axes = tuple(range(1, len(dims) + 1))
    monkeypatch.setenv("KISSFFT_AMD_TILE", name)
    monkeypatch.setenv("KISSFFT_AMD_VERBOSE", "1")
    lib.L.kiss_fft_cleanup()
    try:
        for inv in (False, True):
            say("case nd plane", name, dims, inv)
            want = np.fft.ifftn(cplx(x), axes=axes) * np.prod(dims) if inv else np.fft.fftn(cplx(x), axes=axes)
            got = cplx(nd(lib, dims, x, inv))
            assert rel_l2(got, want) <= ABS_FLOAT_BOUND, (dims, inv)
        monkeypatch.setenv("KISSFFT_AMD_PLANE", "0")
        lib.L.kiss_fft_cleanup()
        got = cplx(nd(lib, dims, x))
        assert rel_l2(got, np.fft.fftn(cplx(x), axes=axes)) <= ABS_FLOAT_BOUND, dims
    finally:
        lib.L.kiss_fft_cleanup()


@pytest.mark.parametrize("A,B,name", SLABS, ids=[p[2] for p in SLABS])
@pytest.mark.parametrize("shape", ["a_m_b", "l_a_m_b", "a_mm_b"])
def test_float_nd_slab(A, B, name, shape, monkeypatch, capfd):
    """A plane variant A x B used as a slab of two non-adjacent axes (the outer axis of
    length A and the last axis of length B, the axes between them fixed per plane), under
    a leading axis or none, both directions, against numpy."""
    dims = {"a_m_b": (A, 3, B), "l_a_m_b": (2, A, 3, B), "a_mm_b": (A, 2, 5, B)}[shape]
    lib = kissfft_amd.load("float")
    x = W.complex_pm1(48, int(np.prod(dims)), 2).reshape((2,) + dims + (2,))
    axes = tuple(range(1, len(dims) + 1))
    monkeypatch.setenv("KISSFFT_AMD_TILE", name)
    monkeypatch.setenv("KISSFFT_AMD_VERBOSE", "1")
    lib.L.kiss_fft_cleanup()
    try:
        for inv in (False, True):
            say("case nd slab", name, dims, inv)
            capfd.readouterr()
            want = np.fft.ifftn(cplx(x), axes=axes) * np.prod(dims) if inv else np.fft.fftn(cplx(x), axes=axes)
            got = cplx(nd(lib, dims, x, inv))
            assert f"plane {A}x{B} inverse={int(inv)}: {name}" in capfd.readouterr().err, "slab pass not used"
            assert rel_l2(got, want) <= ABS_FLOAT_BOUND, (dims, inv)
    finally:
        lib.L.kiss_fft_cleanup()


def test_float_nd_plane_in_place():
    lib = kissfft_amd.load("float")
    dims = (4, 64, 64)
    x = W.complex_pm1(47, int(np.prod(dims)), 3).reshape((3,) + dims + (2,))
    want = np.fft.fftn(cplx(x), axes=(1, 2, 3))
    cfg = lib.fftnd_alloc(dims)
    lib.fftnd_batch(cfg, x, x, 3)
    assert rel_l2(cplx(x), want) <= ABS_FLOAT_BOUND

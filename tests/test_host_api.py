"""CPU tests of the drop-in C-ABI libraries: they load, export every declared symbol,
follow the reference allocation protocol, and build plans (radix schedule, twiddles)
bit-identical to the reference.  No transform runs here (no GPU in this process)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import kissfft_amd
import refkiss
from conftest import ROOT
from oracle import Oracle

DATATYPES = ["float", "int16_t", "int32_t"]
TW_OFFSET = 4 + 4 + 4 * 64  # struct kiss_fft_state: nfft, inverse, factors[64], twiddles


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = "\n".join(l for l in open(h).read().splitlines() if not l.lstrip().startswith("#"))
        for m in re.finditer(r"KISS_FFT_API\s*\*?\s*(\w+)\s*\(", src):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module", params=DATATYPES)
def lib(request):
    return kissfft_amd.load(request.param)


def test_every_declared_symbol_is_exported(lib):
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib.L, n)]
    assert not missing, missing


def test_reference_entry_points_present():
    names = set(declared_symbols())
    for n in ["kiss_fft_alloc", "kiss_fft", "kiss_fft_stride", "kiss_fft_cleanup", "kiss_fft_next_fast_size",
              "kiss_fftr_alloc", "kiss_fftr", "kiss_fftri", "kiss_fftnd_alloc", "kiss_fftnd",
              "kiss_fftndr_alloc", "kiss_fftndr", "kiss_fftndri", "kfc_fft", "kfc_ifft", "kfc_cleanup"]:
        assert n in names


def test_soname_and_internal_symbols_hidden(lib):
    real = os.path.realpath(lib.path)
    assert real.endswith(f"libkissfft-{lib.datatype}.so.131")
    assert not hasattr(lib.L, "kfa_c2c")  # engine seam stays internal


def test_alloc_protocol(lib):
    L = lib.L
    n = ctypes.c_size_t(0)
    assert L.kiss_fft_alloc(1200, 0, None, ctypes.byref(n)) is None
    cb = 2 * np.dtype(lib.scalar).itemsize
    assert n.value == TW_OFFSET + cb * 1200  # sizeof(state) + sizeof(cpx)*(nfft-1)
    buf = ctypes.create_string_buffer(n.value)
    small = ctypes.c_size_t(n.value - 1)
    assert L.kiss_fft_alloc(1200, 0, buf, ctypes.byref(small)) is None
    assert small.value == n.value
    ok = ctypes.c_size_t(n.value)
    assert L.kiss_fft_alloc(1200, 0, buf, ctypes.byref(ok)) == ctypes.addressof(buf)
    assert L.kiss_fft_alloc(0, 0, None, None) is None
    assert L.kiss_fft_alloc(-5, 0, None, None) is None
    assert L.kiss_fftr_alloc(1001, 0, None, None) is None  # odd real size
    assert L.kiss_fftndr_alloc((ctypes.c_int * 2)(4, 7), 2, 0, None, None) is None


@pytest.mark.skipif(not refkiss.available("int16_t"), reason="oracle/_ref not built here")
def test_alloc_sizes_equal_reference(lib):
    R = refkiss.RefKiss(lib.datatype).L
    for n in (1, 2, 7, 96, 1200, 4096):
        a, b = ctypes.c_size_t(0), ctypes.c_size_t(0)
        lib.L.kiss_fft_alloc(n, 0, None, ctypes.byref(a))
        R.kiss_fft_alloc(n, 0, None, ctypes.byref(b))
        assert a.value == b.value
        if n % 2 == 0:
            lib.L.kiss_fftr_alloc(n, 1, None, ctypes.byref(a))
            R.kiss_fftr_alloc(n, 1, None, ctypes.byref(b))
            assert a.value == b.value
    for dims in [(4, 6, 10), (8, 12), (64, 96, 120)]:
        a, b = ctypes.c_size_t(0), ctypes.c_size_t(0)
        d = (ctypes.c_int * len(dims))(*dims)
        lib.L.kiss_fftnd_alloc(d, len(dims), 0, None, ctypes.byref(a))
        R.kiss_fftnd_alloc(d, len(dims), 0, None, ctypes.byref(b))
        assert a.value == b.value


def _cfg_twiddles(lib, cfg_ptr, n):
    nbytes = n * 2 * np.dtype(lib.scalar).itemsize
    raw = (ctypes.c_char * nbytes).from_address(cfg_ptr + TW_OFFSET)
    return np.frombuffer(bytes(raw), lib.scalar).reshape(n, 2)


def _cfg_factors(cfg_ptr):
    raw = (ctypes.c_int * 66).from_address(cfg_ptr)
    out, k = [], 0
    while True:
        p, m = raw[2 + 2 * k], raw[3 + 2 * k]
        out.append((p, m))
        k += 1
        if m == 1:
            break
    return out


def test_host_plan_equals_reference_q15(golden_q15):
    lib = kissfft_amd.load("int16_t")
    for k in golden_q15.files:
        if k.startswith("tw_"):
            _, n, inv = k.split("_")
            cfg = lib.fft_alloc(int(n), bool(int(inv)))
            np.testing.assert_array_equal(_cfg_twiddles(lib, cfg.ptr, int(n)), golden_q15[k], err_msg=k)


def test_host_plan_equals_reference_q31(golden_q31):
    lib = kissfft_amd.load("int32_t")
    for k in golden_q31.files:
        if k.startswith("tw_"):
            _, n, inv = k.split("_")
            cfg = lib.fft_alloc(int(n), bool(int(inv)))
            np.testing.assert_array_equal(_cfg_twiddles(lib, cfg.ptr, int(n)), golden_q31[k], err_msg=k)


def test_host_super_twiddles_equal_reference(golden_q15):
    lib = kissfft_amd.load("int16_t")
    for k in golden_q15.files:
        if k.startswith("stw_"):
            _, n, inv = k.split("_")
            cfg = lib.fftr_alloc(int(n), bool(int(inv)))
            stw_ptr = ctypes.c_void_p.from_address(cfg.ptr + 2 * ctypes.sizeof(ctypes.c_void_p)).value
            cnt = int(n) // 4
            raw = (ctypes.c_char * (cnt * 4)).from_address(stw_ptr)
            got = np.frombuffer(bytes(raw), np.int16).reshape(cnt, 2)
            np.testing.assert_array_equal(got, golden_q15[k], err_msg=k)


def test_host_schedule_and_float_twiddles_equal_oracle():
    lib = kissfft_amd.load("float")
    O = Oracle("float")
    for n in (1, 2, 3, 7, 96, 120, 512, 1000, 1009, 1200, 2048, 3072, 65536):
        for inv in (0, 1):
            cfg = lib.fft_alloc(n, bool(inv))
            assert _cfg_factors(cfg.ptr) == O.factor(n)
            np.testing.assert_array_equal(_cfg_twiddles(lib, cfg.ptr, n), O.twiddles(n, inv))


def test_next_fast_size(lib):
    O = Oracle("float")
    for n in list(range(1, 300)) + [1009, 4097, 65537]:
        assert lib.L.kiss_fft_next_fast_size(n) == O.next_fast_size(n)


@pytest.mark.skipif(kissfft_amd.load("float").device_available(), reason="a GPU is visible")
def test_no_gpu_fails_loudly():
    lib = kissfft_amd.load("float")
    cfg = lib.fft_alloc(64)
    x = np.zeros((64, 2), np.float32)
    y = np.full((64, 2), 7.0, np.float32)
    with pytest.raises(kissfft_amd.KissFFTError, match="no HIP device"):
        lib.fft_batch(cfg, x, y, 1, 64, 64)
    assert np.all(y == 7.0)  # nothing written, no CPU fallback


def test_fastfir_header_maps_real_names(tmp_path):
    """REAL_FASTFIR selects the real filter under the reference's names (the header's
    macro mapping), both spellings compile against the declarations."""
    import subprocess
    src = tmp_path / "ff.c"
    src.write_text('#include "kiss_fastfir.h"\n'
                   "size_t f(kiss_fft_scalar *h, kiss_fft_scalar *x, kiss_fft_scalar *y) {\n"
                   "  size_t nfft = 0, off = 0;\n"
                   "  kiss_fastfir_cfg c = kiss_fastfir_alloc(h, 10, &nfft, 0, 0);\n"
                   "  return kiss_fastfir(c, x, y, 100, &off); }\n")
    inc = os.path.join(ROOT, "include")
    r = subprocess.run(["gcc", "-fsyntax-only", "-Werror", "-DREAL_FASTFIR", "-I", inc, str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run(["gcc", "-fsyntax-only", "-Werror", "-I", inc, str(src)], capture_output=True, text=True)
    assert r.returncode != 0  # without REAL_FASTFIR the names take complex samples


def test_fastfir_fixed_point_refused():
    lib = kissfft_amd.load("int16_t")
    h = np.ones(8, np.int16)
    with pytest.raises(kissfft_amd.KissFFTError):
        lib.fastfir_alloc(h, real=True)


def test_cpp_class_header_compiles():
    """include/kissfft.hh compiles as C++17 with warnings as errors (the test driver)."""
    import subprocess
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "facade_driver.cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_build_rejects_kernels_with_scratch():
    """The build reads the compiler's kernel resource-usage remarks and fails on a kernel
    with scratch memory (a spill or a stack copy), except the general line kernels, whose
    run-time radix needs a run-time-indexed local array."""
    from kissfft_amd import build as B

    tail = " [-Rpass-analysis=kernel-resource-usage]"
    def remarks(name, scratch):
        return (f"x.hpp:1:1: remark: Function Name: {name}{tail}\n"
                f"x.hpp:1:1: remark:     VGPRs: 40{tail}\n"
                f"x.hpp:1:1: remark:     ScratchSize [bytes/lane]: {scratch}{tail}\n")
    text = (remarks("_ZN3kfa6k_fastINS_4Q15PEEEvNS_8FastArgsE", 1624)
            + remarks("_ZN3kfa6k_fastINS_3F32EEEvNS_8FastArgsE", 0)
            + remarks("_ZN3kfa7k_linesINS_3F32EEEvPKNT_6store2E", 96)
            + remarks("_ZN3kfa7k_multiINS_3F32EEEvPKNS_9MultiItemEix", 96))
    assert B._scratch_kernels(text) == [("_ZN3kfa6k_fastINS_4Q15PEEEvNS_8FastArgsE", 1624)]
    assert B._scratch_kernels(remarks("_ZN3kfa6k_tileE", 0)) == []

"""GPU tests of the C++ class interface include/kissfft.hh (the reference's kissfft.hh).

A small driver (tests/cpp/facade_driver.cpp) is compiled against the header and
libkissfft-float and run on files; results are compared with numpy's FP64 transforms,
including the reference's packing of the real transform (dst[0] = (X[0], X[N])).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT
from helpers import ABS_FLOAT_BOUND, rel_l2

pytestmark = pytest.mark.gpu
LIB = os.path.join(ROOT, "kissfft_amd", "lib")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cpp") / "facade_driver")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "facade_driver.cpp"), "-L", LIB, "-lkissfft-float",
                        f"-Wl,-rpath,{LIB}", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def run(driver, tmp_path, args, data: np.ndarray, nout: int) -> np.ndarray:
    fin, fout = tmp_path / "in.raw", tmp_path / "out.raw"
    np.ascontiguousarray(data, np.float32).tofile(fin)
    r = subprocess.run([driver] + [str(a) for a in args] + [str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    y = np.fromfile(fout, np.float32)
    assert y.size == 2 * nout
    return y[0::2] + 1j * y[1::2]


@pytest.mark.parametrize("n", [1200, 1009, 4096, 60])
@pytest.mark.parametrize("inverse", [0, 1])
def test_transform(driver, tmp_path, n, inverse):
    rng = np.random.default_rng(n + inverse)
    s = 3
    x = rng.uniform(-1, 1, (n * s, 2)).astype(np.float32)
    got = run(driver, tmp_path, ["c2c", n, inverse, s], x, n)
    xs = x[::s, 0] + 1j * x[::s, 1]
    want = np.fft.ifft(xs) * n if inverse else np.fft.fft(xs)
    assert rel_l2(got, want) <= ABS_FLOAT_BOUND


@pytest.mark.parametrize("n", [2048, 1000, 7])
@pytest.mark.parametrize("inverse", [0, 1])
def test_transform_real_packing(driver, tmp_path, n, inverse):
    rng = np.random.default_rng(3 * n + inverse)
    x = rng.standard_normal(2 * n).astype(np.float32)
    got = run(driver, tmp_path, ["real", n, inverse], x, n)
    X = np.fft.rfft(x.astype(np.float64))  # 2n reals -> n+1 bins
    if inverse:
        X = np.conj(X)
    want = X[:n].copy()
    want[0] = X[0].real + 1j * X[n].real
    assert rel_l2(got, want) <= ABS_FLOAT_BOUND


def test_transform_many_and_assign(driver, tmp_path):
    rng = np.random.default_rng(5)
    n, howmany = 1000, 37
    x = rng.uniform(-1, 1, (howmany, n, 2)).astype(np.float32)
    got = run(driver, tmp_path, ["many", n, 0, howmany], x, n * howmany).reshape(howmany, n)
    assert rel_l2(got, np.fft.fft(x[..., 0] + 1j * x[..., 1], axis=-1)) <= ABS_FLOAT_BOUND
    x1 = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    got = run(driver, tmp_path, ["assign", n, 1], x1, n)
    assert rel_l2(got, np.fft.ifft(x1[:, 0] + 1j * x1[:, 1]) * n) <= ABS_FLOAT_BOUND

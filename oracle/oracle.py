"""ctypes binding of the oracle libraries (oracle/lib/liboracle-<type>.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker.  The product never imports it.

Complex data is carried as numpy arrays whose last axis is the {re, im} pair of the scalar
type (the kiss_fft_cpx layout, kiss_fft.h:87-90), so one shape convention serves the float,
Q15 (int16) and Q31 (int32) builds.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

SCALARS = {"float": np.float32, "int16_t": np.int16, "int32_t": np.int32}

_c_int_p = ctypes.POINTER(ctypes.c_int)


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


class Oracle:
    """One datatype build of the CPU restatement."""

    def __init__(self, datatype: str = "float"):
        if datatype not in SCALARS:
            raise ValueError(f"unknown datatype {datatype}")
        path = os.path.join(LIBDIR, f"liboracle-{datatype}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        self.datatype = datatype
        self.scalar = SCALARS[datatype]
        L = ctypes.CDLL(path)
        self._L = L
        vp, ci, cl = ctypes.c_void_p, ctypes.c_int, ctypes.c_long
        L.kfo_factor.argtypes = [ci, _c_int_p]
        L.kfo_factor.restype = ci
        L.kfo_twiddles.argtypes = [ci, ci, vp]
        L.kfo_super_twiddles.argtypes = [ci, ci, vp]
        L.kfo_next_fast_size.argtypes = [ci]
        L.kfo_next_fast_size.restype = ci
        L.kfo_fft.argtypes = [ci, ci, vp, ci, vp, cl, cl, cl]
        L.kfo_fftr.argtypes = [ci, vp, vp, cl, cl, cl]
        L.kfo_fftri.argtypes = [ci, vp, vp, cl, cl, cl]
        L.kfo_fftnd.argtypes = [_c_int_p, ci, ci, vp, vp, cl]
        L.kfo_fftndr.argtypes = [_c_int_p, ci, vp, vp, cl]
        L.kfo_fftndri.argtypes = [_c_int_p, ci, vp, vp, cl]

    # -- plan pieces ---------------------------------------------------------------
    def factor(self, n: int) -> list[tuple[int, int]]:
        buf = (ctypes.c_int * 64)()
        cnt = self._L.kfo_factor(n, buf)
        return [(buf[2 * k], buf[2 * k + 1]) for k in range(cnt)]

    def twiddles(self, n: int, inverse: bool) -> np.ndarray:
        out = np.zeros((n, 2), self.scalar)
        self._L.kfo_twiddles(n, int(inverse), _ptr(out))
        return out

    def super_twiddles(self, n: int, inverse: bool) -> np.ndarray:
        out = np.zeros((n // 4, 2), self.scalar)
        self._L.kfo_super_twiddles(n, int(inverse), _ptr(out))
        return out

    def next_fast_size(self, n: int) -> int:
        return self._L.kfo_next_fast_size(n)

    # -- transforms ----------------------------------------------------------------
    def _chk(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x)
        if x.dtype != self.scalar:
            raise TypeError(f"expected {self.scalar}, got {x.dtype}")
        return x

    def fft(self, x: np.ndarray, inverse: bool = False) -> np.ndarray:
        """x: (..., n, 2) -> (..., n, 2); batch over the leading axes."""
        x = self._chk(x)
        n = x.shape[-2]
        howmany = int(np.prod(x.shape[:-2], dtype=np.int64))
        out = np.empty_like(x)
        self._L.kfo_fft(n, int(inverse), _ptr(x), 1, _ptr(out), howmany, n, n)
        return out

    def fft_stride(self, x: np.ndarray, n: int, in_stride: int, inverse: bool = False) -> np.ndarray:
        """kiss_fft_stride on a flat (m, 2) buffer: reads x[k*in_stride], k < n."""
        x = self._chk(x)
        assert x.shape[0] >= (n - 1) * in_stride + 1
        out = np.empty((n, 2), self.scalar)
        self._L.kfo_fft(n, int(inverse), _ptr(x), in_stride, _ptr(out), 1, 0, 0)
        return out

    def fftr(self, x: np.ndarray) -> np.ndarray:
        """x: (..., n) real -> (..., n//2+1, 2)."""
        x = self._chk(x)
        n = x.shape[-1]
        howmany = int(np.prod(x.shape[:-1], dtype=np.int64))
        out = np.empty(x.shape[:-1] + (n // 2 + 1, 2), self.scalar)
        self._L.kfo_fftr(n, _ptr(x), _ptr(out), howmany, n, n // 2 + 1)
        return out

    def fftri(self, X: np.ndarray, n: int) -> np.ndarray:
        """X: (..., n//2+1, 2) -> (..., n) real (unnormalised)."""
        X = self._chk(X)
        assert X.shape[-2] == n // 2 + 1
        howmany = int(np.prod(X.shape[:-2], dtype=np.int64))
        out = np.empty(X.shape[:-2] + (n,), self.scalar)
        self._L.kfo_fftri(n, _ptr(X), _ptr(out), howmany, n // 2 + 1, n)
        return out

    def fftnd(self, x: np.ndarray, dims, inverse: bool = False) -> np.ndarray:
        """x: (..., *dims, 2) -> same shape."""
        x = self._chk(x)
        dims = [int(d) for d in dims]
        nd = len(dims)
        assert list(x.shape[-nd - 1:-1]) == dims
        howmany = int(np.prod(x.shape[:-nd - 1], dtype=np.int64))
        out = np.empty_like(x)
        d = (ctypes.c_int * nd)(*dims)
        self._L.kfo_fftnd(d, nd, int(inverse), _ptr(x), _ptr(out), howmany)
        return out

    def fftndr(self, x: np.ndarray, dims) -> np.ndarray:
        x = self._chk(x)
        dims = [int(d) for d in dims]
        nd = len(dims)
        assert list(x.shape[-nd:]) == dims
        howmany = int(np.prod(x.shape[:-nd], dtype=np.int64))
        out = np.empty(x.shape[:-1] + (dims[-1] // 2 + 1, 2), self.scalar)
        d = (ctypes.c_int * nd)(*dims)
        self._L.kfo_fftndr(d, nd, _ptr(x), _ptr(out), howmany)
        return out

    def fftndri(self, X: np.ndarray, dims) -> np.ndarray:
        X = self._chk(X)
        dims = [int(d) for d in dims]
        nd = len(dims)
        howmany = int(np.prod(X.shape[:-nd - 1], dtype=np.int64))
        out = np.empty(X.shape[:-2] + (dims[-1],), self.scalar)
        d = (ctypes.c_int * nd)(*dims)
        self._L.kfo_fftndri(d, nd, _ptr(X), _ptr(out), howmany)
        return out


def as_complex(x: np.ndarray) -> np.ndarray:
    """(..., 2) pair array -> complex128."""
    x = np.asarray(x, np.float64)
    return x[..., 0] + 1j * x[..., 1]


def from_complex(z: np.ndarray, scalar=np.float32) -> np.ndarray:
    z = np.asarray(z)
    return np.stack([z.real, z.imag], axis=-1).astype(scalar)

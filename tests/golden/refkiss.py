"""ctypes binding of the reference KISS FFT built from its own sources (oracle/_ref/).

Used only by make_golden.py, in the container that holds the reference tree, to write
the golden fixtures, and by the CPU tests that pin the oracle against it when
oracle/_ref/ is present.  The product never loads it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REFDIR = os.path.join(ROOT, "oracle", "_ref")
SCALARS = {"float": np.float32, "int16_t": np.int16, "int32_t": np.int32}

# struct kiss_fft_state layout, _kiss_fft_guts.h:27-32: int nfft, int inverse,
# int factors[64], then the twiddle table.
TW_OFFSET = 4 + 4 + 4 * 64

_libc = ctypes.CDLL(None)
_libc.free.argtypes = [ctypes.c_void_p]


def available(datatype: str = "int16_t") -> bool:
    return os.path.exists(os.path.join(REFDIR, f"libkissfft_ref-{datatype}.so"))


class RefKiss:
    def __init__(self, datatype: str):
        self.datatype = datatype
        self.scalar = SCALARS[datatype]
        L = ctypes.CDLL(os.path.join(REFDIR, f"libkissfft_ref-{datatype}.so"))
        vp, ci = ctypes.c_void_p, ctypes.c_int
        szp = ctypes.POINTER(ctypes.c_size_t)
        ip = ctypes.POINTER(ctypes.c_int)
        for name in ("kiss_fft_alloc", "kiss_fftr_alloc"):
            f = getattr(L, name)
            f.argtypes = [ci, ci, vp, szp]
            f.restype = vp
        for name in ("kiss_fftnd_alloc", "kiss_fftndr_alloc"):
            f = getattr(L, name)
            f.argtypes = [ip, ci, ci, vp, szp]
            f.restype = vp
        for name in ("kiss_fft", "kiss_fftr", "kiss_fftri", "kiss_fftnd", "kiss_fftndr", "kiss_fftndri"):
            getattr(L, name).argtypes = [vp, vp, vp]
        L.kiss_fft_stride.argtypes = [vp, vp, vp, ci]
        L.kiss_fft_next_fast_size.argtypes = [ci]
        L.kiss_fft_next_fast_size.restype = ci
        self.L = L

    def _p(self, a):
        return ctypes.c_void_p(a.ctypes.data)

    def factors(self, n: int):
        cfg = self.L.kiss_fft_alloc(n, 0, None, None)
        raw = (ctypes.c_int * 66).from_address(cfg)
        out, k = [], 0
        while True:
            p, m = raw[2 + 2 * k], raw[3 + 2 * k]
            out.append((p, m))
            k += 1
            if m == 1:
                break
        _libc.free(cfg)
        return out

    def twiddles(self, n: int, inverse: bool) -> np.ndarray:
        cfg = self.L.kiss_fft_alloc(n, int(inverse), None, None)
        nbytes = n * 2 * np.dtype(self.scalar).itemsize
        buf = (ctypes.c_char * nbytes).from_address(cfg + TW_OFFSET)
        tw = np.frombuffer(bytes(buf), self.scalar).reshape(n, 2).copy()
        _libc.free(cfg)
        return tw

    def super_twiddles(self, n: int, inverse: bool) -> np.ndarray:
        cfg = self.L.kiss_fftr_alloc(n, int(inverse), None, None)
        # struct kiss_fftr_state: substate, tmpbuf, super_twiddles (kiss_fftr.c:12-19)
        stw_ptr = ctypes.c_void_p.from_address(cfg + 2 * ctypes.sizeof(ctypes.c_void_p)).value
        cnt = (n // 2) // 2
        nbytes = cnt * 2 * np.dtype(self.scalar).itemsize
        buf = (ctypes.c_char * nbytes).from_address(stw_ptr)
        out = np.frombuffer(bytes(buf), self.scalar).reshape(cnt, 2).copy()
        _libc.free(cfg)
        return out

    def fft(self, x: np.ndarray, inverse: bool = False) -> np.ndarray:
        x = np.ascontiguousarray(x, self.scalar)
        n = x.shape[-2]
        flat = x.reshape(-1, n, 2)
        out = np.empty_like(flat)
        cfg = self.L.kiss_fft_alloc(n, int(inverse), None, None)
        for b in range(flat.shape[0]):
            self.L.kiss_fft(cfg, self._p(flat[b]), self._p(out[b]))
        _libc.free(cfg)
        return out.reshape(x.shape)

    def fft_stride(self, x: np.ndarray, n: int, in_stride: int, inverse: bool = False):
        x = np.ascontiguousarray(x, self.scalar)
        out = np.empty((n, 2), self.scalar)
        cfg = self.L.kiss_fft_alloc(n, int(inverse), None, None)
        self.L.kiss_fft_stride(cfg, self._p(x), self._p(out), in_stride)
        _libc.free(cfg)
        return out

    def fftr(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, self.scalar)
        n = x.shape[-1]
        flat = x.reshape(-1, n)
        out = np.empty((flat.shape[0], n // 2 + 1, 2), self.scalar)
        cfg = self.L.kiss_fftr_alloc(n, 0, None, None)
        for b in range(flat.shape[0]):
            self.L.kiss_fftr(cfg, self._p(flat[b]), self._p(out[b]))
        _libc.free(cfg)
        return out.reshape(x.shape[:-1] + (n // 2 + 1, 2))

    def fftri(self, X: np.ndarray, n: int) -> np.ndarray:
        X = np.ascontiguousarray(X, self.scalar)
        flat = X.reshape(-1, n // 2 + 1, 2)
        out = np.empty((flat.shape[0], n), self.scalar)
        cfg = self.L.kiss_fftr_alloc(n, 1, None, None)
        for b in range(flat.shape[0]):
            self.L.kiss_fftri(cfg, self._p(flat[b]), self._p(out[b]))
        _libc.free(cfg)
        return out.reshape(X.shape[:-2] + (n,))

    def fftnd(self, x: np.ndarray, dims, inverse: bool = False) -> np.ndarray:
        x = np.ascontiguousarray(x, self.scalar)
        dims = [int(d) for d in dims]
        tot = int(np.prod(dims))
        flat = x.reshape(-1, tot, 2)
        out = np.empty_like(flat)
        d = (ctypes.c_int * len(dims))(*dims)
        cfg = self.L.kiss_fftnd_alloc(d, len(dims), int(inverse), None, None)
        for b in range(flat.shape[0]):
            self.L.kiss_fftnd(cfg, self._p(flat[b]), self._p(out[b]))
        _libc.free(cfg)
        return out.reshape(x.shape)

    def fftndr_alloc_ok(self, dims, inverse=False) -> bool:
        d = (ctypes.c_int * len(dims))(*dims)
        cfg = self.L.kiss_fftndr_alloc(d, len(dims), int(inverse), None, None)
        if cfg:
            _libc.free(cfg)
        return bool(cfg)

    def fftndr(self, x: np.ndarray, dims) -> np.ndarray:
        dims = [int(d) for d in dims]
        x = np.ascontiguousarray(x, self.scalar)
        tot = int(np.prod(dims))
        nb = dims[-1] // 2 + 1
        outer = tot // dims[-1]
        flat = x.reshape(-1, tot)
        out = np.empty((flat.shape[0], outer * nb, 2), self.scalar)
        d = (ctypes.c_int * len(dims))(*dims)
        cfg = self.L.kiss_fftndr_alloc(d, len(dims), 0, None, None)
        if not cfg:
            raise RuntimeError(f"reference kiss_fftndr_alloc rejected {dims}")
        for b in range(flat.shape[0]):
            self.L.kiss_fftndr(cfg, self._p(flat[b]), self._p(out[b]))
        _libc.free(cfg)
        return out.reshape(x.shape[:-1] + (nb, 2))

    def fftndri(self, X: np.ndarray, dims) -> np.ndarray:
        dims = [int(d) for d in dims]
        X = np.ascontiguousarray(X, self.scalar)
        tot = int(np.prod(dims))
        nb = dims[-1] // 2 + 1
        outer = tot // dims[-1]
        flat = X.reshape(-1, outer * nb, 2)
        out = np.empty((flat.shape[0], tot), self.scalar)
        d = (ctypes.c_int * len(dims))(*dims)
        cfg = self.L.kiss_fftndr_alloc(d, len(dims), 1, None, None)
        if not cfg:
            raise RuntimeError(f"reference kiss_fftndr_alloc rejected {dims}")
        for b in range(flat.shape[0]):
            self.L.kiss_fftndri(cfg, self._p(flat[b]), self._p(out[b]))
        _libc.free(cfg)
        return out.reshape(X.shape[:-2] + (dims[-1],))

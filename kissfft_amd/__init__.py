"""kissfft_amd -- Python host binding of the MI355X KISS FFT engine.

The product is the C-ABI library libkissfft-<type>.so (kissfft_amd/lib/), a drop-in for
the reference's per-datatype libraries: kiss_fft_alloc / kiss_fft / kiss_fft_stride,
kiss_fftr_alloc / kiss_fftr / kiss_fftri, kiss_fftnd_alloc / kiss_fftnd,
kiss_fftndr_alloc / kiss_fftndr / kiss_fftndri, kfc_*, plus the batched calls of
include/kiss_fft_batch.h.  This module binds those symbols with ctypes, the way a caller
of the reference binds its C API; it holds no compute of its own.

Buffers may be numpy arrays (host memory: the library stages them through HBM) or torch
tensors on the GPU (device memory: used in place).  Complex data is passed as arrays
whose last axis is the {re, im} pair of the scalar type (kiss_fft_cpx layout,
kiss_fft.h:87-90); complex64 numpy/torch arrays have that layout too.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import build as _build

__all__ = ["Library", "Server", "MultiBatch", "FastFir", "load", "SCALARS", "KissFFTError"]

SCALARS = {"float": np.float32, "int16_t": np.int16, "int32_t": np.int32}


class KissFFTError(RuntimeError):
    pass


_libc = ctypes.CDLL(None)
_libc.free.argtypes = [ctypes.c_void_p]


def _addr(a) -> int:
    """Data address of a numpy array or a torch tensor."""
    if a is None:
        return 0
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("buffer must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return a.data_ptr()
    if isinstance(a, int):
        return a
    raise TypeError(f"unsupported buffer type {type(a)}")


class Cfg:
    """An owned cfg pointer (free()'d on collection, as kiss_fft_free does)."""

    def __init__(self, ptr: int, kind: str, **meta):
        if not ptr:
            raise KissFFTError(f"{kind} alloc returned NULL for {meta}")
        self.ptr = ptr
        self.kind = kind
        self.meta = meta

    @property
    def _as_parameter_(self):
        return ctypes.c_void_p(self.ptr)

    def __del__(self):
        # at interpreter exit module globals may already be gone
        if getattr(self, "ptr", 0) and _libc is not None and getattr(_libc, "free", None) is not None:
            _libc.free(self.ptr)
            self.ptr = 0


def _share_hip_runtime() -> None:
    """Load PyTorch (when installed) before the library.

    PyTorch ships its own libamdhip64 with the same soname as ROCm's.  Loaded first, it is
    the one copy of the HIP runtime in the process and the library binds to it, so torch
    tensors' device pointers are valid here.  Loaded the other way round, the process would
    hold two HIP runtimes, and the second one finds no device."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


class Library:
    """ctypes view of one libkissfft-<datatype>.so."""

    def __init__(self, datatype: str = "float"):
        if datatype not in SCALARS:
            raise ValueError(f"datatype must be one of {list(SCALARS)}")
        _share_hip_runtime()
        path = _build.lib_path(datatype)
        if not os.path.exists(path):
            raise KissFFTError(f"{path} is not built: run `python -m kissfft_amd.build` "
                               "(the engine has no CPU fallback)")
        self.datatype = datatype
        self.scalar = SCALARS[datatype]
        self.path = path
        L = ctypes.CDLL(path)
        self.L = L
        vp, ci, ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong
        szp = ctypes.POINTER(ctypes.c_size_t)
        ip = ctypes.POINTER(ctypes.c_int)
        sig = {
            "kiss_fft_alloc": ([ci, ci, vp, szp], vp),
            "kiss_fft": ([vp, vp, vp], None),
            "kiss_fft_stride": ([vp, vp, vp, ci], None),
            "kiss_fft_cleanup": ([], None),
            "kiss_fft_next_fast_size": ([ci], ci),
            "kiss_fftr_alloc": ([ci, ci, vp, szp], vp),
            "kiss_fftr": ([vp, vp, vp], None),
            "kiss_fftri": ([vp, vp, vp], None),
            "kiss_fftnd_alloc": ([ip, ci, ci, vp, szp], vp),
            "kiss_fftnd": ([vp, vp, vp], None),
            "kiss_fftndr_alloc": ([ip, ci, ci, vp, szp], vp),
            "kiss_fftndr": ([vp, vp, vp], None),
            "kiss_fftndri": ([vp, vp, vp], None),
            "kfc_fft": ([ci, vp, vp], None),
            "kfc_ifft": ([ci, vp, vp], None),
            "kfc_cleanup": ([], None),
            "kfc_cached_count": ([], ci),
            "kiss_fft_batch": ([vp, vp, ci, ll, vp, ll, ll], ci),
            "kiss_fftr_batch": ([vp, vp, ll, vp, ll, ll], ci),
            "kiss_fftri_batch": ([vp, vp, ll, vp, ll, ll], ci),
            "kiss_fftnd_batch": ([vp, vp, vp, ll], ci),
            "kiss_fftndr_batch": ([vp, vp, vp, ll], ci),
            "kiss_fftndri_batch": ([vp, vp, vp, ll], ci),
            "kiss_fft_batch_async": ([vp, vp, ci, ll, vp, ll, ll, vp], ci),
            "kiss_fftr_batch_async": ([vp, vp, ll, vp, ll, ll, vp], ci),
            "kiss_fftri_batch_async": ([vp, vp, ll, vp, ll, ll, vp], ci),
            "kiss_fftnd_batch_async": ([vp, vp, vp, ll, vp], ci),
            "kiss_fftndr_batch_async": ([vp, vp, vp, ll, vp], ci),
            "kiss_fftndri_batch_async": ([vp, vp, vp, ll, vp], ci),
            "kiss_fft_batch_noalloc": ([vp, vp, ci, ll, vp, ll, ll, vp], ci),
            "kiss_fft_stream_sync": ([vp], ci),
            "kiss_fft_last_error": ([], ctypes.c_char_p),
            "kiss_fft_device_available": ([], ci),
            "kiss_fft_get_engine_stats": ([vp], None),
            "kiss_fft_server_create": ([ll, ci, ci], vp),
            "kiss_fft_server_submit": ([vp, ci, ci, vp, vp, ll], ll),
            "kiss_fft_server_submit_nd": ([vp, ci, ci, ip, vp, vp, ll], ll),
            "kiss_fft_server_wait": ([vp, ll, ctypes.POINTER(ctypes.c_double)], ci),
            "kiss_fft_server_drain": ([vp], ci),
            "kiss_fft_server_get_stats": ([vp, ctypes.POINTER(ServerStats)], ci),
            "kiss_fft_server_reset_stats": ([vp], None),
            "kiss_fft_server_destroy": ([vp], None),
            "kiss_fft_multi_alloc": ([ctypes.POINTER(MultiItem), ci], vp),
            "kiss_fft_multi_exec": ([vp, vp], ci),
            "kiss_fft_multi_capture": ([vp, vp], ci),
            "kiss_fft_multi_replay": ([vp, vp], ci),
            "kiss_fft_multi_info": ([vp, ctypes.POINTER(ll), ctypes.POINTER(ci), ctypes.POINTER(ci)], ci),
            "kiss_fft_multi_layout": ([vp, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(ci),
                                       ctypes.POINTER(ci)], ci),
            "kiss_fft_multi_fused": ([vp, ctypes.POINTER(ci), ctypes.POINTER(ll), ctypes.POINTER(ci)], ci),
            "kiss_fft_multi_free": ([vp], None),
            "kiss_fastfir_alloc_cpx": ([vp, ctypes.c_size_t, szp, vp, szp], vp),
            "kiss_fastfir_alloc_real": ([vp, ctypes.c_size_t, szp, vp, szp], vp),
            "kiss_fastfir_cpx": ([vp, vp, vp, ctypes.c_size_t, szp], ctypes.c_size_t),
            "kiss_fastfir_real": ([vp, vp, vp, ctypes.c_size_t, szp], ctypes.c_size_t),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res

    # -- errors / status --------------------------------------------------------------
    def last_error(self) -> str:
        return (self.L.kiss_fft_last_error() or b"").decode()

    def check(self, rc: int, what: str) -> None:
        if rc != 0:
            raise KissFFTError(f"{what} failed ({rc}): {self.last_error()}")

    def device_available(self) -> bool:
        return bool(self.L.kiss_fft_device_available())

    def stats(self) -> dict:
        buf = (ctypes.c_longlong * 5)()
        self.L.kiss_fft_get_engine_stats(buf)
        return dict(zip(["plans_built", "plan_hits", "launches", "h2d_bytes", "d2h_bytes"], list(buf)))

    # -- cfgs -------------------------------------------------------------------------
    def fft_alloc(self, nfft: int, inverse: bool = False) -> Cfg:
        return Cfg(self.L.kiss_fft_alloc(nfft, int(inverse), None, None), "kiss_fft", nfft=nfft, inverse=inverse)

    def fftr_alloc(self, nfft: int, inverse: bool = False) -> Cfg:
        return Cfg(self.L.kiss_fftr_alloc(nfft, int(inverse), None, None), "kiss_fftr", nfft=nfft, inverse=inverse)

    def fftnd_alloc(self, dims, inverse: bool = False) -> Cfg:
        d = (ctypes.c_int * len(dims))(*dims)
        return Cfg(self.L.kiss_fftnd_alloc(d, len(dims), int(inverse), None, None), "kiss_fftnd",
                   dims=tuple(dims), inverse=inverse)

    def fftndr_alloc(self, dims, inverse: bool = False) -> Cfg:
        d = (ctypes.c_int * len(dims))(*dims)
        return Cfg(self.L.kiss_fftndr_alloc(d, len(dims), int(inverse), None, None), "kiss_fftndr",
                   dims=tuple(dims), inverse=inverse)

    def alloc_size(self, kind: str, arg, inverse: bool = False) -> int:
        """Bytes a cfg needs (the mem=NULL, lenmem!=NULL query)."""
        n = ctypes.c_size_t(0)
        if kind in ("kiss_fft", "kiss_fftr"):
            getattr(self.L, kind + "_alloc")(arg, int(inverse), None, ctypes.byref(n))
        else:
            d = (ctypes.c_int * len(arg))(*arg)
            getattr(self.L, kind + "_alloc")(d, len(arg), int(inverse), None, ctypes.byref(n))
        return n.value

    # -- reference-shaped single calls (synchronous) -------------------------------------
    def kiss_fft(self, cfg: Cfg, fin, fout) -> None:
        self.L.kiss_fft(cfg, _addr(fin), _addr(fout))

    def kiss_fft_stride(self, cfg: Cfg, fin, fout, in_stride: int) -> None:
        self.L.kiss_fft_stride(cfg, _addr(fin), _addr(fout), in_stride)

    def kiss_fftr(self, cfg: Cfg, timedata, freqdata) -> None:
        self.L.kiss_fftr(cfg, _addr(timedata), _addr(freqdata))

    def kiss_fftri(self, cfg: Cfg, freqdata, timedata) -> None:
        self.L.kiss_fftri(cfg, _addr(freqdata), _addr(timedata))

    def kiss_fftnd(self, cfg: Cfg, fin, fout) -> None:
        self.L.kiss_fftnd(cfg, _addr(fin), _addr(fout))

    def kiss_fftndr(self, cfg: Cfg, timedata, freqdata) -> None:
        self.L.kiss_fftndr(cfg, _addr(timedata), _addr(freqdata))

    def kiss_fftndri(self, cfg: Cfg, freqdata, timedata) -> None:
        self.L.kiss_fftndri(cfg, _addr(freqdata), _addr(timedata))

    # -- batched ----------------------------------------------------------------------
    def fft_batch(self, cfg: Cfg, fin, fout, howmany: int, idist: int, odist: int, in_stride: int = 1,
                  stream=None) -> None:
        if stream is None:
            rc = self.L.kiss_fft_batch(cfg, _addr(fin), in_stride, idist, _addr(fout), odist, howmany)
        else:
            rc = self.L.kiss_fft_batch_async(cfg, _addr(fin), in_stride, idist, _addr(fout), odist, howmany,
                                             _stream_ptr(stream))
        self.check(rc, "kiss_fft_batch")

    def fft_batch_noalloc(self, cfg: Cfg, fin, fout, howmany: int, idist: int, odist: int, in_stride: int = 1,
                          stream=None) -> None:
        """kiss_fft_batch_noalloc: device buffers, nothing allocated (the cfg must have run
        once through fft_batch on this device)."""
        rc = self.L.kiss_fft_batch_noalloc(cfg, _addr(fin), in_stride, idist, _addr(fout), odist, howmany,
                                           _stream_ptr(stream) if stream is not None else None)
        self.check(rc, "kiss_fft_batch_noalloc")

    def fftr_batch(self, cfg: Cfg, timedata, freqdata, howmany: int, idist: int, odist: int, stream=None):
        if stream is None:
            rc = self.L.kiss_fftr_batch(cfg, _addr(timedata), idist, _addr(freqdata), odist, howmany)
        else:
            rc = self.L.kiss_fftr_batch_async(cfg, _addr(timedata), idist, _addr(freqdata), odist, howmany,
                                              _stream_ptr(stream))
        self.check(rc, "kiss_fftr_batch")

    def fftri_batch(self, cfg: Cfg, freqdata, timedata, howmany: int, idist: int, odist: int, stream=None):
        if stream is None:
            rc = self.L.kiss_fftri_batch(cfg, _addr(freqdata), idist, _addr(timedata), odist, howmany)
        else:
            rc = self.L.kiss_fftri_batch_async(cfg, _addr(freqdata), idist, _addr(timedata), odist, howmany,
                                               _stream_ptr(stream))
        self.check(rc, "kiss_fftri_batch")

    def fftnd_batch(self, cfg: Cfg, fin, fout, howmany: int, stream=None):
        if stream is None:
            rc = self.L.kiss_fftnd_batch(cfg, _addr(fin), _addr(fout), howmany)
        else:
            rc = self.L.kiss_fftnd_batch_async(cfg, _addr(fin), _addr(fout), howmany, _stream_ptr(stream))
        self.check(rc, "kiss_fftnd_batch")

    def fftndr_batch(self, cfg: Cfg, timedata, freqdata, howmany: int, stream=None):
        if stream is None:
            rc = self.L.kiss_fftndr_batch(cfg, _addr(timedata), _addr(freqdata), howmany)
        else:
            rc = self.L.kiss_fftndr_batch_async(cfg, _addr(timedata), _addr(freqdata), howmany,
                                                _stream_ptr(stream))
        self.check(rc, "kiss_fftndr_batch")

    def fftndri_batch(self, cfg: Cfg, freqdata, timedata, howmany: int, stream=None):
        if stream is None:
            rc = self.L.kiss_fftndri_batch(cfg, _addr(freqdata), _addr(timedata), howmany)
        else:
            rc = self.L.kiss_fftndri_batch_async(cfg, _addr(freqdata), _addr(timedata), howmany,
                                                 _stream_ptr(stream))
        self.check(rc, "kiss_fftndri_batch")

    def stream_sync(self, stream=None) -> None:
        self.check(self.L.kiss_fft_stream_sync(_stream_ptr(stream) if stream is not None else None),
                   "kiss_fft_stream_sync")

    def server(self, hbm_budget_bytes: int = 0, streams: int = 0, max_group: int = 0) -> "Server":
        return Server(self, hbm_budget_bytes, streams, max_group)

    def fastfir_alloc(self, imp_resp, real: bool = False, nfft: int = 0) -> "FastFir":
        return FastFir(self, imp_resp, real, nfft)

    def multi(self, items) -> "MultiBatch":
        """items: (kind, nfft, fin, fout, howmany) tuples, device buffers."""
        return MultiBatch(self, items)


class FastFir:
    """kiss_fastfir (include/kiss_fastfir.h): overlap-scrap FIR filter of a stream.

    filter(inbuf, outbuf, n_new) follows the reference's kiss_fastfir: the first
    offset + n_new samples of inbuf are filtered, the good output samples are written to
    outbuf (their count is returned), and the unprocessed input is moved to the front of
    inbuf (self.offset samples).  n_new == 0 flushes."""

    def __init__(self, lib: Library, imp_resp, real: bool, nfft: int):
        self.lib = lib
        self.real = real
        h = np.ascontiguousarray(imp_resp, dtype=lib.scalar)
        n = ctypes.c_size_t(nfft)
        nh = h.shape[0]
        fn = lib.L.kiss_fastfir_alloc_real if real else lib.L.kiss_fastfir_alloc_cpx
        self.ptr = fn(h.ctypes.data, nh, ctypes.byref(n), None, None)
        if not self.ptr:
            raise KissFFTError(f"kiss_fastfir_alloc failed: {lib.last_error()}")
        self.nfft = n.value
        self.ngood = self.nfft - nh + 1
        self.offset = 0

    def filter(self, inbuf, outbuf, n_new: int) -> int:
        off = ctypes.c_size_t(self.offset)
        fn = self.lib.L.kiss_fastfir_real if self.real else self.lib.L.kiss_fastfir_cpx
        w = fn(self.ptr, _addr(inbuf), _addr(outbuf), n_new, ctypes.byref(off))
        self.offset = off.value
        return w

    def __del__(self):
        if getattr(self, "ptr", None) and _libc is not None and getattr(_libc, "free", None) is not None:
            _libc.free(self.ptr)
            self.ptr = None


class MultiItem(ctypes.Structure):
    """kiss_fft_multi_item (include/kiss_fft_multi.h)."""
    _fields_ = [("kind", ctypes.c_int), ("nfft", ctypes.c_int), ("fin", ctypes.c_void_p),
                ("fout", ctypes.c_void_p), ("howmany", ctypes.c_longlong)]


class MultiBatch:
    """A kiss_fft_multi: a fixed set of batches of mixed lengths and kinds, run with one call.

    Kinds as for the server (C2C_FWD, C2C_INV, R2C, C2R).  The buffers must stay alive as
    long as the multi-batch is used."""
    C2C_FWD, C2C_INV, R2C, C2R = 0, 1, 2, 3

    def __init__(self, lib: Library, items):
        self.lib = lib
        arr = (MultiItem * len(items))()
        for i, (kind, nfft, fin, fout, howmany) in enumerate(items):
            arr[i] = MultiItem(kind, nfft, _addr(fin), _addr(fout), howmany)
        self._keep = list(items)
        self.ptr = lib.L.kiss_fft_multi_alloc(arr, len(items))
        if not self.ptr:
            raise KissFFTError(f"kiss_fft_multi_alloc failed: {lib.last_error()}")

    def exec(self, stream=None) -> None:
        self.lib.check(self.lib.L.kiss_fft_multi_exec(self.ptr, _stream_ptr(stream)), "kiss_fft_multi_exec")

    def capture(self, stream) -> None:
        self.lib.check(self.lib.L.kiss_fft_multi_capture(self.ptr, _stream_ptr(stream)), "kiss_fft_multi_capture")

    def replay(self, stream=None) -> None:
        self.lib.check(self.lib.L.kiss_fft_multi_replay(self.ptr, _stream_ptr(stream)), "kiss_fft_multi_replay")

    def info(self) -> dict:
        t, g, p = ctypes.c_longlong(), ctypes.c_int(), ctypes.c_int()
        self.lib.L.kiss_fft_multi_info(self.ptr, ctypes.byref(t), ctypes.byref(g), ctypes.byref(p))
        gr, gi, bl, mp = (ctypes.c_int() for _ in range(4))
        self.lib.L.kiss_fft_multi_layout(self.ptr, ctypes.byref(gr), ctypes.byref(gi), ctypes.byref(bl),
                                         ctypes.byref(mp))
        fm, fw, ft = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_int()
        self.lib.L.kiss_fft_multi_fused(self.ptr, ctypes.byref(fm), ctypes.byref(fw), ctypes.byref(ft))
        return {"tiles": t.value, "grid": g.value, "persistent_items": p.value, "groups": gr.value,
                "group_items": gi.value, "bluestein_items": bl.value, "multipass_items": mp.value,
                "fused_groups": fm.value, "fused_workgroups": fw.value, "fused_threads": ft.value}

    def close(self) -> None:
        if getattr(self, "ptr", None):
            self.lib.L.kiss_fft_multi_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServerStats(ctypes.Structure):
    """kiss_fft_server_stats (include/kiss_fft_serve.h)."""
    _fields_ = [(n, ctypes.c_longlong) for n in ("submitted", "completed", "failed", "launches",
                                                 "grouped_requests", "bytes", "budget_bytes",
                                                 "in_use_bytes")] + \
               [(n, ctypes.c_double) for n in ("p50_s", "p99_s", "max_s", "mean_s", "span_s")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class Server:
    """A kiss_fft_server: requests of mixed kind and length, batched on the device.

    Kinds are C2C_FWD, C2C_INV, R2C and C2R.  Buffers are numpy arrays (staged through the
    server's HBM budget) or device tensors (used in place); they must stay alive until the
    request's wait() returns."""
    C2C_FWD, C2C_INV, R2C, C2R = 0, 1, 2, 3

    def __init__(self, lib: Library, hbm_budget_bytes: int = 0, streams: int = 0, max_group: int = 0):
        self.lib = lib
        self.ptr = lib.L.kiss_fft_server_create(int(hbm_budget_bytes), int(streams), int(max_group))
        if not self.ptr:
            raise KissFFTError(f"kiss_fft_server_create failed: {lib.last_error()}")

    def submit(self, kind: int, nfft: int, fin, fout, howmany: int) -> int:
        t = self.lib.L.kiss_fft_server_submit(self.ptr, kind, nfft, _addr(fin), _addr(fout), howmany)
        if t <= 0:
            raise KissFFTError(f"kiss_fft_server_submit failed ({t}): {self.lib.last_error()}")
        return t

    def submit_nd(self, kind: int, dims, fin, fout, howmany: int) -> int:
        d = (ctypes.c_int * len(dims))(*dims)
        t = self.lib.L.kiss_fft_server_submit_nd(self.ptr, kind, len(dims), d, _addr(fin), _addr(fout), howmany)
        if t <= 0:
            raise KissFFTError(f"kiss_fft_server_submit_nd failed ({t}): {self.lib.last_error()}")
        return t

    def wait(self, ticket: int) -> float:
        """Latency in seconds (submit to completion); raises if the request failed."""
        lat = ctypes.c_double()
        rc = self.lib.L.kiss_fft_server_wait(self.ptr, ticket, ctypes.byref(lat))
        if rc != 0:
            raise KissFFTError(f"request {ticket} failed ({rc}): {self.lib.last_error()}")
        return lat.value

    def drain(self) -> None:
        self.lib.check(self.lib.L.kiss_fft_server_drain(self.ptr), "kiss_fft_server_drain")

    def stats(self) -> dict:
        st = ServerStats()
        self.lib.L.kiss_fft_server_get_stats(self.ptr, ctypes.byref(st))
        return st.as_dict()

    def reset_stats(self) -> None:
        self.lib.L.kiss_fft_server_reset_stats(self.ptr)

    def close(self) -> None:
        if getattr(self, "ptr", None):
            self.lib.L.kiss_fft_server_destroy(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _stream_ptr(stream):
    """hipStream_t of a torch.cuda.Stream (or a raw integer handle)."""
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


_LIBS: dict[str, Library] = {}


def load(datatype: str = "float") -> Library:
    if datatype not in _LIBS:
        _LIBS[datatype] = Library(datatype)
    return _LIBS[datatype]

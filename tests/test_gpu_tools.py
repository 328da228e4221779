"""GPU tests of the stream tool fftutil-<type> (the reference's tools/fftutil.c protocol).

test_testkiss_protocol restates the reference harness test/testkiss.py:24-101: random
1-, 2- and 3-D arrays with dims in [2, 4] (even last dimension for real transforms),
uniform (-1, 1), scaled by 32767 (int16) or 2^31-1 (int32), piped through the tool, and
the SNR against numpy's fftn / rfftn must reach 90 dB (10 dB for int16).
"""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT
from oracle import Oracle

pytestmark = pytest.mark.gpu

LIB = os.path.join(ROOT, "kissfft_amd", "lib")
DT = {"float": np.float32, "int16_t": np.int16, "int32_t": np.int32}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_tool(datatype, args, data: bytes, exe="fastfilt") -> bytes:
    cmd = [os.path.join(LIB, f"{exe}-{datatype}")] + args
    r = subprocess.run(cmd, input=data, capture_output=True, timeout=120)
    assert r.returncode == 0, (cmd, r.stderr.decode())
    return r.stdout


@pytest.mark.parametrize("datatype", ["float", "int16_t", "int32_t"])
@pytest.mark.parametrize("real", [False, True])
def test_testkiss_protocol(datatype, real):
    rng = np.random.default_rng(2024 + real)
    minsnr = 10 if datatype == "int16_t" else 90
    for ndim in range(1, 4):
        dims = rng.integers(2, 5, ndim)
        if real:
            dims[-1] = (dims[-1] // 2) * 2
        shape = tuple(int(d) for d in dims)
        x = rng.uniform(-1, 1, shape) if real else rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
        xver = np.fft.rfftn(x) if real else np.fft.fftn(x)
        flat = x.ravel()
        scale = 1.0
        if datatype == "int16_t":
            flat, scale = 32767 * flat, flat.size / 32767.0
        elif datatype == "int32_t":
            flat, scale = 2147483647.0 * flat, flat.size / 2147483647.0
        packed = (flat.astype(np.complex128).view(np.float64) if not real else flat.astype(np.float64))
        args = ["-n", ",".join(str(d) for d in shape)] + (["-R"] if real else [])
        out = np.frombuffer(run_tool(datatype, args, packed.astype(DT[datatype]).tobytes()), DT[datatype])
        out = out.astype(np.float64)
        res = scale * (out[::2] + 1j * out[1::2])
        err = (xver.ravel() - res)
        snr = 10 * math.log10(abs((np.vdot(xver.ravel(), xver.ravel()) + 1e-10) / (np.vdot(err, err) + 1e-10)))
        say(datatype, "real" if real else "complex", shape, f"SNR {snr:.1f} dB")
        assert snr >= minsnr, (datatype, real, shape, snr)


def test_stream_of_records_bit_exact_q15():
    # many records through several batches (-b 64): the same bits as one transform per record
    O = Oracle("int16_t")
    rng = np.random.default_rng(77)
    n, recs = 512, 1000
    x = rng.integers(-32768, 32768, (recs, n, 2)).astype(np.int16)
    raw = x.tobytes() + b"\x01\x02\x03"  # a trailing partial record is dropped
    y = np.frombuffer(run_tool("int16_t", ["-n", str(n), "-b", "64"], raw, exe="fftutil"), np.int16)
    np.testing.assert_array_equal(y.reshape(recs, n, 2), O.fft(x))
    yi = np.frombuffer(run_tool("int16_t", ["-n", str(n), "-i"], x.tobytes()), np.int16)
    np.testing.assert_array_equal(yi.reshape(recs, n, 2), O.fft(x, inverse=True))
    xr = rng.integers(-32768, 32768, (300, 1000)).astype(np.int16)
    yr = np.frombuffer(run_tool("int16_t", ["-n", "1000", "-R", "-b", "7"], xr.tobytes()), np.int16)
    np.testing.assert_array_equal(yr.reshape(300, 501, 2), O.fftr(xr))
    X = rng.integers(-32768, 32768, (50, 12, 10, 2)).astype(np.int16)
    yn = np.frombuffer(run_tool("int16_t", ["-n", "12,10", "-b", "16"], X.tobytes()), np.int16)
    np.testing.assert_array_equal(yn.reshape(50, 12, 10, 2), O.fftnd(X, (12, 10)))


def test_files_and_inverse_real_nd(tmp_path):
    rng = np.random.default_rng(78)
    dims = (16, 24)
    x = rng.standard_normal((5,) + dims).astype(np.float32)
    fin, fmid, fout = tmp_path / "x.raw", tmp_path / "X.raw", tmp_path / "y.raw"
    x.tofile(fin)
    run_tool("float", ["-n", "16,24", "-R", str(fin), str(fmid)], b"")
    X = np.fromfile(fmid, np.float32).reshape(5, 16, 13, 2)
    want = np.fft.rfftn(x.astype(np.float64), axes=(1, 2))
    assert np.linalg.norm((X[..., 0] + 1j * X[..., 1]) - want) / np.linalg.norm(want) <= 1e-6
    run_tool("float", ["-n", "16,24", "-R", "-i", str(fmid), str(fout)], b"")
    y = np.fromfile(fout, np.float32).reshape(x.shape)
    assert np.linalg.norm(y - 16 * 24 * x) / np.linalg.norm(16 * 24 * x) <= 2e-6

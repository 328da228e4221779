"""GPU tests of the multi-batch (include/kiss_fft_multi.h): mixed lengths and kinds in one
call -- compiled-kernel groups, Bluestein items, the persistent general launch, multi-pass
items -- and its hipGraph capture / replay.  Every test runs twice: with the default
layout and with every single-pass item in the persistent launch
(KISSFFT_AMD_MULTI_PERSISTENT=1)."""
import sys

import numpy as np
import pytest

import kissfft_amd
from helpers import ABS_FLOAT_BOUND, cplx, rel_l2
from kissfft_amd import workloads as W
from oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["grouped", "persistent"])
def layout(request, monkeypatch):
    monkeypatch.setenv("KISSFFT_AMD_MULTI_PERSISTENT", "1" if request.param == "persistent" else "0")
    return request.param

C2C, C2C_INV, R2C, C2R = 0, 1, 2, 3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(torch, dev, kind, n, b, seed, dtype):
    """Input / output tensors of one item (complex as trailing pairs)."""
    if dtype == "float":
        shape_in = {C2C: (b, n, 2), C2C_INV: (b, n, 2), R2C: (b, n), C2R: (b, n // 2 + 1, 2)}[kind]
        x = torch.from_numpy(W.uniform_pm1(seed, int(np.prod(shape_in))).reshape(shape_in)).to(dev)
        tdt = torch.float32
    elif dtype == "int32_t":
        shape_in = {C2C: (b, n, 2), C2C_INV: (b, n, 2), R2C: (b, n), C2R: (b, n // 2 + 1, 2)}[kind]
        q = np.random.default_rng(seed).integers(-(1 << 31), 1 << 31, int(np.prod(shape_in)), dtype=np.int64)
        x = torch.from_numpy(q.astype(np.int32).reshape(shape_in)).to(dev)
        tdt = torch.int32
    else:
        shape_in = {C2C: (b, n, 2), C2C_INV: (b, n, 2), R2C: (b, n), C2R: (b, n // 2 + 1, 2)}[kind]
        x = torch.from_numpy(W.q15_uniform(seed, int(np.prod(shape_in))).reshape(shape_in)).to(dev)
        tdt = torch.int16
    shape_out = {C2C: (b, n, 2), C2C_INV: (b, n, 2), R2C: (b, n // 2 + 1, 2), C2R: (b, n)}[kind]
    return x, torch.empty(shape_out, dtype=tdt, device=dev)


def fp64_ref(kind, n, x):
    if kind == C2C:
        return np.fft.fft(cplx(x), axis=-1)
    if kind == C2C_INV:
        return np.fft.ifft(cplx(x), axis=-1) * n
    if kind == R2C:
        return np.fft.rfft(x.astype(np.float64), axis=-1)
    return np.fft.irfft(cplx(x), n=n, axis=-1) * n


FLOAT_ITEMS = [(C2C, 256, 100), (C2C, 1009, 7), (C2C_INV, 1200, 33), (R2C, 4096, 10), (C2R, 4000, 9),
               (C2C, 60, 500), (C2C, 1001, 3), (R2C, 6, 4), (C2C, 1, 3), (C2R, 2, 5), (C2C, 8192, 2),
               (C2C, 65536, 2), (R2C, 40000, 2)]


def test_float_mix(layout):
    import torch
    lib = kissfft_amd.load("float")
    dev = torch.device("cuda:0")
    bufs = [make(torch, dev, k, n, b, 200 + i, "float") for i, (k, n, b) in enumerate(FLOAT_ITEMS)]
    m = lib.multi([(k, n, x, y, b) for (k, n, b), (x, y) in zip(FLOAT_ITEMS, bufs)])
    info = m.info()
    say("multi info", layout, info)
    assert info["multipass_items"] == 2  # 65536 and 40000
    assert info["persistent_items"] + info["group_items"] + info["bluestein_items"] == len(FLOAT_ITEMS) - 2
    if layout == "persistent":
        assert info["persistent_items"] == len(FLOAT_ITEMS) - 2 and info["groups"] == 0
        assert info["grid"] > 0 and info["tiles"] > 0
    else:
        # compiled kernels for 256, 1200 inverse, r2c 4096, c2r 4000 and (compiled at run
        # time) 8192, 1001 = 7 * 11 * 13 and 60; Bluestein for 1009; r2c 6, 1 and c2r 2 (one
        # stage each) in the persistent launch
        assert info["group_items"] == 7 and info["groups"] == 7 and info["bluestein_items"] == 1
        assert info["persistent_items"] == 3
    m.exec()
    for (k, n, b), (x, y) in zip(FLOAT_ITEMS, bufs):
        say("item", k, n, b)
        got = y.cpu().numpy()
        got = got if k == C2R else cplx(got)
        if k == C2C and n == 1:
            np.testing.assert_array_equal(y.cpu().numpy(), x.cpu().numpy())
            continue
        assert rel_l2(got, fp64_ref(k, n, x.cpu().numpy())) <= ABS_FLOAT_BOUND, (k, n)
    # long items take their temporaries from the multi-batch's workspace: the batch
    # captures, and a replay gives the eager results
    eager = [y.clone() for _, y in bufs]
    st = torch.cuda.Stream()
    m.capture(st)
    with torch.cuda.stream(st):
        for _, y in bufs:
            y.fill_(7.0)
    m.replay(st)
    st.synchronize()
    for (k, n, b), (_, y), e in zip(FLOAT_ITEMS, bufs, eager):
        assert torch.equal(y, e), (k, n, b)
    m.close()


Q15_ITEMS = [(C2C, 512, 64), (C2C, 1200, 5), (R2C, 4096, 3), (C2R, 500, 4), (C2C, 1001, 2), (C2C_INV, 60, 9),
             (C2C, 12, 17), (C2C, 65536, 1)]


def test_q15_mix_bit_exact(layout):
    import torch
    lib = kissfft_amd.load("int16_t")
    O = Oracle("int16_t")
    dev = torch.device("cuda:0")
    bufs = [make(torch, dev, k, n, b, 300 + i, "int16_t") for i, (k, n, b) in enumerate(Q15_ITEMS)]
    m = lib.multi([(k, n, x, y, b) for (k, n, b), (x, y) in zip(Q15_ITEMS, bufs)])
    m.exec()
    for (k, n, b), (x, y) in zip(Q15_ITEMS, bufs):
        say("q15 item", k, n, b)
        xs = x.cpu().numpy()
        want = {C2C: lambda: O.fft(xs), C2C_INV: lambda: O.fft(xs, inverse=True), R2C: lambda: O.fftr(xs),
                C2R: lambda: O.fftri(xs, n)}[k]()
        np.testing.assert_array_equal(y.cpu().numpy(), want, err_msg=f"kind {k} n {n}")
    m.close()


def test_graph_capture_and_replays(layout):
    # replays interleaved with eager work on the same stream; every replay is checked
    # against an eager execution of the same multi-batch on the same inputs
    import torch
    lib = kissfft_amd.load("float")
    dev = torch.device("cuda:0")
    items = [(C2C, 256, 40), (C2C, 1009, 3), (R2C, 4096, 6), (C2R, 2000, 5), (C2C_INV, 3072, 4), (C2C, 60, 90)]
    bufs = [make(torch, dev, k, n, b, 400 + i, "float") for i, (k, n, b) in enumerate(items)]
    m = lib.multi([(k, n, x, y, b) for (k, n, b), (x, y) in zip(items, bufs)])
    st = torch.cuda.Stream()
    m.capture(st)
    for rep in range(3):
        with torch.cuda.stream(st):
            for x, _ in bufs:
                x.mul_(1.5).add_(0.25)  # eager work between replays
            for _, y in bufs:
                y.fill_(7.0)
        m.replay(st)
        st.synchronize()
        replayed = [y.clone() for _, y in bufs]
        for _, y in bufs:
            y.fill_(-3.0)
        torch.cuda.synchronize()
        m.exec()  # eager, the calling thread's stream, synchronised
        for (k, n, b), (x, y), r in zip(items, bufs, replayed):
            say("replay", rep, "item", k, n)
            assert torch.equal(r, y), (rep, k, n)
            got = y.cpu().numpy()
            got = got if k == C2R else cplx(got)
            assert rel_l2(got, fp64_ref(k, n, x.cpu().numpy())) <= ABS_FLOAT_BOUND, (rep, k, n)
    m.close()


def test_many_small_batches_grouped():
    """64 batches of 8 (kind, length) keys, counts 1..40: one launch per key, same results
    as the persistent layout (float tolerance) and as the per-batch calls."""
    import torch
    lib = kissfft_amd.load("float")
    dev = torch.device("cuda:0")
    menu = [(C2C, 256), (C2C, 1000), (C2C, 1009), (C2C_INV, 1200), (R2C, 4096), (C2R, 4000), (C2C, 120),
            (C2C_INV, 3072)]
    rng = np.random.default_rng(12)
    items = [(menu[i % 8][0], menu[i % 8][1], int(rng.integers(1, 41))) for i in range(64)]
    bufs = [make(torch, dev, k, n, b, 500 + i, "float") for i, (k, n, b) in enumerate(items)]
    m = lib.multi([(k, n, x, y, b) for (k, n, b), (x, y) in zip(items, bufs)])
    info = m.info()
    say("many small", info)
    assert info["groups"] == 7 and info["group_items"] == 56 and info["bluestein_items"] == 8
    st = torch.cuda.Stream()
    m.capture(st)
    m.replay(st)
    st.synchronize()
    for (k, n, b), (x, y) in zip(items, bufs):
        got = y.cpu().numpy()
        got = got if k == C2R else cplx(got)
        assert rel_l2(got, fp64_ref(k, n, x.cpu().numpy())) <= ABS_FLOAT_BOUND, (k, n, b)
    m.close()


def test_fused_launch_same_results(monkeypatch):
    """The compiled-kernel and Bluestein groups run as one fused launch (a kernel compiled
    at run time from their own bodies, row kernels of other workgroup sizes reshaped to the
    launch's).  Without reshaping the results are bit for bit those of separate launches
    (KISSFFT_AMD_MULTI_FUSED=0); with it, Q15 and Q31 stay bit for bit (and equal to the oracle)
    and float stays within the float bar of both the separate launches and FP64 (a
    reshaped kernel may round a last bit differently); every graph replay equals the eager
    result."""
    import torch
    dev = torch.device("cuda:0")
    menu = [(C2C, 256), (C2C, 1000), (C2C, 1009), (C2C, 1200), (C2C_INV, 3072), (R2C, 4096), (C2R, 4000),
            (C2C, 120)]
    rng = np.random.default_rng(14)
    for dtype in ("float", "int16_t", "int32_t"):
        lib = kissfft_amd.load(dtype)
        items = [(menu[i % 8][0], menu[i % 8][1], int(rng.integers(1, 70))) for i in range(40)]
        if dtype != "float":  # fixed point: no Bluestein; 1009 takes the persistent launch
            items = [it for it in items if it[1] != 1009]
        bufs = [make(torch, dev, k, n, b, 800 + i, dtype) for i, (k, n, b) in enumerate(items)]
        spec = [(k, n, x, y, b) for (k, n, b), (x, y) in zip(items, bufs)]
        monkeypatch.setenv("KISSFFT_AMD_MULTI_FUSED", "0")
        sep = lib.multi(spec)
        assert sep.info()["fused_groups"] == 0
        sep.exec()
        want = [y.clone() for _, y in bufs]
        sep.close()
        monkeypatch.delenv("KISSFFT_AMD_MULTI_FUSED")
        monkeypatch.setenv("KISSFFT_AMD_MULTI_RESHAPE", "0")
        plain = lib.multi(spec)
        say("fused without reshaping", dtype, plain.info())
        for _, y in bufs:
            y.fill_(5)
        torch.cuda.synchronize()
        plain.exec()
        for (k, n, b), (_, y), w in zip(items, bufs, want):
            assert torch.equal(y, w), ("no reshape", dtype, k, n, b)
        plain.close()
        monkeypatch.delenv("KISSFFT_AMD_MULTI_RESHAPE")
        fz = lib.multi(spec)
        info = fz.info()
        say("fused", dtype, info)
        if dtype == "float":
            assert info["fused_groups"] >= 2 and info["fused_workgroups"] > 0 and info["fused_threads"] > 0
        for _, y in bufs:
            y.fill_(7)
        torch.cuda.synchronize()
        fz.exec()
        eager = [y.clone() for _, y in bufs]
        O = Oracle(dtype) if dtype != "float" else None
        for (k, n, b), (x, y), w in zip(items, bufs, want):
            say("fused item", dtype, k, n, b)
            if O is not None:
                assert torch.equal(y, w), (k, n, b)
                xs = x.cpu().numpy()
                ref = {C2C: lambda: O.fft(xs), C2C_INV: lambda: O.fft(xs, inverse=True), R2C: lambda: O.fftr(xs),
                       C2R: lambda: O.fftri(xs, n)}[k]()
                np.testing.assert_array_equal(y.cpu().numpy(), ref, err_msg=f"q15 kind {k} n {n}")
            else:
                got = y.cpu().numpy()
                got = got if k == C2R else cplx(got)
                assert rel_l2(got, fp64_ref(k, n, x.cpu().numpy())) <= ABS_FLOAT_BOUND, (k, n, b)
                w = w.cpu().numpy()
                assert rel_l2(got, w if k == C2R else cplx(w)) <= ABS_FLOAT_BOUND, (k, n, b)
        st = torch.cuda.Stream()
        fz.capture(st)
        for rep in range(3):
            with torch.cuda.stream(st):
                for _, y in bufs:
                    y.fill_(-3)
            fz.replay(st)
            st.synchronize()
            for (k, n, b), (_, y), e in zip(items, bufs, eager):
                assert torch.equal(y, e), (dtype, rep, k, n, b)
            torch.cuda.synchronize()
            fz.exec()  # eager work between the replays
            torch.cuda.synchronize()
        fz.close()


def test_helper_streams_same_results(monkeypatch):
    """KISSFFT_AMD_MULTI_HELPERS=n forks the launches over n streams of the multi-batch and
    joins them on the caller's stream: eager execution and graph replays give the same bits
    as the default single-stream form."""
    import torch
    lib = kissfft_amd.load("float")
    dev = torch.device("cuda:0")
    items = [(C2C, 256, 40), (C2C, 1009, 3), (R2C, 4096, 6), (C2R, 2000, 5), (C2C_INV, 3072, 4), (C2C, 60, 90)]
    bufs = [make(torch, dev, k, n, b, 500 + i, "float") for i, (k, n, b) in enumerate(items)]
    spec = [(k, n, x, y, b) for (k, n, b), (x, y) in zip(items, bufs)]
    monkeypatch.setenv("KISSFFT_AMD_MULTI_HELPERS", "0")
    single = lib.multi(spec)
    single.exec()
    want = [y.clone() for _, y in bufs]
    single.close()
    monkeypatch.setenv("KISSFFT_AMD_MULTI_HELPERS", "2")
    forked = lib.multi(spec)
    for _, y in bufs:
        y.fill_(-3.0)
    torch.cuda.synchronize()
    forked.exec()
    for (k, n, b), (_, y), w in zip(items, bufs, want):
        say("helpers eager", k, n)
        assert torch.equal(y, w), (k, n)
    st = torch.cuda.Stream()
    forked.capture(st)
    for rep in range(2):
        with torch.cuda.stream(st):
            for _, y in bufs:
                y.fill_(7.0)
        forked.replay(st)
        st.synchronize()
        for (k, n, b), (_, y), w in zip(items, bufs, want):
            say("helpers replay", rep, k, n)
            assert torch.equal(y, w), (rep, k, n)
    forked.close()


def test_graph_capture_with_multipass_items():
    """Items longer than one LDS line (float 65536 and 2^20 complex, 2^20 real both ways,
    Q15 65536) take their temporaries from the multi-batch's own workspace, so the batch
    captures; every replay equals an eager execution, and the results match FP64 / the
    oracle."""
    import torch
    dev = torch.device("cuda:0")
    for dtype, items in (("float", [(C2C, 65536, 3), (C2C_INV, 1048576, 2), (R2C, 1048576, 2),
                                    (C2R, 1048576, 2), (C2C, 256, 30), (R2C, 40000, 2)]),
                         ("int16_t", [(C2C, 65536, 3), (C2C_INV, 32768, 2), (C2C, 512, 20)])):
        lib = kissfft_amd.load(dtype)
        bufs = [make(torch, dev, k, n, b, 700 + i, dtype) for i, (k, n, b) in enumerate(items)]
        m = lib.multi([(k, n, x, y, b) for (k, n, b), (x, y) in zip(items, bufs)])
        info = m.info()
        say("multipass capture", dtype, info)
        assert info["multipass_items"] >= 2
        m.exec()
        torch.cuda.synchronize()
        eager = [y.clone() for _, y in bufs]
        st = torch.cuda.Stream()
        m.capture(st)
        for rep in range(3):
            with torch.cuda.stream(st):
                for _, y in bufs:
                    y.fill_(7)
            m.replay(st)
            st.synchronize()
            for (k, n, b), (_, y), e in zip(items, bufs, eager):
                say("replay", dtype, rep, "item", k, n, b)
                assert torch.equal(y, e), (dtype, rep, k, n)
            torch.cuda.synchronize()
            lib.fft_batch(lib.fft_alloc(256), bufs[0][0][:1, :256].contiguous(),
                          torch.empty(1, 256, 2, dtype=bufs[0][0].dtype, device=dev), 1, 256, 256)
        O = Oracle("int16_t") if dtype == "int16_t" else None
        for (k, n, b), (x, y) in zip(items, bufs):
            xs, got = x.cpu().numpy(), y.cpu().numpy()
            if O is not None:
                want = O.fft(xs, inverse=(k == C2C_INV))
                np.testing.assert_array_equal(got, want, err_msg=f"q15 kind {k} n {n}")
            else:
                got = got if k == C2R else cplx(got)
                assert rel_l2(got, fp64_ref(k, n, xs)) <= ABS_FLOAT_BOUND, (k, n)
        m.close()

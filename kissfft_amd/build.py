"""Builds the drop-in libraries libkissfft-<type>.so for gfx950, in-tree.

    python -m kissfft_amd.build            # all three datatypes
    python -m kissfft_amd.build float      # one

Layout (kissfft_amd/lib/):
    obj/kfa_engine.o                 HIP engine (hipcc --offload-arch=gfx950), one object
                                     serving every scalar kind
    libkissfft-<type>.so.131         C host layer of that datatype + the engine; soname
                                     libkissfft-<type>.so.131, as the reference's
                                     Makefile:89-113 names its per-datatype libraries
    libkissfft-<type>.so             symlink to the above
    fftutil-<type>                   the stream tool (tools/fftutil.c protocol), linked
                                     against the library; alias fastfilt-<type>
    libkfa_membench.so               streaming-ceiling kernels for bench.py (not the library)

The host layer is compiled with gcc and without -ffast-math (the fixed-point twiddle
tables depend on exact FP64 rounding).  All HIP code is compiled with packed-FP32
instructions disabled.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")

ARCH = os.environ.get("KISSFFT_AMD_ARCH", "gfx950")
SOVERSION = "131"
DATATYPES = {
    "float": ["-Dkiss_fft_scalar=float"],
    "int16_t": ["-DFIXED_POINT=16"],
    "int32_t": ["-DFIXED_POINT=32"],
}
HOST_SOURCES = ["kf_core.c", "kf_real.c", "kf_nd.c", "kf_ndr.c", "kf_cache.c", "kf_serve.c", "kf_multi.c", "kf_fastfir.c"]
ENGINE_SOURCES = ["kfa_engine.hip", "kfa_fast_reg.hip", "kfa_tile_reg.hip", "kfa_serve.hip", "kfa_multi.hip", "kfa_conv.hip",
                  "kfa_rtc.hip"]
# headers the runtime compiler needs for k_fast and the fused multi-batch launch
# (kfa_rtc.hip), embedded in the library
RTC_HEADERS = ["kfa_fast.hpp", "kfa_arith.hpp", "kfa_lines.hpp", "kfa_q15p.hpp", "kfa_tile.hpp", "kfa_blue.hpp",
               "kfa_q15w.hpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=gnu99", "-fvisibility=hidden", "-DKISS_FFT_SHARED", "-Wall"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd[:3])} ...")
    elif verbose and (r.stderr.strip()):
        # hipcc's host pass warns that it does not know the packed-fp32 feature: harmless
        lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")


# Every kernel must keep its data in registers and LDS: a kernel that spills to scratch
# (private memory in HBM) runs several times slower, and a change that causes one -- an
# argument block copied to the stack, say -- is easy to miss.  The engine is compiled with
# the kernel resource-usage remarks and the build fails on any kernel with scratch.
# The exceptions are the general line kernels (k_lines, and k_multi, its form for the
# multi-batch): their radix schedule is chosen at run time (any radix, the generic DFT
# included), so the butterfly operands sit in a local array indexed at run time (96 bytes
# per lane); they serve the lengths and layouts no compiled kernel covers.
RESOURCE_REMARKS = ["-Xarch_device", "-Rpass-analysis=kernel-resource-usage"]
SCRATCH_ALLOWED = ("_ZN3kfa7k_lines", "_ZN3kfa7k_multi")


def _scratch_kernels(remarks: str) -> list[tuple[str, int]]:
    """(kernel, bytes per lane) for every kernel the remarks report with scratch."""
    out, name = [], None
    for line in remarks.splitlines():
        if "remark: Function Name: " in line:
            name = line.split("remark: Function Name: ", 1)[1].split(" [-Rpass")[0].strip()
        elif "remark:     ScratchSize [bytes/lane]: " in line and name:
            n = int(line.split("ScratchSize [bytes/lane]: ", 1)[1].split()[0])
            if n and not name.startswith(SCRATCH_ALLOWED):
                out.append((name, n))
    return out


def _run_engine(cmd: list[str], verbose: bool) -> int:
    """Compiles one engine source; returns the number of kernels it reports."""
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run([*cmd, *RESOURCE_REMARKS], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd[:3])} ...")
    bad = _scratch_kernels(r.stderr)
    if bad:
        lines = "\n".join(f"  {n} bytes/lane  {k}" for k, n in bad)
        raise RuntimeError(f"kernels use scratch memory (registers spilled or locals on the stack):\n{lines}")
    return r.stderr.count("remark: Function Name: ")


def lib_path(datatype: str) -> str:
    return os.path.join(LIBDIR, f"libkissfft-{datatype}.so")


def build(datatypes=None, verbose: bool = True) -> list[str]:
    datatypes = list(datatypes or DATATYPES)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    engine_objs = [os.path.join(objdir, os.path.splitext(src)[0] + ".o") for src in ENGINE_SOURCES]
    cmds = [[HIPCC, *HIP_FLAGS, "-I", CSRC, "-I", INCLUDE, "-c", os.path.join(CSRC, "engine", src), "-o", obj]
            for src, obj in zip(ENGINE_SOURCES, engine_objs)]
    # the HIP translation units are independent: compile them side by side
    workers = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0") or 0) or (os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as pool:
        nkernels = sum(f.result() for f in [pool.submit(_run_engine, c, verbose) for c in cmds])
    if verbose:
        print(f"{nkernels} kernels; none but the general line kernels use scratch memory", flush=True)
    engine_objs.append(build_rtc_headers(objdir, verbose))
    built = []
    for dt in datatypes:
        host_objs = []
        for src in HOST_SOURCES:
            obj = os.path.join(objdir, f"{os.path.splitext(src)[0]}-{dt}.o")
            _run([CC, *HOST_FLAGS, *DATATYPES[dt], "-I", INCLUDE, "-I", CSRC, "-c",
                  os.path.join(CSRC, "host", src), "-o", obj], verbose)
            host_objs.append(obj)
        soname = f"libkissfft-{dt}.so.{SOVERSION}"
        real = os.path.join(LIBDIR, soname)
        tmp = real + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *host_objs, *engine_objs,
              f"-Wl,-soname,{soname}", "-Wl,-Bsymbolic", "-lpthread", "-lm", "-ldl"], verbose)
        os.replace(tmp, real)
        link = lib_path(dt)
        if os.path.lexists(link):
            os.remove(link)
        os.symlink(soname, link)
        built.append(real)
        built.append(build_fftutil(dt, verbose))
    built.append(build_membench(verbose))
    return built


def build_rtc_headers(objdir: str, verbose: bool = True) -> str:
    """kfa_rtc_headers.o: the kernel headers as string constants, for the runtime compiler
    (kfa_rtc.hip compiles k_fast for lengths outside the built-in registry)."""
    src = os.path.join(objdir, "kfa_rtc_headers.cpp")
    lines = ['// generated by kissfft_amd/build.py from kissfft_amd/csrc/engine: do not edit',
             f'extern "C" const int kfa_rtc_nheaders = {len(RTC_HEADERS)};',
             'extern "C" const char *const kfa_rtc_header_names[] = {' +
             ", ".join(f'"{h}"' for h in RTC_HEADERS) + '};',
             'extern "C" const char *const kfa_rtc_header_srcs[] = {']
    for h in RTC_HEADERS:
        with open(os.path.join(CSRC, "engine", h)) as f:
            text = f.read()
        # one C string literal per header, split into short pieces
        pieces = [text[i:i + 2000] for i in range(0, len(text), 2000)] or [""]
        lines.append("    " + "\n".join(
            '"' + p.replace("\\", "\\\\").replace('"', '\\"').replace("\n", "\\n") + '"' for p in pieces) + ",")
    lines.append("};")
    with open(src, "w") as f:
        f.write("\n".join(lines) + "\n")
    obj = os.path.join(objdir, "kfa_rtc_headers.o")
    _run(["g++", "-O0", "-fPIC", "-fvisibility=hidden", "-c", src, "-o", obj], verbose)
    return obj


def build_membench(verbose: bool = True) -> str:
    """libkfa_membench.so: the streaming-ceiling kernels bench.py reports beside the FFT
    kernels (not part of the FFT library)."""
    out = os.path.join(LIBDIR, "libkfa_membench.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "bench", "kfa_membench.hip")], verbose)
    os.replace(tmp, out)
    return out


def build_fftutil(dt: str, verbose: bool = True) -> str:
    """fftutil-<type>: the reference's stream tool (tools/fftutil.c protocol) over the
    library, plus the name the reference's test harness runs it under (fastfilt-<type>)."""
    exe = os.path.join(LIBDIR, f"fftutil-{dt}")
    _run([CC, "-O2", "-std=gnu99", "-Wall", *DATATYPES[dt], "-I", INCLUDE, os.path.join(CSRC, "tools", "fftutil.c"),
          "-L", LIBDIR, f"-lkissfft-{dt}", "-Wl,-rpath,$ORIGIN", "-o", exe], verbose)
    alias = os.path.join(LIBDIR, f"fastfilt-{dt}")
    if os.path.lexists(alias):
        os.remove(alias)
    os.symlink(os.path.basename(exe), alias)
    return exe


def build_oracle(verbose: bool = True) -> None:
    """Oracle libraries (oracle/lib) and, where the reference tree exists, oracle/_ref."""
    mk = shutil.which("make") or "make"
    _run([mk, "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"], verbose)
    _run([mk, "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], verbose)


if __name__ == "__main__":
    dts = sys.argv[1:] or None
    for p in build(dts):
        print("built", p)

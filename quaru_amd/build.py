"""Build the engine's shared library for gfx950 with hipcc.

``python -m quaru_amd.build`` (or ``build()``) compiles every HIP source in
``quaru_amd/csrc`` into ``quaru_amd/libquaru_hip.so``.  hipcc cross-compiles
for gfx950 without a GPU present.  Packed FP32 instructions are disabled
(``-packed-fp32-ops``) for every HIP translation unit.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libquaru_hip.so")
SOURCES = ["qs_kernels.hip", "qs_sweep.hip", "qs_register.hip", "qs_adjoint.hip",
           "qs_sweep_jit.hip"]
# qs_sweep_launch.hip holds the k_sweep instantiations of one (precision,
# register bits) pair; it is compiled once per pair so that they build in
# parallel.
SWEEP_VARIANTS = [(r, a) for r in ("double", "float") for a in (3, 4, 5)]


def units():
    """(source, extra flags, object name) for every translation unit."""
    out = [("qs_sweep_launch.hip", [f"-DQS_SWEEP_R={r}", f"-DQS_SWEEP_A={a}"],
            f"qs_sweep_launch_{r}_{a}.o") for r, a in SWEEP_VARIANTS]
    return out + [(s, [], s + ".o") for s in SOURCES]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the engine needs ROCm's hipcc to build")


OBJDIR = os.path.join(HERE, "_build")


def _stale(obj: str, src: str) -> bool:
    """An object is stale when it is missing or older than its source or any
    header of csrc (headers are shared, so any change rebuilds every unit)."""
    if not os.path.exists(obj):
        return True
    deps = [src] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    deps.append(os.path.join(os.path.dirname(HERE), "include", "quaru_hip.h"))
    return os.path.getmtime(obj) <= max(os.path.getmtime(d) for d in deps)


def write_kernel_source(objdir: str) -> str:
    """The sweep kernel header as a C++ string literal
    (qs_sweep_kernel_src.inc), embedded in the library for the run-time
    compiles of specialised sweep kernels (qs_sweep_jit.hip)."""
    with open(os.path.join(CSRC, "qs_sweep_kernel.h")) as f:
        text = f.read()
    if ")QSK\"" in text:
        raise RuntimeError("qs_sweep_kernel.h contains the raw-string delimiter")
    out = os.path.join(objdir, "qs_sweep_kernel_src.inc")
    body = 'R"QSK(' + text + ')QSK"\n'
    old = open(out).read() if os.path.exists(out) else None
    if old != body:
        with open(out, "w") as f:
            f.write(body)
    return out


def build(verbose: bool = False, jobs: int = 0, incremental: bool = False) -> str:
    """Compile each translation unit to an object in parallel, then link the
    library.  ``incremental`` keeps objects in quaru_amd/_build and recompiles
    only stale ones (a developer shortcut; the default rebuilds everything)."""
    cc = hipcc()
    jobs = jobs or max(1, min(8, os.cpu_count() or 1))
    if incremental:
        os.makedirs(OBJDIR, exist_ok=True)
        objdir = OBJDIR
    else:
        objdir = tempfile.mkdtemp(prefix="quaru_hip_build_")
    try:
        write_kernel_source(objdir)
        procs = []
        objs = []
        for src, extra, oname in units():
            path = os.path.join(CSRC, src)
            obj = os.path.join(objdir, oname)
            objs.append(obj)
            if incremental and not _stale(obj, path):
                continue
            cmd = [cc, *HIPCC_FLAGS, *extra, "-I", objdir, "-c", path, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                stderr=subprocess.STDOUT, text=True)))
            while sum(p.poll() is None for _, p in procs) >= jobs:
                time.sleep(0.1)
        failed = []
        for src, p in procs:
            out, _ = p.communicate()
            if p.returncode != 0:
                failed.append((src, out))
            elif verbose and out.strip():
                print(out)
        if failed:
            msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
            raise RuntimeError(f"hipcc failed:\n{msg}")
        tmp = LIB + ".tmp"
        cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, LIB)
    finally:
        if not incremental:
            shutil.rmtree(objdir, ignore_errors=True)
    build_programs(verbose=verbose)
    return LIB


# C++ programs over the host API (include/quaru/quaru.hpp): the reference's
# integration tests and examples, linked against the library above.
CPP_DIR = os.path.join(HERE, "cpp")
BIN_DIR = os.path.join(HERE, "bin")
PROGRAMS = ["register_operation_tests", "examples", "adjoint_ansatz"]
# Programs built with hipcc (no library): sweep_memory times the fused
# sweep's load/store path for given tile shapes; qs_sweep_compile is the
# helper process that compiles specialised sweep kernels with hipRTC.
HIP_PROGRAMS = {"sweep_memory": ("sweep_memory.hip", []),
                "qs_sweep_compile": ("qs_sweep_compile.cpp", ["-lhiprtc"])}
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def build_programs(verbose: bool = False) -> list:
    """g++ each C++ program in ``quaru_amd/cpp`` into ``quaru_amd/bin`` (the
    binaries find the library through an rpath relative to themselves), and
    hipcc each HIP program there."""
    os.makedirs(BIN_DIR, exist_ok=True)
    cxx = os.environ.get("CXX", "g++")
    procs = []
    for name in PROGRAMS:
        out = os.path.join(BIN_DIR, name)
        cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", INCLUDE,
               os.path.join(CPP_DIR, name + ".cpp"), "-L", HERE, "-lquaru_hip",
               "-Wl,-rpath,$ORIGIN/..", "-o", out]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((name, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                             stderr=subprocess.STDOUT, text=True)))
    for name, (src, libs) in HIP_PROGRAMS.items():
        cmd = [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", CSRC,
               os.path.join(CPP_DIR, src), *libs, "-o", os.path.join(BIN_DIR, name)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((name, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                             stderr=subprocess.STDOUT, text=True)))
    failed = []
    for name, p in procs:
        log, _ = p.communicate()
        if p.returncode != 0:
            failed.append((name, log))
    if failed:
        raise RuntimeError("g++ failed:\n" + "\n".join(f"--- {n}\n{o}" for n, o in failed))
    return [os.path.join(BIN_DIR, n) for n in PROGRAMS + list(HIP_PROGRAMS)]


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))

"""Process exit while sweep kernels compile at run time (no GPU needed).
Each kernel compiles with hipRTC in a helper process that the library
starts; at exit the library stops the helpers still running and waits for
none.  Compiling in the process itself went wrong at exit: hipRTC's
exit-time teardown ran under compiles still in flight, and the process
crashed or hung.  The compiler cache is switched off so that every compile
really runs."""
import os
import subprocess
import time

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "quaru_amd", "csrc")
LIBDIR = os.path.join(ROOT, "quaru_amd")
SRC = os.path.join(ROOT, "tests", "cpp", "jit_exit.cpp")
HIPCC = "/opt/rocm/bin/hipcc"

pytestmark = pytest.mark.skipif(
    not (os.path.exists(HIPCC) and os.path.exists(os.path.join(LIBDIR, "libquaru_hip.so"))),
    reason="needs hipcc and the built library")


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("jit_exit") / "jit_exit")
    subprocess.run([HIPCC, "-std=c++17", "-O1", "-I", CSRC, SRC, "-L", LIBDIR, "-lquaru_hip",
                    f"-Wl,-rpath,{LIBDIR}", "-o", out],
                   check=True, capture_output=True, timeout=600)
    return out


def _run(prog, *args, **env_extra):
    env = dict(os.environ, AMD_COMGR_CACHE="0", QS_SWEEP_JIT="1", **env_extra)
    t0 = time.perf_counter()
    p = subprocess.run([prog, *args], env=env, capture_output=True, text=True, timeout=120)
    return p, time.perf_counter() - t0


def test_compiles_succeed(prog):
    p, _ = _run(prog, "wait")
    assert p.returncode == 0, p.stdout + p.stderr
    assert "compiled 8 failed 0" in p.stdout, p.stdout + p.stderr


def test_exit_with_compiles_in_flight(prog):
    for _ in range(6):
        p, dt = _run(prog)
        assert p.returncode == 0, f"exit status {p.returncode}\n{p.stdout}{p.stderr}"
        assert "main returns" in p.stdout
        # the program sleeps 0.5 s; the exit itself waits for no compile
        assert dt < 10.0, f"exit took {dt:.1f} s"


def test_exit_status_kept(prog):
    """The exit status of a program that exits with compiles running."""
    for status in ("0", "3"):
        p, dt = _run(prog, "status", status)
        assert p.returncode == int(status), f"exit status {p.returncode}\n{p.stdout}{p.stderr}"
        assert "main returns" in p.stdout
        assert dt < 10.0


def test_register_patterns_match_the_run_time_test(tmp_path):
    """The register-slot patterns the specialised kernels take as constants
    (slot mask and values) select exactly the registers the generic kernel's
    run-time test selects, for random slot qubits and patterns."""
    out = str(tmp_path / "jit_patterns")
    src = os.path.join(ROOT, "tests", "cpp", "jit_patterns.cpp")
    subprocess.run([HIPCC, "-std=c++17", "-O1", "-I", CSRC, src, "-L", LIBDIR, "-lquaru_hip",
                    f"-Wl,-rpath,{LIBDIR}", "-o", out],
                   check=True, capture_output=True, timeout=600)
    p = subprocess.run([out], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.startswith("checked "), p.stdout

"""The C++ host API on the GPU: the reference's integration tests
(tests/register_operation_tests.rs) and examples (grovers.rs,
quantum_teleportation.rs, superdense_coding.rs, the QFT and modular
multiplication of shors.rs) compiled as C++ programs over
include/quaru/quaru.hpp, run in complex128 and complex64."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "quaru_amd", "bin")


def _run(name: str, dtype: str) -> str:
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
    print(f"running {name} {dtype}", flush=True)
    p = subprocess.run([exe, dtype], capture_output=True, text=True, timeout=100)
    print(p.stdout[-4000:])
    if p.returncode != 0:
        raise AssertionError(f"{name} {dtype} exited {p.returncode}:\n{p.stdout}\n{p.stderr}")
    return p.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_cpp_register_operation_tests(dtype):
    out = _run("register_operation_tests", dtype)
    assert "FAIL" not in out and out.rstrip().splitlines()[-1].startswith("OK")


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_cpp_examples(dtype):
    out = _run("examples", dtype)
    assert "FAIL" not in out and out.rstrip().splitlines()[-1].startswith("OK")


def test_cpp_programs_built():
    """The programs link against the library (no GPU needed to check)."""
    for name in ("register_operation_tests", "examples"):
        assert os.path.exists(os.path.join(BIN, name)), f"{name} not built"


def test_cpp_header_is_self_contained(tmp_path):
    """include/quaru/quaru.hpp compiles on its own with -Wall -Wextra -Werror."""
    src = tmp_path / "t.cpp"
    src.write_text("#include <quaru/quaru.hpp>\nint main() { return 0; }\n")
    p = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr

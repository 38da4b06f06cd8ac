"""The C++ host API on the GPU: the reference's integration tests
(tests/register_operation_tests.rs) and examples (grovers.rs,
quantum_teleportation.rs, superdense_coding.rs, the QFT and modular
multiplication of shors.rs) compiled as C++ programs over
include/quaru/quaru.hpp, run in complex128 and complex64."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "quaru_amd", "bin")


def _run(name: str, dtype: str, *args, stdin=None) -> str:
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
    print(f"running {name} {' '.join(args)} {dtype}", flush=True)
    cmd = [exe, *args, dtype] if not args else [exe, *args]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=100, input=stdin)
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


def _sections(out: str) -> dict:
    sec, cur = {}, None
    for line in out.split():
        if line in ("energies", "grads", "vjp"):
            cur = sec.setdefault(line, [])
        else:
            cur.append(float(line))
    return {k: np.array(v) for k, v in sec.items()}


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,ckpt,tol", [("c128", "0", 1e-12), ("c128", "1", 1e-12),
                                            ("c64", "1", 1e-5)])
def test_cpp_adjoint_golden_4q(golden, dtype, ckpt, tol):
    """quaru::adjoint (Function-style forward/backward in quaru.hpp) on the
    4-qubit golden: energies and parameter-shift gradients written by numpy
    (tests/golden/make_golden.py)."""
    g = golden("adj_4q.npz")
    th = np.asarray(g["thetas"], dtype=np.float64)
    text = "\n".join(" ".join(f"{x:.17g}" for x in row) for row in th)
    out = _sections(_run("adjoint_ansatz", dtype, str(int(g["n"])), str(int(g["layers"])),
                         dtype, ckpt, stdin=text))
    grads = out["grads"].reshape(th.shape)
    assert np.max(np.abs(out["energies"] - g["energies"])) < tol
    assert np.max(np.abs(grads - g["grads"])) < tol
    w = np.arange(1, th.shape[0] + 1, dtype=np.float64)
    assert np.max(np.abs(out["vjp"].reshape(th.shape) - w[:, None] * g["grads"])) < tol * th.shape[0]


def test_cpp_programs_built():
    """The programs link against the library (no GPU needed to check)."""
    for name in ("register_operation_tests", "examples", "adjoint_ansatz"):
        assert os.path.exists(os.path.join(BIN, name)), f"{name} not built"


def test_cpp_header_is_self_contained(tmp_path):
    """include/quaru/quaru.hpp compiles on its own with -Wall -Wextra -Werror."""
    src = tmp_path / "t.cpp"
    src.write_text("#include <quaru/quaru.hpp>\nint main() { return 0; }\n")
    p = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr

"""CPU check of the sweep planner's op-merging rules (qs_sweep_merge.h).

tests/cpp/merge_check.cpp applies random gate runs one by one and as the
merged op list the sweep encodes (multiplexed 2x2 ops in place of the gates
they absorb); both must agree to 1e-12.  A copy of the rules with one ordering
check removed must fail the same program, so the check has teeth.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "quaru_amd", "csrc")
SRC = os.path.join(ROOT, "tests", "cpp", "merge_check.cpp")
HIPCC = "/opt/rocm/bin/hipcc"

pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")


def build_and_run(include_dir, out_dir):
    exe = os.path.join(out_dir, "merge_check")
    subprocess.run([HIPCC, "-std=c++17", "-O1", "-I", include_dir, SRC, "-o", exe],
                   check=True, capture_output=True, timeout=300)
    return subprocess.run([exe], capture_output=True, text=True, timeout=300)


def test_merged_ops_equal_gates_one_by_one(tmp_path):
    res = build_and_run(CSRC, str(tmp_path))
    assert res.returncode == 0, res.stdout + res.stderr
    ok, trials, merged = res.stdout.split()
    assert ok == "ok" and int(merged) > int(trials)  # merges happen, many of them


def test_checker_catches_a_dropped_ordering_update(tmp_path):
    planted = tmp_path / "planted"
    planted.mkdir()
    shutil.copy(os.path.join(CSRC, "qs_internal.h"), planted)
    text = open(os.path.join(CSRC, "qs_sweep_merge.h")).read()
    line = "            lastx[t] = j;\n"
    assert text.count(line) == 1
    (planted / "qs_sweep_merge.h").write_text(text.replace(line, ""))
    res = build_and_run(str(planted), str(tmp_path))
    assert res.returncode == 1 and "max diff" in res.stdout

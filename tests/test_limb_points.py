"""The MSM kernels' limb-form point arithmetic (G1L, G2L: bucket sums, merges and reductions
held in 29-bit limbs; P1L, P2L: the constant-work complete formulas, bn254.cuh) against the
word-form formulas of the same header, on the host: mixed-addition chains with doublings and
cancellations, general additions (p + q, p + p, p + (-p)), every value's bound after every
step, the complete formulas coordinate by coordinate (every input at 2q - 1 included), and
G1's b3 step on values below 7q. The word-form formulas are the ones the GPU MSM tests pin
against the oracle; this pins the limb forms' rare branches and bounds without a GPU. Built
host-only with hipcc at -O0 (seconds)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is not installed")
def test_limb_points_match_word_form(tmp_path):
    exe = tmp_path / "limb_points"
    subprocess.run([HIPCC, "-O0", "-std=c++17", "--cuda-host-only", "-x", "hip",
                    os.path.join(HERE, "host", "limb_points.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=600)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok "), r.stdout

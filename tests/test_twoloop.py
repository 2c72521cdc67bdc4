"""The two-loop recursion on coefficients (superdca_amd/csrc/lbfgs_coeffs.hpp), which turns
the batched L-BFGS's inner products into the next search direction, against the recursion
on explicit vectors (Nocedal & Wright, Algorithm 7.4) and against the secant equation
H y_k = s_k of the newest pair, for pair slots in order and for a wrapped ring of slots
(tests/twoloop_check.cpp drives it; the pairs' s.y matrix is not symmetric, so a
transposed product would show)."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "superdca_amd", "csrc")


def test_coefficient_recursion_matches_vectors_and_secant_equation(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = str(tmp_path / "twoloop_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "twoloop_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    rows = [line.split() for line in out if line.strip()]
    assert [int(r[1]) for r in rows] == [0, 1, 3, 10, 5, 8]
    for r in rows:
        assert float(r[2]) < 1e-12, r  # direction vs the vector recursion (relative, max norm)
        assert float(r[3]) < 1e-12, r  # -H y_newest = -s_newest

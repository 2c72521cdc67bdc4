"""The host line search of the batched L-BFGS (superdca_amd/csrc/linesearch.hpp) on the
test functions of More and Thuente, "Line search algorithms with guaranteed sufficient
decrease", ACM TOMS 20 (1994) 286-307, section 5 (tests/linesearch_check.cpp drives it).

SuperDCA's optimizer uses the same method through CppNumericalSolvers, which is not
vendored (plmDCA.hpp:218-221,304), so its trajectory cannot be compared; the paper's
Tables 1-6 (final step to two significant digits and the number of function
evaluations, for starting steps 1e-3, 1e-1, 10 and 1000) pin this implementation instead,
and every final step must satisfy the strong Wolfe conditions."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "superdca_amd", "csrc")

# (function, alpha0) -> (final step as printed in the paper, evaluations)
PAPER = {
    (1, 1e-3): (1.4, 6), (1, 1e-1): (1.4, 3), (1, 1e1): (10, 1), (1, 1e3): (37, 4),
    (2, 1e-3): (1.6, 12), (2, 1e-1): (1.6, 8), (2, 1e1): (1.6, 8), (2, 1e3): (1.6, 11),
    (3, 1e-3): (1.0, 12), (3, 1e-1): (1.0, 12), (3, 1e1): (1.0, 10), (3, 1e3): (1.0, 13),
    (4, 1e-3): (0.085, 4), (4, 1e-1): (0.10, 1), (4, 1e1): (0.35, 3), (4, 1e3): (0.83, 4),
    (5, 1e-3): (0.075, 6), (5, 1e-1): (0.078, 3), (5, 1e1): (0.073, 7), (5, 1e3): (0.076, 8),
    (6, 1e-3): (0.93, 13), (6, 1e-1): (0.93, 11), (6, 1e1): (0.92, 8), (6, 1e3): (0.92, 11),
}
# (ftol, gtol) of each function's table
CONSTANTS = {1: (1e-3, 0.1), 2: (0.1, 0.1), 3: (0.1, 0.1), 4: (1e-3, 1e-3), 5: (1e-3, 1e-3), 6: (1e-3, 1e-3)}


def sig2(x: float) -> float:
    return float(f"{x:.2g}")


@pytest.fixture(scope="module")
def runs(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = str(tmp_path_factory.mktemp("ls") / "linesearch_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "linesearch_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    res = {}
    for line in out.splitlines():
        t = line.split()
        fid, a0 = int(t[0]), float(t[1])
        stp, f, g, f0, g0 = map(float, t[2:7])
        res[(fid, a0)] = {"stp": stp, "f": f, "g": g, "f0": f0, "g0": g0, "nev": int(t[7]), "res": int(t[8])}
    return res


def test_every_run_converges_to_a_strong_wolfe_point(runs):
    assert set(runs) == set(PAPER)
    for (fid, a0), r in runs.items():
        ftol, gtol = CONSTANTS[fid]
        assert r["res"] == 1, (fid, a0)  # Converged
        assert r["f"] <= r["f0"] + ftol * r["stp"] * r["g0"], (fid, a0)
        assert abs(r["g"]) <= gtol * abs(r["g0"]), (fid, a0)


def test_steps_and_evaluations_match_the_paper(runs):
    for key, (stp, nev) in PAPER.items():
        r = runs[key]
        assert r["nev"] == nev, (key, r["nev"], nev)
        assert sig2(r["stp"]) == pytest.approx(stp), (key, r["stp"], stp)

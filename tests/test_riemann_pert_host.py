"""CPU: the perturbation-form Riemann solvers of the fused path
(snapy_linfel_amd/csrc/riemann_pert.hpp) equal the total-state solvers
(riemann_impl.hpp) minus the background flux (0, p_f, 0, 0, 0) in FP64 to
1e-11 of the largest flux, over subsonic, supersonic and LLF-fallback
states; in FP32, a 1 Pa pressure perturbation on a 1e5 Pa hydrostatic
background gives flux perturbations within 1e-6 of FP64 (per component,
against its largest size), where the total-state FP32 solver is off by more
than 1e-3.  The device headers are compiled for the host with hipcc."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "host_checks", "riemann_pert_check.cpp")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("rpert") / "riemann_pert_check")
    subprocess.run([hipcc, "-std=c++17", "-O2", "-I", os.path.join(ROOT, "snapy_linfel_amd", "csrc"),
                    "-I", os.path.join(ROOT, "include"), SRC, "-o", exe],
                   check=True, capture_output=True, timeout=300)
    return exe


@pytest.mark.parametrize("solver", [0, 1, 2], ids=["roe", "lmars", "hllc"])
def test_perturbation_form_equals_total_form(checker, solver):
    r = subprocess.run([checker, str(solver), "20000"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    w = r.stdout.split()
    fp64, pert, total = float(w[1]), float(w[3]), float(w[5])
    assert fp64 < 1e-11
    assert pert < 1e-6
    assert total > 1e-3

"""CPU: the hand-written reverse sweeps of the Roe and LMARS fluxes
(snapy_linfel_amd/csrc/riemann_adj.hpp) equal the transposed Jacobian that
forward-mode duals give on the same solver code (riemann_impl.hpp), to
roundoff, over random states that reach every branch (Roe: LLF fallback, left
and right one-sided fluxes, Roe flux; LMARS: both upwind sides).  The device
headers are compiled for the host with hipcc."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "host_checks", "riemann_adj_check.cpp")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("radj") / "riemann_adj_check")
    subprocess.run([hipcc, "-std=c++17", "-O2", "-I", os.path.join(ROOT, "snapy_linfel_amd", "csrc"),
                    "-I", os.path.join(ROOT, "include"), SRC, "-o", exe],
                   check=True, capture_output=True, timeout=300)
    return exe


@pytest.mark.parametrize("solver,name", [(0, "roe"), (1, "lmars")])
def test_reverse_flux_matches_duals(checker, solver, name):
    r = subprocess.run([checker, str(solver), "40000"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    worst, counts = r.stdout.split()[1], [int(x) for x in r.stdout.split()[3::2]]
    assert float(worst) < 1e-10
    if name == "roe":
        assert min(counts) > 100, r.stdout  # every branch reached

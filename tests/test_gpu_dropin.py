"""The C++ drop-in API (include/schwinger_mg.hpp) used the way the reference's
callers use its headers: tests/cpp/dropin_test.cpp builds a hierarchy with
build_mg_hierarchy, solves with cg_solve / gcr_solve_mg / cg_solve_precond and
computes forces, each checked against the oracle on the same inputs."""
import os
import subprocess

import pytest

from helpers import banner

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "dropin_test")


def test_cpp_dropin_against_oracle():
    banner("C++ drop-in caller")
    assert os.path.exists(EXE), "tests/cpp/dropin_test is built by __graft_entry__.build()"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0 and "ALL OK" in r.stdout


LAMBDA_EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "lambda_callers")


def test_cpp_lambda_callers_against_oracle():
    """OpApply and preconditioner lambdas, as the reference's callers write them
    (src/hmc_forces.cpp:113-129, src/mode_mg_hmc.cpp:92-97), through cg_solve,
    cg_solve_precond, fcg_solve, fgmres_solve_generic, mr_smooth_op,
    estimate_lambda_max and SparseCoarseOp::build: equal to the header's own operator
    objects and to the oracle."""
    banner("C++ lambda callers")
    assert os.path.exists(LAMBDA_EXE), "tests/cpp/lambda_callers is built by __graft_entry__.build()"
    r = subprocess.run([LAMBDA_EXE], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0 and "ALL OK" in r.stdout

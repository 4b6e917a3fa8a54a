"""CPU checks: golden fixture of the oracle pass, the C-ABI library exports,
error paths that need no GPU, and the bench CLI contract."""
import ctypes as ct
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from dftfe_amd import fe
from oracle import chfsi_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "cube20_p4_n64_oracle.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_problem_checksums(golden):
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_golden import problem_checksums

    got = problem_checksums(fe.build_workload("cube20_p4_n64"))
    for k, v in golden["checksums"].items():
        assert got[k] == pytest.approx(v, rel=1e-12, abs=1e-12), k


@pytest.mark.parametrize("name", ["cube30_p6_n1024", "cube24_p5_n256"])
def test_problem_checksums_full_size(name):
    """The full-size workloads regenerate the inputs their fixtures were made from."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_golden import fixture_path, problem_checksums

    with open(fixture_path(name)) as f:
        gold = json.load(f)
    got = problem_checksums(fe.build_workload(name), N=gold["states"])
    for k, v in gold["checksums"].items():
        assert got[k] == pytest.approx(v, rel=1e-12, abs=1e-12), k


def test_oracle_pass_matches_golden(golden):
    prob = fe.build_workload("cube20_p4_n64")
    ref = O.OracleOperator(prob)
    lo, up, theta, fnorm = O.lanczos_bounds(ref, prob.lanczos_start, 10)
    assert (lo, up) == (golden["lanczos"]["lower"], golden["lanczos"]["upper"])
    b = golden["bounds"]
    assert O.first_call_bounds(lo, up, 64, prob.M) == pytest.approx((b["a0"], b["bLow"], b["bUp"]))
    lam, res, _ = O.chfsi_pass(ref, prob.initial_guess(), b["a0"], b["bLow"], b["bUp"], 20, 64)
    assert np.abs(lam - golden["pass1"]["eigenvalues"]).max() < 1e-11
    assert np.abs(res - golden["pass1"]["residuals"]).max() < 1e-10


def _header_functions():
    with open(os.path.join(ROOT, "include", "chfsi.h")) as f:
        txt = f.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(chfsi_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    from dftfe_amd import lib as L

    lib = L.load()
    names = _header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), f"{n} declared in chfsi.h but not exported"
        assert n in L.SIGNATURES, f"{n} has no ctypes signature"


def test_error_paths_without_gpu():
    from dftfe_amd import lib as L

    lib = L.load()
    h = ct.c_void_p()
    st = lib.chfsi_operator_create(None, ct.byref(h))
    assert st == 1 and b"null" in lib.chfsi_last_error()
    d = L.OperatorDesc()
    d.n_dofs = 0
    assert lib.chfsi_operator_create(ct.byref(d), ct.byref(h)) == 1
    assert lib.chfsi_solver_reinit_bounds(None, 0.0, 1.0, 2.0) == 1
    p = L.Params()
    lib.chfsi_default_params(ct.byref(p))
    assert p.cheby_block_size == 400 and p.use_shep == 1
    assert abs(p.first_scf_degree_factor - 1.34) < 1e-15


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr

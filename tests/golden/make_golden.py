"""Writes golden fixtures of oracle ChFSI passes (the reference ships no vectors
for this path; these pin the oracle's own output so later changes to it are
caught, and give the GPU tests at the production shapes a checked answer
without re-running the oracle's pass on the GPU box).

    python tests/golden/make_golden.py              # every fixture
    python tests/golden/make_golden.py cube20_p4_n64

Fixtures (tests/golden/<workload>_oracle.json):
  cube20_p4_n64    20 bohr, 8^3 cells, p=4: two passes of 64 states, degree 20
  cube30_p6_n1024  30 bohr, 16^3 cells, p=6 (884,736 DOFs): one pass of the
                   first 64 states (initial_guess(N=64)) in blocks of 32, degree 30
  cube24_p5_n256   24 bohr, 12^3 cells, p=5 (216,000 DOFs): one pass of 64
                   states in blocks of 32, degree 20
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from dftfe_amd import fe  # noqa: E402
from oracle import chfsi_oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

# workload -> (states, block, degree, passes) of the fixture's pass(es)
SPECS = {
    "cube20_p4_n64": (64, 64, 20, 2),
    "cube30_p6_n1024": (64, 32, 30, 1),
    "cube24_p5_n256": (64, 32, 20, 1),
}


def fixture_path(name):
    return os.path.join(HERE, f"{name}_oracle.json")


def problem_checksums(prob, N=None):
    """Sums of the seeded inputs; N = states of the start block (default: the workload's)."""
    return {
        "M": int(prob.M), "C": int(prob.C), "n_pairs": int(len(prob.pair_cell)),
        "mass_sum": float(prob.mass.sum()),
        "V_sum": float(prob.V_q.sum()), "V_min": float(prob.V_q.min()),
        "pairC_abs_sum": float(np.abs(prob.pair_C).sum()),
        "proj_V_sum": float(prob.proj_V.sum()),
        "X0_sum": float(prob.initial_guess(N=N).sum()),
        "lanczos_start_sum": float(prob.lanczos_start.sum()),
    }


def make(name):
    N, block, m, passes = SPECS[name]
    t0 = time.time()
    prob = fe.build_workload(name)
    ref = O.OracleOperator(prob)
    lo, up, theta, fnorm = O.lanczos_bounds(ref, prob.lanczos_start, 10)
    a0, bLow, bUp = O.first_call_bounds(lo, up, N, prob.M)
    X0 = prob.initial_guess(N=N)
    lam, res, X = O.chfsi_pass(ref, X0, a0, bLow, bUp, m, block)
    out = {
        "workload": name,
        "checksums": problem_checksums(prob, N=N if N != prob.wl.N else None),
        "H_e_cell0_trace": float(np.trace(ref.H_e[0])),
        "H_e_abs_sum": float(np.abs(ref.H_e).sum()),
        "lanczos": {"steps": 10, "theta_min": float(theta[0]), "theta_max": float(theta[-1]),
                    "fnorm": float(fnorm), "lower": lo, "upper": up},
        "bounds": {"a0": a0, "bLow": bLow, "bUp": bUp},
        "states": N, "degree": m, "block": block,
        "pass1": {"eigenvalues": lam.tolist(), "residuals": res.tolist()},
    }
    if passes > 1:
        lam2, res2, _ = O.chfsi_pass(ref, X, *O.next_bounds(lam), bUp, m, block)
        out["pass2"] = {"eigenvalues": lam2.tolist(), "residuals": res2.tolist()}
    path = fixture_path(name)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path} ({time.time() - t0:.0f} s)")


def main(argv):
    for name in (argv or list(SPECS)):
        make(name)


if __name__ == "__main__":
    main(sys.argv[1:])

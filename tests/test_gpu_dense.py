"""GPU parity of the library's FP64 MFMA dense kernels (split-K gram for the
overlap and X^T H~X, row-panel rotations) against the oracle and against the
rocBLAS route (dense_rocblas), through the C ABI.

Workload: 20 bohr, 8^3 cells, p=4 (32,768 DOFs).  Column counts that are not
multiples of the 128-wide tiles (72, 200), that span several tiles (300) and a
multiple of the 16-wide K slabs (256: the 16-byte load path of the rotation).
Bars: CholGS 1e-10 relative to the oracle, Ritz values 1e-10 Ha, a whole pass
on either route within 1e-10 Ha of the oracle.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dftfe_amd import fe  # noqa: E402
from oracle import chfsi_oracle as O  # noqa: E402


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def small():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dftfe_amd.chfsi import KohnShamOperatorDevice

    prob = fe.build_workload("cube20_p4_n64")
    return prob, O.OracleOperator(prob), KohnShamOperatorDevice(prob)


@pytest.mark.parametrize("N", [72, 200, 256, 300])
def test_cholgs_and_rayleigh_ritz(small, N):
    from dftfe_amd.chfsi import cholgs, rayleigh_ritz

    prob, ref, op = small
    X = prob.sqrt_mass[:, None] * prob.initial_guess(N=N, seed_offset=N)
    Xd = torch.tensor(X, device="cuda")
    cholgs(op, Xd)
    Xo = O.cholgs(X)
    G = Xd.T @ Xd
    assert float((G - torch.eye(N, dtype=G.dtype, device=G.device)).abs().max()) < 1e-12
    assert _rel(Xd.cpu().numpy(), Xo) < 1e-10
    lam = rayleigh_ritz(op, Xd, 64)
    lam_o, Xr = O.rayleigh_ritz(ref, Xo, 64)
    assert np.abs(lam - lam_o).max() < 1e-10
    # rotated block: X^T H~ X diagonal with the Ritz values
    Hp = (Xd.T @ op.apply(Xd)).cpu().numpy()
    assert np.abs(Hp - np.diag(lam)).max() < 1e-10 * max(1.0, np.abs(lam).max())


def test_pass_library_kernels_and_rocblas(small):
    """One FP64 pass of 200 states in blocks of 64 on each dense route (library
    kernels, rocBLAS, library grams with rocBLAS rotations), all within 1e-10 Ha
    of the oracle."""
    from dftfe_amd.chfsi import ChebyshevSolverDevice

    prob, ref, op = small
    N, B, m = 200, 64, 20
    lo, up, _, _ = O.lanczos_bounds(ref, prob.lanczos_start, 10)
    bounds = O.first_call_bounds(lo, up, N, prob.M)
    X = prob.initial_guess(N=N, seed_offset=7)
    lam_o, res_o, _ = O.chfsi_pass(ref, X, *bounds, m, B)
    for rb in (0, 1, 2):
        solver = ChebyshevSolverDevice(*bounds, cheby_block_size=B, cheby_degree=m,
                                       reuse_lanczos_upper=True, filter_precision=0,
                                       dense_rocblas=rb)
        lam, res, _ = solver.solve(op, torch.tensor(X, device="cuda"))
        err = float(np.abs(lam - lam_o).max())
        print(f"dense_rocblas={rb}: max|dlambda| {err:.2e} Ha")
        assert err <= 1e-10
        assert np.abs(res - res_o).max() <= 1e-8

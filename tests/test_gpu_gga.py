"""GPU parity of the GGA cell Hamiltonian rebuild (chfsi_operator_set_potential_gga)
against the oracle's cell_hamiltonians with the GGA vector.

Workload: periodic 20 bohr cube, 8^3 cells, p=4, random V_eff and random
d(E_xc)/d(sigma) grad rho at every quadrature point.  Tolerances: cell
matrices 1e-13 relative (max-norm), FP64 HX 1e-13 relative.
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
def gga():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dftfe_amd.chfsi import KohnShamOperatorDevice

    prob = fe.build_workload("cube20_p4_n64")
    op = KohnShamOperatorDevice(prob)
    rng = np.random.default_rng(11)
    V = prob.V_q + 0.1 * rng.standard_normal(prob.V_q.shape)
    g = 0.05 * rng.standard_normal((prob.C, prob.nqc, 3)) * prob.JxW[None, :, None]
    return prob, op, V, g


def test_gga_cell_matrices_match_oracle(gga):
    prob, op, V, g = gga
    print(f"gga cell matrices C={prob.C} ne={prob.ne} nq={prob.nqc}", flush=True)
    op.set_potential_gga(V * prob.JxW[None, :], g)
    H = op.cell_matrices().cpu().numpy()
    want = O.cell_hamiltonians(prob, V_q=V, gJxW=g)
    assert _rel(H, want) < 1e-13
    assert np.abs(H - H.transpose(0, 2, 1)).max() < 1e-13 * np.abs(H).max()


def test_gga_hx_matches_oracle(gga):
    prob, op, V, g = gga
    op.set_potential_gga(V * prob.JxW[None, :], g)
    ref = O.OracleOperator(prob, H_e=O.cell_hamiltonians(prob, V_q=V, gJxW=g))
    X = prob.initial_guess(N=48, seed_offset=7)
    Xd = torch.tensor(X, device="cuda")
    Yd = torch.zeros_like(Xd)
    op.HX(Xd, False, 1.0, Yd)
    assert _rel(Yd.cpu().numpy(), ref.apply(X)) < 1e-13


def test_gga_zero_vector_equals_lda(gga):
    """gJxW = 0 adds only zero terms: the matrices equal the LDA build to
    rounding (the GGA build sums its quadrature terms in another grouping), and
    an LDA rebuild afterwards restores the LDA matrices bit for bit."""
    prob, op, V, g = gga
    vJxW = V * prob.JxW[None, :]
    op.set_potential(vJxW)
    H_lda = op.cell_matrices().cpu().numpy()
    op.set_potential_gga(vJxW, np.zeros_like(g))
    H0 = op.cell_matrices().cpu().numpy()
    assert _rel(H0, H_lda) < 1e-15
    op.set_potential_gga(vJxW, g)
    op.set_potential(vJxW)
    assert np.array_equal(op.cell_matrices().cpu().numpy(), H_lda)


def test_gga_rejects_bad_shape(gga):
    prob, op, V, g = gga
    with pytest.raises(ValueError):
        op.set_potential_gga(V * prob.JxW[None, :], g[:, :, :2])


def test_gga_bf16_copy_matches_separate_conversion(gga):
    """The build writes the BF16 cell matrices itself; they equal the separate
    FP64 -> BF16 conversion of an upload of the same FP64 matrices bit for bit
    (seen through the BF16 apply), and the FP32 copy likewise."""
    prob, op, V, g = gga
    op.set_potential_gga(V * prob.JxW[None, :], g)
    X = torch.tensor(prob.initial_guess(N=32, seed_offset=5), device="cuda", dtype=torch.float32)
    Y16, Y32 = torch.zeros_like(X), torch.zeros_like(X)
    op.HX(X, False, 1.0, Y16, bf16=True)
    op.HX(X, False, 1.0, Y32)
    H = op.cell_matrices()
    op.set_cell_matrices(H)
    Z16, Z32 = torch.zeros_like(X), torch.zeros_like(X)
    op.HX(X, False, 1.0, Z16, bf16=True)
    op.HX(X, False, 1.0, Z32)
    torch.cuda.synchronize()
    assert torch.equal(Y16, Z16)
    assert torch.equal(Y32, Z32)

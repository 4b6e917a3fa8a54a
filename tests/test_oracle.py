"""Pins of the CPU oracle (no GPU).

The reference ships no unit-level golden vectors for this path (its tests are
end-to-end DFT goldens that need deal.II, ELPA, libxc and UPF data), so every
oracle function is pinned here by mathematics: closed forms, brute force on
tiny meshes, independently assembled matrices and invariants.  Each check is
chosen so that a dropped term, a wrong index or a transposed operand fails.
"""
import math

import numpy as np
import pytest
import scipy.linalg as sla

from dftfe_amd import fe
from oracle import chfsi_oracle as O


@pytest.fixture(scope="module")
def brute():
    prob = fe.build_workload("brute")
    ref = O.OracleOperator(prob)
    H = ref.sparse_matrix().toarray()
    return prob, ref, H


# ---------------------------------------------------------------- 1-D rules
@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_gll_rule_exactness(p):
    x, w = fe.gll_rule(p)
    assert x[0] == 0.0 and x[-1] == 1.0 and np.all(np.diff(x) > 0)
    for k in range(2 * p):                       # exact to degree 2p-1
        assert abs((w * x ** k).sum() - 1.0 / (k + 1)) < 1e-14
    # not exact at degree 2p (Lobatto rules lose two degrees)
    assert abs((w * x ** (2 * p)).sum() - 1.0 / (2 * p + 1)) > 1e-8


@pytest.mark.parametrize("nq", [2, 5, 8])
def test_gauss_rule_exactness(nq):
    x, w = fe.gauss_rule(nq)
    for k in range(2 * nq):
        assert abs((w * x ** k).sum() - 1.0 / (k + 1)) < 1e-14


def test_lagrange_tables():
    x, _ = fe.gll_rule(4)
    phi, dphi = fe.lagrange_tables(x, x)
    assert np.allclose(phi, np.eye(5), atol=1e-13)          # cardinal property
    t = np.linspace(0, 1, 17)
    phi, dphi = fe.lagrange_tables(x, t)
    assert np.allclose(phi.sum(1), 1.0) and np.allclose(dphi.sum(1), 0.0, atol=1e-11)
    # derivative against a central difference
    h = 1e-6
    pp, _ = fe.lagrange_tables(x, t + h)
    pm, _ = fe.lagrange_tables(x, t - h)
    assert np.allclose((pp - pm) / (2 * h), dphi, atol=1e-6)


# ---------------------------------------------------------------- cell data
def test_stiffness_two_ways(brute):
    """Tensor-factored stiffness (fe) == point-by-point gradient quadrature (oracle)."""
    prob, ref, _ = brute
    G = O.shape_gradients(prob)
    K = np.einsum("q,qid,qjd->ij", prob.JxW, G, G)
    assert np.abs(K - prob.K_e).max() < 1e-12 * np.abs(K).max()
    assert np.abs(K.sum(1)).max() < 1e-12                    # constants in the kernel
    ev = np.linalg.eigvalsh(K)
    assert ev[0] > -1e-12 and np.sum(ev < 1e-10) == 1         # one zero mode


def test_cell_hamiltonian_constant_and_linear_potential(brute):
    """With V = v0: sum_ij H_e = v0 |cell|.  With V = x: sum_j H_e[i,j] = int x N_i,
    checked against the 1-D closed form of the tensor basis."""
    prob, _, _ = brute
    v0 = -1.7
    H = O.cell_hamiltonians(prob, V_q=np.full((prob.C, prob.nqc), v0), cells=[0, 5])
    assert abs(H[0].sum() - v0 * prob.h ** 3) < 1e-12
    xq = prob.quad_points([3])[0]
    H = O.cell_hamiltonians(prob, V_q=np.tile(xq[:, 0], (prob.C, 1)), cells=[3])[0]
    # int x N_i over the cell = prod of 1-D integrals; x = x0 + h*t
    pp = prob.wl.p + 1
    gx, gw = fe.gauss_rule(prob.wl.nq + 2)
    phi, _ = fe.lagrange_tables(prob.gll_x, gx)
    I0 = (phi * gw[:, None]).sum(0) * prob.h
    x0 = prob.cell_coords[3, 0] * prob.h
    I1 = (phi * (gw * (x0 + prob.h * gx))[:, None]).sum(0) * prob.h
    a, b, c = np.arange(prob.ne) % pp, (np.arange(prob.ne) // pp) % pp, np.arange(prob.ne) // pp ** 2
    want = I1[a] * I0[b] * I0[c]
    assert np.abs(H.sum(1) - want).max() < 1e-12


def test_product_shape_gradients_match_oracle(brute):
    """The product's vectorised gradient table == the oracle's point-by-point one."""
    prob, _, _ = brute
    assert np.abs(prob.Gq - O.shape_gradients(prob)).max() < 1e-13


def _assemble(prob, He):
    A = np.zeros((prob.M, prob.M))
    for c in range(prob.C):
        m = prob.cell_map[c]
        np.add.at(A, (m[:, None], m[None, :]), He[c])
    return A


def test_cell_hamiltonian_gga_term(brute):
    """GGA cell term 2 int g.(grad N_i N_j + N_i grad N_j), pinned without its formula:
    - it is symmetric and linear in g;
    - for constant g it is 2 int g.grad(N_i N_j), whose assembled sum over the
      periodic mesh vanishes (a wrong relative sign of the two halves does not);
    - its row sums are 2 g_d int dN_i/dx_d (sum_j N_j = 1), a boundary integral:
      the 1-D factor phi_a(1) - phi_a(0) = delta_{a,p} - delta_{a,0} times the
      integrals of the other two 1-D factors (pins the factor 2, the sign and
      the direction index d)."""
    prob, _, _ = brute
    C, nqc, pp = prob.C, prob.nqc, prob.wl.p + 1
    zero = np.zeros((C, nqc))
    H0 = O.cell_hamiltonians(prob, V_q=zero)
    rng = np.random.default_rng(3)
    g1 = rng.standard_normal((C, nqc, 3)) * prob.JxW[None, :, None]
    g2 = rng.standard_normal((C, nqc, 3)) * prob.JxW[None, :, None]
    T1 = O.cell_hamiltonians(prob, V_q=zero, gJxW=g1) - H0
    T2 = O.cell_hamiltonians(prob, V_q=zero, gJxW=g2) - H0
    T12 = O.cell_hamiltonians(prob, V_q=zero, gJxW=0.5 * g1 - 2.0 * g2) - H0
    scale = np.abs(T1).max()
    assert np.abs(T1 - T1.transpose(0, 2, 1)).max() < 1e-14 * scale
    assert np.abs(T12 - (0.5 * T1 - 2.0 * T2)).max() < 1e-13 * scale

    gvec = np.array([0.3, -0.7, 1.1])
    gc = np.broadcast_to(gvec[None, None, :] * prob.JxW[None, :, None], (C, nqc, 3))
    Tc = O.cell_hamiltonians(prob, V_q=zero, gJxW=gc) - H0
    assert np.abs(Tc).max() > 0.1
    assert np.abs(_assemble(prob, Tc)).max() < 1e-12 * np.abs(Tc).max()

    I0 = prob.gll_w * prob.h                  # int phi_b over [0, h] (GLL weights, exact)
    edge = np.zeros(pp)
    edge[-1], edge[0] = 1.0, -1.0              # phi_a(h) - phi_a(0)
    i = np.arange(prob.ne)
    a, b, c = i % pp, (i // pp) % pp, i // (pp * pp)
    want = 2.0 * (gvec[0] * edge[a] * I0[b] * I0[c] + gvec[1] * I0[a] * edge[b] * I0[c]
                  + gvec[2] * I0[a] * I0[b] * edge[c])
    for k in (0, C - 1):
        assert np.abs(Tc[k].sum(1) - want).max() < 1e-12


def test_mass_vector():
    for mode in ("unique", "dealii"):
        prob = fe.build_workload("brute", mode)
        assert abs(prob.mass.sum() - prob.wl.L ** 3) < 1e-10
        assert np.all(prob.mass[~prob.constrained] > 0)
        assert np.all(prob.sqrt_mass[prob.constrained] == 0)
        assert np.all(prob.inv_sqrt_mass[prob.constrained] == 0)


# ---------------------------------------------------------------- operator
def test_hx_equals_assembled_matrix(brute):
    """Cell-batched HX (gather, GEMM, projectors, scatter) == independently
    assembled sparse H~ (COO sum of cell blocks + C V C^T)."""
    prob, ref, H = brute
    X = prob.initial_guess(N=5)
    assert np.abs(ref.apply(X) - H @ X).max() < 1e-13 * np.abs(H @ X).max()
    assert np.abs(H - H.T).max() < 1e-13 * np.abs(H).max()


def test_hx_flags(brute):
    prob, ref, H = brute
    X = prob.initial_guess(N=3)
    D = prob.initial_guess(N=3, seed_offset=1)
    s = prob.inv_sqrt_mass[:, None]
    assert np.allclose(ref.HX(X, False, 1.0, D), s * D + H @ X, atol=1e-12)
    assert np.allclose(ref.HX(X, True, 0.3, D), D + 0.3 * (H @ X), atol=1e-12)


def test_dealii_layout_spectrum(brute):
    """Duplicated periodic DOFs with constraints: same spectrum plus one zero per slave."""
    prob, _, H = brute
    pd = fe.build_workload("brute", "dealii")
    Hd = O.OracleOperator(pd).sparse_matrix().toarray()
    eu = np.linalg.eigvalsh(H)
    ed = np.linalg.eigvalsh(Hd)
    nz = len(pd.slave_rows)
    # remove nz zeros from the dealii spectrum
    idx = np.argsort(np.abs(ed))[:nz]
    assert np.abs(ed[idx]).max() < 1e-10
    ed = np.sort(np.delete(ed, idx))
    assert np.abs(ed - eu).max() < 1e-10 * np.abs(eu).max()


def test_free_particle_spectrum():
    """V = 0, no projectors: lowest eigenvalue exactly 0 (eigenvector M^{1/2} 1),
    next shell six-fold degenerate near (1/2)(2 pi / L)^2."""
    prob = fe.build_workload("brute")
    prob.V_q[:] = 0.0
    prob.pair_cell = prob.pair_cell[:0]
    prob.pair_atom = prob.pair_atom[:0]
    prob.pair_C = prob.pair_C[:0]
    ref = O.OracleOperator(prob)
    H = ref.sparse_matrix().toarray()
    ev, vec = np.linalg.eigh(H)
    assert abs(ev[0]) < 1e-12
    v = prob.sqrt_mass / np.linalg.norm(prob.sqrt_mass)
    assert abs(abs(v @ vec[:, 0]) - 1) < 1e-12
    k1 = 0.5 * (2 * math.pi / prob.wl.L) ** 2
    assert np.ptp(ev[1:7]) < 1e-10 * k1
    assert abs(ev[1] - k1) < 1e-3 * k1
    assert ev[7] > ev[6] + 0.5 * k1


def test_nonlocal_term_psd(brute):
    prob, ref, H = brute
    p0 = fe.build_workload("brute")
    p0.pair_cell = p0.pair_cell[:0]
    p0.pair_atom = p0.pair_atom[:0]
    p0.pair_C = p0.pair_C[:0]
    Hl = O.OracleOperator(p0).sparse_matrix().toarray()
    ev = np.linalg.eigvalsh(H - Hl)
    assert ev[0] > -1e-12 and ev[-1] > 1e-3
    assert np.linalg.matrix_rank(H - Hl, tol=1e-10 * ev[-1]) <= prob.n_proj


# ---------------------------------------------------------------- filter
def test_chebyshev_filter_closed_form():
    """On a diagonal operator every component is multiplied by
    T_m((l-c)/e) / T_m((a0-c)/e)."""
    lam = np.linspace(-1.0, 50.0, 40)
    m, a, b, a0 = 20, 5.0, 41.0, -1.0

    def hx(src, flag, scalar, dst):
        return dst + scalar * lam[:, None] * src

    Y = O.chebyshev_filter(hx, np.eye(40), m, a, b, a0)
    want = O.chebyshev_scalar(lam, m, a, b, a0)
    assert np.allclose(np.diag(Y), want, rtol=1e-10, atol=1e-12)
    assert np.abs(Y - np.diag(np.diag(Y))).max() == 0.0
    # textbook T_20 check of the closed form itself
    t = np.array([0.3])
    assert abs(O.chebyshev_scalar(np.array([(a + b) / 2 + 0.3 * (b - a) / 2]), m, a, b, a0)[0]
               * O.chebyshev_scalar(np.array([a0]), m, a, b, a0)[0]
               - np.polynomial.chebyshev.chebval(t, [0] * 20 + [1])[0]
               / np.polynomial.chebyshev.chebval((a0 - (a + b) / 2) / ((b - a) / 2), [0] * 20 + [1])) < 1e-12


def _dense_operator(n=60, lo=-0.07, hi=315.0, seed=5):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.concatenate([np.linspace(lo, 0.93, 20), np.geomspace(1.2, hi, n - 20)])
    A = (Q * lam) @ Q.T
    A = 0.5 * (A + A.T)

    def hx(src, flag, scalar, dst):
        return dst + scalar * (A @ src)
    return A, lam, Q, hx


def test_chebyshev_filter_defect_equals_plain():
    """The defect form X D_m + E_m is the same polynomial filter: in FP64 it
    agrees with chebyshev_filter to rounding, for a random block and for the
    eigenbasis (where E = 0 and D_m is the closed form)."""
    A, lam, Q, hx = _dense_operator()
    m, a, b, a0 = 30, 0.93, 315.0, -0.07
    X = np.random.default_rng(1).standard_normal((60, 8))
    Yp = O.chebyshev_filter(hx, X, m, a, b, a0)
    Yd = O.chebyshev_filter_defect(hx, X, m, a, b, a0)
    assert np.abs(Yd - Yp).max() / np.abs(Yp).max() < 1e-11
    Yd = O.chebyshev_filter_defect(hx, Q[:, :8], m, a, b, a0)
    want = Q[:, :8] * O.chebyshev_scalar(lam[:8], m, a, b, a0)[None, :]
    assert np.abs(Yd - want).max() < 1e-10


def test_chebyshev_filter_defect_fp32_rounding_follows_residual():
    """Near converged vectors the FP32 defect form is far more accurate than
    the FP32 plain recurrence (same operator, same rounding of every stored
    iterate): its rounding scales with |R|, the plain one with |X|."""
    A, lam, Q, hx = _dense_operator()
    m, a, b, a0 = 30, 0.93, 315.0, -0.07
    rng = np.random.default_rng(2)
    X = Q[:, :16] + 1e-6 * rng.standard_normal((60, 16))
    Y64 = O.chebyshev_filter(hx, X, m, a, b, a0)

    def plain32(X):
        f = np.float32
        e, c = (b - a) / 2, (b + a) / 2
        sigma = e / (a0 - c)
        gamma = 2 / sigma
        Xa = X.astype(f)
        Y = (f(sigma / e) * (hx(Xa.astype(float), False, 1.0, np.zeros(X.shape)).astype(f)
                             + f(-c) * Xa)).astype(f)
        for _ in range(2, m + 1):
            s2 = 1 / (gamma - sigma)
            a1, a2 = 2 * s2 / e, -(sigma * s2)
            Xa = (f(a2) * Xa + f(-c * a1) * Y).astype(f)
            Xa = hx(Y.astype(float), True, a1, Xa.astype(float)).astype(f)
            Xa, Y = Y, Xa
            sigma = s2
        return Y.astype(float)

    err_plain = np.abs(plain32(X) - Y64).max() / np.abs(Y64).max()
    err_def = np.abs(O.chebyshev_filter_defect(hx, X, m, a, b, a0, dtype=np.float32) - Y64).max() \
        / np.abs(Y64).max()
    assert err_plain > 1e-8
    assert err_def < err_plain / 100, (err_def, err_plain)


# ---------------------------------------------------------------- dense steps
def test_cholgs_orthonormal_and_span(brute):
    prob, _, _ = brute
    X = prob.initial_guess(N=12)
    Q = O.cholgs(X)
    assert np.abs(Q.T @ Q - np.eye(12)).max() < 1e-12
    P = Q @ Q.T
    assert np.abs(P @ X - X).max() < 1e-10 * np.abs(X).max()
    # lower-triangular relation: X = Q L^T
    Lt = Q.T @ X
    assert np.abs(np.tril(Lt, -1)).max() < 1e-10


def test_rayleigh_ritz_routes(brute):
    prob, ref, H = brute
    X = O.cholgs(prob.sqrt_mass[:, None] * prob.initial_guess(N=12))
    lam, Xr = O.rayleigh_ritz(ref, X, 5)
    assert np.allclose(lam, np.linalg.eigvalsh(X.T @ H @ X), atol=1e-12)
    R = Xr.T @ H @ Xr
    assert np.abs(R - np.diag(lam)).max() < 1e-11
    lam2, _ = O.rayleigh_ritz_gep(ref, prob.sqrt_mass[:, None] * prob.initial_guess(N=12), 5)
    assert np.abs(lam2 - lam).max() < 1e-10
    r = O.residual_norms(ref, Xr, lam, 5)
    assert np.allclose(r, np.linalg.norm(H @ Xr - Xr * lam, axis=0), atol=1e-12)


def test_rayleigh_ritz_spectrum_split(brute):
    """Spectrum-split Rayleigh-Ritz on a non-orthonormal block: the kept
    eigenvalues are the top N - n_core eigenvalues of the generalised problem
    (X^T H X, X^T X) solved by LAPACK's sygvd; Xfrac is orthonormal and
    diagonalises H with those values; it spans the top columns of the full
    Rayleigh-Ritz; and X comes back orthonormalised but not rotated."""
    prob, ref, H = brute
    X = prob.sqrt_mass[:, None] * prob.initial_guess(N=12, seed_offset=7)
    n_core = 7
    lam, Xo, Xf = O.rayleigh_ritz_spectrum_split(ref, X, n_core, 5)
    gen = sla.eigh(X.T @ H @ X, X.T @ X, eigvals_only=True)
    assert lam.shape == (5,)
    assert np.abs(lam - gen[n_core:]).max() < 1e-10 * np.abs(gen).max()
    assert np.abs(Xf.T @ Xf - np.eye(5)).max() < 1e-12
    assert np.abs(Xf.T @ H @ Xf - np.diag(lam)).max() < 1e-10
    lam_full, Xr = O.rayleigh_ritz(ref, O.cholgs(X), 5)
    P = Xr[:, n_core:] @ Xr[:, n_core:].T
    assert np.abs(P @ Xf - Xf).max() < 1e-10
    assert np.abs(Xo - O.cholgs(X)).max() < 1e-12
    # a wrong split index (keeping the bottom instead of the top) fails
    assert np.abs(lam_full[:5] - lam).max() > 1e-3


def test_chfsi_pass_spectrum_split_matches_full_pass(brute):
    """The split pass filters exactly as the full pass; its eigenvalues and
    residual norms are the top N - n_core of the full pass's, its X is the full
    pass's filtered block after CholGS only, and Xfrac spans the top Ritz vectors."""
    prob, ref, H = brute
    N, n_core = 12, 8
    lo, up, _, _ = O.lanczos_bounds(ref, prob.lanczos_start, 10)
    a0, bLow, bUp = O.first_call_bounds(lo, up, N, prob.M, reproducible=True)
    X = prob.initial_guess(N=N, seed_offset=3)
    lam, res, Xf_full = O.chfsi_pass(ref, X, a0, bLow, bUp, 6, 4)
    lam_s, res_s, Xs, Xfrac = O.chfsi_pass_spectrum_split(ref, X, a0, bLow, bUp, 6, 4, n_core)
    assert np.abs(lam_s - lam[n_core:]).max() < 1e-10
    assert np.abs(res_s - res[n_core:]).max() < 1e-9
    # physical basis: M-orthonormal columns, same span as the full pass's top vectors
    Mm = prob.mass
    assert np.abs(Xs.T @ (Mm[:, None] * Xs) - np.eye(N)).max() < 1e-10
    P = Xf_full[:, n_core:] @ (Xf_full[:, n_core:].T * Mm[None, :])
    assert np.abs(P @ Xfrac - Xfrac).max() < 1e-8 * np.abs(Xfrac).max()


def test_lanczos_bounds_bracket(brute):
    prob, ref, H = brute
    ev = np.linalg.eigvalsh(H)
    lo, up, theta, fnorm = O.lanczos_bounds(ref, prob.lanczos_start, 10)
    assert up >= ev[-1]
    assert theta[0] >= ev[0] - 1e-10 and theta[-1] <= ev[-1] + 1e-10
    assert lo == math.floor(theta[0])


def test_chfsi_converges_to_exact(brute):
    """Repeated passes with the reference's bound bookkeeping converge to the
    lowest N eigenvalues of the assembled operator."""
    prob, ref, H = brute
    N = 16
    exact = np.linalg.eigvalsh(H)[:N]
    lo, up, _, _ = O.lanczos_bounds(ref, prob.lanczos_start, 10)
    a0, bLow, bUp = O.first_call_bounds(lo, up, N, prob.M, reproducible=True)
    X = prob.initial_guess(N=N + 4)
    lam = None
    for _ in range(12):
        lam, res, X = O.chfsi_pass(ref, X, a0, bLow, bUp, 12, 8)
        a0, bLow = O.next_bounds(lam)
    assert np.abs(lam[:N] - exact).max() < 1e-8


# ---------------------------------------------------------------- derivatives
def test_hx_adjoint_v_finite_differences(brute):
    """dL/dV of L = <Ybar, H~(V) X> equals central differences (H is linear in V,
    so the difference quotient is exact up to round-off)."""
    prob, ref, _ = brute
    X = prob.initial_guess(N=3, seed_offset=40)
    Yb = prob.initial_guess(N=3, seed_offset=41)
    g = O.hx_adjoint_v(ref, Yb, X, 0.7)
    rng = np.random.default_rng(5)
    for _ in range(4):
        c, q = rng.integers(prob.C), rng.integers(prob.nqc)
        h = 1e-3
        Vp, Vm = prob.V_q.copy(), prob.V_q.copy()
        Vp[c, q] += h
        Vm[c, q] -= h
        Lp = (Yb * (0.7 * O.with_potential(prob, Vp).apply(X))).sum()
        Lm = (Yb * (0.7 * O.with_potential(prob, Vm).apply(X))).sum()
        assert abs((Lp - Lm) / (2 * h) - g[c, q]) < 1e-9 * max(1.0, abs(g[c, q]))


def test_filter_backward_finite_differences(brute):
    prob, ref, _ = brute
    m, a, b, a0 = 6, 8.0, 60.0, -2.0
    X0 = prob.sqrt_mass[:, None] * prob.initial_guess(N=2, seed_offset=42)
    Yb = prob.initial_guess(N=2, seed_offset=43)

    def L(op, X):
        return (Yb * O.chebyshev_filter(lambda s, f, sc, d: op.HX(s, f, sc, d), X, m, a, b, a0)).sum()

    Xbar, g = O.filter_backward(ref, X0, Yb, m, a, b, a0)
    rng = np.random.default_rng(6)
    for _ in range(3):
        c, q = rng.integers(prob.C), rng.integers(prob.nqc)
        h = 1e-4
        Vp, Vm = prob.V_q.copy(), prob.V_q.copy()
        Vp[c, q] += h
        Vm[c, q] -= h
        fd = (L(O.with_potential(prob, Vp), X0) - L(O.with_potential(prob, Vm), X0)) / (2 * h)
        assert abs(fd - g[c, q]) < 1e-6 * max(1.0, np.abs(g).max())
    D = prob.initial_guess(N=2, seed_offset=44)
    h = 1e-6
    fd = (L(ref, X0 + h * D) - L(ref, X0 - h * D)) / (2 * h)
    assert abs(fd - (Xbar * D).sum()) < 1e-6 * abs(fd)


def test_density_integrates_to_occupation(brute):
    """With M-orthonormal FE-basis vectors, int rho = sum_i w_i up to the
    lumped-vs-Gauss quadrature difference; exact for the constant vector."""
    prob, ref, _ = brute
    one = np.ones((prob.M, 1))
    rho = O.density_at_quadrature(prob, one, np.array([1.0]), spin_factor=1.0)
    assert abs((rho * prob.JxW[None, :]).sum() - prob.wl.L ** 3) < 1e-10


# ----------------------------------------------------------------------------
# Fermi-Dirac occupations (utils/dftUtils.cc:56-64)
# ----------------------------------------------------------------------------


def test_fermi_dirac_closed_forms():
    kT, mu = 0.01, -0.3
    x = np.linspace(-0.5, 0.5, 2001)
    f_plus, f_minus = O.fermi_dirac(mu + x, mu, kT), O.fermi_dirac(mu - x, mu, kT)
    # f(mu) = 1/2 exactly, particle-hole symmetry f(mu + x) + f(mu - x) = 1
    assert O.fermi_dirac(np.array([mu]), mu, kT)[0] == 0.5
    assert np.abs(f_plus + f_minus - 1.0).max() <= 2.3e-16
    # limits without overflow; monotone decreasing
    far = O.fermi_dirac(np.array([mu - 1e3, mu + 1e3, mu - 1e6 * kT, mu + 1e6 * kT]), mu, kT)
    assert np.all(np.isfinite(far))
    assert far[0] == 1.0 and far[2] == 1.0 and far[1] < 1e-300 and far[3] == 0.0
    assert np.all(np.diff(f_plus) <= 0)
    # the textbook form 1 / (1 + e^t) where it does not overflow (|t| <= 30)
    t = np.linspace(-30, 30, 601)
    ref = 1.0 / (1.0 + np.exp(t))
    got = O.fermi_dirac(mu + t * kT, mu, kT)
    assert np.abs(got - ref).max() < 1e-15
    assert np.abs(got / ref - 1.0).max() < 1e-13
    # slope at the chemical potential: -1 / (4 kT)
    h = 1e-6
    slope = (O.fermi_dirac(np.array([mu + h]), mu, kT) - O.fermi_dirac(np.array([mu - h]), mu, kT))
    assert slope[0] / (2 * h) == pytest.approx(-1.0 / (4 * kT), rel=1e-8)
    # the reference's branch form at its switch point and either side
    for fac in (-3.0, -1e-9, 0.0, 1e-9, 3.0):
        want = np.exp(-fac) / (1 + np.exp(-fac)) if fac >= 0 else 1.0 / (1 + np.exp(fac))
        assert O.fermi_dirac(np.array([mu + fac * kT]), mu, kT)[0] == pytest.approx(want,
                                                                                   rel=1e-15)


def test_fermi_occupations_matches_oracle():
    """The product's bisection on the chemical potential (dftfe_amd.autograd,
    CPU torch): occupations equal the oracle's Fermi-Dirac at the returned mu,
    and they sum to the requested count."""
    torch = pytest.importorskip("torch")
    from dftfe_amd.autograd import fermi_occupations

    rng = np.random.default_rng(5)
    eps = np.sort(rng.uniform(-1.0, 2.0, 256))
    for n_occ, kT in ((204.8, 0.01), (100.0, 0.001), (10.5, 0.05)):
        occ, mu = fermi_occupations(torch.tensor(eps), n_occ, kT)
        occ = occ.numpy()
        assert np.abs(occ - O.fermi_dirac(eps, mu, kT)).max() < 1e-15
        assert occ.sum() == pytest.approx(n_occ, abs=1e-9)
        assert np.all((occ >= 0) & (occ <= 1))

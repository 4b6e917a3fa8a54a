"""ORACLE -- CPU FP64 restatement of the reference ChFSI pass.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module, and only as the checker.
The product path (dftfe_amd) never imports it and has no CPU fallback.

Every function cites the reference file:line whose behaviour it restates.
The reference itself cannot be built here (deal.II, p4est, ELPA, libxc are
absent), so each function is pinned by the mathematics instead (closed forms,
brute force, invariants); see tests/test_oracle.py.

Layout: node-major multivectors, X[node, vec] (numpy row-major == the
reference's X[node*N + vec]).
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

# ----------------------------------------------------------------------------
# Cell Hamiltonians
# ----------------------------------------------------------------------------


def shape_gradients(prob) -> np.ndarray:
    """3-D shape gradients at Gauss points, (nqc, ne, 3), physical units.

    Built point by point from the 1-D tables (not from the tensor-factored
    stiffness the product uses) so that the stiffness below is an independent
    computation.  Reference: utils/FEBasisOperations.cc:1135-1330 (stiffness
    from shape gradients and JxW)."""
    nq, pp = prob.wl.nq, prob.wl.p + 1
    phi, dphi, h = prob.phi, prob.dphi, prob.h
    G = np.zeros((prob.nqc, prob.ne, 3))
    for i in range(prob.ne):
        a, b, c = i % pp, (i // pp) % pp, i // (pp * pp)
        for q in range(prob.nqc):
            qx, qy, qz = q % nq, (q // nq) % nq, q // (nq * nq)
            G[q, i, 0] = dphi[qx, a] * phi[qy, b] * phi[qz, c] / h
            G[q, i, 1] = phi[qx, a] * dphi[qy, b] * phi[qz, c] / h
            G[q, i, 2] = phi[qx, a] * phi[qy, b] * dphi[qz, c] / h
    return G


def cell_hamiltonians(prob, V_q: Optional[np.ndarray] = None, cells=None,
                      gJxW: Optional[np.ndarray] = None) -> np.ndarray:
    """H_e[i,j] = 1/2 int grad N_i . grad N_j + sum_q vJxW_q N_i(q) N_j(q).

    Reference: src/dftOperator/hamiltonianMatrixCalculatorFlattenedDevice.cc:61-116
    (hamMatrixKernelLDA, real case); CPU hamiltonianMatrixCalculator.cc:21-985.

    gJxW (C, nqc, 3), the GGA vector d(E_xc)/d(sigma) grad rho JxW, adds
    2 sum_q sum_d gJxW[q, d] (dN_i/dx_d N_j + N_i dN_j/dx_d) at each point,
    term by term as hamMatrixKernelGGAMemOpt
    (hamiltonianMatrixCalculatorFlattenedDevice.cc:283-456, val at :429-441).
    """
    V_q = prob.V_q if V_q is None else V_q
    cells = np.arange(prob.C) if cells is None else np.asarray(cells)
    G = shape_gradients(prob)
    K = np.einsum("q,qid,qjd->ij", prob.JxW, G, G)
    # graded meshes: a cell of edge a = s h has JxW s^3 JxW_h and gradients
    # grad_h / s, so its stiffness integral is s K_h
    # (per-cell cellShapeFunctionGradientIntegral, hamiltonianMatrixCalculatorFlattenedDevice.cc:110-111;
    #  inverse Jacobian of Cartesian cells, :408-420)
    scale = prob.cell_scale if getattr(prob, "cell_size", None) is not None else np.ones(prob.C)
    vJxW = V_q[cells] * prob.JxW[None, :] * (scale[cells] ** 3)[:, None]
    Nt = prob.Nq.T
    shift = 0.0
    if prob.kpoint is not None:
        if getattr(prob, "cell_size", None) is not None:
            raise ValueError("Bloch cell terms are built for uniform meshes only")
        shift = bloch_cell_terms(prob, G)
    H = np.empty((len(cells), prob.ne, prob.ne), dtype=np.result_type(np.float64, shift))
    for k in range(len(cells)):
        sc = scale[cells[k]]
        H[k] = 0.5 * sc * K + (Nt * vJxW[k][None, :]) @ prob.Nq + shift
        if gJxW is not None:
            U = np.einsum("qd,qid->qi", gJxW[cells[k]], G / sc)  # g . grad N_i at q
            H[k] += 2.0 * (U.T @ prob.Nq + prob.Nq.T @ U)
    return H


def bloch_cell_terms(prob, G: Optional[np.ndarray] = None) -> np.ndarray:
    """k-dependent part of the cell Hamiltonian acting on the periodic part u of
    psi = exp(i k.x) u:  1/2 |k|^2 int N_i N_j  -  i int N_i (k . grad N_j).

    Reference: src/dftOperator/hamiltonianMatrixCalculator.cc:562-571 and
    :790-797 (real part, kSquareTimesHalf * NiNj), :577-735 and :851-872
    (imaginary part from d_invJacKPointTimesJxW, shapeFunctionDataCalculator.cc:190-224),
    stored [iNode][jNode] and applied transposed by the cell GEMM
    (matrixVectorProductImplementations.cc:66-78).  Quadrature sums point by
    point from the 3-D shape gradients."""
    G = shape_gradients(prob) if G is None else G
    k = prob.kpoint
    Mq = (prob.Nq * prob.JxW[:, None]).T @ prob.Nq                  # int N_i N_j
    kG = G @ k                                                      # (nqc, ne) k . grad N_j
    Dq = (prob.Nq * prob.JxW[:, None]).T @ kG                       # int N_i k.grad N_j
    return 0.5 * float(k @ k) * Mq - 1j * Dq


# ----------------------------------------------------------------------------
# Operator
# ----------------------------------------------------------------------------


class OracleOperator:
    """FP64 H~ = M^{-1/2} H M^{-1/2} on node-major blocks.

    HX follows src/dftOperator/kohnShamDFTOperator.cc:950-1044 (CPU) /
    kohnShamDFTOperatorDevice.cc:3765-3860 (device): scale src by M^{-1/2},
    distribute constraints, per-cell dense GEMM (matrixVectorProductImplementations.cc:96-169),
    nonlocal C V C^T (computeNonLocalHamiltonianTimesXMemoryOpt.cc:269-503),
    distribute_slave_to_master, scale dst by M^{-1/2}.
    """

    def __init__(self, prob, H_e: Optional[np.ndarray] = None, cell_chunk: int = 256):
        self.prob = prob
        self.H_e = cell_hamiltonians(prob) if H_e is None else H_e
        self.cell_chunk = cell_chunk
        C, ne, M = prob.C, prob.ne, prob.M
        rows = np.arange(C * ne)
        # gather matrix A: (C*ne) x M with A[c*ne+i, map[c,i]] = 1
        self.A = sp.csr_matrix((np.ones(C * ne), (rows, prob.cell_map.ravel())), shape=(C * ne, M))
        for i in range(ne):
            if len(np.unique(prob.cell_map[:, i])) != C:
                raise ValueError("local node index maps two cells to one DOF")
        self.s = prob.inv_sqrt_mass
        self.sq = prob.sqrt_mass
        # nonlocal: pair p -> projector ids 4*atom + alpha
        self.pair_proj = (4 * prob.pair_atom[:, None] + np.arange(4)[None, :])
        # projector columns grouped per cell: Cc[c, i, s], slot s -> projector
        self.Cc = None
        if len(prob.pair_cell):
            cnt = np.bincount(prob.pair_cell, minlength=C)
            S = 4 * int(cnt.max())
            self.Cc = np.zeros((C, ne, S), dtype=prob.pair_C.dtype)
            self.slot_proj = np.full((C, S), prob.n_proj, dtype=np.int64)
            fill = np.zeros(C, dtype=np.int64)
            for q, c in enumerate(prob.pair_cell):
                s0 = 4 * fill[c]
                self.Cc[c, :, s0:s0 + 4] = prob.pair_C[q]
                self.slot_proj[c, s0:s0 + 4] = self.pair_proj[q]
                fill[c] += 1
            valid = self.slot_proj.ravel() < prob.n_proj
            cols = np.nonzero(valid)[0]
            self.R = sp.csr_matrix((np.ones(len(cols)), (self.slot_proj.ravel()[valid], cols)),
                                   shape=(prob.n_proj, C * S))

    # -- constraints (utils/constraintMatrixInfo.cc; device
    #    utils/constraintMatrixInfoDevice.cc:544-588 distribute, :596-816 slave->master,
    #    :818-855 set_zero).  Affine rows x_s = sum_j c_sj x_j (periodic: one
    #    master, coefficient 1; hanging nodes: the coarse face or edge nodes
    #    with the coarse basis values as coefficients).
    def _constraint_matrix(self):
        """Sparse (S x M) matrix of the constraint coefficients, row r for slave_rows[r]."""
        if not hasattr(self, "_Pc"):
            p = self.prob
            rows, ptr, cols, vals = p.constraint_csr()
            r = np.repeat(np.arange(len(rows)), np.diff(ptr))
            self._Pc = sp.csr_matrix((vals, (r, cols)), shape=(len(rows), p.M))
        return self._Pc

    def distribute(self, X):
        p = self.prob
        if len(p.slave_rows):
            X[p.slave_rows] = self._constraint_matrix() @ X

    def distribute_slave_to_master(self, Y):
        p = self.prob
        if len(p.slave_rows):
            Ys = Y[p.slave_rows].copy()
            Y[p.slave_rows] = 0.0
            Y += self._constraint_matrix().T @ Ys

    def set_zero(self, X):
        p = self.prob
        if len(p.slave_rows):
            X[p.slave_rows] = 0.0

    def h_fe(self, Xs: np.ndarray) -> np.ndarray:
        """Unscaled FE H (local + nonlocal) on an already-distributed block."""
        p = self.prob
        B = Xs.shape[1]
        dt = np.result_type(Xs.dtype, self.H_e.dtype)
        Xc = Xs[p.cell_map].astype(dt, copy=False)                   # (C, ne, B) gather
        Yc = np.empty_like(Xc)
        for c0 in range(0, p.C, self.cell_chunk):
            c1 = min(p.C, c0 + self.cell_chunk)
            Yc[c0:c1] = np.matmul(self.H_e[c0:c1], Xc[c0:c1])
        if self.Cc is not None:
            # P_alpha = sum over the atom's cells of C_e[:, alpha]^H X_e, then
            # Y_e += C_e V P for every cell in the support
            # (complex: computeNonLocalHamiltonianTimesXMemoryOpt.cc:25-264).
            S = self.Cc.shape[2]
            CcH = self.Cc.transpose(0, 2, 1)
            if np.iscomplexobj(CcH):
                CcH = CcH.conj()
            Pc = np.matmul(CcH, Xc)                                  # (C, S, B)
            W = p.proj_V[:, None] * (self.R @ Pc.reshape(p.C * S, B))
            W = np.vstack([W, np.zeros((1, B), dtype=W.dtype)])       # row for empty slots
            Yc += np.matmul(self.Cc, W[self.slot_proj])               # (C, ne, B)
        # scatter-add (matrixVectorProductImplementations.cc:127-168 adds each
        # cell's rows into the global vector); looping over the local node index
        # keeps each vectorised add free of repeated DOFs.
        Y = np.zeros((p.M, B), dtype=Yc.dtype)
        for i in range(p.ne):
            Y[p.cell_map[:, i]] += Yc[:, i, :]
        return Y

    def HX(self, src: np.ndarray, scale_flag: bool, scalar: float, dst: np.ndarray) -> np.ndarray:
        """dst <- (scale_flag ? dst : M^{-1/2} dst) + scalar * H~ src
        (kohnShamDFTOperatorDevice.cc:3790-3859; returns the new dst)."""
        Xs = (scalar * self.s)[:, None] * src
        self.distribute(Xs)
        base = dst * self.sq[:, None] if scale_flag else dst.copy()
        Y = base + self.h_fe(Xs)
        self.distribute_slave_to_master(Y)
        return self.s[:, None] * Y

    def apply(self, X: np.ndarray) -> np.ndarray:
        """H~ X."""
        Xs = self.s[:, None] * X
        self.distribute(Xs)
        Y = self.h_fe(Xs)
        self.distribute_slave_to_master(Y)
        return self.s[:, None] * Y

    def sparse_matrix(self) -> sp.csr_matrix:
        """Assembled H~ as a sparse matrix (for brute-force checks)."""
        p = self.prob
        C, ne = p.C, p.ne
        r = np.repeat(p.cell_map, ne, axis=1).ravel()
        c = np.tile(p.cell_map, (1, ne)).ravel()
        H = sp.coo_matrix((self.H_e.reshape(C, ne * ne).ravel(), (r, c)), shape=(p.M, p.M)).tocsr()
        if len(p.pair_cell):
            # assembled projector vectors c_alpha (M x n_proj)
            rows = np.repeat(p.cell_map[p.pair_cell], 4, axis=1).ravel()
            cols = np.repeat(self.pair_proj[:, None, :], ne, axis=1).ravel()
            Cg = sp.coo_matrix((p.pair_C.ravel(), (rows, cols)), shape=(p.M, p.n_proj)).tocsr()
            H = H + Cg @ sp.diags(p.proj_V) @ Cg.conj().T
        if len(p.slave_rows):
            # P^T H P with P the prolongation: identity on unconstrained rows,
            # the constraint coefficients on slave rows (assembled from the
            # CSR arrays directly, not through distribute)
            keep = np.nonzero(~p.constrained)[0]
            rows, ptr, cols, vals = p.constraint_csr()
            pr = np.concatenate([keep, np.repeat(rows, np.diff(ptr))])
            pc = np.concatenate([keep, cols])
            pv = np.concatenate([np.ones(len(keep)), vals])
            P = sp.csr_matrix((pv, (pr, pc)), shape=(p.M, p.M))
            H = (P.T @ H @ P).tocsr()
            mask = np.zeros(p.M)
            mask[keep] = 1.0
            H = sp.diags(mask) @ H @ sp.diags(mask)
        S = sp.diags(self.s)
        return (S @ H @ S).tocsr()


# ----------------------------------------------------------------------------
# Lanczos bounds
# ----------------------------------------------------------------------------


def lanczos_bounds(op, v0: np.ndarray, steps: int, reproducible: bool = False):
    """k-step Lanczos on H~: returns (floor(theta_min), ceil(theta_max + |f|/10)),
    or + |f| when reproducible.  Reference: src/linAlg/linearAlgebraOperationsOpt.cc:2891-3023
    (device twin linearAlgebraOperationsDevice.cc:340-527)."""
    v = v0.copy()
    op.set_zero(v[:, None]) if v.ndim == 1 else None
    v = v / np.linalg.norm(v)
    f = op.apply(v[:, None])[:, 0]
    alpha = np.vdot(v, f).real
    f = f - alpha * v
    T = np.zeros((steps, steps))
    T[0, 0] = alpha
    for j in range(1, steps):
        beta = np.linalg.norm(f)
        v_old = v
        v = f / beta
        f = op.apply(v[:, None])[:, 0]
        f = f - beta * v_old
        alpha = np.vdot(v, f).real
        f = f - alpha * v
        T[j, j - 1] = beta
        T[j - 1, j] = beta
        T[j, j] = alpha
    theta = np.linalg.eigvalsh(T)
    fnorm = np.linalg.norm(f)
    lower = math.floor(theta[0])
    upper = math.ceil(theta[-1] + (fnorm if reproducible else fnorm / 10.0))
    return float(lower), float(upper), theta, fnorm


def first_call_bounds(lower: float, upper: float, N: int, M_global: int, reproducible: bool = False):
    """bLow = a0 + (bUp - a0) * N / M * (10 or 200).  Reference:
    src/solvers/eigenSolvers/chebyshevOrthogonalizedSubspaceIterationSolverDevice.cc:264-271."""
    a0, bUp = lower, upper
    bLow = a0 + (bUp - a0) * N / M_global * (10.0 if reproducible else 200.0)
    return a0, bLow, bUp


# ----------------------------------------------------------------------------
# Chebyshev filter
# ----------------------------------------------------------------------------


def chebyshev_filter(apply_hx, X: np.ndarray, m: int, a: float, b: float, a0: float) -> np.ndarray:
    """Scaled three-term Chebyshev recurrence of degree m on [a, b], a0-scaled.

    Reference: src/linAlg/linearAlgebraOperationsOpt.cc:276-352.  `apply_hx(src,
    scale_flag, scalar, dst)` has the HX(src, scaleFlag, scalar, dst) semantics."""
    e = (b - a) / 2.0
    c = (b + a) / 2.0
    sigma = e / (a0 - c)
    sigma1 = sigma
    gamma = 2.0 / sigma1
    Xa = X.copy()
    Y = apply_hx(Xa, False, 1.0, np.zeros_like(Xa))
    alpha1, alpha2 = sigma1 / e, -c
    Y = alpha1 * (Y + alpha2 * Xa)
    for _degree in range(2, m + 1):
        sigma2 = 1.0 / (gamma - sigma)
        alpha1, alpha2 = 2.0 * sigma2 / e, -(sigma * sigma2)
        Xa = alpha2 * Xa + (-c * alpha1) * Y
        Xa = apply_hx(Y, True, alpha1, Xa)
        Xa, Y = Y, Xa
        sigma = sigma2
    return Y


def chebyshev_filter_defect(apply_hx, X: np.ndarray, m: int, a: float, b: float, a0: float,
                            dtype=np.float64) -> np.ndarray:
    """The filter of chebyshev_filter written as Y_k = X diag(D_k) + E_k (no
    reference counterpart: the device's defect form of the low-precision
    filter).  theta are the columns' Rayleigh quotients and R = H~X - X theta
    (both FP64); D_k = p_k(theta) follows the scalar recurrence and E_k the same
    three-term recurrence in `dtype` with the source term alpha1 D_k R:
        D_1 = (sigma/e)(theta - c),  E_1 = (sigma/e) R,
        D_k+1 = alpha1 (theta - c) D_k + alpha2 D_k-1,
        E_k+1 = alpha1 (H~ - c) E_k + alpha2 E_k-1 + alpha1 D_k R.
    Equal to chebyshev_filter in exact arithmetic; with dtype float32 the
    rounding of E scales with |R|, not with |X|."""
    e = (b - a) / 2.0
    c = (b + a) / 2.0
    sigma = e / (a0 - c)
    gamma = 2.0 / sigma
    HX = apply_hx(X, False, 1.0, np.zeros_like(X))
    theta = (X * HX).sum(axis=0) / (X * X).sum(axis=0)
    R = HX - X * theta[None, :]
    Rd = R.astype(dtype)
    D_prev, D = np.ones(X.shape[1]), (sigma / e) * (theta - c)
    E_prev = np.zeros(X.shape, dtype=dtype)
    E = (dtype(sigma / e) * Rd).astype(dtype)

    def hx_low(src, scalar, dst):
        # H~ applied in FP64 to the stored values, result rounded to dtype
        return apply_hx(src.astype(np.float64), True, scalar, dst.astype(np.float64)).astype(dtype)

    for _degree in range(2, m + 1):
        sigma2 = 1.0 / (gamma - sigma)
        alpha1, alpha2 = 2.0 * sigma2 / e, -(sigma * sigma2)
        En = (dtype(alpha2) * E_prev + dtype(-c * alpha1) * E
              + (alpha1 * D).astype(dtype)[None, :] * Rd).astype(dtype)
        En = hx_low(E, alpha1, En)
        D_prev, D = D, alpha1 * (theta - c) * D + alpha2 * D_prev
        E_prev, E = E, En
        sigma = sigma2
    return X * D[None, :] + E.astype(np.float64)


def chebyshev_scalar(lam: np.ndarray, m: int, a: float, b: float, a0: float) -> np.ndarray:
    """Closed form of the filter on an eigen-component: T_m((l-c)/e) / T_m((a0-c)/e)."""
    e = (b - a) / 2.0
    c = (b + a) / 2.0

    def T(m, t):
        t = np.asarray(t, dtype=np.float64)
        out = np.empty_like(t)
        inside = np.abs(t) <= 1
        out[inside] = np.cos(m * np.arccos(t[inside]))
        ti = t[~inside]
        out[~inside] = np.sign(ti) ** m * np.cosh(m * np.arccosh(np.abs(ti)))
        return out

    return T(m, (lam - c) / e) / T(m, np.array([(a0 - c) / e]))[0]


# ----------------------------------------------------------------------------
# CholGS, Rayleigh-Ritz, residuals
# ----------------------------------------------------------------------------


def cholgs(X: np.ndarray) -> np.ndarray:
    """S = X^T X, S = L L^T, X <- X L^{-T}.  Reference: src/linAlg/pseudoGS.cc:29-262;
    device src/linAlg/pseudoGSDevice.cc:121-368."""
    XH = X.conj().T
    S = XH @ X
    L = np.linalg.cholesky(S)
    return sla.solve_triangular(L, XH, lower=True).conj().T


def rayleigh_ritz(op, X: np.ndarray, block: int):
    """Hp = X^T H~ X (blocked HX), symmetrised; Hp = Q diag(l) Q^T; X <- X Q.
    Reference: src/linAlg/rayleighRitzDevice.cc:80-353 (SHEP) and the XtHX
    blocks of kohnShamDFTOperatorDevice.cc:4000-4157."""
    N = X.shape[1]
    XH = X.conj().T
    Hp = np.zeros((N, N), dtype=X.dtype)
    for j in range(0, N, block):
        HXb = op.apply(X[:, j:j + block])
        Hp[:, j:j + block] = XH @ HXb
    Hp = 0.5 * (Hp + Hp.conj().T)
    lam, Q = np.linalg.eigh(Hp)
    return lam, X @ Q


def rayleigh_ritz_gep(op, X: np.ndarray, block: int):
    """CPU-default generalized route: S = L L^T, Hhat = L^{-1} Hp L^{-T}, eig,
    X <- X L^{-T} Q.  Reference: src/linAlg/linearAlgebraOperationsOpt.cc:613-969."""
    N = X.shape[1]
    XH = X.conj().T
    S = XH @ X
    Hp = np.zeros((N, N), dtype=X.dtype)
    for j in range(0, N, block):
        Hp[:, j:j + block] = XH @ op.apply(X[:, j:j + block])
    Hp = 0.5 * (Hp + Hp.conj().T)
    L = np.linalg.cholesky(S)
    Linv = sla.solve_triangular(L, np.eye(N), lower=True)
    lam, Q = np.linalg.eigh(Linv @ Hp @ Linv.conj().T)
    return lam, X @ (Linv.conj().T @ Q)


def residual_norms(op, X: np.ndarray, lam: np.ndarray, block: int) -> np.ndarray:
    """r_j = || H~ x_j - lambda_j x_j ||_2.  Reference:
    src/linAlg/linearAlgebraOperationsOpt.cc:2394-2503; device :4610-4766."""
    N = X.shape[1]
    r = np.zeros(N)
    for j in range(0, N, block):
        R = op.apply(X[:, j:j + block]) - X[:, j:j + block] * lam[None, j:j + block]
        r[j:j + block] = np.sqrt((R.real ** 2 + R.imag ** 2).sum(axis=0))
    return r


# ----------------------------------------------------------------------------
# One ChFSI pass
# ----------------------------------------------------------------------------


def chfsi_pass(op, X: np.ndarray, a0: float, bLow: float, bUp: float, m: int, block: int,
               route: str = "shep"):
    """One pass in the physical FE basis.  Reference sequence:
    src/solvers/eigenSolvers/chebyshevOrthogonalizedSubspaceIterationSolverDevice.cc:155-736
    (M^{1/2} scale :358-363, block filter :376-526, CholGS+RR :609-643,
    residual :668-713, M^{-1/2} unscale :719-733).

    Returns (eigenvalues, residual norms, X)."""
    p = op.prob
    Xt = op.sq[:, None] * X
    N = X.shape[1]

    def hx(src, scale_flag, scalar, dst):
        return op.HX(src, scale_flag, scalar, dst)

    for j in range(0, N, block):
        Xt[:, j:j + block] = chebyshev_filter(hx, Xt[:, j:j + block], m, bLow, bUp, a0)
    if route == "shep":
        Xt = cholgs(Xt)
        lam, Xt = rayleigh_ritz(op, Xt, block)
    else:
        lam, Xt = rayleigh_ritz_gep(op, Xt, block)
    res = residual_norms(op, Xt, lam, block)
    return lam, res, op.s[:, None] * Xt


def rayleigh_ritz_spectrum_split(op, X: np.ndarray, n_core: int, block: int):
    """Generalised Rayleigh-Ritz keeping only the top N - n_core pairs:
    X <- X L^{-T} (S = X^T X = L L^T), Hp = X^T H~ X, Hp = Q diag(l) Q^T, and
    Xfrac = X Q[:, n_core:].  Returns (l[n_core:], X orthonormalised, Xfrac).
    Reference: src/linAlg/rayleighRitzDevice.cc:822-1278
    (rayleighRitzGEPSpectrumSplitDirect; eigenvalues by index n_core..N-1 at
    :1401-1406, rotation of the top states only at :1430-1441)."""
    X = cholgs(X)
    N = X.shape[1]
    XH = X.conj().T
    Hp = np.zeros((N, N), dtype=X.dtype)
    for j in range(0, N, block):
        Hp[:, j:j + block] = XH @ op.apply(X[:, j:j + block])
    Hp = 0.5 * (Hp + Hp.conj().T)
    lam, Q = np.linalg.eigh(Hp)
    return lam[n_core:], X, X @ Q[:, n_core:]


def chfsi_pass_spectrum_split(op, X: np.ndarray, a0: float, bLow: float, bUp: float, m: int,
                              block: int, n_core: int):
    """One pass with the spectrum split, in the physical FE basis: filter as
    chfsi_pass, then the split Rayleigh-Ritz (no separate CholGS), residuals of
    the top pairs only, and both X and Xfrac unscaled.  Reference:
    chebyshevOrthogonalizedSubspaceIterationSolverDevice.cc:584-607 (split
    branch), :676-690 (residuals on Xfrac), :718-732 (unscaling).
    Returns (eigenvalues, residual norms, X, Xfrac)."""
    Xt = op.sq[:, None] * X
    N = X.shape[1]

    def hx(src, scale_flag, scalar, dst):
        return op.HX(src, scale_flag, scalar, dst)

    for j in range(0, N, block):
        Xt[:, j:j + block] = chebyshev_filter(hx, Xt[:, j:j + block], m, bLow, bUp, a0)
    lam, Xt, Xf = rayleigh_ritz_spectrum_split(op, Xt, n_core, block)
    res = residual_norms(op, Xf, lam, block)
    return lam, res, op.s[:, None] * Xt, op.s[:, None] * Xf


def next_bounds(lam: np.ndarray):
    """After a pass: a0 = lambda_0, bLow = lambda_{N-1} (src/dft/kohnShamEigenSolve.cc:489-498)."""
    return float(lam[0]), float(lam[-1])


# ----------------------------------------------------------------------------
# Derivatives with respect to the local potential (no reference counterpart;
# restated from the definitions H_e = 1/2 K + N^T diag(V JxW) N and the
# three-term recurrence above)
# ----------------------------------------------------------------------------


def quad_values(prob, X: np.ndarray) -> np.ndarray:
    """Values at the quadrature points, (C, nqc, B): psi(c, q, v) = sum_i N_i(q) X[g(c,i), v]."""
    return np.matmul(prob.Nq[None, :, :], X[prob.cell_map])


def hx_adjoint_v(op, Ybar: np.ndarray, X: np.ndarray, scalar: float = 1.0) -> np.ndarray:
    """dL/dV(c, q) for Y = scalar * H~(V) X and Ybar = dL/dY:
    scalar * JxW_q * sum_v (N S Ybar_v)(q) (N S X_v)(q)."""
    p = op.prob
    Ys = op.s[:, None] * Ybar
    Xs = op.s[:, None] * X
    op.distribute(Ys)
    op.distribute(Xs)
    a = quad_values(p, Ys)
    b = quad_values(p, Xs)
    return scalar * p.jxw_cells * (a * b).sum(axis=2)


def filter_backward(op, X0: np.ndarray, Ybar: np.ndarray, m: int, a: float, b: float, a0: float):
    """Reverse pass of chebyshev_filter with all iterates stored:
    returns (dL/dX0, dL/dV(c, q))."""
    e = (b - a) / 2.0
    c = (b + a) / 2.0
    sigma = e / (a0 - c)
    gamma = 2.0 / sigma
    a1 = [0.0, sigma / e]
    a2 = [0.0, 0.0]
    for _k in range(2, m + 1):
        sigma2 = 1.0 / (gamma - sigma)
        a1.append(2.0 * sigma2 / e)
        a2.append(-(sigma * sigma2))
        sigma = sigma2
    Xs = [X0]
    Xs.append(a1[1] * (op.apply(X0) - c * X0))
    for k in range(2, m + 1):
        Xs.append(a1[k] * (op.apply(Xs[k - 1]) - c * Xs[k - 1]) + a2[k] * Xs[k - 2])
    Lam = [np.zeros_like(X0) for _ in range(m + 1)]
    Lam[m] = Ybar.copy()
    g = np.zeros((op.prob.C, op.prob.nqc))
    for k in range(m, 0, -1):
        g += hx_adjoint_v(op, Lam[k], Xs[k - 1], a1[k])
        Lam[k - 1] += a1[k] * (op.apply(Lam[k]) - c * Lam[k])
        if k >= 2:
            Lam[k - 2] += a2[k] * Lam[k]
    return Lam[0], g


def with_potential(prob, V_q: np.ndarray):
    """Operator of the same problem with another local potential (cell matrices rebuilt)."""
    import copy

    p2 = copy.copy(prob)
    p2.V_q = V_q
    return OracleOperator(p2)


# ----------------------------------------------------------------------------
# Density and occupations
# ----------------------------------------------------------------------------


def fermi_dirac(eps: np.ndarray, mu: float, kT: float) -> np.ndarray:
    """Stable Fermi-Dirac occupation (utils/dftUtils.cc:56-64)."""
    t = (eps - mu) / kT
    return np.where(t <= 0, 1.0 / (1.0 + np.exp(np.minimum(t, 0.0))),
                    np.exp(-np.maximum(t, 0.0)) / (1.0 + np.exp(-np.maximum(t, 0.0))))


def density_at_quadrature(prob, X: np.ndarray, occ: np.ndarray, spin_factor: float = 2.0,
                          kweight: float = 1.0) -> np.ndarray:
    """rho(q) = sum_i f_i w_k s |psi_i(q)|^2, psi interpolated from the nodes
    after the constrained rows are distributed from their masters.
    Reference: src/dft/densityCalculator.cc:212-254, :622-689.  Shape (C, nqc)."""
    rho = np.zeros((prob.C, prob.nqc))
    w = spin_factor * kweight * occ
    if len(prob.slave_rows):
        rows, ptr, cols, vals = prob.constraint_csr()
        r = np.repeat(np.arange(len(rows)), np.diff(ptr))
        X = X.copy()
        X[rows] = sp.csr_matrix((vals, (r, cols)), shape=(len(rows), prob.M)) @ X
    for c0 in range(0, prob.C, 256):
        c1 = min(prob.C, c0 + 256)
        Xc = X[prob.cell_map[c0:c1]]                      # (k, ne, N)
        psi = np.einsum("qi,kiv->kqv", prob.Nq, Xc)
        rho[c0:c1] = np.einsum("kqv,v->kq", (psi * psi.conj()).real, w)
    return rho

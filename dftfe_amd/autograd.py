"""torch.autograd bridge of the differentiable ChFSI pass.

The FE Hamiltonian apply and the Chebyshev filter are custom autograd
Functions whose forward and backward both run through libchfsi.so:

    HXFunction      Y = H~(V) X           backward: Xbar = H~ Ybar,
                                                    Vbar = adjoint of HX w.r.t. V(q)
    FilterFunction  Y = p_m(H~(V)) X      backward: checkpointed reverse recurrence

CholGS and Rayleigh-Ritz are composed from torch.linalg (FP64) so the band
energy E = sum_i s f_i eps_i of one pass is differentiable with respect to the
local potential at every quadrature point.  The operator must hold the cell
matrices of the V it is called with (KohnShamOperatorDevice.set_potential).

No reference counterpart: the reference has no derivative of the eigensolver;
its density calculator (src/dft/densityCalculator.cc:37-689) supplies the
Hellmann-Feynman gradient sum_i s f_i |psi_i(q)|^2 JxW(q) this is checked
against at convergence.
"""
from __future__ import annotations

import math

import torch


class HXFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, V, X, op):
        ctx.op = op
        Xc = X.contiguous()
        ctx.save_for_backward(Xc)
        Y = torch.zeros_like(Xc)
        op.HX(Xc, False, 1.0, Y)
        return Y

    @staticmethod
    def backward(ctx, Ybar):
        (X,) = ctx.saved_tensors
        op = ctx.op
        Ybar = Ybar.contiguous()
        Xbar = torch.zeros_like(X)
        op.HX(Ybar, False, 1.0, Xbar)
        gV = op.hx_adjoint_v(Ybar, X)
        return gV, Xbar, None


class FilterFunction(torch.autograd.Function):
    """precision 0 (FP64) or 1 (FP32 recurrence in both directions, FP64
    accumulation of the potential gradient)."""

    @staticmethod
    def forward(ctx, V, X, op, m, a, b, a0, checkpoint_every, precision=0):
        from .chfsi import chebyshev_filter

        ctx.op = op
        ctx.args = (m, a, b, a0, checkpoint_every, precision)
        X0 = X.contiguous()
        ctx.save_for_backward(X0)
        Y = X0.clone()
        chebyshev_filter(op, Y, m, a, b, a0, precision)
        return Y

    @staticmethod
    def backward(ctx, Ybar):
        (X0,) = ctx.saved_tensors
        m, a, b, a0, every, precision = ctx.args
        Xbar, gV = ctx.op.filter_backward(X0, Ybar.contiguous(), m, a, b, a0, every, precision)
        return gV, Xbar, None, None, None, None, None, None, None


def fermi_occupations(eps: torch.Tensor, n_occ: float, kT: float, iters: int = 200):
    """Fermi-Dirac occupations f_i in [0, 1] with sum_i f_i = n_occ (bisection
    on the chemical potential; stable form of utils/dftUtils.cc:56-64)."""
    eps = eps.detach()
    lo, hi = float(eps.min()) - 50 * kT, float(eps.max()) + 50 * kT

    def occ(mu):
        t = (eps - mu) / kT
        return torch.where(t <= 0, 1.0 / (1.0 + torch.exp(t.clamp(max=0))),
                           torch.exp(-t.clamp(min=0)) / (1.0 + torch.exp(-t.clamp(min=0))))

    for _ in range(iters):
        mu = 0.5 * (lo + hi)
        if float(occ(mu).sum()) > n_occ:
            hi = mu
        else:
            lo = mu
    mu = 0.5 * (lo + hi)
    return occ(mu), mu


def band_energy(op, V: torch.Tensor, X0: torch.Tensor, bounds, m: int, occ: torch.Tensor,
                block: int, spin: float = 2.0, checkpoint_every: int = 5, precision: int = 0):
    """One ChFSI pass from the FE-basis start X0 with fixed occupations:
    returns (E_band, eigenvalues, Ritz vectors in the Loewdin basis).
    Differentiable in V (C x n_q) through the filter and Rayleigh-Ritz.
    precision selects the filter arithmetic (0 FP64, 1 FP32 with FP64
    gradient accumulation); CholGS and Rayleigh-Ritz stay FP64."""
    a0, bLow, bUp = bounds
    sq = torch.as_tensor(op.prob.sqrt_mass, device=X0.device)
    Xt = (sq[:, None] * X0).contiguous()
    N = Xt.shape[1]
    cols = []
    for j in range(0, N, block):
        cols.append(FilterFunction.apply(V, Xt[:, j:j + block], op, m, bLow, bUp, a0,
                                         checkpoint_every, precision))
    Y = torch.cat(cols, dim=1)
    S = Y.T @ Y
    Lc = torch.linalg.cholesky(S)
    # Y L^{-T} through the explicit N x N inverse: a triangular solve with the
    # M-wide right-hand side exceeds the BLAS library's TRSM workspace
    eye = torch.eye(N, dtype=Y.dtype, device=Y.device)
    Linv = torch.linalg.solve_triangular(Lc, eye, upper=False)
    Q = Y @ Linv.T
    HQ = torch.cat([HXFunction.apply(V, Q[:, j:j + block], op) for j in range(0, N, block)],
                   dim=1)
    Hp = Q.T @ HQ
    Hp = 0.5 * (Hp + Hp.T)
    eps = torch.linalg.eigvalsh(Hp)
    E = spin * (occ * eps).sum()
    return E, eps, Q


def hellmann_feynman_gradient(op, X_fe: torch.Tensor, occ: torch.Tensor, spin: float = 2.0):
    """sum_i s f_i |psi_i(q)|^2 JxW(c, q) from FE-basis eigenvectors (density route)."""
    rho = op.density(X_fe, spin * occ)
    jxw = torch.as_tensor(op.prob.jxw_cells, device=rho.device)
    return rho * jxw


__all__ = ["HXFunction", "FilterFunction", "band_energy", "fermi_occupations",
           "hellmann_feynman_gradient"]

"""Host-side mirror of the reference's device operator and ChFSI solver classes.

    KohnShamOperatorDevice   ~ operatorDFTDeviceClass / kohnShamDFTOperatorDeviceClass
                               (include/operatorDevice.h:43-420)
    ChebyshevSolverDevice    ~ chebyshevOrthogonalizedSubspaceIterationSolverDevice
                               (include/chebyshevOrthogonalizedSubspaceIterationSolverDevice.h:37-166)
    free functions           ~ linearAlgebraOperationsDevice::* (include/linearAlgebraOperationsDevice.h)

Device memory and streams come from PyTorch (plumbing only); every compute call
goes through libchfsi.so.  Multivectors are node-major torch tensors of shape
(M, N) (X[node, vec]) on the GPU.
"""
from __future__ import annotations

import ctypes as ct
from typing import Optional

import numpy as np
import torch

from . import lib as L


def _stream() -> ct.c_void_p:
    return ct.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: torch.Tensor) -> ct.c_void_p:
    return ct.c_void_p(t.data_ptr())


def _as(arr, dtype):
    return np.ascontiguousarray(arr, dtype=dtype)


class KohnShamOperatorDevice:
    """Device FE Hamiltonian H~ = M^{-1/2} H M^{-1/2} of one rank's full domain."""

    def __init__(self, prob, device: str = "cuda:0", kpoints=None):
        """kpoints: optional (n_k, 3) Bloch wave vectors; the operator then also
        holds the complex cell matrices and projectors of every k-point
        (d_cellHamiltonianMatrix[kpointSpinIndex] of the reference)."""
        self.lib = L.load()
        self.prob = prob
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        keep = {}

        def arr(name, a, dtype, ctype):
            keep[name] = _as(a, dtype)
            return keep[name].ctypes.data_as(ct.POINTER(ctype))

        d = L.OperatorDesc()
        d.n_dofs = prob.M
        d.n_cells = prob.C
        d.nodes_per_cell = prob.ne
        d.quad_per_cell = prob.nqc
        d.cell_map = arr("map", prob.cell_map, np.int64, ct.c_int64)
        d.cell_color = arr("col", prob.cell_color, np.int8, ct.c_int8)
        d.sqrt_mass = arr("sq", prob.sqrt_mass, np.float64, ct.c_double)
        d.inv_sqrt_mass = arr("isq", prob.inv_sqrt_mass, np.float64, ct.c_double)
        d.shape_values = arr("N", prob.Nq, np.float64, ct.c_double)
        d.stiffness = arr("K", prob.K_e, np.float64, ct.c_double)
        graded = getattr(prob, "cell_size", None) is not None
        if graded:
            # per-cell geometry: each cell's stiffness integral, and the
            # reference-cell gradients with the cells' (Cartesian) inverse Jacobians
            sc = prob.cell_scale
            d.cell_stiffness = arr("Kc", sc[:, None, None] * prob.K_e[None], np.float64,
                                   ct.c_double)
            if prob.Gq is not None:
                d.shape_gradients = arr("Gq", prob.Gq * prob.h, np.float64, ct.c_double)
                d.inverse_jacobians = arr("iJ", np.repeat((1.0 / prob.cell_size)[:, None], 3,
                                                          axis=1), np.float64, ct.c_double)
                d.jacobian_kind = 3
        elif prob.Gq is not None:
            d.shape_gradients = arr("Gq", prob.Gq, np.float64, ct.c_double)
        d.n_proj = prob.n_proj
        d.n_pairs = len(prob.pair_cell)
        d.proj_per_pair = 4
        if len(prob.pair_cell):
            d.pair_cell = arr("pc", prob.pair_cell, np.int32, ct.c_int32)
            d.pair_proj0 = arr("pp", 4 * prob.pair_atom, np.int32, ct.c_int32)
            d.pair_C = arr("pC", prob.pair_C, np.float64, ct.c_double)
            d.proj_V = arr("pV", prob.proj_V, np.float64, ct.c_double)
        if len(prob.slave_rows):
            rows, ptr, cols, vals = prob.constraint_csr()
            d.n_constrained = len(rows)
            d.constraint_rows = arr("cr", rows, np.int64, ct.c_int64)
            d.constraint_ptr = arr("cp", ptr, np.int64, ct.c_int64)
            d.constraint_cols = arr("cc", cols, np.int64, ct.c_int64)
            d.constraint_vals = arr("cv", vals, np.float64, ct.c_double)
        if graded:
            d.jxw = arr("jxw", prob.jxw_cells, np.float64, ct.c_double)
        else:
            d.jxw = arr("jxw", np.tile(prob.JxW, prob.C), np.float64, ct.c_double)
        self.kpoints = None
        if kpoints is not None:
            from . import fe

            kp = np.atleast_2d(_as(kpoints, np.float64))
            self.kpoints = kp
            d.n_kpoints = len(kp)
            d.kpoints = arr("kp", kp, np.float64, ct.c_double)
            d.cell_mass = arr("Mc", prob.M_c, np.float64, ct.c_double)
            d.cell_grad = arr("Gc", prob.G_c, np.float64, ct.c_double)
            if len(prob.pair_cell):
                zc = []
                for k in kp:
                    pa, pc, pC = fe.projectors(prob, k)
                    if not (np.array_equal(pa, prob.pair_atom) and np.array_equal(pc, prob.pair_cell)):
                        raise ValueError("k-point projector pairs differ from the real ones")
                    zc.append(pC)
                zc = np.ascontiguousarray(np.stack(zc), dtype=np.complex128)
                d.pair_C_kpoints = arr("zC", zc.view(np.float64), np.float64, ct.c_double)
        h = ct.c_void_p()
        L.check(self.lib.chfsi_operator_create(ct.byref(d), ct.byref(h)))
        self.handle = h
        self.M = prob.M
        self.set_potential(prob.vJxW)

    def __del__(self):
        if getattr(self, "handle", None) and self.handle.value:
            self.lib.chfsi_operator_destroy(self.handle)
            self.handle = None

    # -- setup --------------------------------------------------------------
    def set_potential(self, vJxW: np.ndarray) -> None:
        """computeHamiltonianMatricesAllkpt: cell matrices from V_eff JxW on device."""
        v = torch.as_tensor(_as(vJxW, np.float64), device=self.device)
        L.check(self.lib.chfsi_operator_set_potential(self.handle, _ptr(v), _stream()))
        torch.cuda.current_stream().synchronize()

    def set_potential_gga(self, vJxW: np.ndarray, gJxW: np.ndarray) -> None:
        """hamMatrixKernelGGAMemOpt: cell matrices from V_eff JxW and the
        (C, n_q, 3) GGA vector d(E_xc)/d(sigma) grad rho JxW, on device."""
        p = self.prob
        g = _as(gJxW, np.float64)
        if g.shape != (p.C, p.nqc, 3):
            raise ValueError(f"gJxW must have shape {(p.C, p.nqc, 3)}, got {g.shape}")
        v = torch.as_tensor(_as(vJxW, np.float64), device=self.device)
        gt = torch.as_tensor(g, device=self.device)
        L.check(self.lib.chfsi_operator_set_potential_gga(self.handle, _ptr(v), _ptr(gt),
                                                          _stream()))
        torch.cuda.current_stream().synchronize()

    def set_cell_matrices(self, H: torch.Tensor) -> None:
        L.check(self.lib.chfsi_operator_set_cell_matrices(self.handle, _ptr(H), _stream()))

    def cell_matrices(self) -> torch.Tensor:
        p = self.prob
        H = torch.empty((p.C, p.ne, p.ne), dtype=torch.float64, device=self.device)
        L.check(self.lib.chfsi_operator_get_cell_matrices(self.handle, _ptr(H), _stream()))
        return H

    def set_hx_variant(self, precision: int, variant: int) -> None:
        L.check(self.lib.chfsi_operator_set_hx_variant(self.handle, precision, variant))

    def set_proj_variant(self, precision: int, variant: int) -> None:
        L.check(self.lib.chfsi_operator_set_proj_variant(self.handle, precision, variant))

    def autotune(self, B: int) -> dict:
        """Time every HX tile variant on a B-column block and keep the fastest."""
        b64, b32 = ct.c_int32(), ct.c_int32()
        t = (ct.c_float * 48)()
        L.check(self.lib.chfsi_operator_autotune(self.handle, B, ct.byref(b64), ct.byref(b32), t,
                                                 _stream()))
        return {"fp64": b64.value, "fp32": b32.value,
                "fp64_ms": [x for x in t[0:16] if x >= 0],
                "fp32_ms": [x for x in t[24:40] if x >= 0],
                "fp64_proj_ms": [x for x in t[16:24] if x >= 0],
                "fp32_proj_ms": [x for x in t[40:48] if x >= 0]}

    def getSqrtMassVec(self) -> int:
        return self.lib.chfsi_operator_sqrt_mass(self.handle)

    def getInvSqrtMassVec(self) -> int:
        return self.lib.chfsi_operator_inv_sqrt_mass(self.handle)

    # -- applies ------------------------------------------------------------
    def HX(self, src: torch.Tensor, scale_flag: bool, scalar: float, dst: torch.Tensor,
           bf16: bool = False) -> None:
        """dst <- (scale_flag ? dst : M^{-1/2} dst) + scalar H~ src (in place on dst).
        float64 blocks run the FP64 apply, float32 blocks the FP32 apply, or with
        bf16=True the BF16-input / FP32-accumulate apply."""
        B = src.shape[1]
        if bf16:
            if src.dtype != torch.float32:
                raise TypeError("the BF16 apply takes float32 blocks")
            L.check(self.lib.chfsi_hx_bf16(self.handle, _ptr(src), src.stride(0), _ptr(dst),
                                           dst.stride(0), B, int(scale_flag), float(scalar),
                                           _stream()))
        elif src.dtype == torch.float64:
            L.check(self.lib.chfsi_hx(self.handle, _ptr(src), src.stride(0), _ptr(dst),
                                      dst.stride(0), B, int(scale_flag), float(scalar), _stream()))
        elif src.dtype == torch.float32:
            L.check(self.lib.chfsi_hx_fp32(self.handle, _ptr(src), src.stride(0), _ptr(dst),
                                           dst.stride(0), B, int(scale_flag), float(scalar),
                                           _stream()))
        else:
            raise TypeError("HX supports float64 and float32 blocks")

    # -- complex (Bloch k-point) form -------------------------------------------
    def reinitkPointSpinIndex(self, k: int, spin: int = 0) -> None:
        """Selects the k-point the complex calls act on (spin is not modelled)."""
        L.check(self.lib.chfsi_operator_set_kpoint(self.handle, int(k)))

    def HX_complex(self, src: torch.Tensor, scale_flag: bool, scalar: float,
                   dst: torch.Tensor) -> None:
        """Complex HX at the active k-point: complex128 -> FP64, complex64 -> FP32."""
        if src.dtype == torch.complex128:
            prec = L.FP64
        elif src.dtype == torch.complex64:
            prec = L.FP32
        else:
            raise TypeError("complex HX takes complex128 or complex64 blocks")
        L.check(self.lib.chfsi_hx_complex(self.handle, _ptr(src), src.stride(0), _ptr(dst),
                                          dst.stride(0), src.shape[1], int(scale_flag),
                                          float(scalar), prec, _stream()))

    def set_hx_variant_complex(self, precision: int, variant: int) -> None:
        L.check(self.lib.chfsi_operator_set_hx_variant_complex(self.handle, precision, variant))

    def set_proj_variant_complex(self, variant: int) -> None:
        """Complex projection: 0 the cell form, 1 owned DOF rows (both precisions)."""
        L.check(self.lib.chfsi_operator_set_proj_variant_complex(self.handle, variant))

    def autotune_complex(self, Xs, kidx, bounds, m: int = 2) -> dict:
        """Time the complex tile variants on a batch and keep the fastest per precision."""
        kb, keep = _kbatch(Xs, kidx, bounds, m, L.FP64)
        b64, b32 = ct.c_int32(), ct.c_int32()
        t = (ct.c_float * 16)()
        L.check(self.lib.chfsi_operator_autotune_complex(self.handle, ct.byref(kb), ct.byref(b64),
                                                         ct.byref(b32), t, _stream()))
        return {"fp64": b64.value, "fp32": b32.value,
                "fp64_ms": [x for x in t[0:8] if x >= 0], "fp32_ms": [x for x in t[8:16] if x >= 0]}

    def HXCheby(self, src: torch.Tensor, dst: torch.Tensor) -> None:
        """dst <- dst + H src (unscaled FE H; src already M^{-1/2}-scaled)."""
        L.check(self.lib.chfsi_hx_cheby(self.handle, _ptr(src), src.stride(0), _ptr(dst),
                                        dst.stride(0), src.shape[1], _stream()))

    def apply(self, X: torch.Tensor) -> torch.Tensor:
        """H~ X as a new tensor."""
        Y = torch.zeros_like(X)
        self.HX(X, False, 1.0, Y)
        return Y

    def distribute(self, X: torch.Tensor) -> None:
        L.check(self.lib.chfsi_constraints_distribute(self.handle, _ptr(X), X.stride(0),
                                                      X.shape[1], _stream()))

    def distribute_slave_to_master(self, X: torch.Tensor) -> None:
        L.check(self.lib.chfsi_constraints_slave_to_master(self.handle, _ptr(X), X.stride(0),
                                                           X.shape[1], _stream()))

    def set_zero(self, X: torch.Tensor) -> None:
        L.check(self.lib.chfsi_constraints_set_zero(self.handle, _ptr(X), X.stride(0),
                                                    X.shape[1], _stream()))


    # -- density and derivatives ------------------------------------------------
    def density(self, X: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
        """rho(c, q) = sum_v w_v psi_v(q)^2 for FE-basis wavefunctions X (M x N)."""
        p = self.prob
        rho = torch.empty((p.C, p.nqc), dtype=torch.float64, device=self.device)
        w = w.to(device=self.device, dtype=torch.float64).contiguous()
        L.check(self.lib.chfsi_density(self.handle, _ptr(X), X.stride(0), X.shape[1], _ptr(w),
                                       _ptr(rho), 0, _stream()))
        return rho

    def density_complex(self, X: torch.Tensor, w: torch.Tensor,
                        rho: Optional[torch.Tensor] = None) -> torch.Tensor:
        """rho(c, q) (+)= sum_v w_v |psi_v(q)|^2 for complex FE-basis wavefunctions
        X (M x N, complex128); with rho given, the k-point's contribution is added."""
        p = self.prob
        acc = rho is not None
        if rho is None:
            rho = torch.empty((p.C, p.nqc), dtype=torch.float64, device=self.device)
        w = w.to(device=self.device, dtype=torch.float64).contiguous()
        L.check(self.lib.chfsi_density_complex(self.handle, _ptr(X), X.stride(0), X.shape[1],
                                               _ptr(w), _ptr(rho), int(acc), _stream()))
        return rho

    def hx_adjoint_v(self, Ybar: torch.Tensor, X: torch.Tensor, scalar: float = 1.0,
                     g: Optional[torch.Tensor] = None) -> torch.Tensor:
        """dL/dV(c, q) for Y = scalar H~(V) X given Ybar = dL/dY (accumulates into g)."""
        p = self.prob
        acc = g is not None
        if g is None:
            g = torch.empty((p.C, p.nqc), dtype=torch.float64, device=self.device)
        L.check(self.lib.chfsi_hx_adjoint_v(self.handle, _ptr(Ybar), Ybar.stride(0), _ptr(X),
                                            X.stride(0), X.shape[1], float(scalar), _ptr(g),
                                            int(acc), _stream()))
        return g

    def filter_backward(self, X0: torch.Tensor, Ybar: torch.Tensor, m: int, a: float, b: float,
                        a0: float, checkpoint_every: int = 5, precision: int = L.FP64):
        """Reverse pass of the degree-m filter: returns (dL/dX0, dL/dV).
        precision FP32: FP32 iterates and adjoints, FP64 gradient accumulation."""
        p = self.prob
        Xbar = torch.empty_like(X0)
        g = torch.zeros((p.C, p.nqc), dtype=torch.float64, device=self.device)
        L.check(self.lib.chfsi_chebyshev_filter_backward(
            self.handle, _ptr(X0), X0.stride(0), _ptr(Ybar), Ybar.stride(0), X0.shape[1], m, a, b,
            a0, checkpoint_every, _ptr(Xbar), Xbar.stride(0), _ptr(g), precision, _stream()))
        return Xbar, g


# -- linearAlgebraOperationsDevice mirrors ------------------------------------


def chebyshev_filter(op: KohnShamOperatorDevice, X: torch.Tensor, m: int, a: float, b: float,
                     a0: float, precision: int = L.FP64) -> None:
    """In-place degree-m filter of a (M x B) block view of a node-major array."""
    L.check(op.lib.chfsi_chebyshev_filter(op.handle, _ptr(X), X.stride(0), X.shape[1], m, a, b,
                                          a0, precision, _stream()))


def _kbatch(Xs, kidx, bounds, m: int, precision: int):
    """chfsi_kbatch for the list of complex128 blocks Xs at k-points kidx;
    bounds: list of (a, b, a0) per entry.  Returns (struct, arrays to keep alive)."""
    n = len(Xs)
    for X in Xs:
        if X.dtype != torch.complex128 or X.stride(1) != 1:
            raise TypeError("batched filter blocks must be complex128 row-major views")
    kk = np.asarray(kidx, dtype=np.int32)
    ptrs = (ct.c_void_p * n)(*[X.data_ptr() for X in Xs])
    ld = np.asarray([X.stride(0) for X in Xs], dtype=np.int64)
    nv = np.asarray([X.shape[1] for X in Xs], dtype=np.int32)
    bnd = np.asarray(bounds, dtype=np.float64).reshape(n, 3)
    a, b, a0 = (np.ascontiguousarray(bnd[:, i]) for i in range(3))
    kb = L.KBatch()
    kb.n = n
    kb.kpoint = kk.ctypes.data_as(L.c_int32_p)
    kb.X = ct.cast(ptrs, ct.POINTER(ct.c_void_p))
    kb.ld = ld.ctypes.data_as(L.c_int64_p)
    kb.n_vec = nv.ctypes.data_as(L.c_int32_p)
    kb.a = a.ctypes.data_as(L.c_double_p)
    kb.b = b.ctypes.data_as(L.c_double_p)
    kb.a0 = a0.ctypes.data_as(L.c_double_p)
    kb.degree = m
    kb.precision = precision
    return kb, (kk, ptrs, ld, nv, a, b, a0)


def chebyshev_filter_complex(op: KohnShamOperatorDevice, X: torch.Tensor, m: int, a: float,
                             b: float, a0: float, precision: int = L.FP64) -> None:
    """In-place degree-m filter of a complex128 block at the active k-point."""
    L.check(op.lib.chfsi_chebyshev_filter_complex(op.handle, _ptr(X), X.stride(0), X.shape[1], m,
                                                  a, b, a0, precision, _stream()))


def batch_filter(op: KohnShamOperatorDevice, Xs, kidx, bounds, m: int,
                 precision: int = L.FP64) -> None:
    """All entries filtered together with plain launches (one launch per colour
    per degree for the whole batch)."""
    kb, keep = _kbatch(Xs, kidx, bounds, m, precision)
    L.check(op.lib.chfsi_batch_filter(op.handle, ct.byref(kb), _stream()))


class BatchFilter:
    """The batched filter captured once into a HIP graph and replayed."""

    def __init__(self, op: KohnShamOperatorDevice, Xs, kidx, bounds, m: int,
                 precision: int = L.FP64):
        self.lib = op.lib
        self.op = op
        self.n = len(Xs)
        kb, self._keep = _kbatch(Xs, kidx, bounds, m, precision)
        h = ct.c_void_p()
        L.check(self.lib.chfsi_batch_filter_create(op.handle, ct.byref(kb), _stream(),
                                                   ct.byref(h)))
        self.handle = h

    def set_bounds(self, bounds) -> None:
        bnd = np.asarray(bounds, dtype=np.float64).reshape(self.n, 3)
        a, b, a0 = (np.ascontiguousarray(bnd[:, i]) for i in range(3))
        L.check(self.lib.chfsi_batch_filter_set_bounds(self.handle, a.ctypes.data_as(L.c_double_p),
                                                       b.ctypes.data_as(L.c_double_p),
                                                       a0.ctypes.data_as(L.c_double_p)))

    def run(self) -> None:
        """Replay the captured graph on the current stream."""
        L.check(self.lib.chfsi_batch_filter_run(self.handle, _stream()))

    def launch(self) -> None:
        """Issue the same launches directly (no graph)."""
        L.check(self.lib.chfsi_batch_filter_launch(self.handle, _stream()))

    def __del__(self):
        if getattr(self, "handle", None) and self.handle.value:
            self.lib.chfsi_batch_filter_destroy(self.handle)
            self.handle = None


def lanczos_bounds_complex(op: KohnShamOperatorDevice, v0: torch.Tensor, steps: int,
                           reproducible: bool = False):
    lo, up = ct.c_double(), ct.c_double()
    L.check(op.lib.chfsi_lanczos_bounds_complex(op.handle, _ptr(v0), steps, int(reproducible),
                                                ct.byref(lo), ct.byref(up), _stream()))
    return lo.value, up.value


def cholgs_complex(op: KohnShamOperatorDevice, X: torch.Tensor) -> None:
    L.check(op.lib.chfsi_cholgs_complex(op.handle, _ptr(X), X.shape[0], X.shape[1], _stream()))


def rayleigh_ritz_complex(op: KohnShamOperatorDevice, X: torch.Tensor, block: int) -> np.ndarray:
    N = X.shape[1]
    ev = np.zeros(N)
    L.check(op.lib.chfsi_rayleigh_ritz_complex(op.handle, _ptr(X), X.shape[0], N, block,
                                               ev.ctypes.data_as(L.c_double_p), _stream()))
    return ev


def residual_norms_complex(op: KohnShamOperatorDevice, X: torch.Tensor, lam: np.ndarray,
                           block: int) -> np.ndarray:
    N = X.shape[1]
    lam = _as(lam, np.float64)
    r = np.zeros(N)
    L.check(op.lib.chfsi_residual_norms_complex(op.handle, _ptr(X), X.shape[0], N, block,
                                                lam.ctypes.data_as(L.c_double_p),
                                                r.ctypes.data_as(L.c_double_p), _stream()))
    return r


def lanczos_bounds(op: KohnShamOperatorDevice, v0: torch.Tensor, steps: int,
                   reproducible: bool = False):
    lo, up = ct.c_double(), ct.c_double()
    L.check(op.lib.chfsi_lanczos_bounds(op.handle, _ptr(v0), steps, int(reproducible),
                                        ct.byref(lo), ct.byref(up), _stream()))
    return lo.value, up.value


def cholgs(op: KohnShamOperatorDevice, X: torch.Tensor) -> None:
    L.check(op.lib.chfsi_cholgs(op.handle, _ptr(X), X.shape[0], X.shape[1], _stream()))


def rayleigh_ritz(op: KohnShamOperatorDevice, X: torch.Tensor, block: int) -> np.ndarray:
    N = X.shape[1]
    ev = np.zeros(N)
    L.check(op.lib.chfsi_rayleigh_ritz(op.handle, _ptr(X), X.shape[0], N, block,
                                       ev.ctypes.data_as(L.c_double_p), _stream()))
    return ev


def rayleigh_ritz_spectrum_split(op: KohnShamOperatorDevice, X: torch.Tensor, n_core: int,
                                 block: int, Xfrac: Optional[torch.Tensor] = None,
                                 mixed_xthx: bool = False):
    """rayleighRitzGEPSpectrumSplitDirect (rayleighRitzDevice.cc:822-1278):
    orthonormalise X in place (Cholesky), then keep the top N - n_core Ritz
    pairs of X^T H~ X.  Returns (eigenvalues, Xfrac) with Xfrac = X Q[:, n_core:]
    (M x (N - n_core), written into the given tensor when one is passed)."""
    M, N = X.shape
    nfr = N - n_core
    if Xfrac is None:
        Xfrac = torch.empty(M, nfr, dtype=X.dtype, device=X.device)
    ev = np.zeros(nfr)
    L.check(op.lib.chfsi_rayleigh_ritz_spectrum_split(
        op.handle, _ptr(X), M, N, n_core, _ptr(Xfrac), Xfrac.stride(0), block,
        ev.ctypes.data_as(L.c_double_p), int(mixed_xthx), _stream()))
    return ev, Xfrac


def residual_norms(op: KohnShamOperatorDevice, X: torch.Tensor, lam: np.ndarray,
                   block: int) -> np.ndarray:
    N = X.shape[1]
    lam = _as(lam, np.float64)
    r = np.zeros(N)
    L.check(op.lib.chfsi_residual_norms(op.handle, _ptr(X), X.shape[0], N, block,
                                        lam.ctypes.data_as(L.c_double_p),
                                        r.ctypes.data_as(L.c_double_p), _stream()))
    return r


class ChebyshevSolverDevice:
    """chebyshevOrthogonalizedSubspaceIterationSolverDevice: one ChFSI pass per solve()."""

    def __init__(self, a0: float = 0.0, bLow: float = 0.0, bUp: float = 0.0, *,
                 cheby_block_size: int = 400, cheby_degree: int = 0, lanczos_steps: int = 20,
                 reproducible: bool = False, filter_precision: int = L.FP32,
                 first_scf_degree_factor: float = 1.34, reuse_lanczos_upper: bool = False,
                 fp64_tail_degrees: int = 0, mixed_dense: bool = False,
                 filter_columns: int = 0, mixed_prec_switch_residual: float = 0.0,
                 mixed_prec_defect_filter: bool = False, use_shep: bool = True,
                 lanczos_full_residual_margin: bool = False,
                 is_pseudopotential: bool = False, mixed_prec_fp64_top_columns: int = 0,
                 dense_rocblas: int = 0):
        """mixed_dense: FP32 off-diagonal blocks in the overlap, CholGS rotation,
        X^T H X and Rayleigh-Ritz rotation when a pass runs with
        use_mixed_prec_overall (the reference's USE MIXED PREC CGS O / CGS SR /
        XTHX / RR_SR together).  filter_columns: columns one filter recurrence
        runs on, whole blocks of cheby_block_size (0: all, within a 16 GiB
        scratch); every column's result is the same whatever the grouping.
        mixed_prec_switch_residual > 0: once a pass's residual norms over its
        lowest 90% of states are all below it, later passes filter in FP64.
        mixed_prec_defect_filter:
        run the FP32 filter in defect form (X p(theta) + E, E in FP32), whose
        rounding scales with the residual.  use_shep=False: Rayleigh-Ritz in
        the generalised form without CholGS (rayleighRitzGEP).
        lanczos_full_residual_margin: bUp = ceil(theta_max + |f|) (the
        reference's reproducible-output margin) instead of |f| / 10.
        is_pseudopotential: gates the first-SCF degree factor and the halving
        of a looked-up degree for all-electron runs.
        mixed_prec_fp64_top_columns: with a lower filter precision, the last
        k columns (highest Ritz values) are filtered in FP64.
        dense_rocblas: FP64 dense steps on the library's MFMA kernels (0),
        on rocBLAS (1 or True), or the library's gram with rocBLAS rotations
        (2)."""
        self.lib = L.load()
        p = L.Params()
        self.lib.chfsi_default_params(ct.byref(p))
        p.cheby_block_size = cheby_block_size
        p.cheby_degree = cheby_degree
        p.lanczos_steps = lanczos_steps
        p.reproducible = int(reproducible)
        p.filter_precision = filter_precision
        p.first_scf_degree_factor = first_scf_degree_factor
        p.reuse_lanczos_upper = int(reuse_lanczos_upper)
        p.fp64_tail_degrees = fp64_tail_degrees
        p.filter_columns = filter_columns
        p.mixed_prec_switch_residual = mixed_prec_switch_residual
        p.mixed_prec_defect_filter = int(mixed_prec_defect_filter)
        p.use_shep = int(use_shep)
        p.lanczos_full_residual_margin = int(lanczos_full_residual_margin)
        p.is_pseudopotential = int(is_pseudopotential)
        p.mixed_prec_fp64_top_columns = int(mixed_prec_fp64_top_columns)
        p.dense_rocblas = int(dense_rocblas)
        for f in ("mixed_prec_overlap", "mixed_prec_cgs_rotation", "mixed_prec_xthx",
                  "mixed_prec_rr_rotation", "mixed_prec_xthx_spectrum_split"):
            setattr(p, f, int(mixed_dense))
        self.params = p
        h = ct.c_void_p()
        L.check(self.lib.chfsi_solver_create(ct.byref(p), a0, bLow, bUp, ct.byref(h)))
        self.handle = h
        self.last_timings_ms = None

    def __del__(self):
        if getattr(self, "handle", None) and self.handle.value:
            self.lib.chfsi_solver_destroy(self.handle)
            self.handle = None

    def set_filter_precision(self, precision: int, fp64_tail_degrees: int = 0) -> None:
        """Per-pass precision schedule: filter precision used when solve() is
        called with use_mixed_prec_overall=True."""
        self.params.filter_precision = precision
        self.params.fp64_tail_degrees = fp64_tail_degrees
        L.check(self.lib.chfsi_solver_set_params(self.handle, ct.byref(self.params)))

    def reinitSpectrumBounds(self, a0: float, bLow: float, bUp: float) -> None:
        L.check(self.lib.chfsi_solver_reinit_bounds(self.handle, a0, bLow, bUp))

    def bounds(self):
        a, b, c = ct.c_double(), ct.c_double(), ct.c_double()
        L.check(self.lib.chfsi_solver_get_bounds(self.handle, ct.byref(a), ct.byref(b),
                                                 ct.byref(c)))
        return a.value, b.value, c.value

    def last_filter_precision(self) -> int:
        """Filter precision of the last pass (L.FP64 once the residual switch fired)."""
        p = ct.c_int32()
        L.check(self.lib.chfsi_solver_last_filter_precision(self.handle, ct.byref(p)))
        return p.value

    def set_lanczos_start(self, v0: Optional[torch.Tensor]) -> None:
        if v0 is None:
            L.check(self.lib.chfsi_solver_set_lanczos_start(self.handle, None, 0))
        else:
            L.check(self.lib.chfsi_solver_set_lanczos_start(self.handle, _ptr(v0), v0.numel()))

    def solve(self, op: KohnShamOperatorDevice, X: torch.Tensor, *,
              is_first_filtering_call: bool = False, compute_residual: bool = True,
              use_mixed_prec_overall: bool = False, is_first_scf: bool = False):
        """One pass.  Returns (eigenvalues, residual norms, bUp); X is overwritten."""
        M, N = X.shape
        ev = np.zeros(N)
        res = np.zeros(N)
        bup = ct.c_double()
        t = np.zeros(6)
        L.check(self.lib.chfsi_solve(self.handle, op.handle, _ptr(X), M, N,
                                     ev.ctypes.data_as(L.c_double_p),
                                     res.ctypes.data_as(L.c_double_p),
                                     int(is_first_filtering_call), int(compute_residual),
                                     int(use_mixed_prec_overall), int(is_first_scf),
                                     ct.byref(bup), t.ctypes.data_as(L.c_double_p), _stream()))
        self.last_timings_ms = dict(zip(["lanczos", "filter", "cholgs", "rayleigh_ritz",
                                         "residual", "total"], t.tolist()))
        return ev, res, bup.value

    def solve_no_rr(self, op: KohnShamOperatorDevice, X: torch.Tensor, n_passes: int = 1, *,
                    scale_degree: bool = False, use_mixed_prec_overall: bool = False) -> None:
        """solveNoRR: n_passes of filter + CholGS on X (FE basis, overwritten)
        without Rayleigh-Ritz; a0 and bLow as last set, bUp refreshed by Lanczos
        unless reuse_lanczos_upper."""
        M, N = X.shape
        L.check(self.lib.chfsi_solve_no_rr(self.handle, op.handle, _ptr(X), M, N, n_passes,
                                           int(scale_degree), int(use_mixed_prec_overall),
                                           _stream()))

    def solve_spectrum_split(self, op: KohnShamOperatorDevice, X: torch.Tensor, n_core: int,
                             Xfrac: Optional[torch.Tensor] = None, *,
                             is_first_filtering_call: bool = False,
                             compute_residual: bool = True,
                             use_mixed_prec_overall: bool = False, is_first_scf: bool = False):
        """One pass with the spectrum split (the solve branch taken when the
        caller asks for fewer eigenvalues than wavefunctions): X is filtered and
        orthonormalised in place, the top N - n_core Ritz vectors go to Xfrac.
        Returns (eigenvalues, residual norms, bUp, Xfrac), N - n_core values each.
        The caller's bookkeeping for the next pass is bLow = eigenvalues[-1]
        with a0 kept (kohnShamEigenSolve.cc:483-498)."""
        M, N = X.shape
        nfr = N - n_core
        if Xfrac is None:
            Xfrac = torch.empty(M, nfr, dtype=X.dtype, device=X.device)
        ev = np.zeros(nfr)
        res = np.zeros(nfr)
        bup = ct.c_double()
        t = np.zeros(6)
        L.check(self.lib.chfsi_solve_spectrum_split(
            self.handle, op.handle, _ptr(X), M, N, n_core, _ptr(Xfrac), Xfrac.stride(0),
            ev.ctypes.data_as(L.c_double_p), res.ctypes.data_as(L.c_double_p),
            int(is_first_filtering_call), int(compute_residual), int(use_mixed_prec_overall),
            int(is_first_scf), ct.byref(bup), t.ctypes.data_as(L.c_double_p), _stream()))
        self.last_timings_ms = dict(zip(["lanczos", "filter", "cholgs", "rayleigh_ritz",
                                         "residual", "total"], t.tolist()))
        return ev, res, bup.value, Xfrac

    def solve_complex(self, op: KohnShamOperatorDevice, X: torch.Tensor, *,
                      is_first_filtering_call: bool = False, compute_residual: bool = True,
                      use_mixed_prec_overall: bool = False, is_first_scf: bool = False):
        """One pass on complex128 X at the operator's active k-point."""
        if X.dtype != torch.complex128:
            raise TypeError("solve_complex takes complex128 wavefunctions")
        M, N = X.shape
        ev = np.zeros(N)
        res = np.zeros(N)
        bup = ct.c_double()
        t = np.zeros(6)
        L.check(self.lib.chfsi_solve_complex(self.handle, op.handle, _ptr(X), M, N,
                                             ev.ctypes.data_as(L.c_double_p),
                                             res.ctypes.data_as(L.c_double_p),
                                             int(is_first_filtering_call), int(compute_residual),
                                             int(use_mixed_prec_overall), int(is_first_scf),
                                             ct.byref(bup), t.ctypes.data_as(L.c_double_p),
                                             _stream()))
        self.last_timings_ms = dict(zip(["lanczos", "filter", "cholgs", "rayleigh_ritz",
                                         "residual", "total"], t.tolist()))
        return ev, res, bup.value

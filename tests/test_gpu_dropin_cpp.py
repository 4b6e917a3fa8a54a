"""GPU parity of the C++ drop-in classes (include/dftfe_chfsi.hpp):
kohnShamDFTOperatorDeviceClass and chebyshevOrthogonalizedSubspaceIterationSolverDevice
driven from C++ (dftfe_amd/bin/dropin_solve, source tests/cpp/dropin_solve.cpp)
as DFT-FE's kohnShamEigenSpaceCompute drives them, checked against the oracle.

Workload: periodic 20 bohr cube, 8^3 cells, p=4 (32,768 DOFs), 8 wells with
s+p projectors, 64 wavefunctions in blocks of 32, degree 20, seeded bounds from
10 Lanczos steps.  Tolerances: H~X 1e-13 relative; eigenvalues 1e-10 Ha;
residual norms 1e-8; the no-Rayleigh-Ritz pass (filter + CholGS) 1e-9 relative.
"""
import os
import subprocess

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dftfe_amd import fe  # noqa: E402
from oracle import chfsi_oracle as O  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "dftfe_amd", "bin", "dropin_solve")
N, BLOCK, DEGREE, N_SPLIT = 64, 32, 20, 24


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _write(d, name, a, dtype):
    np.ascontiguousarray(a, dtype=dtype).tofile(os.path.join(d, name))


def _read(d, name, shape):
    return np.fromfile(os.path.join(d, name), dtype=np.float64).reshape(shape)


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(DRIVER):
        raise RuntimeError(f"{DRIVER} is not built (run dftfe_amd.build)")
    prob = fe.build_workload("cube20_p4_n64")
    ref = O.OracleOperator(prob)
    lo, up, _, _ = O.lanczos_bounds(ref, prob.lanczos_start, 10)
    bounds = O.first_call_bounds(lo, up, N, prob.M)
    X0 = prob.initial_guess(N=N)
    g = 0.05 * np.random.default_rng(4).standard_normal((prob.C, prob.nqc, 3)) \
        * prob.JxW[None, :, None]
    d = str(tmp_path_factory.mktemp("dropin"))
    meta = dict(M=prob.M, C=prob.C, ne=prob.ne, nq=prob.nqc, n_proj=prob.n_proj,
                n_pairs=len(prob.pair_cell), N=N, B=BLOCK, m=DEGREE, a0=bounds[0],
                bLow=bounds[1], bUp=bounds[2], n_split=N_SPLIT)
    with open(os.path.join(d, "meta.txt"), "w") as f:
        for k, v in meta.items():
            f.write(f"{k} {v!r}\n")
    _write(d, "map.i64", prob.cell_map, np.int64)
    _write(d, "color.i8", prob.cell_color, np.int8)
    _write(d, "sqrt_mass.f64", prob.sqrt_mass, np.float64)
    _write(d, "inv_sqrt_mass.f64", prob.inv_sqrt_mass, np.float64)
    _write(d, "Nq.f64", prob.Nq, np.float64)
    _write(d, "K.f64", prob.K_e, np.float64)
    _write(d, "Gq.f64", prob.Gq, np.float64)
    _write(d, "jxw.f64", np.tile(prob.JxW, prob.C), np.float64)
    _write(d, "pair_cell.i32", prob.pair_cell, np.int32)
    _write(d, "pair_proj0.i32", 4 * prob.pair_atom, np.int32)
    _write(d, "pair_C.f64", prob.pair_C, np.float64)
    _write(d, "proj_V.f64", prob.proj_V, np.float64)
    _write(d, "vJxW.f64", prob.vJxW, np.float64)
    _write(d, "gJxW.f64", g, np.float64)
    _write(d, "X0.f64", X0, np.float64)
    print(f"dropin_solve on cube20_p4_n64 N={N} B={BLOCK} m={DEGREE}", flush=True)
    res = subprocess.run([DRIVER, d], capture_output=True, text=True, timeout=120)
    if res.returncode != 0:
        raise RuntimeError(f"dropin_solve failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    return prob, ref, bounds, X0, g, d


def test_dropin_hx(run):
    prob, ref, _, X0, _, d = run
    assert _rel(_read(d, "hx.f64", (prob.M, N)), ref.apply(X0)) < 1e-13


def test_dropin_solve_two_passes(run):
    """solve, the caller's bound update (a0 = lowest, bLow = highest Ritz value),
    solve again: eigenvalues and residual norms of both passes."""
    prob, ref, bounds, X0, _, d = run
    lam_o, res_o, X_o = O.chfsi_pass(ref, X0, *bounds, DEGREE, BLOCK)
    assert np.abs(_read(d, "eig1.f64", (N,)) - lam_o).max() <= 1e-10
    assert np.abs(_read(d, "res1.f64", (N,)) - res_o).max() <= 1e-8
    a0, bLow = O.next_bounds(lam_o)
    lam2_o, res2_o, _ = O.chfsi_pass(ref, X_o, a0, bLow, bounds[2], DEGREE, BLOCK)
    assert np.abs(_read(d, "eig2.f64", (N,)) - lam2_o).max() <= 1e-10
    assert np.abs(_read(d, "res2.f64", (N,)) - res2_o).max() <= 1e-8


def test_dropin_mixed_overall_default_params_stays_fp64(run):
    """useMixedPrecOverall = true with the reference's default parameters
    (USE MIXED PREC CHEBY false) must run the FP64 filter: FP64 parity."""
    prob, ref, bounds, X0, _, d = run
    lam_o, _, _ = O.chfsi_pass(ref, X0, *bounds, DEGREE, BLOCK)
    assert np.abs(_read(d, "eig_mixed_default.f64", (N,)) - lam_o).max() <= 1e-10


def test_dropin_gep_route(run):
    """useSubspaceProjectedSHEPGPU = false: Rayleigh-Ritz in the generalised form
    (rayleighRitzGEP) on the filtered, unorthonormalised block."""
    prob, ref, bounds, X0, _, d = run
    lam_o, res_o, _ = O.chfsi_pass(ref, X0, *bounds, DEGREE, BLOCK, route="gep")
    assert np.abs(_read(d, "eig_gep.f64", (N,)) - lam_o).max() <= 1e-10
    assert np.abs(_read(d, "res_gep.f64", (N,)) - res_o).max() <= 1e-8


def test_dropin_solve_no_rr(run):
    """solveNoRR, one pass: the filtered block after CholGS, in the FE basis
    (CholGS is unique: X L^{-T} with the Cholesky factor of X^T X)."""
    prob, ref, bounds, X0, _, d = run
    Xt = ref.sq[:, None] * X0

    def hx(src, scale_flag, scalar, dst):
        return ref.HX(src, scale_flag, scalar, dst)

    for j in range(0, N, BLOCK):
        Xt[:, j:j + BLOCK] = O.chebyshev_filter(hx, Xt[:, j:j + BLOCK], DEGREE, bounds[1],
                                                bounds[2], bounds[0])
    want = ref.s[:, None] * O.cholgs(Xt)
    assert _rel(_read(d, "Xnr.f64", (prob.M, N)), want) < 1e-9


def test_dropin_spectrum_split(run):
    """eigenValues sized to the kept states selects the spectrum-split route."""
    prob, ref, bounds, X0, _, d = run
    lam_o, _, _, Xf_o = O.chfsi_pass_spectrum_split(ref, X0, *bounds, DEGREE, BLOCK, N - N_SPLIT)
    assert np.abs(_read(d, "eig_split.f64", (N_SPLIT,)) - lam_o).max() <= 1e-10
    Xf = _read(d, "Xfrac.f64", (prob.M, N_SPLIT))
    # Ritz vectors up to sign: |<x, x_o>| = 1 in the mass inner product
    ov = np.abs(np.einsum("mi,mi->i", Xf * prob.mass[:, None], Xf_o))
    assert np.abs(ov - 1.0).max() < 1e-8


def test_dropin_gga_rebuild(run):
    prob, _, _, X0, g, d = run
    ref_g = O.OracleOperator(prob, H_e=O.cell_hamiltonians(prob, gJxW=g))
    assert _rel(_read(d, "hx_gga.f64", (prob.M, N)), ref_g.apply(X0)) < 1e-13


DRIVER_C = os.path.join(ROOT, "dftfe_amd", "bin", "dropin_solve_complex")


def test_dropin_complex_kpoint(tmp_path):
    """The classes built with USE_COMPLEX (dataTypes::number = std::complex<double>)
    on the 20 bohr / p=4 operator with 6 seeded k-points: HX at k-point 3 and
    two solves with the caller's bound update against the oracle's complex
    passes (32 states, blocks of 16, degree 20)."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(DRIVER_C):
        raise RuntimeError(f"{DRIVER_C} is not built (run dftfe_amd.build)")
    prob = fe.build_workload("cube20_p4_n64")
    kpts, _ = fe.kpoint_set(prob, n=6)
    kidx, Nc, Bc, m = 3, 32, 16, 20
    ref = O.OracleOperator(fe.with_kpoint(prob, kpts[kidx]))
    lo, up, _, _ = O.lanczos_bounds(ref, prob.lanczos_start.astype(np.complex128), 10)
    bounds = O.first_call_bounds(lo, up, Nc, prob.M)
    X0 = fe.complex_guess(prob, Nc, seed_offset=17)
    d = str(tmp_path)
    meta = dict(M=prob.M, C=prob.C, ne=prob.ne, nq=prob.nqc, n_proj=prob.n_proj,
                n_pairs=len(prob.pair_cell), N=Nc, B=Bc, m=m, nk=len(kpts), kidx=kidx,
                a0=bounds[0], bLow=bounds[1], bUp=bounds[2])
    with open(os.path.join(d, "meta.txt"), "w") as f:
        for k, v in meta.items():
            f.write(f"{k} {v!r}\n")
    _write(d, "map.i64", prob.cell_map, np.int64)
    _write(d, "color.i8", prob.cell_color, np.int8)
    _write(d, "sqrt_mass.f64", prob.sqrt_mass, np.float64)
    _write(d, "inv_sqrt_mass.f64", prob.inv_sqrt_mass, np.float64)
    _write(d, "Nq.f64", prob.Nq, np.float64)
    _write(d, "K.f64", prob.K_e, np.float64)
    _write(d, "jxw.f64", np.tile(prob.JxW, prob.C), np.float64)
    _write(d, "pair_cell.i32", prob.pair_cell, np.int32)
    _write(d, "pair_proj0.i32", 4 * prob.pair_atom, np.int32)
    _write(d, "pair_C.f64", prob.pair_C, np.float64)
    _write(d, "proj_V.f64", prob.proj_V, np.float64)
    _write(d, "vJxW.f64", prob.vJxW, np.float64)
    _write(d, "kpts.f64", kpts, np.float64)
    _write(d, "Mc.f64", prob.M_c, np.float64)
    _write(d, "Gc.f64", prob.G_c, np.float64)
    zc = np.stack([fe.projectors(prob, k)[2] for k in kpts]).astype(np.complex128)
    _write(d, "pair_C_k.f64", zc.view(np.float64), np.float64)
    _write(d, "Xc0.c128", X0, np.complex128)
    print(f"dropin_solve_complex on cube20_p4_n64 k={kidx} N={Nc}", flush=True)
    res = subprocess.run([DRIVER_C, d], capture_output=True, text=True, timeout=120)
    if res.returncode != 0:
        raise RuntimeError(f"dropin_solve_complex failed ({res.returncode}):\n{res.stdout}\n"
                           f"{res.stderr}")
    hx = np.fromfile(os.path.join(d, "hx_c.c128"), dtype=np.complex128).reshape(prob.M, Nc)
    assert _rel(hx, ref.apply(X0)) < 1e-13
    lam_o, res_o, X_o = O.chfsi_pass(ref, X0, *bounds, m, Bc)
    assert np.abs(_read(d, "eigc1.f64", (Nc,)) - lam_o).max() <= 1e-10
    assert np.abs(_read(d, "resc1.f64", (Nc,)) - res_o).max() <= 1e-8
    lam2_o, res2_o, _ = O.chfsi_pass(ref, X_o, *O.next_bounds(lam_o), bounds[2], m, Bc)
    assert np.abs(_read(d, "eigc2.f64", (Nc,)) - lam2_o).max() <= 1e-10
    assert np.abs(_read(d, "resc2.f64", (Nc,)) - res2_o).max() <= 1e-8

"""GPU parity of the HIP ChFSI path against the CPU FP64 oracle.

Every compute call goes through libchfsi.so (the C ABI); the oracle is the
checker.  Workload: periodic 20 bohr cube, 8^3 cells, p=4 (32,768 DOFs),
8 Gaussian wells with s+p projectors, 64 wavefunctions, degree 20.
Tolerances: FP64 HX 1e-13 relative (max-norm), eigenvalues 1e-10 Ha,
FP32 HX 2e-6 relative.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dftfe_amd import fe  # noqa: E402
from oracle import chfsi_oracle as O  # noqa: E402


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def small():
    _dev()
    from dftfe_amd.chfsi import KohnShamOperatorDevice

    prob = fe.build_workload("cube20_p4_n64")
    ref = O.OracleOperator(prob)
    op = KohnShamOperatorDevice(prob)
    return prob, ref, op


@pytest.fixture(scope="module")
def small_dealii():
    _dev()
    from dftfe_amd.chfsi import KohnShamOperatorDevice

    prob = fe.build_workload("cube20_p4_n64", mode="dealii")
    ref = O.OracleOperator(prob)
    op = KohnShamOperatorDevice(prob)
    return prob, ref, op


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_cell_matrices_match_oracle(small):
    prob, ref, op = small
    H = op.cell_matrices().cpu().numpy()
    assert _rel(H, ref.H_e) < 1e-13


@pytest.mark.parametrize("B", [64, 40, 1, 200])
def test_hx_fp64_matches_oracle(small, B):
    prob, ref, op = small
    X = prob.initial_guess(N=B, seed_offset=B)
    Xd = torch.tensor(X, device="cuda")
    Yd = torch.zeros_like(Xd)
    op.HX(Xd, False, 1.0, Yd)
    Y = Yd.cpu().numpy()
    assert _rel(Y, ref.apply(X)) < 1e-13


@pytest.mark.parametrize("precision,variant", [(0, v) for v in range(8)] + [(1, v) for v in range(12)])
def test_hx_tile_variants(small, precision, variant):
    """Every HX tile variant gives the oracle's H~X (FP64: 1e-13, FP32: 2e-6)."""
    prob, ref, op = small
    X = prob.initial_guess(N=72, seed_offset=21)
    dt = torch.float64 if precision == 0 else torch.float32
    Xd = torch.tensor(X, device="cuda", dtype=dt)
    Yd = torch.zeros_like(Xd)
    op.set_hx_variant(precision, variant)
    try:
        op.HX(Xd, False, 1.0, Yd)
    finally:
        op.set_hx_variant(precision, 0)
    want = ref.apply(X if precision == 0 else X.astype(np.float32).astype(np.float64))
    assert _rel(Yd.double().cpu().numpy(), want) < (1e-13 if precision == 0 else 2e-6)


@pytest.mark.parametrize("B", [2, 71, 130])
@pytest.mark.parametrize("variant", [6, 7])
def test_hx_fp64_ring_widths(small, variant, B):
    """The FP64 ring (variants 6, 7) at the narrowest even width, at an odd
    width (it takes the register-staged kernel) and across a partial
    64-column tile: the oracle's H~X to 1e-13."""
    prob, ref, op = small
    X = prob.initial_guess(N=B, seed_offset=B + 31)
    Xd = torch.tensor(X, device="cuda")
    Yd = torch.zeros_like(Xd)
    op.set_hx_variant(0, variant)
    try:
        op.HX(Xd, False, 1.0, Yd)
    finally:
        op.set_hx_variant(0, 0)
    assert _rel(Yd.cpu().numpy(), ref.apply(X)) < 1e-13


@pytest.mark.parametrize("B", [72, 256, 70])
@pytest.mark.parametrize("variant", [5, 6, 11])
def test_hx_ring_variants_bitwise(small, variant, B):
    """The direct-to-LDS ring variants (FP32 5, 6, 11) add the same products in
    the same order as the register-staged kernel (variant 2), and every other add
    is an explicit FMA, so a degree-4 FP32 filter (scaled input, first-touch
    Chebyshev combine, nonlocal slices) gives bitwise-equal blocks: B = 72
    leaves a partial column tile, B = 256 full tiles, and B = 70 (not a multiple
    of 4) takes the register-staged kernel instead."""
    from dftfe_amd import lib as L
    from dftfe_amd.chfsi import chebyshev_filter

    prob, ref, op = small
    X = torch.tensor(prob.sqrt_mass[:, None] * prob.initial_guess(N=B, seed_offset=B + 3),
                     device="cuda")
    out = {}
    for v in (2, variant):
        Xd = X.clone()
        op.set_hx_variant(L.FP32, v)
        try:
            chebyshev_filter(op, Xd, 4, 5.0, 41.0, 0.0, precision=L.FP32)
        finally:
            op.set_hx_variant(L.FP32, 0)
        out[v] = Xd
    assert torch.isfinite(out[2]).all()
    assert torch.equal(out[2], out[variant])


@pytest.mark.parametrize("B", [72, 256, 70, 520])
def test_projection_ring_bitwise(small, B):
    """The direct-to-LDS FP32 projection (variant 3) forms the same products in
    the same order as variant 0, so a degree-4 FP32 filter gives bitwise-equal
    blocks; B = 70 (not a multiple of 4) takes variant 0 itself, B = 520 the
    256-column tiles with a partial last tile."""
    from dftfe_amd import lib as L
    from dftfe_amd.chfsi import chebyshev_filter

    prob, ref, op = small
    X = torch.tensor(prob.sqrt_mass[:, None] * prob.initial_guess(N=B, seed_offset=B + 5),
                     device="cuda")
    out = {}
    for v in (0, 3):
        Xd = X.clone()
        op.set_proj_variant(L.FP32, v)
        try:
            chebyshev_filter(op, Xd, 4, 5.0, 41.0, 0.0, precision=L.FP32)
        finally:
            op.set_proj_variant(L.FP32, 0)
        out[v] = Xd
    assert torch.isfinite(out[0]).all()
    assert torch.equal(out[0], out[3])


@pytest.mark.parametrize("precision,variant", [(p, v) for p in (0, 1) for v in range(6)])
def test_projection_variants(small, precision, variant):
    """Every projection-kernel variant gives the oracle's H~X."""
    prob, ref, op = small
    X = prob.initial_guess(N=72, seed_offset=22)
    dt = torch.float64 if precision == 0 else torch.float32
    Xd = torch.tensor(X, device="cuda", dtype=dt)
    Yd = torch.zeros_like(Xd)
    op.set_proj_variant(precision, variant)
    try:
        op.HX(Xd, False, 1.0, Yd)
    finally:
        op.set_proj_variant(precision, 0)
    want = ref.apply(X if precision == 0 else X.astype(np.float32).astype(np.float64))
    assert _rel(Yd.double().cpu().numpy(), want) < (1e-13 if precision == 0 else 2e-6)


@pytest.mark.parametrize("name,mode", [("tiny", "unique"), ("brute", "unique"),
                                       ("cube20_p4_n64", "dealii")])
def test_projection_owned_rows(name, mode):
    """Projection variants 4 (direct-to-LDS rings) and 5 (register-staged) form
    C^T x over DOF rows with the projector values assembled over the cell
    nodes, each row read by the cells that cover its projectors.  They give the
    oracle's H~X on a mesh of 2 cells per side (a cell's opposite faces map to
    the same DOFs), on p = 3, and with the periodic rows kept as constraint
    rows (dealii layout), in both precisions, for one column (B = 1, the
    register-staged fallback), B = 6 (FP64 ring, FP32 fallback) and B = 40."""
    _dev()
    from dftfe_amd.chfsi import KohnShamOperatorDevice

    prob = fe.build_workload(name, mode=mode)
    assert len(prob.pair_cell) > 0
    ref = O.OracleOperator(prob)
    op = KohnShamOperatorDevice(prob)
    for B in (1, 6, 40):
        X = prob.initial_guess(N=B, seed_offset=70 + B)
        for prec, dt, tol in ((0, torch.float64, 1e-13), (1, torch.float32, 2e-6)):
            want = ref.apply(X if prec == 0 else X.astype(np.float32).astype(np.float64))
            Xd = torch.tensor(X, device="cuda", dtype=dt)
            for v in (4, 5):
                Yd = torch.zeros_like(Xd)
                op.set_proj_variant(prec, v)
                try:
                    op.HX(Xd, False, 1.0, Yd)
                finally:
                    op.set_proj_variant(prec, 0)
                err = _rel(Yd.double().cpu().numpy(), want)
                assert err < tol, (name, mode, B, prec, v, err)


@pytest.mark.parametrize("n,atoms", [(2, 70), (4, 40)])
def test_crowded_cells(n, atoms):
    """Cells overlapping more than 128 projector slots (70 wells on 2^3 cells:
    276-280 slots per cell, three groups; 40 wells on 4^3 cells: 128-160): the
    operator hands each further group of 32 pairs to an extra cell id with a
    zero cell matrix in a later colour.  Every FP64 / FP32 tile variant and
    projection variant, the one-column kernel and the BF16 apply give the
    oracle's H~X."""
    _dev()
    from dftfe_amd.chfsi import KohnShamOperatorDevice

    wl = fe.Workload(f"crowd{n}", 8.0, n, 2, 3, atoms, 16, 16, 8, seed=3030 + n)
    prob = fe.build_problem(wl)
    slots = np.bincount(prob.pair_cell, minlength=prob.C) * 4
    assert slots.max() > 128
    ref = O.OracleOperator(prob)
    op = KohnShamOperatorDevice(prob)
    bad = []
    for N in (24, 1):
        X = prob.initial_guess(N=N, seed_offset=5 + N)
        want = {0: ref.apply(X), 1: ref.apply(X.astype(np.float32).astype(np.float64))}
        for prec, nvar, kind in ((0, 8, "hx"), (1, 12, "hx"), (0, 6, "proj"), (1, 6, "proj")):
            setter = op.set_hx_variant if kind == "hx" else op.set_proj_variant
            dt = torch.float64 if prec == 0 else torch.float32
            Xd = torch.tensor(X, device="cuda", dtype=dt)
            for v in range(nvar):
                Yd = torch.zeros_like(Xd)
                setter(prec, v)
                try:
                    op.HX(Xd, False, 1.0, Yd)
                finally:
                    setter(prec, 0)
                err = _rel(Yd.double().cpu().numpy(), want[prec])
                if not err < (1e-13 if prec == 0 else 2e-6):
                    bad.append((N, kind, prec, v, err))
        Xd = torch.tensor(X, device="cuda", dtype=torch.float32)
        for v in range(6):
            Yd = torch.zeros_like(Xd)
            op.set_hx_variant(2, v)
            try:
                op.HX(Xd, False, 1.0, Yd, bf16=True)
            finally:
                op.set_hx_variant(2, 0)
            err = _rel(Yd.double().cpu().numpy(), want[1])
            if not err < 3e-2:
                bad.append((N, "bf16", 2, v, err))
    assert not bad, bad
    # the k-point kernels carry at most 128 slots per cell: refused, not wrong
    from dftfe_amd.lib import ChfsiError

    kpts, _ = fe.kpoint_set(prob, n=2)
    with pytest.raises(ChfsiError, match="k-points"):
        KohnShamOperatorDevice(prob, kpoints=kpts)


def test_hx_strided_block_and_flags(small):
    """HX on a column block of a wider array, with scale_flag / scalar semantics."""
    prob, ref, op = small
    X = prob.initial_guess(N=96, seed_offset=5)
    D = prob.initial_guess(N=96, seed_offset=6)
    for scale_flag, scalar in [(False, 1.0), (True, 0.37), (False, -2.5)]:
        Xd = torch.tensor(X, device="cuda")
        Dd = torch.tensor(D, device="cuda")
        op.HX(Xd[:, 16:80], scale_flag, scalar, Dd[:, 8:72])
        want = D.copy()
        want[:, 8:72] = ref.HX(X[:, 16:80], scale_flag, scalar, D[:, 8:72])
        got = Dd.cpu().numpy()
        assert _rel(got, want) < 1e-13
        assert np.array_equal(got[:, :8], D[:, :8]) and np.array_equal(got[:, 72:], D[:, 72:])


def test_hx_cheby_semantics(small):
    """HXCheby: dst += H src with the unscaled FE H."""
    prob, ref, op = small
    X = prob.initial_guess(N=64, seed_offset=9)
    D = prob.initial_guess(N=64, seed_offset=10)
    Xd = torch.tensor(X, device="cuda")
    Dd = torch.tensor(D, device="cuda")
    op.HXCheby(Xd, Dd)
    want = D + ref.h_fe(X)
    assert _rel(Dd.cpu().numpy(), want) < 1e-13


def test_hx_deterministic(small):
    prob, ref, op = small
    Xd = torch.tensor(prob.initial_guess(N=64), device="cuda")
    Y1 = op.apply(Xd)
    Y2 = op.apply(Xd)
    assert torch.equal(Y1, Y2)


def test_hx_fp32(small):
    prob, ref, op = small
    X = prob.initial_guess(N=64, seed_offset=3)
    Xd = torch.tensor(X, device="cuda", dtype=torch.float32)
    Yd = torch.zeros_like(Xd)
    op.HX(Xd, False, 1.0, Yd)
    assert _rel(Yd.double().cpu().numpy(), ref.apply(X.astype(np.float32).astype(np.float64))) < 2e-6


def test_hx_bf16(small):
    """BF16 cell matrices and inputs, FP32 accumulation: error at the BF16
    rounding level (2^-8 relative per operand), and clearly not FP32-exact."""
    prob, ref, op = small
    X = prob.initial_guess(N=64, seed_offset=3)
    Xd = torch.tensor(X, device="cuda", dtype=torch.float32)
    Yd = torch.zeros_like(Xd)
    op.HX(Xd, False, 1.0, Yd, bf16=True)
    err = _rel(Yd.double().cpu().numpy(), ref.apply(X.astype(np.float32).astype(np.float64)))
    print(f"bf16 HX max-norm relative error {err:.3e}")
    assert 1e-5 < err < 3e-2


@pytest.mark.parametrize("B", [64, 72, 8])
def test_hx_bf16_ring_matches_register_kernel(small, B):
    """K = 32 BF16 ring (variants 0, 2-5: stage, wave and column counts) against the
    register-staged K = 16 form (variant 1): FP32 accumulation rounding apart.
    Cells here hold 16 projector slots (one 16-wide slice), so every stage
    count is odd and the last stage's second half is the zero half."""
    from dftfe_amd import lib as L

    prob, ref, op = small
    X = prob.initial_guess(N=B, seed_offset=B + 11)
    Xd = torch.tensor(X, device="cuda", dtype=torch.float32)
    out = {}
    for v in (0, 1, 2, 3, 4, 5):
        Yd = torch.zeros_like(Xd)
        op.set_hx_variant(L.BF16, v)
        try:
            op.HX(Xd, False, 1.0, Yd, bf16=True)
        finally:
            op.set_hx_variant(L.BF16, 0)
        out[v] = Yd.double().cpu().numpy()
    for v in (0, 2, 3, 4, 5):
        assert np.isfinite(out[v]).all()
        assert _rel(out[v], out[1]) < 1e-5, v


def test_filter_mixed_with_fp64_tail(small):
    """FP32 / BF16 filters, optionally finished with FP64 degrees: eigenvalue
    differences to the FP64 oracle after one pass stay at the precision's level
    (measured: FP32 3.1e-8, FP32 + 4 FP64 degrees 2.9e-8, BF16 5.6e-4,
    BF16 + 6 FP64 degrees 6.1e-4 Ha: rounding in the undamped low end of the
    spectrum accumulates over the whole recurrence, so a short FP64 tail does
    not remove it)."""
    from dftfe_amd.chfsi import ChebyshevSolverDevice

    prob, ref, op = small
    lo, up, _, _ = O.lanczos_bounds(ref, prob.lanczos_start, 10)
    bounds = O.first_call_bounds(lo, up, 64, prob.M)
    X = prob.initial_guess()
    lam_o, _, _ = _oracle_pass(prob, ref, X, bounds, 20)
    errs = {}
    for prec, tail in [(1, 0), (1, 4), (2, 0), (2, 6)]:
        solver = ChebyshevSolverDevice(*bounds, cheby_block_size=64, cheby_degree=20,
                                       reuse_lanczos_upper=True, filter_precision=prec,
                                       fp64_tail_degrees=tail)
        lam, _, _ = solver.solve(op, torch.tensor(X, device="cuda"), use_mixed_prec_overall=True)
        errs[(prec, tail)] = np.abs(lam - lam_o).max()
    print("max|dlambda| vs FP64 oracle by (precision, fp64 tail):", errs)
    assert errs[(1, 0)] < 1e-6 and errs[(1, 4)] < 1e-6
    assert errs[(2, 0)] < 1e-2 and errs[(2, 6)] < 1e-2


def test_hx_symmetric(small):
    """<Y, H~ X> = <H~ Y, X> on the device operator."""
    prob, ref, op = small
    X = torch.tensor(prob.initial_guess(N=8, seed_offset=1), device="cuda")
    Y = torch.tensor(prob.initial_guess(N=8, seed_offset=2), device="cuda")
    a = (Y * op.apply(X)).sum().item()
    b = (op.apply(Y) * X).sum().item()
    assert abs(a - b) < 1e-11 * abs(a)


def test_lanczos_bounds_match_oracle(small):
    from dftfe_amd.chfsi import lanczos_bounds

    prob, ref, op = small
    v0 = prob.lanczos_start
    lo, up = lanczos_bounds(op, torch.tensor(v0, device="cuda"), 10)
    olo, oup, theta, fnorm = O.lanczos_bounds(ref, v0, 10)
    assert (lo, up) == (olo, oup)


@pytest.mark.parametrize("name", ["cube20_p4_n64", "cube18_p6_n64"])
def test_single_column_symmetric_read(name):
    """The one-column kernel reads only the slices on and below the diagonal of
    device-built (exactly symmetric) cell matrices; after set_cell_matrices
    (caller's matrices, not assumed symmetric) it reads them whole.  Both give
    the oracle's H~x (FP64 1e-13, FP32 2e-6) and agree to rounding, at p = 4
    and p = 6."""
    _dev()
    from dftfe_amd.chfsi import KohnShamOperatorDevice

    prob = fe.build_workload(name)
    ref = O.OracleOperator(prob)
    op = KohnShamOperatorDevice(prob)
    X = prob.initial_guess(N=1, seed_offset=91)
    out = {}
    for form in ("sym", "full"):
        if form == "full":
            op.set_cell_matrices(op.cell_matrices())
        for dt, tol in ((torch.float64, 1e-13), (torch.float32, 2e-6)):
            Xd = torch.tensor(X, device="cuda", dtype=dt)
            Yd = torch.zeros_like(Xd)
            op.HX(Xd, False, 1.0, Yd)
            want = ref.apply(X if dt == torch.float64 else X.astype(np.float32).astype(np.float64))
            y = Yd.double().cpu().numpy()
            assert _rel(y, want) < tol, (form, dt)
            out[(form, dt)] = y
    assert _rel(out[("sym", torch.float64)], out[("full", torch.float64)]) < 1e-14
    assert _rel(out[("sym", torch.float32)], out[("full", torch.float32)]) < 1e-6


def test_hx_fp32_single_column_matches_oracle(small):
    """One column takes the single-column kernel (cell matrices streamed as a
    matrix-vector product) in FP32 too."""
    prob, ref, op = small
    X = prob.initial_guess(N=1, seed_offset=3)
    Xd = torch.tensor(X, device="cuda", dtype=torch.float32)
    Yd = torch.zeros_like(Xd)
    op.HX(Xd, False, 1.0, Yd)
    want = ref.apply(X.astype(np.float32).astype(np.float64))
    assert _rel(Yd.double().cpu().numpy(), want) < 2e-6


@pytest.mark.parametrize("N", [64, 1])
def test_filter_fp64_matches_oracle(small, N):
    """N = 1 runs the whole recurrence (scaled apply, first-touch Chebyshev
    combine with the source term) through the single-column kernel."""
    from dftfe_amd.chfsi import chebyshev_filter

    prob, ref, op = small
    X = prob.sqrt_mass[:, None] * prob.initial_guess(N=N)
    a, b, a0 = 5.0, 41.0, 0.0
    Xd = torch.tensor(X, device="cuda")
    chebyshev_filter(op, Xd, 20, a, b, a0)
    want = O.chebyshev_filter(lambda s, f, sc, d: ref.HX(s, f, sc, d), X, 20, a, b, a0)
    assert _rel(Xd.cpu().numpy(), want) < 1e-11


def test_cholgs_and_rr(small):
    from dftfe_amd.chfsi import cholgs, rayleigh_ritz, residual_norms

    prob, ref, op = small
    X = prob.sqrt_mass[:, None] * prob.initial_guess()
    Xd = torch.tensor(X, device="cuda")
    cholgs(op, Xd)
    Xo = O.cholgs(X)
    G = Xd.T @ Xd
    assert torch.allclose(G, torch.eye(64, device="cuda", dtype=torch.float64), atol=1e-12)
    assert _rel(Xd.cpu().numpy(), Xo) < 1e-10
    lam = rayleigh_ritz(op, Xd, 64)
    lam_o, Xo2 = O.rayleigh_ritz(ref, Xo, 64)
    assert np.abs(lam - lam_o).max() < 1e-10
    res = residual_norms(op, Xd, lam, 64)
    res_o = O.residual_norms(ref, Xo2, lam_o, 64)
    assert np.abs(res - res_o).max() < 1e-9


def _oracle_pass(prob, ref, X, bounds, m):
    return O.chfsi_pass(ref, X, bounds[0], bounds[1], bounds[2], m, 64)


def test_pass_fp64_matches_oracle(small):
    from dftfe_amd.chfsi import ChebyshevSolverDevice

    prob, ref, op = small
    lo, up, _, _ = O.lanczos_bounds(ref, prob.lanczos_start, 10)
    bounds = O.first_call_bounds(lo, up, 64, prob.M)
    X = prob.initial_guess()
    lam_o, res_o, X_o = _oracle_pass(prob, ref, X, bounds, 20)
    solver = ChebyshevSolverDevice(*bounds, cheby_block_size=64, cheby_degree=20,
                                   lanczos_steps=10, reuse_lanczos_upper=True)
    Xd = torch.tensor(X, device="cuda")
    lam, res, bup = solver.solve(op, Xd)
    assert np.abs(lam - lam_o).max() <= 1e-10
    assert np.abs(res - res_o).max() <= 1e-8
    # first filtering call: Lanczos on the device reproduces the oracle bounds
    solver2 = ChebyshevSolverDevice(cheby_block_size=64, cheby_degree=20, lanczos_steps=10)
    solver2.set_lanczos_start(torch.tensor(prob.lanczos_start, device="cuda"))
    Xd2 = torch.tensor(X, device="cuda")
    lam2, _, bup2 = solver2.solve(op, Xd2, is_first_filtering_call=True)
    assert solver2.bounds() == pytest.approx(bounds, abs=1e-12)
    assert np.abs(lam2 - lam_o).max() <= 1e-10


def test_pass_mixed_fp32(small):
    """FP32 filter, FP64 CholGS/RR/residual: eigenvalue error vs the FP64 oracle."""
    from dftfe_amd.chfsi import ChebyshevSolverDevice

    prob, ref, op = small
    lo, up, _, _ = O.lanczos_bounds(ref, prob.lanczos_start, 10)
    bounds = O.first_call_bounds(lo, up, 64, prob.M)
    X = prob.initial_guess()
    lam_o, res_o, _ = _oracle_pass(prob, ref, X, bounds, 20)
    solver = ChebyshevSolverDevice(*bounds, cheby_block_size=64, cheby_degree=20,
                                   reuse_lanczos_upper=True)
    Xd = torch.tensor(X, device="cuda")
    lam, res, _ = solver.solve(op, Xd, use_mixed_prec_overall=True)
    err = np.abs(lam - lam_o).max()
    print(f"mixed-precision single pass max|dlambda| = {err:.3e} Ha")
    assert err < 1e-7


def test_pass_blocks_and_mixed_dense(small):
    """N = 64 in blocks of 16: the block-triangular overlap / rotation match the
    oracle (FP64); the mixed-precision dense steps (FP32 off-diagonal blocks,
    FP64 filter) stay within FP32 rounding of it."""
    from dftfe_amd import lib as L
    from dftfe_amd.chfsi import ChebyshevSolverDevice

    prob, ref, op = small
    lo, up, _, _ = O.lanczos_bounds(ref, prob.lanczos_start, 10)
    bounds = O.first_call_bounds(lo, up, 64, prob.M)
    X = prob.initial_guess()
    lam_o, res_o, _ = O.chfsi_pass(ref, X, *bounds, 20, 16)
    for mixed in (False, True):
        solver = ChebyshevSolverDevice(*bounds, cheby_block_size=16, cheby_degree=20,
                                       reuse_lanczos_upper=True, filter_precision=L.FP64,
                                       mixed_dense=mixed)
        Xd = torch.tensor(X, device="cuda")
        lam, res, _ = solver.solve(op, Xd, use_mixed_prec_overall=True)
        err = float(np.abs(lam - lam_o).max())
        print(f"blocks of 16, mixed dense={mixed}: max|dlambda| = {err:.3e} Ha")
        assert err < (1e-6 if mixed else 1e-10)


def test_dealii_layout_matches_unique(small, small_dealii):
    """Duplicated periodic DOFs + constraints give the same HX on shared DOFs
    and the same Ritz values as the folded numbering."""
    prob_u, ref_u, op_u = small
    prob_d, ref_d, op_d = small_dealii
    X_d = prob_d.initial_guess(N=32, seed_offset=4)
    Xd = torch.tensor(X_d, device="cuda")
    Y = op_d.apply(Xd).cpu().numpy()
    assert _rel(Y, ref_d.apply(X_d)) < 1e-13
    assert np.abs(Y[prob_d.slave_rows]).max() == 0.0


def test_constraint_kernels(small_dealii):
    prob, ref, op = small_dealii
    X = prob.initial_guess(N=16, seed_offset=12)
    X[prob.slave_rows] = np.random.default_rng(0).normal(size=(len(prob.slave_rows), 16))
    Xd = torch.tensor(X, device="cuda")
    op.distribute(Xd)
    want = X.copy()
    ref.distribute(want)
    assert np.array_equal(Xd.cpu().numpy(), want)
    Xd = torch.tensor(X, device="cuda")
    op.distribute_slave_to_master(Xd)
    want = X.copy()
    ref.distribute_slave_to_master(want)
    assert _rel(Xd.cpu().numpy(), want) < 1e-15
    Xd = torch.tensor(X, device="cuda")
    op.set_zero(Xd)
    want = X.copy()
    ref.set_zero(want)
    assert np.array_equal(Xd.cpu().numpy(), want)


def test_not_spd_error(small):
    """A zero column makes the overlap singular: CHFSI_ERR_NOT_SPD, no crash."""
    from dftfe_amd.chfsi import cholgs
    from dftfe_amd.lib import ChfsiError

    prob, ref, op = small
    X = prob.initial_guess(N=8)
    X[:, 3] = 0.0
    Xd = torch.tensor(X, device="cuda")
    with pytest.raises(ChfsiError) as ei:
        cholgs(op, Xd)
    assert ei.value.status == 4

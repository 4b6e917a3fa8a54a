"""GPU parity at the production FE degrees on small meshes.

The HX, projection, cell-build and quadrature-product kernels are instantiated
per cell size: p=6 (343 nodes padded to 352, 22 row blocks, 8^3 Gauss points)
and p=5 (216 nodes, 14 row blocks, 6^3 points) are the instantiations behind
the 884,736-DOF and 216,000-DOF workloads.  These meshes (18 bohr, 6^3 cells,
8 wells with s+p projectors) run the same instantiations at sizes where the
oracle applies H~ in well under a second, so every entry point is compared
with the oracle directly.

Bars: cell matrices and FP64 HX 1e-13 relative (max-norm); FP32 HX 2e-6; FP64
filter 1e-11; density and V-adjoint 1e-12; filter backward 1e-11; Lanczos
bounds exact; pass eigenvalues <= 1e-10 Ha.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dftfe_amd import fe  # noqa: E402
from oracle import chfsi_oracle as O  # noqa: E402

SHAPES = ["cube18_p6_n64", "cube18_p5_n64"]


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module", params=SHAPES)
def shp(request):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dftfe_amd.chfsi import KohnShamOperatorDevice

    print(f"\nbuilding {request.param}", flush=True)
    prob = fe.build_workload(request.param)
    ref = O.OracleOperator(prob)
    op = KohnShamOperatorDevice(prob)
    lo, up, _, _ = O.lanczos_bounds(ref, prob.lanczos_start, 10)
    prob.bounds = O.first_call_bounds(lo, up, prob.wl.N, prob.M)
    return prob, ref, op


def test_cell_matrices(shp):
    prob, ref, op = shp
    assert prob.ne in (216, 343)
    H = op.cell_matrices().cpu().numpy()
    assert _rel(H, ref.H_e) < 1e-13


@pytest.mark.parametrize("B", [64, 40, 1, 200])
def test_hx_fp64(shp, B):
    prob, ref, op = shp
    X = prob.initial_guess(N=B, seed_offset=B)
    Xd = torch.tensor(X, device="cuda")
    Yd = torch.zeros_like(Xd)
    op.HX(Xd, False, 1.0, Yd)
    assert _rel(Yd.cpu().numpy(), ref.apply(X)) < 1e-13


def test_hx_tile_and_projection_variants(shp):
    """Every HX tile variant (FP64 0-4, FP32 0-11) and every projection variant
    (0-5 per precision) gives the oracle's H~X."""
    prob, ref, op = shp
    X = prob.initial_guess(N=72, seed_offset=21)
    want = {0: ref.apply(X), 1: ref.apply(X.astype(np.float32).astype(np.float64))}
    tol = {0: 1e-13, 1: 2e-6}
    bad = []
    for prec, nvar, setter, reset in ((0, 5, op.set_hx_variant, 0), (1, 12, op.set_hx_variant, 0),
                                      (0, 6, op.set_proj_variant, 0),
                                      (1, 6, op.set_proj_variant, 0)):
        dt = torch.float64 if prec == 0 else torch.float32
        for v in range(nvar):
            Xd = torch.tensor(X, device="cuda", dtype=dt)
            Yd = torch.zeros_like(Xd)
            setter(prec, v)
            try:
                op.HX(Xd, False, 1.0, Yd)
            finally:
                setter(prec, reset)
            err = _rel(Yd.double().cpu().numpy(), want[prec])
            if not err < tol[prec]:
                bad.append((setter.__name__, prec, v, err))
    assert not bad, bad


@pytest.mark.parametrize("variant", [5, 6, 11])
def test_hx_ring_variants_bitwise(shp, variant):
    """The direct-to-LDS FP32 ring variants add the same products in the same
    order as the register-staged variant 2: a degree-4 FP32 filter on 256
    columns gives bitwise-equal blocks."""
    from dftfe_amd import lib as L
    from dftfe_amd.chfsi import chebyshev_filter

    prob, ref, op = shp
    X = torch.tensor(prob.sqrt_mass[:, None] * prob.initial_guess(N=256, seed_offset=7),
                     device="cuda")
    out = {}
    for v in (2, variant):
        Xd = X.clone()
        op.set_hx_variant(L.FP32, v)
        try:
            chebyshev_filter(op, Xd, 4, prob.bounds[1], prob.bounds[2], prob.bounds[0],
                             precision=L.FP32)
        finally:
            op.set_hx_variant(L.FP32, 0)
        out[v] = Xd
    assert torch.isfinite(out[2]).all()
    assert torch.equal(out[2], out[variant])


def test_hx_fp32_and_bf16(shp):
    prob, ref, op = shp
    X = prob.initial_guess(N=64, seed_offset=3)
    want = ref.apply(X.astype(np.float32).astype(np.float64))
    Xd = torch.tensor(X, device="cuda", dtype=torch.float32)
    Yd = torch.zeros_like(Xd)
    op.HX(Xd, False, 1.0, Yd)
    assert _rel(Yd.double().cpu().numpy(), want) < 2e-6
    Yd.zero_()
    op.HX(Xd, False, 1.0, Yd, bf16=True)
    err = _rel(Yd.double().cpu().numpy(), want)
    assert 1e-5 < err < 3e-2
    X1 = Xd[:, :1].contiguous()
    Y1 = torch.zeros_like(X1)
    op.HX(X1, False, 1.0, Y1)
    assert _rel(Y1.double().cpu().numpy(), ref.apply(X[:, :1].astype(np.float32).astype(np.float64))) < 2e-6


@pytest.mark.parametrize("B", [64, 200])
def test_hx_bf16_ring_matches_register_kernel(shp, B):
    """The K = 32 BF16 ring (variants 0, 2-5) and the register-staged K = 16
    kernel (variant 1) round the same operands to BF16 and accumulate in FP32
    in different orders: equal to FP32 accumulation rounding, and all at the
    BF16 level of the oracle."""
    from dftfe_amd import lib as L

    prob, ref, op = shp
    X = prob.initial_guess(N=B, seed_offset=B + 9)
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
    want = ref.apply(X.astype(np.float32).astype(np.float64))
    for v in (0, 2, 3, 4, 5):
        assert np.isfinite(out[v]).all()
        assert _rel(out[v], out[1]) < 1e-5, v
        assert 1e-5 < _rel(out[v], want) < 3e-2, v


def test_hx_strided_flags_and_cheby(shp):
    prob, ref, op = shp
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
    Xd = torch.tensor(X[:, :64].copy(), device="cuda")
    Dd = torch.tensor(D[:, :64].copy(), device="cuda")
    op.HXCheby(Xd, Dd)
    assert _rel(Dd.cpu().numpy(), D[:, :64] + ref.h_fe(X[:, :64])) < 1e-13


@pytest.mark.parametrize("N", [32, 1])
def test_filter_fp64(shp, N):
    """The workload's own degree (30 at p=6, 20 at p=5) on [bLow, bUp] from
    the oracle's Lanczos bounds."""
    from dftfe_amd.chfsi import chebyshev_filter

    prob, ref, op = shp
    a0, bLow, bUp = prob.bounds
    X = prob.sqrt_mass[:, None] * prob.initial_guess(N=N, seed_offset=40)
    Xd = torch.tensor(X, device="cuda")
    chebyshev_filter(op, Xd, prob.wl.m, bLow, bUp, a0)
    want = O.chebyshev_filter(ref.HX, X, prob.wl.m, bLow, bUp, a0)
    assert _rel(Xd.cpu().numpy(), want) < 1e-11


def test_lanczos_bounds(shp):
    from dftfe_amd.chfsi import lanczos_bounds

    prob, ref, op = shp
    v0 = prob.lanczos_start
    lo, up = lanczos_bounds(op, torch.tensor(v0, device="cuda"), 10)
    olo, oup, _, _ = O.lanczos_bounds(ref, v0, 10)
    assert (lo, up) == (olo, oup)


def test_pass_fp64(shp):
    """One pass of 64 states in blocks of 32 at the workload's degree, then a
    second pass with the caller's bound update, against the oracle."""
    from dftfe_amd.chfsi import ChebyshevSolverDevice

    prob, ref, op = shp
    wl = prob.wl
    bounds = prob.bounds
    X = prob.initial_guess()
    lam_o, res_o, X_o = O.chfsi_pass(ref, X, *bounds, wl.m, wl.B)
    solver = ChebyshevSolverDevice(*bounds, cheby_block_size=wl.B, cheby_degree=wl.m,
                                   reuse_lanczos_upper=True, filter_precision=0)
    Xd = torch.tensor(X, device="cuda")
    lam, res, _ = solver.solve(op, Xd)
    assert np.abs(lam - lam_o).max() <= 1e-10
    assert np.abs(res - res_o).max() <= 1e-8
    a0, bLow = O.next_bounds(lam_o)
    lam_o2, res_o2, _ = O.chfsi_pass(ref, X_o, a0, bLow, bounds[2], wl.m, wl.B)
    solver.reinitSpectrumBounds(*O.next_bounds(lam), bounds[2])
    lam2, res2, _ = solver.solve(op, Xd)
    assert np.abs(lam2 - lam_o2).max() <= 1e-10
    assert np.abs(res2 - res_o2).max() <= 1e-8


def test_density_and_adjoint(shp):
    prob, ref, op = shp
    X = prob.initial_guess(N=40, seed_offset=50)
    w = np.linspace(0.5, 2.0, 40)
    rho = op.density(torch.tensor(X, device="cuda"), torch.tensor(w)).cpu().numpy()
    assert _rel(rho, O.density_at_quadrature(prob, X, w, spin_factor=1.0)) < 1e-12
    Yb = prob.initial_guess(N=40, seed_offset=52)
    g = op.hx_adjoint_v(torch.tensor(Yb, device="cuda"), torch.tensor(X, device="cuda"), 0.3)
    assert _rel(g.cpu().numpy(), O.hx_adjoint_v(ref, Yb, X, 0.3)) < 1e-12


def test_filter_backward(shp):
    prob, ref, op = shp
    a0, a, b = prob.bounds
    m = 7
    X0 = prob.sqrt_mass[:, None] * prob.initial_guess(N=24, seed_offset=53)
    Yb = prob.initial_guess(N=24, seed_offset=54)
    Xbar, g = op.filter_backward(torch.tensor(X0, device="cuda"), torch.tensor(Yb, device="cuda"),
                                 m, a, b, a0, checkpoint_every=3)
    Xo, go = O.filter_backward(ref, X0, Yb, m, a, b, a0)
    assert _rel(Xbar.cpu().numpy(), Xo) < 1e-11
    assert _rel(g.cpu().numpy(), go) < 1e-11


def test_gga_build(shp):
    prob, ref, op = shp
    rng = np.random.default_rng(11)
    V = prob.V_q + 0.1 * rng.standard_normal(prob.V_q.shape)
    g = 0.05 * rng.standard_normal((prob.C, prob.nqc, 3)) * prob.JxW[None, :, None]
    try:
        op.set_potential_gga(V * prob.JxW[None, :], g)
        H = op.cell_matrices().cpu().numpy()
        assert _rel(H, O.cell_hamiltonians(prob, V_q=V, gJxW=g)) < 1e-13
        op.set_potential(V * prob.JxW[None, :])
        H = op.cell_matrices().cpu().numpy()
        assert _rel(H, O.cell_hamiltonians(prob, V_q=V)) < 1e-13
    finally:
        op.set_potential(prob.vJxW)


def test_dealii_layout(shp):
    """Duplicated periodic DOFs with slave -> master constraints at this degree."""
    from dftfe_amd.chfsi import KohnShamOperatorDevice

    prob, _, _ = shp
    pd = fe.build_workload(prob.wl.name, mode="dealii")
    rd = O.OracleOperator(pd)
    od = KohnShamOperatorDevice(pd)
    X = pd.initial_guess(N=32, seed_offset=4)
    Y = od.apply(torch.tensor(X, device="cuda")).cpu().numpy()
    assert _rel(Y, rd.apply(X)) < 1e-13
    assert np.abs(Y[pd.slave_rows]).max() == 0.0

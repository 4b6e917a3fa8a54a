"""The no-allocation HX entry point (chfsi_hx_noalloc) against chfsi_hx.

chfsi_hx_noalloc runs the same kernels as chfsi_hx with the projection scratch
in a fixed device array instead of operator-owned buffers, so its results are
the same bits.  Two calls in a row (the second on a narrower block, which
leaves the first call's data in the shared array) are each compared with
chfsi_hx through the C ABI; a block too wide for the fixed array is rejected.
"""
import ctypes as ct

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from dftfe_amd import fe  # noqa: E402


def test_hx_noalloc_matches_hx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dftfe_amd import lib as L
    from dftfe_amd.chfsi import KohnShamOperatorDevice

    print("chfsi_hx_noalloc: cube20_p4_n64, B = 96 then 40", flush=True)
    prob = fe.build_workload("cube20_p4_n64")
    op = KohnShamOperatorDevice(prob)
    lib = op.lib
    st = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    for B, scale_flag, scalar in ((96, 0, 1.0), (40, 1, -0.75)):
        X = torch.tensor(prob.initial_guess(N=B, seed_offset=B), device="cuda")
        D = torch.tensor(prob.initial_guess(N=B, seed_offset=B + 1), device="cuda")
        want = D.clone()
        got = D.clone()
        L.check(lib.chfsi_hx(op.handle, ct.c_void_p(X.data_ptr()), B,
                             ct.c_void_p(want.data_ptr()), B, B, scale_flag, scalar, st))
        L.check(lib.chfsi_hx_noalloc(op.handle, ct.c_void_p(X.data_ptr()), B,
                                     ct.c_void_p(got.data_ptr()), B, B, scale_flag, scalar, st))
        torch.cuda.synchronize()
        assert torch.isfinite(got).all()
        assert torch.equal(got, want), float((got - want).abs().max())
    # a block whose projection scratch exceeds the fixed array is refused
    assert len(prob.pair_cell) > 0  # the size check precedes every launch only with projectors
    Bbig = 1 << 14
    Xb = torch.zeros((prob.M, 8), device="cuda", dtype=torch.float64)
    st_code = lib.chfsi_hx_noalloc(op.handle, ct.c_void_p(Xb.data_ptr()), Bbig,
                                   ct.c_void_p(Xb.data_ptr()), Bbig, Bbig, 0, 1.0, st)
    assert st_code == 1 and b"fixed scratch" in lib.chfsi_last_error()

#!/usr/bin/env python
"""ChFSI pass benchmark on one MI355X.

Workload (default cube30_p6_n1024): periodic 30 bohr cube, 16^3 cells, FE
degree p=6 (343x343 cell Hamiltonians, 884,736 DOFs), Gauss 8^3 quadrature,
108 Gaussian wells with s+p projectors, 1024 FP64 wavefunctions in blocks of
256, Chebyshev degree 30, Lanczos bounds from 10 steps on the first call.

A step is one ChFSI pass (filter + CholGS + Rayleigh-Ritz + residuals).  The
timed region runs K passes in the mixed-precision form (FP32 filter on MFMA
in defect form, FP64 CholGS / RR / residuals) and then K passes of the plain
FP64 form from the same start, with the reference's bound bookkeeping between
passes (a0 = lambda_0, bLow = lambda_{N-1}, bUp refreshed by Lanczos with the
full |f| margin: bUp = ceil(theta_max + |f|), the reference's
reproducible-output margin, because the default |f|/10 after 10 Lanczos
steps leaves bUp (315) below the top of this operator's spectrum (315.2), and
the filter then amplifies that eigenvector every pass).

A convergence record (default 20 passes, untimed) runs FP64, the timed mixed
form and the plain FP32 filter side by side from the same start and reports,
per pass, the eigenvalue, density and band-energy differences to FP64.

Prints one JSON line: value = mixed-precision pass throughput in algorithmic
GFLOP/s (filter + dense flops of one pass / pass time), plus pass times, HX
GFLOP/s, roofline of the fused HX and the CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
PEAK_HBM_GBPS = 8000.0
# mixed passes: the columns of the top quarter of the states keep an FP64
# filter (their filter gain is the smallest, so the FP32 rounding of the
# defect form is largest relative to them); -1 = N / 4
FP64_TOP_DEFAULT = -1
METRIC = "ChFSI pass time + HX GFLOP/s, mixed vs FP64 GPU at 1e-8 Ha eigval error"


def hx_flops(prob, B):
    """Algorithmic flops of one block HX: dense cell GEMMs + C^T X and C (V C^T X)."""
    return 2.0 * prob.C * prob.ne ** 2 * B + 16.0 * len(prob.pair_cell) * prob.ne * B


def hx_bytes(prob, B, s):
    """Algorithmic HBM bytes of one fused block HX with element size s: read X,
    read the cell matrices, read+write Y, read the index map."""
    return prob.M * B * s + prob.C * prob.ne ** 2 * s + 2 * prob.M * B * s + prob.C * prob.ne * 4


def measured_hx_traffic(workload, width, variant, proj_variant, precision="fp32"):
    """HBM bytes of one fused HX on `width` columns (8 colour launches of the
    given tile variant + the given projection variant + reduction) from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE table committed under profiles/ for this
    workload and width; None when no measurement covers the configuration."""
    path = os.path.join(ROOT, "profiles", f"hx_traffic_{workload}_w{width}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)[precision]
    tile = t["tile_variants"].get(str(variant))
    proj = t.get("projection_variants", {}).get(str(proj_variant))
    if tile is None or proj is None:
        return None
    n = t["colour_launches_per_hx"]
    return int(n * (tile["fetch_bytes"] + tile["write_bytes"]) + proj["fetch_bytes"]
               + proj["write_bytes"] + t["reduction"]["fetch_bytes"]
               + t["reduction"]["write_bytes"])


def dense_without_hx(phases, hx64_ms):
    """CholGS + Rayleigh-Ritz + residual of one pass less the two FP64 HX on
    all columns that Rayleigh-Ritz and the residual apply (ms)."""
    if "cholgs" not in phases:
        return None
    return round(phases["cholgs"] + phases["rayleigh_ritz"] + phases["residual"] - 2 * hx64_ms, 2)


def pass_flops(prob, N, B, m, lanczos_steps):
    """Algorithmic flops of one pass: filter, Lanczos, CholGS (syrk + trsm), RR
    (N/B HX + lower X^T HX + rotation + eig), residual (N/B HX)."""
    M = prob.M
    nblk = -(-N // B)
    f = m * nblk * hx_flops(prob, B)
    f += lanczos_steps * hx_flops(prob, 1)
    f += 2.0 * M * N * N                       # syrk (lower) + trsm
    f += nblk * hx_flops(prob, B) + 1.0 * M * N * N + 2.0 * M * N * N + 9.0 * N ** 3
    f += nblk * hx_flops(prob, B) + 3.0 * M * N
    return f


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed ChFSI passes per precision")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cube30_p6_n1024")
    ap.add_argument("--hx-reps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 passes")
    ap.add_argument("--no-kpoints", action="store_true",
                    help="skip the batched k-point filter measurement")
    ap.add_argument("--kpoint-reps", type=int, default=3)
    ap.add_argument("--no-diff", action="store_true",
                    help="skip the differentiation leg (24 bohr, 12^3 cells, p=5, 256 bands)")
    ap.add_argument("--converge-passes", type=int, default=20,
                    help="convergence record: FP64, the mixed form and the plain FP32 filter "
                         "side by side from the same start for this many passes, with per-pass "
                         "eigenvalue, density and band-energy differences to FP64 (untimed; "
                         "0 to skip)")
    ap.add_argument("--lanczos-margin", choices=("full", "tenth"), default="full",
                    help="Lanczos upper-bound margin of every pass: |f| (default) or |f|/10")
    ap.add_argument("--mixed-form", choices=("defect", "plain"), default="defect",
                    help="FP32 filter form of the timed mixed passes")
    ap.add_argument("--fp64-top", type=int, default=FP64_TOP_DEFAULT,
                    help="mixed passes: the top states' columns filtered in FP64 (-1: N/4)")
    ap.add_argument("--hx-only", action="store_true",
                    help="setup, tile autotune and the HX timing only (for counter runs), "
                         "plus a 1 GiB device copy that calibrates the traffic counters")
    ap.add_argument("--pass-only", action="store_true",
                    help="setup, tile autotune and three passes of the timed mixed form only, "
                         "the last between two 1-element fill kernels that mark it in a "
                         "kernel trace")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "sharding of the path is not built.", file=sys.stderr)
        sys.exit(2)

    import torch
    from dftfe_amd import fe
    from dftfe_amd import lib as L
    from dftfe_amd.chfsi import ChebyshevSolverDevice, KohnShamOperatorDevice

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    wl = fe.WORKLOADS[args.workload]
    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)  # noqa: E731

    t0 = time.time()
    prob = fe.build_problem(wl)
    op = KohnShamOperatorDevice(prob)
    torch.cuda.synchronize()
    log(f"setup {time.time() - t0:.1f}s: M={prob.M} cells={prob.C} ne={prob.ne} "
        f"pairs={len(prob.pair_cell)} N={wl.N} B={wl.B} m={wl.m}")

    N, B, m = wl.N, wl.B, wl.m
    if args.fp64_top < 0:
        args.fp64_top = N // 4
    # columns the solver filters in one recurrence (and applies H~ to at once in
    # Rayleigh-Ritz and the residuals): its default rule, whole blocks within a
    # 16 GiB scratch of 8 bytes per DOF and column
    W = min(N, max(B, ((16 << 30) // (8 * prob.M)) // B * B))
    tune = op.autotune(W)
    log(f"autotune (width {W}): {tune}")
    X0 = torch.tensor(prob.initial_guess(), device=dev)
    v0 = torch.tensor(prob.lanczos_start, device=dev)

    # ---- fused HX timing with HIP events on the stream the library launches on
    def time_hx(dtype, width, bf16=False):
        Xb = torch.randn(prob.M, width, device=dev, dtype=dtype)
        Yb = torch.zeros_like(Xb)
        for _ in range(3):
            op.HX(Xb, False, 1.0, Yb, bf16=bf16)
        st = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(args.hx_reps):
            op.HX(Xb, True, 1.0, Yb, bf16=bf16)
        e1.record(st)
        e1.synchronize()
        return e0.elapsed_time(e1) / args.hx_reps

    def time_hx_bf16():
        """BF16 apply on W columns: the LDS-ring kernels (variants 0, 2-5) and
        the register-staged one (variant 1); the fastest stays selected."""
        t = []
        for v in range(6):
            op.set_hx_variant(L.BF16, v)
            t.append(time_hx(torch.float32, W, bf16=True))
        op.set_hx_variant(L.BF16, int(np.argmin(t)))
        return t

    if args.hx_only:
        hx32_ms = time_hx(torch.float32, W)
        hx64_ms = time_hx(torch.float64, W)
        hxb_ms = time_hx_bf16()
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        F = hx_flops(prob, W)
        print(json.dumps({"width": W, "hx_fp32_ms": round(hx32_ms, 4),
                          "hx_fp64_ms": round(hx64_ms, 4),
                          "hx_bf16_ms_variants": [round(x, 4) for x in hxb_ms],
                          "hx_fp32_tflops": round(F / hx32_ms / 1e9, 2),
                          "hx_fp64_tflops": round(F / hx64_ms / 1e9, 2),
                          "calibration_copy_bytes": 4 << 28, "autotune": tune}), flush=True)
        return

    finals = {}

    full_margin = args.lanczos_margin == "full"
    head_defect = args.mixed_form == "defect"

    def make_solver(dense=False, defect=False, top=0, rocblas=0, shep=True):
        s_ = ChebyshevSolverDevice(cheby_block_size=B, cheby_degree=m,
                                   lanczos_steps=wl.lanczos_steps, mixed_dense=dense,
                                   mixed_prec_defect_filter=defect,
                                   lanczos_full_residual_margin=full_margin,
                                   mixed_prec_fp64_top_columns=top, dense_rocblas=rocblas,
                                   use_shep=shep)
        s_.set_lanczos_start(v0)
        return s_

    def run_passes(mixed: bool, npass: int, dense: bool = False, defect: bool = False,
                   label=None, top: int = 0, rocblas: int = 0, shep: bool = True,
                   precision=None):
        """npass passes from the same start with the reference's bookkeeping
        (a0 = lowest, bLow = top Ritz value, kohnShamEigenSolve.cc:483-498).
        With a label, the last Ritz vectors and values are kept for the error
        checks; precision: the filter precision of the mixed passes."""
        solver = make_solver(dense, defect, top, rocblas, shep)
        if precision is not None:
            solver.set_filter_precision(precision)
        X = X0.clone()
        lams, ress, phases = [], [], []
        torch.cuda.synchronize()
        t = time.perf_counter()
        for k in range(npass):
            lam, res, _ = solver.solve(op, X, is_first_filtering_call=(k == 0),
                                       use_mixed_prec_overall=mixed)
            a0, bLow = float(lam[0]), float(lam[-1])
            _, _, bUp = solver.bounds()
            solver.reinitSpectrumBounds(a0, bLow, bUp)
            lams.append(lam)
            ress.append(res)
            phases.append(dict(solver.last_timings_ms,
                               filter_precision=solver.last_filter_precision()))
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) / npass * 1e3
        if label:
            finals[label] = (X, lams[-1])
        return ms, lams, ress, phases

    if args.pass_only:
        solver = make_solver(False, head_defect, args.fp64_top)
        X = X0.clone()
        marker = torch.zeros(1, device=dev, dtype=torch.int32)
        for k in range(3):
            if k == 2:
                torch.cuda.synchronize()
                marker.fill_(1)
                t = time.perf_counter()
            lam, _, _ = solver.solve(op, X, is_first_filtering_call=(k == 0),
                                     use_mixed_prec_overall=True)
            _, _, bUp = solver.bounds()
            solver.reinitSpectrumBounds(float(lam[0]), float(lam[-1]), bUp)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) * 1e3
        marker.fill_(2)
        torch.cuda.synchronize()
        print(json.dumps({"mixed_pass_ms": round(ms, 2), "phase_ms": solver.last_timings_ms}),
              flush=True)
        return

    for _ in range(args.warmup):
        run_passes(True, 1, defect=head_defect, top=args.fp64_top)
        if not args.no_fp64:
            run_passes(False, 1)
    log("warmup done")

    mixed_ms, lam_mx, res_mx, ph_mx = run_passes(True, args.steps, defect=head_defect,
                                                 top=args.fp64_top, label="mixed")
    log(f"mixed ({args.mixed_form}, top {args.fp64_top} in FP64): {mixed_ms:.1f} ms/pass; "
        f"last max residual {res_mx[-1].max():.3e}")
    fp64_ms, lam_64, res_64, ph_64 = (None, None, None, None)
    if not args.no_fp64:
        fp64_ms, lam_64, res_64, ph_64 = run_passes(False, args.steps, label="fp64")
        log(f"fp64: {fp64_ms:.1f} ms/pass; last max residual {res_64[-1].max():.3e}")
    # the mixed form with the filter bulk in BF16 (the schedule's first passes)
    bf_ms, _, _, ph_bf = run_passes(True, 2, defect=head_defect, top=args.fp64_top,
                                    precision=L.BF16)
    log(f"mixed form with BF16 filter: {bf_ms:.1f} ms/pass; phases {ph_bf[-1]}")
    # the headline mixed passes with the dense FP64 steps on rocBLAS (Tensile
    # DGEMM per block) instead of the library's MFMA kernels, and with the
    # generalised Rayleigh-Ritz (SUBSPACE PROJ SHEP GPU = false: no CholGS,
    # one rotation by L^-T Q)
    rb_ms, lam_rb, _, ph_rb = run_passes(True, 2, defect=head_defect, top=args.fp64_top,
                                         rocblas=1)
    log(f"mixed, dense steps on rocBLAS: {rb_ms:.1f} ms/pass; phases {ph_rb[-1]}")
    hy_ms, _, _, ph_hy = run_passes(True, 2, defect=head_defect, top=args.fp64_top, rocblas=2)
    log(f"mixed, library grams + rocBLAS rotations: {hy_ms:.1f} ms/pass; phases {ph_hy[-1]}")
    gep_ms, lam_gep, _, ph_gep = run_passes(True, 2, defect=head_defect, top=args.fp64_top,
                                            shep=False)
    log(f"mixed, generalised Rayleigh-Ritz: {gep_ms:.1f} ms/pass; phases {ph_gep[-1]}")
    # the same mixed passes with the reference's opt-in mixed-precision dense
    # steps (FP32 off-diagonal blocks of the overlap, CholGS rotation, X^T H X
    # and Rayleigh-Ritz rotation)
    md_ms, lam_md, _, ph_md = run_passes(True, args.steps, dense=True)
    log(f"mixed + mixed dense: {md_ms:.1f} ms/pass")

    # the same mixed passes with the spectrum split: the top B Ritz pairs go to
    # a separate block (n_core = N - B), X is orthonormalised but not rotated,
    # and the next pass keeps a0 and takes bLow from the top eigenvalue.  The
    # filter only rescales columns when a0 changes, so the filtered subspace is
    # the full passes' one and the kept eigenvalues match the top B of the full
    # mixed passes up to the FP32 rounding of the filter, which differs between
    # the rotated and the unrotated block.
    def run_split_passes(npass: int):
        n_core = N - B
        solver = make_solver(defect=head_defect)
        X = X0.clone()
        Xf = torch.empty(prob.M, N - n_core, dtype=X.dtype, device=dev)
        lams, phases = [], []
        torch.cuda.synchronize()
        t = time.perf_counter()
        for k in range(npass):
            lam, _, _, _ = solver.solve_spectrum_split(op, X, n_core, Xf,
                                                       is_first_filtering_call=(k == 0),
                                                       use_mixed_prec_overall=True)
            a0, _, bUp = solver.bounds()
            solver.reinitSpectrumBounds(a0, float(lam[-1]), bUp)
            lams.append(lam)
            phases.append(solver.last_timings_ms)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / npass * 1e3, lams, phases, n_core

    # the mixed passes with the plain FP32 filter (no defect form, no FP64 columns)
    df_ms, lam_df, _, ph_df = run_passes(True, args.steps, defect=not head_defect,
                                         label="other")
    errors = density_band_errors(prob, op, finals, log) if "fp64" in finals else None
    finals.clear()
    log(f"mixed, {'plain' if head_defect else 'defect-form'} filter: {df_ms:.1f} ms/pass")

    conv = None
    if args.converge_passes > 0:
        conv = convergence_record(prob, op, X0, make_solver, args.converge_passes,
                                  head_defect, args.fp64_top, log)

    sp_ms, lam_sp, ph_sp, n_core = run_split_passes(args.steps)
    sp_err = [float(np.abs(a - b[n_core:]).max()) for a, b in zip(lam_sp, lam_mx)]
    log(f"mixed, spectrum split (top {N - n_core} kept): {sp_ms:.1f} ms/pass; "
        f"max|dlambda| vs full mixed {max(sp_err):.2e}")
    del X0

    hx32_ms = time_hx(torch.float32, W)
    hx64_ms = time_hx(torch.float64, W)
    F = hx_flops(prob, W)
    hx32_gf = F / hx32_ms / 1e6
    hx64_gf = F / hx64_ms / 1e6
    log(f"HX {W} columns: fp32 {hx32_ms:.3f} ms = {hx32_gf:.0f} GFLOP/s; "
        f"fp64 {hx64_ms:.3f} ms = {hx64_gf:.0f} GFLOP/s")
    hxb_ms = time_hx_bf16()
    hxb_best = min(hxb_ms)
    log(f"HX {W} columns bf16 variants (ms): {[round(x, 3) for x in hxb_ms]} "
        f"({hx32_ms / hxb_best:.2f}x fp32)")
    hx32_blk = time_hx(torch.float32, B)
    hx64_blk = time_hx(torch.float64, B)
    log(f"HX {B} columns: fp32 {hx32_blk:.3f} ms, fp64 {hx64_blk:.3f} ms")

    PF = pass_flops(prob, N, B, m, wl.lanczos_steps)
    value = PF / mixed_ms / 1e6

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed_ms, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 filter / fp64 CholGS,RR,residual",
        "data": "synthetic (seeded FE problem and start vectors)",
        "config": {"workload": args.workload, "dofs": prob.M, "cells": prob.C,
                   "fe_degree": wl.p, "wavefunctions": N, "block": B, "filter_columns": W,
                   "cheby_degree": m,
                   "atoms": wl.n_atoms, "passes": args.steps, "parallelism": "single GPU"},
        "mixed_form": {"filter": f"fp32 {args.mixed_form}", "fp64_top_columns": args.fp64_top,
                       "lanczos_margin": args.lanczos_margin},
        "mixed_pass_ms": round(mixed_ms, 2),
        "mixed_dense_pass_ms": round(md_ms, 2),
        "bf16_pass_ms": round(bf_ms, 2),
        "phase_ms_bf16_last": ph_bf[-1],
        "bf16_over_fp64_speedup": round(fp64_ms / bf_ms, 3) if fp64_ms else None,
        "dense_rocblas_pass_ms": round(rb_ms, 2),
        "phase_ms_dense_rocblas_last": ph_rb[-1],
        "max_eig_err_dense_rocblas_vs_mixed_Ha": [float(np.abs(a - b).max())
                                                  for a, b in zip(lam_rb, lam_mx)],
        "dense_hybrid_pass_ms": round(hy_ms, 2),
        "phase_ms_dense_hybrid_last": ph_hy[-1],
        "gep_pass_ms": round(gep_ms, 2),
        "phase_ms_gep_last": ph_gep[-1],
        "max_eig_err_gep_vs_mixed_Ha": [float(np.abs(a - b).max())
                                        for a, b in zip(lam_gep, lam_mx)],
        "dense_fp64_ms_without_hx": dense_without_hx(ph_mx[-1], hx64_ms),
        ("mixed_plain_pass_ms" if head_defect else "mixed_defect_pass_ms"): round(df_ms, 2),
        ("phase_ms_mixed_plain_last" if head_defect else "phase_ms_mixed_defect_last"): ph_df[-1],
        "phase_ms_mixed_dense_last": ph_md[-1],
        "spectrum_split_pass_ms": round(sp_ms, 2),
        "spectrum_split_kept": N - n_core,
        "phase_ms_spectrum_split_last": ph_sp[-1],
        "max_eig_err_spectrum_split_vs_mixed_top_Ha": sp_err,
        "fp64_pass_ms": round(fp64_ms, 2) if fp64_ms else None,
        "mixed_over_fp64_speedup": round(fp64_ms / mixed_ms, 3) if fp64_ms else None,
        "hx_columns": W,
        "hx_fp32_ms": round(hx32_ms, 4),
        "hx_fp32_gflops": round(hx32_gf, 1),
        "hx_fp64_ms": round(hx64_ms, 4),
        "hx_fp64_gflops": round(hx64_gf, 1),
        "hx_block_columns": B,
        "hx_fp32_ms_block": round(hx32_blk, 4),
        "hx_fp64_ms_block": round(hx64_blk, 4),
        "pass_gflop": round(PF / 1e9, 1),
        "hx_tile_autotune": tune,
        "phase_ms_mixed_last": ph_mx[-1],
        "phase_ms_fp64_last": ph_64[-1] if ph_64 else None,
        "max_residual_mixed": [float(r.max()) for r in res_mx],
    }
    if lam_64 is not None:
        out["max_eig_err_mixed_dense_vs_fp64_Ha"] = [float(np.abs(a - b).max())
                                                    for a, b in zip(lam_md, lam_64)]
        out["max_eig_err_mixed_vs_fp64_Ha"] = [float(np.abs(a - b).max())
                                              for a, b in zip(lam_mx, lam_64)]
        out["max_eig_err_mixed_plain_vs_fp64_Ha" if head_defect else
            "max_eig_err_mixed_defect_vs_fp64_Ha"] = [float(np.abs(a - b).max())
                                                      for a, b in zip(lam_df, lam_64)]
        out["max_residual_fp64"] = [float(r.max()) for r in res_64]
    pm = tune.get("fp32_proj_ms") or [0.0]
    proj32 = int(np.argmin(pm))  # the projection variant the autotuner kept
    proj64 = int(np.argmin(tune.get("fp64_proj_ms") or [0.0]))
    vb = int(np.argmin(hxb_ms))
    # BF16 apply: FP32 multivectors (read, read + write Y), BF16 cell matrices
    bf16_bytes = 3 * prob.M * W * 4 + prob.C * prob.ne ** 2 * 2 + prob.C * prob.ne * 4
    bf16_traffic = measured_hx_traffic(args.workload, W, vb, proj32, "bf16")
    out["roofline"] = {
        "kernel": f"fused FP32 HX on {W} columns (8 colour launches of tile variant "
                  f"{tune['fp32']} + projection variant {proj32} + reduction)",
        "bound": "mfma",
        "achieved": round(F / hx32_ms / 1e9, 2),
        "peak": PEAK_TFLOPS["fp32"],
        "unit": "TFLOP/s",
        "frac": round(F / hx32_ms / 1e9 / PEAK_TFLOPS["fp32"], 4),
        "traffic": measured_hx_traffic(args.workload, W, tune["fp32"], proj32),
        "traffic_source": f"profiles/hx_traffic_{args.workload}_w{W}.json, FP32 tile variant "
                          f"{tune['fp32']}, projection variant {proj32} "
                          f"(rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0)",
        "algorithmic_bytes": hx_bytes(prob, W, 4),
        "fp64": {"achieved": round(F / hx64_ms / 1e9, 2), "peak": PEAK_TFLOPS["fp64"],
                 "frac": round(F / hx64_ms / 1e9 / PEAK_TFLOPS["fp64"], 4),
                 "tile_variant": tune["fp64"], "projection_variant": proj64,
                 "traffic": measured_hx_traffic(args.workload, W, tune["fp64"], proj64, "fp64")},
        # the BF16 apply moves FP32 multivectors: its bound is HBM, not the
        # BF16 MFMA rate (the MFMA view is kept beside it)
        "bf16": {"bound": "hbm", "achieved": round(bf16_bytes / hxb_best / 1e6, 1),
                 "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                 "frac": round(bf16_bytes / hxb_best / 1e6 / PEAK_HBM_GBPS, 4),
                 "algorithmic_bytes": bf16_bytes, "traffic": bf16_traffic,
                 "traffic_GBps": round(bf16_traffic / hxb_best / 1e6, 1) if bf16_traffic else None,
                 "mfma_tflops": round(F / hxb_best / 1e9, 2), "mfma_peak": PEAK_TFLOPS["bf16"],
                 "ms_variants": [round(x, 4) for x in hxb_ms], "variant": vb,
                 "speedup_vs_fp32": round(hx32_ms / hxb_best, 3)},
    }

    if errors:
        out["density_band_energy_vs_fp64"] = errors
    out["cell_rebuild"] = cell_rebuild(prob, op, log)
    out["density"] = density_leg(prob, op, torch.tensor(prob.initial_guess(), device=dev), log)
    if not args.no_diff:
        out["differentiation"] = differentiation(log)
    if conv is not None:
        out["convergence"] = conv
    if not args.no_kpoints:
        out["kpoint_batch"] = kpoint_batch(args.kpoint_reps, log)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(prob, B, log)
    print(json.dumps(out), flush=True)


def convergence_record(prob, op, X0, make_solver, P, head_defect, fp64_top, log):
    """P passes of FP64, of the timed mixed form, of the plain FP32 filter and of
    the mixed form with BF16 filters in passes 1-2, side by side from the same
    start with the same bookkeeping.  Per pass and
    mixed form: max |dlambda| over all states and over the lowest 90%, the
    quadrature-weighted L2 norm of the density difference and the band-energy
    difference (Fermi-Dirac occupations of 0.8 N electron pairs at kT = 0.01 Ha
    from the FP64 eigenvalues, spin factor 2), against the FP64 passes."""
    import torch
    from dftfe_amd.autograd import fermi_occupations

    from dftfe_amd import lib as L

    N = X0.shape[1]
    nl = int(0.9 * N)
    # bf16_schedule: the mixed form with the bulk of the filter in BF16 in
    # passes 1-2 (BF16 cell matrices and inputs, FP32 accumulation), FP32 after
    forms = {"fp64": (False, make_solver()),
             "mixed": (True, make_solver(defect=head_defect, top=fp64_top)),
             "plain_fp32": (True, make_solver()),
             "bf16_schedule": (True, make_solver(defect=head_defect, top=fp64_top))}
    others = ("mixed", "plain_fp32", "bf16_schedule")
    Xs = {k: X0.clone() for k in forms}
    jxw = torch.tensor(prob.JxW, device="cuda")[None, :]
    rec = {"passes": P, "lowest": nl, "fp64_max_residual": [], "fp64_max_residual_lowest": []}
    for k in others:
        rec[k] = {"max_eig_diff_all_Ha": [], "max_eig_diff_lowest_Ha": [], "density_l2": [],
                  "band_energy_Ha": []}
    for p in range(P):
        lam = {}
        for k, (mixed, solver) in forms.items():
            if k == "bf16_schedule":
                solver.set_filter_precision(L.BF16 if p < 2 else L.FP32)
            l, r, _ = solver.solve(op, Xs[k], is_first_filtering_call=(p == 0),
                                   use_mixed_prec_overall=mixed)
            _, _, bUp = solver.bounds()
            solver.reinitSpectrumBounds(float(l[0]), float(l[-1]), bUp)
            lam[k] = l
            if k == "fp64":
                rec["fp64_max_residual"].append(float(r.max()))
                rec["fp64_max_residual_lowest"].append(float(r[:nl].max()))
        occ, _ = fermi_occupations(torch.tensor(lam["fp64"], device="cuda"), 0.8 * N, 0.01)
        w = (2.0 * occ).cpu()
        rho64 = op.density(Xs["fp64"], w)
        for k in others:
            d = op.density(Xs[k], w) - rho64
            r_ = rec[k]
            r_["max_eig_diff_all_Ha"].append(float(np.abs(lam[k] - lam["fp64"]).max()))
            r_["max_eig_diff_lowest_Ha"].append(float(np.abs(lam[k][:nl] - lam["fp64"][:nl]).max()))
            r_["density_l2"].append(float(torch.sqrt((d * d * jxw).sum())))
            r_["band_energy_Ha"].append(abs(float((w.numpy() * (lam[k] - lam["fp64"])).sum())))
    for k in others:
        e = rec[k]["max_eig_diff_all_Ha"]
        rho = rec[k]["density_l2"]
        ok = [i + 1 for i in range(P) if all(x <= 1e-8 for x in e[i:])]
        okr = [i + 1 for i in range(P) if all(x <= 1e-7 for x in rho[i:])]
        rec[k]["eig_1e-8_from_pass"] = ok[0] if ok else None
        rec[k]["density_1e-7_from_pass"] = okr[0] if okr else None
        log(f"convergence ({k}): max|dlambda| <= 1e-8 Ha from pass {rec[k]['eig_1e-8_from_pass']}, "
            f"density L2 <= 1e-7 from pass {rec[k]['density_1e-7_from_pass']}; last "
            f"{e[-1]:.2e} Ha / {rho[-1]:.2e}")
    return rec


def cell_rebuild(prob, op, log, reps=5):
    """Per-SCF rebuild of every cell Hamiltonian on the device from V_eff JxW
    (LDA) and from V_eff JxW plus the GGA vector (C x n_q x 3), timed with HIP
    events on the stream the library launches on; the LDA matrices of the
    workload are restored afterwards.  Flops: the executed MFMA work of the
    symmetric build (16 x 16 node-pair blocks with ib <= jb), 2 per pair and
    quadrature term, FP64."""
    import torch
    from dftfe_amd.chfsi import _ptr, _stream

    dev = torch.device("cuda:0")
    v = torch.tensor(prob.vJxW, device=dev)
    g = 0.01 * torch.randn(prob.C, prob.nqc, 3, dtype=torch.float64, device=dev) \
        * torch.tensor(prob.JxW, device=dev)[None, :, None]
    lib = op.lib
    calls = {
        "lda": lambda: lib.chfsi_operator_set_potential(op.handle, _ptr(v), _stream()),
        "gga": lambda: lib.chfsi_operator_set_potential_gga(op.handle, _ptr(v), _ptr(g),
                                                            _stream()),
    }
    res = {}
    for name, call in calls.items():
        for _ in range(2):
            if call() != 0:
                raise RuntimeError(f"cell rebuild ({name}) failed")
        st = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            call()
        e1.record(st)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        # MFMA work the kernel issues: the 16 x 16 node-pair blocks with
        # ib <= jb (the transposed blocks are written from the same sums),
        # 2 flops per pair and quadrature term (1 term per point for LDA, 4
        # for GGA); padding not counted
        nb = (prob.ne + 15) // 16
        pairs = nb * (nb + 1) // 2 * 256
        fl = 2 * pairs * prob.nqc * (1 if name == "lda" else 4) * prob.C
        tf = fl / ms / 1e9
        assert tf <= PEAK_TFLOPS["fp64"], f"{name} build rate {tf:.1f} above the FP64 peak"
        res[f"{name}_ms"] = round(ms, 3)
        res[f"{name}_fp64_tflops"] = round(tf, 2)
    op.set_potential(prob.vJxW)
    log(f"cell rebuild: {res}")
    return res


def density_band_errors(prob, op, finals, log):
    """After the timed passes: density and band energy of each mixed-precision
    form against the FP64 passes from the same start, with the same
    occupations (Fermi-Dirac, kT = 0.01 Ha, 0.8 N electron pairs, from the
    FP64 eigenvalues; spin factor 2).  Density error: quadrature-weighted L2
    norm of the difference, relative to the L2 norm of the FP64 density."""
    import torch
    from dftfe_amd.autograd import fermi_occupations

    X64, lam64 = finals["fp64"]
    N = X64.shape[1]
    occ, _ = fermi_occupations(torch.tensor(lam64, device="cuda"), 0.8 * N, 0.01)
    w = (2.0 * occ).cpu()
    jxw = torch.tensor(prob.JxW, device="cuda")[None, :]
    rho64 = op.density(X64, w)
    E64 = float((w.numpy() * lam64).sum())
    n64 = float(torch.sqrt((rho64 * rho64 * jxw).sum()))
    res = {}
    for label in ("mixed", "other"):
        if label not in finals:
            continue
        X, lam = finals[label]
        d = op.density(X, w) - rho64
        res[label] = {"density_l2": float(torch.sqrt((d * d * jxw).sum())),
                      "density_l2_rel": float(torch.sqrt((d * d * jxw).sum())) / n64,
                      "band_energy_Ha": abs(float((w.numpy() * lam).sum()) - E64)}
    log(f"errors vs fp64 after the timed passes: {res}")
    return res


def density_leg(prob, op, X, log, reps=5):
    """Density at the quadrature points from all N wavefunctions of the
    workload (computeRhoFromPSI's contraction, one interpolation per column:
    2 n_q n_e flops per cell and column, FP64 MFMA), timed with HIP events on
    the stream the library launches on."""
    import torch

    N = X.shape[1]
    w = torch.full((N,), 2.0, dtype=torch.float64)
    for _ in range(2):
        op.density(X, w)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        op.density(X, w)
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    fl = 2 * prob.C * prob.nqc * prob.ne * N
    res = {"columns": N, "ms": round(ms, 3), "fp64_tflops": round(fl / ms / 1e9, 2)}
    log(f"density: {res}")
    return res


def differentiation(log, reps=2):
    """Differentiation form on its own workload: periodic 24 bohr cube, 12^3
    cells, p=5 (216,000 DOFs), 32 wells, 256 bands in blocks of 256, degree from
    the workload, Fermi-Dirac occupations at kT = 0.01 Ha.  After two ChFSI
    passes, times the band energy's forward (filter with checkpoints every 5
    degrees, CholGS, projected Hamiltonian, eigenvalues) and its backward
    (checkpointed filter adjoint and the HX adjoint in V), FP64 filter and FP32
    filter with FP64 gradient accumulation, with HIP events on torch's stream
    (the library launches on it).  Reports the relative max-norm difference of
    the two gradients."""
    import torch
    from dftfe_amd import fe
    from dftfe_amd.autograd import band_energy, fermi_occupations
    from dftfe_amd.chfsi import ChebyshevSolverDevice, KohnShamOperatorDevice

    prob = fe.build_workload("cube24_p5_n256")
    wl = prob.wl
    op = KohnShamOperatorDevice(prob)
    solver = ChebyshevSolverDevice(cheby_block_size=wl.B, cheby_degree=wl.m, lanczos_steps=10)
    solver.set_lanczos_start(torch.tensor(prob.lanczos_start, device="cuda"))
    X = torch.tensor(prob.initial_guess(), device="cuda")
    for k in range(2):
        lam, _, _ = solver.solve(op, X, is_first_filtering_call=(k == 0))
        _, _, bUp = solver.bounds()
        solver.reinitSpectrumBounds(float(lam[0]), float(lam[-1]), bUp)
    occ, _ = fermi_occupations(torch.tensor(lam, device="cuda"), 0.8 * wl.N, 0.01)
    st = torch.cuda.current_stream()
    out = {"workload": "cube24_p5_n256", "dofs": prob.M, "bands": wl.N, "degree": wl.m,
           "checkpoint_every": 5}
    grads = {}
    for prec, name in ((0, "fp64"), (1, "fp32")):
        fwd, bwd = [], []
        for _ in range(reps + 1):
            V = torch.tensor(prob.V_q, device="cuda", requires_grad=True)
            e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            e0.record(st)
            E, _, _ = band_energy(op, V, X.clone(), solver.bounds(), wl.m, occ, wl.B,
                                  precision=prec)
            e1.record(st)
            E.backward()
            e2.record(st)
            e2.synchronize()
            fwd.append(e0.elapsed_time(e1))
            bwd.append(e1.elapsed_time(e2))
            grads[name] = V.grad.detach()
        out[f"{name}_forward_ms"] = round(float(np.median(fwd[1:])), 2)
        out[f"{name}_backward_ms"] = round(float(np.median(bwd[1:])), 2)
    out["grad_fp32_vs_fp64_rel"] = float((grads["fp32"] - grads["fp64"]).abs().max()
                                         / grads["fp64"].abs().max())
    log(f"differentiation: {out}")
    return out


def kpoint_batch(reps, log):
    """Batched k-points: periodic 20 bohr cube, 8^3 cells, p=4 (32,768 DOFs),
    8 wells, 16 seeded k-points in the first Brillouin zone, complex
    wavefunctions with band counts N_k uniform in {384..1024}, Chebyshev degree
    20.  Times one filter pass over every k-point three ways on the same stream:
    the 16 k-points one after the other with plain launches, all k-points in
    one batch with plain launches, and the batch replayed as one HIP graph."""
    import torch

    from dftfe_amd import fe
    from dftfe_amd import lib as L
    from dftfe_amd.chfsi import BatchFilter, KohnShamOperatorDevice, lanczos_bounds_complex

    prob = fe.build_workload("cube20_p4_n64")
    kpts, nvec = fe.kpoint_set(prob, n=16)
    op = KohnShamOperatorDevice(prob, kpoints=kpts)
    m = 20
    dev = "cuda"
    sq = torch.as_tensor(prob.sqrt_mass, device=dev)
    gen = torch.Generator(device=dev).manual_seed(5)
    Xs = [torch.randn(prob.M, int(n), dtype=torch.complex128, device=dev, generator=gen)
          * sq[:, None] for n in nvec]
    v0 = torch.as_tensor(prob.lanczos_start.astype(np.complex128), device=dev)
    bounds = []
    for k, n in enumerate(nvec):
        op.reinitkPointSpinIndex(k)
        lo, up = lanczos_bounds_complex(op, v0, 10)
        # first-call interval with the REPRODUCIBLE OUTPUT factor (10): the
        # default factor 200 exceeds bUp when N_k / M > 1/200
        bounds.append((lo + (up - lo) * n / prob.M * 10.0, up, lo))
    kidx = list(range(len(nvec)))
    tune = op.autotune_complex(Xs, kidx, bounds)
    log(f"complex tile autotune: {tune}")
    flops = m * sum(4.0 * hx_flops(prob, int(n)) for n in nvec)
    st = torch.cuda.current_stream()

    def timeit(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            fn()
        e1.record(st)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    res = {"workload": "cube20_p4, 16 k-points", "dofs": prob.M, "degree": m,
           "n_vec": [int(n) for n in nvec], "filter_gflop": round(flops / 1e9, 1),
           "autotune": tune}
    for prec, name in ((L.FP64, "complex_fp64"), (L.FP32, "complex_fp32")):
        # both projection forms (0: cell nodes, 1: owned DOF rows, the
        # default); the line reports the faster
        per = {}
        for pv in (0, 1):
            op.set_proj_variant_complex(pv)
            batch = BatchFilter(op, Xs, kidx, bounds, m, prec)
            single = [BatchFilter(op, [Xs[k]], [k], [bounds[k]], m, prec) for k in kidx]
            t_seq = timeit(lambda: [b.launch() for b in single])
            t_plain = timeit(batch.launch)
            t_graph = timeit(batch.run)
            del batch, single
            per[pv] = (t_seq, t_plain, t_graph)
        op.set_proj_variant_complex(1)
        pv = min(per, key=lambda v: per[v][2])
        t_seq, t_plain, t_graph = per[pv]
        log(f"k-point filter {name}: sequential {t_seq:.1f} ms, batched {t_plain:.1f} ms, "
            f"graph {t_graph:.1f} ms = {flops / t_graph / 1e9:.1f} TFLOP/s (projection "
            f"variant {pv}; graph with the cell form {per[0][2]:.1f} ms, owned rows "
            f"{per[1][2]:.1f} ms)")
        res[name] = {"sequential_ms": round(t_seq, 2), "batched_ms": round(t_plain, 2),
                     "graph_ms": round(t_graph, 2),
                     "graph_over_sequential_speedup": round(t_seq / t_graph, 3),
                     "graph_tflops": round(flops / t_graph / 1e9, 2),
                     "projection_variant": pv,
                     "graph_ms_by_projection_variant": [round(per[v][2], 2) for v in (0, 1)]}
    del op, Xs
    torch.cuda.empty_cache()
    return res


def cpu_baseline(prob, B, log):
    """The reference CPU ChFSI path restated in numpy (the oracle: per-cell
    batched GEMM HX, blocked filter, Cholesky-Gram-Schmidt, Rayleigh-Ritz,
    residuals; OpenBLAS FP64), timed on this host's threads:
      * value: one FP64 pass of the 20 bohr, 8^3-cell, p=4 workload (64 states,
        degree 20, 10 Lanczos steps; chebyshevOrthogonalizedSubspaceIterationSolver.cc:115-607),
        median of 3, as pass throughput (the same flop count as the GPU's
        pass metric);
      * beside it: two HX applications of one 256-column block of the
        benchmark workload itself."""
    import platform

    import threadpoolctl
    from dftfe_amd import fe
    from oracle import chfsi_oracle as O

    info = threadpoolctl.threadpool_info()
    blas = next((i for i in info if i.get("user_api") == "blas"), {})
    threads = max((i.get("num_threads", 1) for i in info), default=1)
    cpu = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    small = fe.build_workload("cube20_p4_n64")
    swl = small.wl
    ref_s = O.OracleOperator(small)
    X0 = small.initial_guess()
    times = []
    for _ in range(3):
        t = time.perf_counter()
        lo, up, _, _ = O.lanczos_bounds(ref_s, small.lanczos_start, swl.lanczos_steps)
        a0, bLow, bUp = O.first_call_bounds(lo, up, swl.N, small.M)
        O.chfsi_pass(ref_s, X0, a0, bLow, bUp, swl.m, swl.B)
        times.append(time.perf_counter() - t)
    pass_s = float(np.median(times))
    PF = pass_flops(small, swl.N, swl.B, swl.m, swl.lanczos_steps)
    log(f"cpu FP64 pass (20 bohr, p=4, 64 states): {pass_s * 1e3:.0f} ms "
        f"= {PF / pass_s / 1e9:.2f} GFLOP/s on {threads} threads")
    lam_o, _, _ = O.chfsi_pass(ref_s, X0, a0, bLow, bUp, swl.m, swl.B)

    # the same pass on the GPU (FP64, same start and bounds), median of 5
    import torch
    from dftfe_amd.chfsi import ChebyshevSolverDevice, KohnShamOperatorDevice

    op_s = KohnShamOperatorDevice(small)
    gt, lam_g = [], None
    for _ in range(6):
        solver = ChebyshevSolverDevice(a0, bLow, bUp, cheby_block_size=swl.B,
                                       cheby_degree=swl.m, reuse_lanczos_upper=True)
        Xd = torch.tensor(X0, device="cuda")
        torch.cuda.synchronize()
        t = time.perf_counter()
        lam_g, _, _ = solver.solve(op_s, Xd)
        torch.cuda.synchronize()
        gt.append(time.perf_counter() - t)
    gpu_s = float(np.median(gt[1:]))
    gpu_err = float(np.abs(lam_g - lam_o).max())
    log(f"gpu FP64 pass on the same workload: {gpu_s * 1e3:.2f} ms; max|dlambda| vs the CPU "
        f"pass {gpu_err:.1e} Ha")

    t = time.time()
    ref = O.OracleOperator(prob)
    log(f"cpu oracle setup {time.time() - t:.1f}s")
    X = prob.initial_guess(N=B, seed_offset=99)
    reps = 2
    t = time.perf_counter()
    for _ in range(reps):
        ref.h_fe(X)
    dt = (time.perf_counter() - t) / reps
    gf = hx_flops(prob, B) / dt / 1e9
    log(f"cpu HX {dt:.2f}s = {gf:.1f} GFLOP/s on {threads} threads")
    return {"value": round(PF / pass_s / 1e9, 3), "unit": "GFLOP/s (FP64 ChFSI pass)",
            "cores": threads, "kind": "port",
            "pass_ms": round(pass_s * 1e3, 1), "pass_gflop": round(PF / 1e9, 3),
            "gpu_fp64_pass_ms_same_workload": round(gpu_s * 1e3, 2),
            "gpu_vs_cpu_max_eig_diff_Ha": gpu_err,
            "cpu_model": cpu, "blas": f"{blas.get('internal_api', '?')} {blas.get('version', '')} "
                                      f"({blas.get('architecture', '')})".strip(),
            "hx_benchmark_workload_gflops": round(gf, 2), "hx_seconds": round(dt, 3),
            "sample": "one FP64 ChFSI pass (Lanczos 10 steps, degree-20 filter, CholGS, "
                      "Rayleigh-Ritz, residuals) of the 20 bohr / 8^3-cell / p=4 workload with "
                      "64 states, median of 3, through the numpy/OpenBLAS oracle; plus 2 HX "
                      f"applications (B={B}) of the benchmark workload"}


if __name__ == "__main__":
    main()

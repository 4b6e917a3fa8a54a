#!/usr/bin/env python
"""ChFSI pass benchmark on one MI355X.

Workload (default cube30_p6_n1024): periodic 30 bohr cube, 16^3 cells, FE
degree p=6 (343x343 cell Hamiltonians, 884,736 DOFs), Gauss 8^3 quadrature,
108 Gaussian wells with s+p projectors, 1024 FP64 wavefunctions in blocks of
256, Chebyshev degree 30, Lanczos bounds from 10 steps on the first call.

A step is one ChFSI pass (filter + CholGS + Rayleigh-Ritz + residuals).  The
timed region runs K passes in the mixed-precision form (FP32 filter on MFMA,
FP64 CholGS / RR / residuals) and then K passes of the plain FP64 form from
the same start, with the reference's bound bookkeeping between passes
(a0 = lambda_0, bLow = lambda_{N-1}, bUp refreshed by Lanczos).

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

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}
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
    ap.add_argument("--converge-passes", type=int, default=0,
                    help="also run FP64, mixed and mixed defect-form passes from the same start "
                         "this many times and report their eigenvalue differences (untimed)")
    ap.add_argument("--monotone-blow", action="store_true",
                    help="convergence leg: bLow never rises above its previous value")
    ap.add_argument("--hx-only", action="store_true",
                    help="setup, tile autotune and the HX timing only (for counter runs), "
                         "plus a 1 GiB device copy that calibrates the traffic counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "sharding of the path is not built.", file=sys.stderr)
        sys.exit(2)

    import torch
    from dftfe_amd import fe
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
    # columns the solver filters in one recurrence (and applies H~ to at once in
    # Rayleigh-Ritz and the residuals): its default rule, whole blocks within a
    # 16 GiB scratch of 8 bytes per DOF and column
    W = min(N, max(B, ((16 << 30) // (8 * prob.M)) // B * B))
    tune = op.autotune(W)
    log(f"autotune (width {W}): {tune}")
    X0 = torch.tensor(prob.initial_guess(), device=dev)
    v0 = torch.tensor(prob.lanczos_start, device=dev)

    # ---- fused HX timing with HIP events on the stream the library launches on
    def time_hx(dtype, width):
        Xb = torch.randn(prob.M, width, device=dev, dtype=dtype)
        Yb = torch.zeros_like(Xb)
        for _ in range(3):
            op.HX(Xb, False, 1.0, Yb)
        st = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(args.hx_reps):
            op.HX(Xb, True, 1.0, Yb)
        e1.record(st)
        e1.synchronize()
        return e0.elapsed_time(e1) / args.hx_reps

    if args.hx_only:
        hx32_ms = time_hx(torch.float32, W)
        hx64_ms = time_hx(torch.float64, W)
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        F = hx_flops(prob, W)
        print(json.dumps({"width": W, "hx_fp32_ms": round(hx32_ms, 4),
                          "hx_fp64_ms": round(hx64_ms, 4),
                          "hx_fp32_tflops": round(F / hx32_ms / 1e9, 2),
                          "hx_fp64_tflops": round(F / hx64_ms / 1e9, 2),
                          "calibration_copy_bytes": 4 << 28, "autotune": tune}), flush=True)
        return

    finals = {}

    def run_passes(mixed: bool, npass: int, dense: bool = False, defect: bool = False,
                   monotone: bool = False, label=None):
        """npass passes from the same start with the reference's bookkeeping
        (a0 = lowest, bLow = top Ritz value, kohnShamEigenSolve.cc:483-498), or
        with bLow never rising above its previous value (monotone).  With a
        label, the last Ritz vectors and values are kept for the error checks."""
        solver = ChebyshevSolverDevice(cheby_block_size=B, cheby_degree=m,
                                       lanczos_steps=wl.lanczos_steps, mixed_dense=dense,
                                       mixed_prec_defect_filter=defect)
        solver.set_lanczos_start(v0)
        X = X0.clone()
        lams, ress, phases = [], [], []
        torch.cuda.synchronize()
        t = time.perf_counter()
        for k in range(npass):
            lam, res, _ = solver.solve(op, X, is_first_filtering_call=(k == 0),
                                       use_mixed_prec_overall=mixed)
            a0, bLow = float(lam[0]), float(lam[-1])
            if monotone and k > 0:
                bLow = min(bLow, solver.bounds()[1])
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

    for _ in range(args.warmup):
        run_passes(True, 1)
        if not args.no_fp64:
            run_passes(False, 1)
    log("warmup done")

    mixed_ms, lam_mx, res_mx, ph_mx = run_passes(True, args.steps, label="mixed")
    log(f"mixed: {mixed_ms:.1f} ms/pass; last max residual {res_mx[-1].max():.3e}")
    fp64_ms, lam_64, res_64, ph_64 = (None, None, None, None)
    if not args.no_fp64:
        fp64_ms, lam_64, res_64, ph_64 = run_passes(False, args.steps, label="fp64")
        log(f"fp64: {fp64_ms:.1f} ms/pass; last max residual {res_64[-1].max():.3e}")
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
        solver = ChebyshevSolverDevice(cheby_block_size=B, cheby_degree=m,
                                       lanczos_steps=wl.lanczos_steps)
        solver.set_lanczos_start(v0)
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

    # the mixed passes with the FP32 filter in defect form (X p(theta) + E)
    df_ms, lam_df, res_df, ph_df = run_passes(True, args.steps, defect=True, label="defect")
    errors = density_band_errors(prob, op, finals, log) if "fp64" in finals else None
    finals.clear()
    log(f"mixed, defect-form filter: {df_ms:.1f} ms/pass")

    conv = None
    if args.converge_passes > 0:
        # the same start, P passes of FP64, of the plain FP32 filter and of its
        # defect form: how far the mixed eigenvalues sit from the FP64 ones as
        # the passes converge (all states, and the lowest 90%, the states an SCF
        # occupies; the top ones sit next to the filter's cut)
        P = args.converge_passes
        mono = args.monotone_blow
        _, lam_c64, res_c64, _ = run_passes(False, P, monotone=mono)
        nlow = int(0.9 * N)
        conv = {"passes": P, "lowest": nlow, "monotone_bLow": mono,
                "max_residual_fp64_all": [float(r.max()) for r in res_c64],
                "max_residual_fp64_lowest": [float(r[:nlow].max()) for r in res_c64]}
        for name, defect in (("plain", False), ("defect", True)):
            _, lam_c, res_c, _ = run_passes(True, P, defect=defect, monotone=mono)
            conv[name] = {
                "max_eig_diff_all_Ha": [float(np.abs(a - b).max()) for a, b in zip(lam_c, lam_c64)],
                "max_eig_diff_lowest_Ha": [float(np.abs(a[:nlow] - b[:nlow]).max())
                                           for a, b in zip(lam_c, lam_c64)],
                "max_residual_lowest": [float(r[:nlow].max()) for r in res_c],
                "max_residual_top": [float(r[nlow:].max()) for r in res_c]}
            log(f"converge ({name}): {P} passes; last max|dlambda| all "
                f"{conv[name]['max_eig_diff_all_Ha'][-1]:.2e}, lowest {nlow} "
                f"{conv[name]['max_eig_diff_lowest_Ha'][-1]:.2e}")

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
        "mixed_pass_ms": round(mixed_ms, 2),
        "mixed_dense_pass_ms": round(md_ms, 2),
        "mixed_defect_pass_ms": round(df_ms, 2),
        "phase_ms_mixed_defect_last": ph_df[-1],
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
        out["max_eig_err_mixed_defect_vs_fp64_Ha"] = [float(np.abs(a - b).max())
                                                     for a, b in zip(lam_df, lam_64)]
        out["max_residual_fp64"] = [float(r.max()) for r in res_64]
    pm = tune.get("fp32_proj_ms") or [0.0]
    proj32 = int(np.argmin(pm))  # the projection variant the autotuner kept
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
                 "frac": round(F / hx64_ms / 1e9 / PEAK_TFLOPS["fp64"], 4)},
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


def cell_rebuild(prob, op, log, reps=5):
    """Per-SCF rebuild of every cell Hamiltonian on the device from V_eff JxW
    (LDA) and from V_eff JxW plus the GGA vector (C x n_q x 3), timed with HIP
    events on the stream the library launches on; the LDA matrices of the
    workload are restored afterwards.  Flops: 2 n_e^2 n_q per cell (LDA),
    4 n_e^2 n_q (GGA); FP64 on the vector units."""
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
        fl = (2 if name == "lda" else 4) * prob.C * prob.ne ** 2 * prob.nqc
        res[f"{name}_ms"] = round(ms, 3)
        res[f"{name}_fp64_tflops"] = round(fl / ms / 1e9, 2)
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
    for label in ("mixed", "defect"):
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
        batch = BatchFilter(op, Xs, kidx, bounds, m, prec)
        single = [BatchFilter(op, [Xs[k]], [k], [bounds[k]], m, prec) for k in kidx]
        t_seq = timeit(lambda: [b.launch() for b in single])
        t_plain = timeit(batch.launch)
        t_graph = timeit(batch.run)
        del batch, single
        log(f"k-point filter {name}: sequential {t_seq:.1f} ms, batched {t_plain:.1f} ms, "
            f"graph {t_graph:.1f} ms = {flops / t_graph / 1e9:.1f} TFLOP/s")
        res[name] = {"sequential_ms": round(t_seq, 2), "batched_ms": round(t_plain, 2),
                     "graph_ms": round(t_graph, 2),
                     "graph_over_sequential_speedup": round(t_seq / t_graph, 3),
                     "graph_tflops": round(flops / t_graph / 1e9, 2)}
    del op, Xs
    torch.cuda.empty_cache()
    return res


def cpu_baseline(prob, B, log):
    """The FP64 oracle's HX (numpy / OpenBLAS, per-cell batched GEMM) on the
    same workload, one B-column block, timed on this host's cores."""
    import threadpoolctl
    from oracle import chfsi_oracle as O

    t = time.time()
    ref = O.OracleOperator(prob)
    log(f"cpu oracle setup {time.time() - t:.1f}s")
    X = prob.initial_guess(N=B, seed_offset=99)
    reps = 2
    t = time.perf_counter()
    for _ in range(reps):
        ref.h_fe(X)
    dt = (time.perf_counter() - t) / reps
    threads = max((i.get("num_threads", 1) for i in threadpoolctl.threadpool_info()), default=1)
    gf = hx_flops(prob, B) / dt / 1e9
    log(f"cpu HX {dt:.2f}s = {gf:.1f} GFLOP/s on {threads} threads")
    return {"value": round(gf, 2), "unit": "GFLOP/s (HX)", "cores": threads, "kind": "port",
            "hx_seconds": round(dt, 3),
            "sample": f"{reps} full HX applications (B={B}) of the same workload through the "
                      "numpy/OpenBLAS FP64 oracle (per-cell batched GEMM, gather/scatter, "
                      "projectors)"}


if __name__ == "__main__":
    main()

/*
 * chfsi.h -- C ABI of the MI355X (gfx950) Chebyshev-filtered subspace
 * iteration library (libchfsi.so).
 *
 * Drop-in boundary for one MPI rank's full-domain ChFSI pass of DFT-FE.
 * Every entry point below replaces one member of the reference's operator /
 * solver classes; the citation next to each declaration names it.
 *
 *   operatorDFTDeviceClass              include/operatorDevice.h:43-420
 *   kohnShamDFTOperatorDeviceClass      include/kohnShamDFTOperatorDevice.h:43-782
 *   chebyshevOrthogonalizedSubspaceIterationSolverDevice
 *                                       include/chebyshevOrthogonalizedSubspaceIterationSolverDevice.h:37-166
 *   linearAlgebraOperationsDevice::*    include/linearAlgebraOperationsDevice.h:68-402
 *   dftUtils::constraintMatrixInfoDevice include/constraintMatrixInfoDevice.h:36-211
 *
 * Conventions
 *   - Multivectors are node-major, X[node * ld + vec] (the reference's
 *     flattened layout), resident in device memory owned by the caller.
 *   - All calls are asynchronous on the given stream (hipStream_t passed as
 *     void*; NULL = the null stream).  Calls that return host values
 *     (eigenvalues, residual norms, bounds) synchronise that stream first.
 *   - Errors are returned as chfsi_status; chfsi_last_error() gives a
 *     thread-local message.  Nothing calls exit().
 *   - Handles are not thread-safe; use one handle per host thread.
 */
#ifndef CHFSI_H
#define CHFSI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  CHFSI_OK = 0,
  CHFSI_ERR_ARG = 1,      /* invalid argument or size */
  CHFSI_ERR_HIP = 2,      /* HIP runtime error */
  CHFSI_ERR_BLAS = 3,     /* rocBLAS / rocSOLVER error */
  CHFSI_ERR_NOT_SPD = 4,  /* overlap matrix not positive definite (pseudoGSDevice.cc:297-301) */
  CHFSI_ERR_EIG = 5,      /* dense eigensolver failed */
  CHFSI_ERR_STATE = 6,    /* call out of order (e.g. no potential set) */
  CHFSI_ERR_NOMEM = 7,    /* device allocation failed */
  CHFSI_ERR_NAN = 8       /* non-finite value detected */
} chfsi_status;

typedef enum {
  CHFSI_FP64 = 0, /* filter in FP64 (the drop-in path) */
  CHFSI_FP32 = 1, /* filter bulk in FP32 on MFMA, FP32 accumulation */
  CHFSI_BF16 = 2  /* BF16 cell matrices and inputs, FP32 accumulation */
} chfsi_precision;

typedef struct chfsi_operator chfsi_operator;
typedef struct chfsi_solver chfsi_solver;

/* Description of the discretised Kohn-Sham operator on one rank.  Host
 * pointers; copied to the device by chfsi_operator_create.  Mirrors what
 * kohnShamDFTOperatorDeviceClass::init/reinit gather from deal.II
 * (kohnShamDFTOperatorDevice.cc:492-933, :938-1031). */
typedef struct {
  int64_t n_dofs;            /* M: rows of the node-major multivector          */
  int32_t n_cells;           /* C                                              */
  int32_t nodes_per_cell;    /* n_e = (p+1)^3                                   */
  int32_t quad_per_cell;     /* n_q                                            */
  const int64_t *cell_map;   /* C*n_e: DOF of local node i of cell c
                                (vectorUtilities.cc:398-500)                   */
  const int8_t *cell_color;  /* C: colour >= 0; cells of one colour share no DOF.  A cell
                                overlapping more than 128 projector slots takes further
                                colours (j * n_colours + colour for its j-th extra group of
                                slots); all colours together must stay below 128 */
  const double *sqrt_mass;   /* M: M^{1/2}, 0 on constrained rows              */
  const double *inv_sqrt_mass; /* M: M^{-1/2}, 0 on constrained rows           */
  const double *shape_values;  /* n_q*n_e: N_i(q), row q                        */
  const double *stiffness;     /* n_e*n_e: int grad N_i . grad N_j (all cells)  */
  /* nonlocal projectors (computeNonLocalHamiltonianTimesXMemoryOptBatchGEMMDevice.cc) */
  int32_t n_proj;            /* total projectors                               */
  int32_t n_pairs;           /* (atom, cell) pairs in projector support         */
  int32_t proj_per_pair;     /* projectors per atom (4: s + 3 p)               */
  const int32_t *pair_cell;  /* n_pairs                                        */
  const int32_t *pair_proj0; /* n_pairs: first projector id of the pair's atom */
  const double *pair_C;      /* n_pairs*n_e*proj_per_pair: C_e[i][alpha]        */
  const double *proj_V;      /* n_proj: coupling constants                     */
  /* affine constraints, CSR over constrained rows (constraintMatrixInfoDevice.cc:446-542).
   * n_constrained == 0 means none.  Rows that are all single-master, unit
   * coefficient (periodic identities) are folded into the cell map; any other
   * row (hanging nodes: several masters, fractional coefficients) makes every
   * apply distribute before the cell products and send the constrained rows'
   * sums to their masters after them (kohnShamDFTOperatorDevice.cc:3806-3838).
   * Masters must not be constrained themselves. */
  int64_t n_constrained;
  const int64_t *constraint_rows;  /* n_constrained slave rows                  */
  const int64_t *constraint_ptr;   /* n_constrained+1                           */
  const int64_t *constraint_cols;  /* masters                                   */
  const double *constraint_vals;   /* coefficients                              */
  /* quadrature weights JxW per (cell, quadrature point), C*n_q; nullable when
   * neither the density nor the potential adjoint is used */
  const double *jxw;
  /* Bloch k-points (complex form; kohnShamDFTOperatorDeviceClass keeps one
   * set of cell matrices per k-point, d_cellHamiltonianMatrix[kpointSpinIndex],
   * hamiltonianMatrixCalculator.cc:562-872).  n_kpoints == 0: real operator only.
   * The operator acts on the periodic part u of psi = exp(i k.x) u:
   *   H_k,e = H_e + |k|^2/2 Mc - i sum_d k_d Gd,   nonlocal C_k V C_k^H. */
  int32_t n_kpoints;
  const double *kpoints;          /* 3*n_kpoints, Cartesian (bohr^-1)               */
  const double *cell_mass;        /* n_e*n_e: int N_i N_j (all cells)                */
  const double *cell_grad;        /* 3*n_e*n_e: [d][i][j] = int N_i dN_j/dx_d         */
  const double *pair_C_kpoints;   /* n_kpoints*n_pairs*n_e*proj_per_pair complex,
                                     interleaved (re, im): C_e at each k-point       */
  /* physical shape gradients at the quadrature points, n_q*n_e*3, [q][i][d]
   * (the reference-cell gradients times the inverse Jacobian of the Cartesian
   * cells, hamiltonianMatrixCalculatorFlattenedDevice.cc:415-427); nullable,
   * needed only by chfsi_operator_set_potential_gga */
  const double *shape_gradients;
  /* per-cell geometry (graded or non-Cartesian meshes); nullable.
   * cell_stiffness: C*n_e*n_e, each cell's int grad N_i . grad N_j (the
   * reference's cellShapeFunctionGradientIntegral,
   * hamiltonianMatrixCalculatorFlattenedDevice.cc:110-111); when set it
   * replaces `stiffness`.
   * inverse_jacobians with jacobian_kind 3 (C*3 diagonals, Cartesian cells)
   * or 9 (C*9 row-major matrices): shape_gradients are then the reference
   * cell's and the GGA build applies each cell's inverse Jacobian
   * (:402-427); jacobian_kind 0: shape_gradients are physical and shared. */
  const double *cell_stiffness;
  const double *inverse_jacobians;
  int32_t jacobian_kind;
} chfsi_operator_desc;

/* Batch of k-point blocks filtered together (complex form). */
typedef struct {
  int32_t n;                /* entries (at most 64)                                   */
  const int32_t *kpoint;    /* n: operator k-point of each entry                      */
  void *const *X;           /* n: device pointers, complex double node-major M x ld   */
  const int64_t *ld;        /* n: leading dimensions (complex elements)               */
  const int32_t *n_vec;     /* n: columns filtered (band counts N_k)                  */
  const double *a;          /* n: lower end of the damped interval (bLow)             */
  const double *b;          /* n: upper bound (bUp)                                   */
  const double *a0;         /* n: scaling point (lowest Ritz value)                   */
  int32_t degree;           /* Chebyshev degree m                                     */
  int32_t precision;        /* CHFSI_FP64: complex double; CHFSI_FP32: complex float
                               filter between one cast in and one cast out           */
} chfsi_kbatch;

typedef struct chfsi_batch chfsi_batch;

/* Solver parameters: the ChFSI knobs of dftParameters
 * (utils/dftParameters.cc:840-1021). */
typedef struct {
  int32_t cheby_block_size;        /* CHEBY WFC BLOCK SIZE                      */
  int32_t cheby_degree;            /* CHEBYSHEV POLYNOMIAL DEGREE (0: table)    */
  int32_t lanczos_steps;           /* reference: 20, or 40 when reproducible    */
  int32_t reproducible;            /* REPRODUCIBLE OUTPUT                       */
  int32_t reuse_lanczos_upper;     /* REUSE LANCZOS UPPER BOUND                 */
  double first_scf_degree_factor;  /* CHEBYSHEV FILTER POLYNOMIAL DEGREE SCALING
                                      FACTOR FIRST SCF (1.34)                   */
  int32_t filter_precision;        /* chfsi_precision for the filter bulk when
                                      use_mixed_prec_overall is set (default
                                      CHFSI_FP64, as USE MIXED PREC CHEBY is false
                                      by default, utils/dftParameters.cc:1262)  */
  int32_t use_shep;                /* SUBSPACE PROJ SHEP GPU: 1 = CholGS + RR
                                      (rayleighRitz), 0 = Rayleigh-Ritz in the
                                      generalised form without CholGS
                                      (rayleighRitzGEP, rayleighRitzDevice.cc:355-819) */
  int32_t fp64_tail_degrees;       /* last degrees of a mixed-precision filter run
                                      in FP64 (0: whole filter in filter_precision) */
  /* Mixed-precision dense steps, applied when use_mixed_prec_overall is set:
   * diagonal blocks (cheby_block_size wide) in FP64, off-diagonal blocks in
   * FP32 (linearAlgebraOperationsDevice.cc:3831-4609, :2243-3077,
   * kohnShamDFTOperatorDevice.cc:4161-; parameters utils/dftParameters.cc:942-966) */
  int32_t mixed_prec_overlap;      /* USE MIXED PREC CGS O: S = X^T X            */
  int32_t mixed_prec_cgs_rotation; /* USE MIXED PREC CGS SR: X <- X L^{-T}        */
  int32_t mixed_prec_xthx;         /* USE MIXED PREC XTHX: Hp = X^T H~ X          */
  int32_t mixed_prec_rr_rotation;  /* USE MIXED PREC RR_SR: X <- X Q              */
  int32_t mixed_prec_xthx_spectrum_split; /* USE MIXED PREC XTHX SPECTRUM SPLIT:
                                      column blocks within the first n_core states
                                      of the spectrum-split X^T H~ X in FP32      */
  int32_t filter_columns;          /* columns filtered by one recurrence: a multiple
                                      of cheby_block_size (0: all N, within a 16 GiB
                                      filter scratch).  No reference counterpart: the
                                      reference filters one block at a time to bound
                                      memory; every column's arithmetic is the same */
  double mixed_prec_switch_residual; /* > 0: once a pass's residual norms over its
                                      lowest 90% of states are all below this, later
                                      passes filter in FP64 even with
                                      use_mixed_prec_overall (a per-pass form of the
                                      reference's useMixedPrecOverall = !scfConverged,
                                      dft.cc); 0: never.  Reset by a first filtering
                                      call */
  int32_t mixed_prec_defect_filter; /* with the FP32 filter: run it in defect form,
                                      Y = X diag(p(theta)) + E with theta the
                                      columns' Rayleigh quotients and R = H~X - X
                                      theta from one FP64 apply, E filtered in FP32
                                      (same recurrence plus the source p_k(theta) R),
                                      so the FP32 rounding scales with the residual.
                                      No reference counterpart; 0: plain FP32 */
  int32_t is_pseudopotential;      /* PSEUDOPOTENTIAL CALCULATION (default false,
                                      utils/dftParameters.cc:1173): the first-SCF
                                      degree factor applies only when set, and a
                                      degree taken from the table is halved when not
                                      set (…SolverDevice.cc:307-326; CGS is the only
                                      orthogonalisation on the device)             */
  int32_t lanczos_full_residual_margin; /* 1: bUp = ceil(theta_max + |f|), the margin
                                      the reference takes under REPRODUCIBLE OUTPUT,
                                      without that flag's other effects (Lanczos
                                      steps, first-call bLow factor); 0: |f| / 10
                                      unless reproducible
                                      (linearAlgebraOperationsDevice.cc:510-525).
                                      With 10 Lanczos steps the |f| / 10 margin can
                                      leave bUp below the top of the spectrum, and
                                      the filter then amplifies the states above
                                      bUp every pass                               */
  int32_t mixed_prec_fp64_top_columns; /* with a lower filter precision: the last k
                                      columns (the highest Ritz values of the
                                      previous pass) are filtered in FP64.  Their
                                      filter gain is the smallest, so in the FP32
                                      forms their rounding, relative to the
                                      column, is the largest.  No reference
                                      counterpart; 0: none                         */
  int32_t dense_rocblas;           /* FP64 dense steps: 0: the library's own FP64
                                      MFMA kernels (split-K lower-triangle gram for
                                      the overlap and X^T H~ X, row-panel
                                      rotations); 1: rocBLAS dgemm for all of them;
                                      2: the library's gram with rocBLAS dgemm for
                                      the rotations.  The mixed-precision dense
                                      options use rocBLAS either way             */
} chfsi_params;

const char *chfsi_last_error(void);
void chfsi_default_params(chfsi_params *params);

/* Chebyshev degree of a solve pass (…SolverDevice.cc:29-46 order_lookup,
 * :96-105 setChebyshevOrder, :307-326): params->cheby_degree, or when it is 0
 * the table entry for bUp, halved (truncated) unless is_pseudopotential; then
 * times first_scf_degree_factor (truncated) when is_first_scf and
 * is_pseudopotential.  Host only; no device needed. */
chfsi_status chfsi_chebyshev_degree(const chfsi_params *params, double bUp, int32_t is_first_scf,
                                    int32_t *degree);

/* ---------------- operator (operatorDFTDeviceClass) ---------------- */

/* init + reinit + computeMassVector (operatorDevice.h:58-59, :139-140, :160-164) */
chfsi_status chfsi_operator_create(const chfsi_operator_desc *desc, chfsi_operator **op);
void chfsi_operator_destroy(chfsi_operator *op);

/* Cell Hamiltonian build from V_eff*JxW at quadrature points (device, C*n_q):
 * H_e = 1/2 K_e + N^T diag(vJxW_c) N.  Replaces computeHamiltonianMatricesAllkpt
 * / hamMatrixKernelLDA (hamiltonianMatrixCalculatorFlattenedDevice.cc:61-116, :1073-1375). */
chfsi_status chfsi_operator_set_potential(chfsi_operator *op, const double *vJxW_dev, void *stream);

/* GGA form of the same rebuild: gJxW_dev (device, C*n_q*3, [c][q][d]) is
 * d(E_xc)/d(sigma) * grad rho * JxW at each quadrature point and
 *   H_e = 1/2 K_e + N^T diag(vJxW_c) N
 *         + 2 sum_q gJxW(q) . (grad N_i(q) N_j(q) + N_i(q) grad N_j(q)).
 * Needs shape_gradients in the description.  Replaces hamMatrixKernelGGAMemOpt
 * (hamiltonianMatrixCalculatorFlattenedDevice.cc:283-456, real case). */
chfsi_status chfsi_operator_set_potential_gga(chfsi_operator *op, const double *vJxW_dev,
                                              const double *gJxW_dev, void *stream);

/* Direct upload of FP64 cell matrices (device, C*n_e*n_e, row i contiguous). */
chfsi_status chfsi_operator_set_cell_matrices(chfsi_operator *op, const double *H_dev, void *stream);

/* Copy the FP64 cell matrices out (device, C*n_e*n_e) -- for checks. */
chfsi_status chfsi_operator_get_cell_matrices(chfsi_operator *op, double *H_dev, void *stream);

/* Tile-shape selection of the fused HX kernel (no reference counterpart: the
 * reference calls hipBLAS gemmStridedBatched).  chfsi_operator_autotune times
 * every tile variant of both precisions on a B-column block, keeps the fastest,
 * then times the projection-kernel variants with it and keeps the fastest.
 * times_ms (48 entries) [0..15]: FP64 tile variants, [16..23]: FP64 projection
 * variants, [24..39]: FP32 tile variants, [40..47]: FP32 projection variants
 * (ms per HX; -1 where a variant does not exist).  Projection variants 4
 * (direct-to-LDS rings) and 5 (register-staged) read each DOF row once per
 * covering cell (owned rows with assembled projector values) instead of once
 * per cell node.  precision CHFSI_BF16 selects
 * the BF16 apply's kernel: K = 32 MFMA on the direct-to-LDS ring with 0
 * (default) 8 waves and 2 stages (two workgroups per CU), 2: 8 waves and 3
 * stages, 3: 16 waves and 3 stages, 4 / 5: 16 waves, 128-column tiles, 3 / 2
 * stages; 1: the register-staged K = 16 form. */
chfsi_status chfsi_operator_set_hx_variant(chfsi_operator *op, int32_t precision, int32_t variant);
chfsi_status chfsi_operator_set_proj_variant(chfsi_operator *op, int32_t precision,
                                             int32_t variant);
chfsi_status chfsi_operator_autotune(chfsi_operator *op, int32_t B, int32_t *best64,
                                     int32_t *best32, float *times_ms, void *stream);

/* getSqrtMassVec / getInvSqrtMassVec (operatorDevice.h:68-72): device pointers. */
const double *chfsi_operator_sqrt_mass(const chfsi_operator *op);
const double *chfsi_operator_inv_sqrt_mass(const chfsi_operator *op);
int64_t chfsi_operator_n_dofs(const chfsi_operator *op);

/* HX (operatorDevice.h:176-185; kohnShamDFTOperatorDevice.cc:3765-3860):
 *   dst <- (scale_flag ? dst : M^{-1/2} dst) + scalar * M^{-1/2} H M^{-1/2} src
 * src, dst: node-major FP64 blocks of B columns with leading dimensions ld_src, ld_dst.
 * src is not modified (the reference restores it when doUnscalingSrc). */
chfsi_status chfsi_hx(chfsi_operator *op, const double *src, int64_t ld_src, double *dst,
                      int64_t ld_dst, int32_t B, int32_t scale_flag, double scalar, void *stream);

/* chfsi_hx without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels
 * and results (bit for bit) as chfsi_hx; the nonlocal projection scratch,
 * n_slots*B + n_proj*B doubles (plus 256-byte alignment), lives in a fixed
 * 64 MiB device array per GPU instead of operator-owned buffers, and larger
 * blocks are rejected with CHFSI_ERR_ARG.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (other streams, other threads) race. */
chfsi_status chfsi_hx_noalloc(chfsi_operator *op, const double *src, int64_t ld_src,
                              double *dst, int64_t ld_dst, int32_t B, int32_t scale_flag,
                              double scalar, void *stream);

/* Same operator with FP32 multivectors, FP32 cell matrices, FP32 MFMA. */
chfsi_status chfsi_hx_fp32(chfsi_operator *op, const float *src, int64_t ld_src, float *dst,
                           int64_t ld_dst, int32_t B, int32_t scale_flag, float scalar,
                           void *stream);

/* Same operator with BF16 cell matrices and BF16-rounded inputs, FP32
 * accumulation (MFMA) and FP32 multivectors: the loose-tolerance early-pass form. */
chfsi_status chfsi_hx_bf16(chfsi_operator *op, const float *src, int64_t ld_src, float *dst,
                           int64_t ld_dst, int32_t B, int32_t scale_flag, float scalar,
                           void *stream);

/* HXCheby (operatorDevice.h:188-197; kohnShamDFTOperatorDevice.cc:3874-3997):
 *   dst <- dst + H src   (unscaled FE H; src already carries its M^{-1/2} scaling). */
chfsi_status chfsi_hx_cheby(chfsi_operator *op, const double *src, int64_t ld_src, double *dst,
                            int64_t ld_dst, int32_t B, void *stream);

/* Constraint kernels (constraintMatrixInfoDevice.cc:544-588, :596-816, :818-855). */
chfsi_status chfsi_constraints_distribute(chfsi_operator *op, double *X, int64_t ld, int32_t B, void *stream);
chfsi_status chfsi_constraints_slave_to_master(chfsi_operator *op, double *X, int64_t ld, int32_t B, void *stream);
chfsi_status chfsi_constraints_set_zero(chfsi_operator *op, double *X, int64_t ld, int32_t B, void *stream);

/* ---------------- density and derivatives ---------------- */

/* Density at quadrature points, the contraction of computeRhoFromPSI
 * (src/dft/densityCalculator.cc:212-254, :622-689):
 *   rho[c*n_q + q] (+)= sum_v w[v] (sum_i N_i(q) X[g(c,i), v])^2
 * X: FE-basis wavefunctions (device, ld >= N); w: device weights
 * (spin factor * k weight * occupation); rho: device C*n_q. */
chfsi_status chfsi_density(chfsi_operator *op, const double *X, int64_t ld, int32_t N,
                           const double *w_dev, double *rho_dev, int32_t accumulate,
                           void *stream);

/* The same contraction for complex (Bloch) wavefunctions, |psi|^2 = Re^2 + Im^2
 * (densityCalculator.cc:212-254 with dataTypes::number complex): X is complex
 * double interleaved (re, im), ld and N in complex elements, w holds N real
 * weights (spin factor * k weight * occupation); contributions of several
 * k-points add with accumulate = 1. */
chfsi_status chfsi_density_complex(chfsi_operator *op, const void *X, int64_t ld, int32_t N,
                                   const double *w_dev, double *rho_dev, int32_t accumulate,
                                   void *stream);

/* Adjoint of HX with respect to the local potential at the quadrature points
 * (no reference counterpart; differentiation harness).  For Y = scalar * H~ X,
 *   g[c*n_q + q] (+)= scalar * JxW(c,q) * sum_v (N(q).(M^{-1/2} Ybar_v)) (N(q).(M^{-1/2} X_v))
 * is dL/dV(c,q) given Ybar = dL/dY.  The adjoint with respect to X is HX itself
 * (H~ is symmetric). */
chfsi_status chfsi_hx_adjoint_v(chfsi_operator *op, const double *Ybar, int64_t ldy,
                                const double *X, int64_t ldx, int32_t B, double scalar,
                                double *g_dev, int32_t accumulate, void *stream);

/* Reverse pass of chfsi_chebyshev_filter (no reference counterpart).
 * Given the filter input X0 and Ybar = dL/dY of its output, writes
 * Xbar = dL/dX0 and accumulates dL/dV(c,q) into g.  Forward iterates are
 * recomputed from checkpoints stored every `checkpoint_every` degrees.
 * precision CHFSI_FP64, or CHFSI_FP32: iterates and adjoints in FP32 on the
 * FP32 apply, the gradient accumulated in FP64. */
chfsi_status chfsi_chebyshev_filter_backward(chfsi_operator *op, const double *X0, int64_t ldx,
                                             const double *Ybar, int64_t ldy, int32_t B,
                                             int32_t m, double a, double b, double a0,
                                             int32_t checkpoint_every, double *Xbar,
                                             int64_t ldxb, double *g_dev, int32_t precision,
                                             void *stream);

/* ---------------- linear algebra (linearAlgebraOperationsDevice) ---------------- */

/* chebyshevFilter (linearAlgebraOperationsDevice.cc:531-727): in-place degree-m
 * scaled Chebyshev filter of the B-column block X (already M^{1/2}-scaled).
 * precision selects the arithmetic of the recurrence. */
chfsi_status chfsi_chebyshev_filter(chfsi_operator *op, double *X, int64_t ld, int32_t B, int32_t m,
                                    double a, double b, double a0, int32_t precision, void *stream);

/* lanczosLowerUpperBoundEigenSpectrum (linearAlgebraOperationsDevice.cc:340-527).
 * v0: device start vector of length M (normalised and constrained rows zeroed
 * internally).  Returns floor(theta_min), ceil(theta_max + |f|/10 or |f|). */
chfsi_status chfsi_lanczos_bounds(chfsi_operator *op, const double *v0_dev, int32_t steps,
                                  int32_t reproducible, double *lower, double *upper, void *stream);

/* pseudoGramSchmidtOrthogonalization (pseudoGSDevice.cc:81-463): X <- X L^{-T}, S = X^T X = L L^T. */
chfsi_status chfsi_cholgs(chfsi_operator *op, double *X, int64_t M, int32_t N, void *stream);

/* rayleighRitz (rayleighRitzDevice.cc:80-353): Hp = X^T H~ X, eig, X <- X Q.
 * eig_host: N ascending eigenvalues. */
chfsi_status chfsi_rayleigh_ritz(chfsi_operator *op, double *X, int64_t M, int32_t N, int32_t block,
                                 double *eig_host, void *stream);

/* computeEigenResidualNorm (linearAlgebraOperationsDevice.cc:4610-4766). */
chfsi_status chfsi_residual_norms(chfsi_operator *op, const double *X, int64_t M, int32_t N,
                                  int32_t block, const double *eig_host, double *res_host,
                                  void *stream);

/* ---------------- solver (chebyshevOrthogonalizedSubspaceIterationSolverDevice) ---------------- */

/* constructor (…SolverDevice.h:48-54) */
chfsi_status chfsi_solver_create(const chfsi_params *params, double a0, double bLow, double bUp,
                                 chfsi_solver **solver);
void chfsi_solver_destroy(chfsi_solver *solver);
/* Replace the parameters of an existing solver (per-pass precision schedules). */
chfsi_status chfsi_solver_set_params(chfsi_solver *solver, const chfsi_params *params);
/* reinitSpectrumBounds (…SolverDevice.h:121-124) */
chfsi_status chfsi_solver_reinit_bounds(chfsi_solver *solver, double a0, double bLow, double bUp);
chfsi_status chfsi_solver_get_bounds(const chfsi_solver *solver, double *a0, double *bLow, double *bUp);
/* Filter precision (chfsi_precision) of the solver's last pass.  No reference
 * counterpart: reports the mixed_prec_switch_residual decision. */
chfsi_status chfsi_solver_last_filter_precision(const chfsi_solver *solver, int32_t *precision);
/* Supplies the Lanczos start vector used on a first filtering call (device,
 * length M); without it a seeded uniform [0,1) vector is generated. */
chfsi_status chfsi_solver_set_lanczos_start(chfsi_solver *solver, const double *v0_dev,
                                            int64_t M);

/* solve (…SolverDevice.h:61-77; .cc:155-736): one ChFSI pass on X (device,
 * node-major M x N, FE basis).  Overwrites X with the Ritz vectors, writes N
 * ascending eigenvalues and residual norms (host), returns the upper bound of
 * the unwanted spectrum in *bUp_out.  Per-phase device times (ms) go to
 * timings_ms[0..5] = {lanczos, filter, cholgs, rayleigh_ritz, residual, total}
 * when timings_ms is not NULL.  Work is ordered on `stream`; the solver's own
 * streams (the filter's FP64 top columns, rotation copy-backs) start after
 * an event on it and are joined back into it before the call returns. */
chfsi_status chfsi_solve(chfsi_solver *solver, chfsi_operator *op, double *X, int64_t M,
                         int32_t N, double *eig_host, double *res_host,
                         int32_t is_first_filtering_call, int32_t compute_residual,
                         int32_t use_mixed_prec_overall, int32_t is_first_scf, double *bUp_out,
                         double *timings_ms, void *stream);

/* solveNoRR (…SolverDevice.h:83-95; .cc:742-1080): n_passes of filter + CholGS
 * on X (device, node-major M x N, FE basis) without Rayleigh-Ritz, X returned
 * in the FE basis.  Lanczos refreshes only bUp (unless reuse_lanczos_upper);
 * a0 and bLow stay as last set.  scale_degree multiplies the degree by
 * first_scf_degree_factor (the reference does so for pseudopotential runs). */
chfsi_status chfsi_solve_no_rr(chfsi_solver *solver, chfsi_operator *op, double *X, int64_t M,
                               int32_t N, int32_t n_passes, int32_t scale_degree,
                               int32_t use_mixed_prec_overall, void *stream);

/* rayleighRitzGEPSpectrumSplitDirect (rayleighRitzDevice.cc:822-1278): on X
 * (device, node-major M x N, Loewdin basis) S = X^T X = L L^T and
 * X <- X L^{-T}, then Hp = X^T H~ X, its eigendecomposition, and only the top
 * N - n_core pairs kept: Xfrac (device, M x (N - n_core), leading dimension
 * ld_frac) = X Q[:, n_core:N] and the N - n_core eigenvalues ascending in
 * eig_host.  X is left orthonormal and unrotated.  mixed_xthx: column blocks
 * of Hp within the first n_core states from FP32 copies. */
chfsi_status chfsi_rayleigh_ritz_spectrum_split(chfsi_operator *op, double *X, int64_t M,
                                                int32_t N, int32_t n_core, double *Xfrac,
                                                int64_t ld_frac, int32_t block, double *eig_host,
                                                int32_t mixed_xthx, void *stream);

/* solve with the spectrum split (…SolverDevice.cc:155-736, the branch taken
 * when eigenValues.size() != N): Lanczos and filter as chfsi_solve, then the
 * spectrum-split Rayleigh-Ritz above (no separate CholGS), residual norms of
 * the N - n_core Ritz pairs in Xfrac, and X and Xfrac both returned in the FE
 * basis.  eig_host and res_host hold N - n_core values.  As in the reference's
 * caller (kohnShamEigenSolve.cc:483-498), the next pass sets bLow to the last
 * eigenvalue and keeps a0. */
chfsi_status chfsi_solve_spectrum_split(chfsi_solver *solver, chfsi_operator *op, double *X,
                                        int64_t M, int32_t N, int32_t n_core, double *Xfrac,
                                        int64_t ld_frac, double *eig_host, double *res_host,
                                        int32_t is_first_filtering_call,
                                        int32_t compute_residual,
                                        int32_t use_mixed_prec_overall, int32_t is_first_scf,
                                        double *bUp_out, double *timings_ms, void *stream);

/* ---------------- complex (Bloch k-point) form ----------------
 * Complex multivectors are std::complex<double> (or <float>) node-major,
 * X[node * ld + vec], ld and widths counted in complex elements.  The calls
 * below act on the operator's active k-point (chfsi_operator_set_kpoint),
 * like the reference's operator after reinitkPointSpinIndex; the solver
 * routines are the USE_COMPLEX instantiations of the real ones. */

/* reinitkPointSpinIndex (operatorDevice.h:148-149; kohnShamDFTOperatorDevice.cc:1037-1057). */
chfsi_status chfsi_operator_set_kpoint(chfsi_operator *op, int32_t kpoint);
int32_t chfsi_operator_n_kpoints(const chfsi_operator *op);

/* HX of the active k-point (kohnShamDFTOperatorDevice.cc:3765-3860, complex):
 *   dst <- (scale_flag ? dst : M^{-1/2} dst) + scalar * M^{-1/2} H_k M^{-1/2} src.
 * precision CHFSI_FP64: complex double; CHFSI_FP32: complex float (both multivectors). */
chfsi_status chfsi_hx_complex(chfsi_operator *op, const void *src, int64_t ld_src, void *dst,
                              int64_t ld_dst, int32_t B, int32_t scale_flag, double scalar,
                              int32_t precision, void *stream);

/* chebyshevFilter of a complex double block at the active k-point
 * (linearAlgebraOperationsDevice.cc:531-727, complex); precision as chfsi_chebyshev_filter. */
chfsi_status chfsi_chebyshev_filter_complex(chfsi_operator *op, void *X, int64_t ld, int32_t B,
                                            int32_t m, double a, double b, double a0,
                                            int32_t precision, void *stream);

/* Batched filter over k-points (no reference counterpart: the reference
 * filters one k-point after the other, src/dft/dft.cc:3077-3130 calling
 * kohnShamEigenSolve.cc:335-366).  Every degree is one projection launch,
 * one reduction launch and eight colour launches covering all entries.
 *   chfsi_batch_filter         -- plain launches, immediately;
 *   chfsi_batch_filter_create  -- prepares tables and scratch and captures the
 *                                 whole filter into one HIP graph on `stream`;
 *   chfsi_batch_filter_set_bounds -- new (a, b, a0) per entry (graph kept);
 *   chfsi_batch_filter_run     -- replays the graph on `stream`;
 *   chfsi_batch_filter_launch  -- issues the prepared launches without the graph.
 * The X pointers, widths and degree are fixed at creation. */
chfsi_status chfsi_batch_filter(chfsi_operator *op, const chfsi_kbatch *batch, void *stream);
chfsi_status chfsi_batch_filter_create(chfsi_operator *op, const chfsi_kbatch *batch,
                                       void *stream, chfsi_batch **out);
chfsi_status chfsi_batch_filter_set_bounds(chfsi_batch *bt, const double *a, const double *b,
                                           const double *a0);
chfsi_status chfsi_batch_filter_run(chfsi_batch *bt, void *stream);
/* The same launches issued directly on `stream` (no graph). */
chfsi_status chfsi_batch_filter_launch(chfsi_batch *bt, void *stream);
void chfsi_batch_filter_destroy(chfsi_batch *bt);

/* Tile variant of the complex HX kernel (as chfsi_operator_set_hx_variant). */
chfsi_status chfsi_operator_set_hx_variant_complex(chfsi_operator *op, int32_t precision,
                                                   int32_t variant);

/* Complex projection (no reference counterpart; the reference's projection is
 * hipBLAS gemmBatched + a reduction gemm, computeNonLocalHamiltonianTimesX-
 * MemoryOptBatchGEMMDevice.cc:48-82): 1 (default) each DOF row once per
 * covering cell, with each k-point's projector values assembled over the cell
 * nodes; 0 the cell form, once per cell node.  Both precisions, every complex
 * call and batch made after the switch. */
chfsi_status chfsi_operator_set_proj_variant_complex(chfsi_operator *op, int32_t variant);

/* Tile-variant timing of the complex HX on a batch (as chfsi_operator_autotune):
 * keeps the fastest variant per precision; times_ms[0..7] FP64, [8..15] FP32. */
chfsi_status chfsi_operator_autotune_complex(chfsi_operator *op, const chfsi_kbatch *batch,
                                             int32_t *best64, int32_t *best32, float *times_ms,
                                             void *stream);

/* Complex counterparts of the dense steps at the active k-point
 * (pseudoGSDevice.cc:81-463, rayleighRitzDevice.cc:80-353,
 * linearAlgebraOperationsDevice.cc:340-527 and :4610-4766, complex):
 * S = X^H X = L L^H, X <- X L^{-H};  Hp = X^H H~ X, X <- X Q;  residual norms. */
chfsi_status chfsi_lanczos_bounds_complex(chfsi_operator *op, const void *v0_dev, int32_t steps,
                                          int32_t reproducible, double *lower, double *upper,
                                          void *stream);
chfsi_status chfsi_cholgs_complex(chfsi_operator *op, void *X, int64_t M, int32_t N,
                                  void *stream);
chfsi_status chfsi_rayleigh_ritz_complex(chfsi_operator *op, void *X, int64_t M, int32_t N,
                                         int32_t block, double *eig_host, void *stream);
chfsi_status chfsi_residual_norms_complex(chfsi_operator *op, const void *X, int64_t M,
                                          int32_t N, int32_t block, const double *eig_host,
                                          double *res_host, void *stream);

/* solve with complex double X at the active k-point (…SolverDevice.cc:155-736,
 * USE_COMPLEX); arguments as chfsi_solve. */
chfsi_status chfsi_solve_complex(chfsi_solver *solver, chfsi_operator *op, void *X, int64_t M,
                                 int32_t N, double *eig_host, double *res_host,
                                 int32_t is_first_filtering_call, int32_t compute_residual,
                                 int32_t use_mixed_prec_overall, int32_t is_first_scf,
                                 double *bUp_out, double *timings_ms, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CHFSI_H */

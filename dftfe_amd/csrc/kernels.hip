// gfx950 kernels of the ChFSI hot path.
//
// The FE Hamiltonian apply is one fused kernel per colour class of cells:
//   1. gather the cell's nodal rows of the source block through the cell->DOF
//      map into LDS (scaled by M^{-1/2}), tile by tile along K;
//   2. MFMA (16x16x4, FP64 or FP32) of the dense cell Hamiltonian against the
//      gathered block, accumulated in registers; the nonlocal term C V C^T is
//      the same GEMM continued over extra K tiles (the cell's projector
//      columns against the V-scaled projections W);
//   3. epilogue: scaled scatter into the destination with the Chebyshev
//      three-term combine applied exactly once per DOF (first touch in colour
//      order).  Cells of one colour share no DOF, so the scatter is a plain
//      read-modify-write: no atomics, deterministic results.
// Reference structure being replaced (three kernels and several HBM sweeps per
// HX): src/dftOperator/matrixVectorProductImplementationsDevice.cc:27-117,
// computeNonLocalHamiltonianTimesXMemoryOptBatchGEMMDevice.cc:27-283,
// src/linAlg/linearAlgebraOperationsDevice.cc:37-104 (combinedDeviceKernel).

#include "kernels.hpp"
#include "mfma.hpp"
#include "scatter.hpp"

namespace chfsi {

// ---------------------------------------------------------------------------
// Fused HX over one colour class
// ---------------------------------------------------------------------------
// Work item = (cell, column tile).  The 1-D grid is laid out so that the
// column tiles of one cell are consecutive items inside one XCD's slice
// (workgroup w runs on XCD w % 8): the cell matrix then stays in that XCD's L2
// while its tiles are processed.
//
// 8 waves; wave = (wr, wc): wc selects CBW column blocks of 16, wr selects the
// row blocks rb = wr + WR*j.  MFMA operand mapping (16x16x4): lane l supplies
// A[l%16][k] and B[k][l%16] with k = 4*(l/16) + r for the r-th of four MFMAs
// over a 16-wide K slice; result lane l holds column l%16, rows acc_row().
// Eight-wave FP32 tiles of up to 64 columns fit two workgroups per CU in LDS;
// with one column block per wave their registers are held to 128 so that both
// are resident.  NB = 1 keeps one LDS buffer of each slice (two barriers per
// slice instead of one): the FP64 tiles then fit two workgroups per CU, so one
// workgroup's gather and scatter overlap the other's MFMAs.
template <typename T, int NW, int TN, int WC, int NRB, int NB = 2>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(
    NB == 1 ? NW / 2
            : (NW == 8 && TN <= 64 && TN == 16 * WC && sizeof(T) == 4 ? 4 : 1)))) void hx_kernel(
    HxLaunch<T> a) {
  using v4 = typename V4<T>::t;
  constexpr int NT = 64 * NW;
  constexpr int WR = NW / WC;
  constexpr int RBW = (NRB + WR - 1) / WR;
  constexpr int RS = row_stride<T>();
  constexpr int ROWS_MAX = NRB * 16;
  constexpr int CBW = TN / 16 / WC;
  constexpr int KPT = kTile * TN / NT;
  static_assert(KPT >= 1, "column tile too narrow");
  static_assert(CBW >= 1 && TN % (16 * WC) == 0, "bad column split");
  constexpr int HCH = (ROWS_MAX * 4 + NT - 1) / NT;
  constexpr int NLMAX = 128;             // projector slots per cell
  constexpr int KMAX = ROWS_MAX + NLMAX;  // K rows: cell nodes, then projector slots

  // K slices of the cell matrix and of the gathered block (NB buffers)
  static_assert(NB == 1 || NB == 2, "one or two slice buffers");
  __shared__ __align__(16) T Hs[NB][ROWS_MAX * RS];
  __shared__ __align__(16) T Xs[NB][TN * RS];
  // per-cell tables: DOF and first-touch bit of each local node (epilogue);
  // element offset of the source row and input scale of every K row (a DOF
  // row of src for the cell's nodes, a row of W for its projector slots;
  // scale 0 pads)
  __shared__ uint32_t smap[ROWS_MAX];
  __shared__ T sout[ROWS_MAX];
  __shared__ int64_t soff[KMAX];
  __shared__ T sscl[KMAX];

  const int wid = blockIdx.x;
  const int local = wid >> 3;
  if (local >= a.per_xcd) return;
  const int item = (wid & 7) * a.per_xcd + local;
  if (item >= a.ncells * a.ntiles) return;
  const int cell = a.cells[item / a.ntiles];
  const int v0 = (item % a.ntiles) * TN;

  const int tid = threadIdx.x;
  __builtin_assume(tid < NT);
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wc = wave % WC;
  const int wr = wave / WC;
  constexpr int nrb = NRB;
  constexpr int nkt = NRB;

  int nl_begin = 0, nslot = 0;
  if (a.nl_off) {
    nl_begin = a.nl_off[cell];
    nslot = a.nl_off[cell + 1] - nl_begin;
  }
  const int ntot = nkt + (nslot >> 4);

  const uint32_t *cmap = a.map + (int64_t)cell * a.ne;
  const T *Hcell = a.Ht + (int64_t)cell * a.ne_pad * a.ne_pad;
  const T *NLcell = a.nl_C ? a.nl_C + (int64_t)nl_begin * a.ne_pad : Hcell;

  // gather tables (padding rows and empty slots read row 0 with scale 0)
  for (int i = tid; i < KMAX; i += NT) {
    int64_t off = 0;
    T sc = T(0);
    if (i < ROWS_MAX) {
      uint32_t w = 0;
      T so = T(1);
      if (i < a.ne) {
        w = cmap[i];
        const uint32_t r = w & 0x7fffffffu;
        off = (int64_t)r * a.ld_src;
        sc = a.s_in ? a.s_in[r] : T(1);
        if (a.s_out) so = a.s_out[r];
      }
      smap[i] = w;
      sout[i] = so;
    } else {
      const int sl = i - ROWS_MAX;
      const int pj = sl < nslot ? a.slot_proj[nl_begin + sl] : -1;
      if (pj >= 0) {
        off = (int64_t)pj * a.ldW;
        sc = T(1);
      }
    }
    soff[i] = off;
    sscl[i] = sc;
  }

  const int gv = tid % TN;
  const int gk = (tid / TN) * KPT;
  const bool vok = (v0 + gv) < a.B;
  const int vcl = vok ? v0 + gv : 0;
  constexpr int nchunk = ROWS_MAX * 4;

  v4 acc[RBW][CBW];
#pragma unroll
  for (int j = 0; j < RBW; ++j)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = v4{0, 0, 0, 0};

  // register staging of the next slice: loaded while slice t is multiplied
  v4 hregA[HCH];
  T xregA[KPT], xsclA[KPT];

  // Branch-free global loads into registers: indices are clamped to valid
  // addresses and out-of-range values are multiplied by zero; the cell-matrix
  // or projector tile and the src or W rows are selected per slice without a
  // branch, so all loads of a slice are in flight together and overlap the
  // MFMAs of the previous slice.
  const T vmask = vok ? T(1) : T(0);
  auto load_tile = [&](int t, v4 (&hreg)[HCH], T (&xreg)[KPT], T (&xscl)[KPT]) {
    const bool own = t < nkt;
    const T *tile_src = own ? Hcell + (int64_t)t * ROWS_MAX * 16
                            : NLcell + (int64_t)(t - nkt) * ROWS_MAX * 16;
#pragma unroll
    for (int j = 0; j < HCH; ++j) {
      const int chunk = (tid + NT * j < nchunk) ? tid + NT * j : nchunk - 1;
      hreg[j] = reinterpret_cast<const v4 *>(tile_src)[chunk];
    }
    const T *xb = (own ? a.src : a.W) + vcl;
#pragma unroll
    for (int kk = 0; kk < KPT; ++kk) {
      const int k = t * 16 + gk + kk;
      xreg[kk] = xb[soff[k]];
      xscl[kk] = vmask * sscl[k];
    }
  };
  auto store_tile = [&](int buf, const v4 (&hreg)[HCH], const T (&xreg)[KPT],
                        const T (&xscl)[KPT]) {
#pragma unroll
    for (int j = 0; j < HCH; ++j) {
      const int chunk = tid + NT * j;
      if (chunk < nchunk)
        *reinterpret_cast<v4 *>(&Hs[buf][(chunk >> 2) * RS + (chunk & 3) * 4]) = hreg[j];
    }
#pragma unroll
    for (int kk = 0; kk < KPT; ++kk) Xs[buf][gv * RS + gk + kk] = xscl[kk] * xreg[kk];
  };

  // MFMAs of slice t from LDS buffer t & 1
  auto compute = [&](int buf) {
    v4 b4[CBW];
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int col = (wc * CBW + cb) * 16 + (lane & 15);
      b4[cb] = *reinterpret_cast<const v4 *>(&Xs[buf][col * RS + (lane >> 4) * 4]);
    }
    // Row blocks in groups of G: the G accumulators are independent, so the
    // MFMAs are issued r-major (no back-to-back dependence on one accumulator)
    // while the next group's operands are read from LDS.
    // (FP64 MFMA takes twice as long per instruction, so two chains suffice.)
    constexpr int GMAX = sizeof(T) == 8 ? 2 : 4;
    constexpr int G = RBW < GMAX ? RBW : GMAX;
    const T *hb = &Hs[buf][(lane & 15) * RS + (lane >> 4) * 4];
    v4 a4[G], a4n[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int rb = wr + WR * g;
      a4[g] = *reinterpret_cast<const v4 *>(hb + (rb < nrb ? rb : 0) * 16 * RS);
    }
#pragma unroll
    for (int j0 = 0; j0 < RBW; j0 += G) {
      if (j0 + G < RBW) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const int jn = j0 + G + g;
          const int rb = wr + WR * jn;
          if (jn < RBW) a4n[g] = *reinterpret_cast<const v4 *>(hb + (rb < nrb ? rb : 0) * 16 * RS);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const int j = j0 + g;
          if (j < RBW && wr + WR * j < nrb) {
#pragma unroll
            for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = mfma16(a4[g][r], b4[cb][r], acc[j][cb]);
          }
        }
#pragma unroll
      for (int g = 0; g < G; ++g) a4[g] = a4n[g];
    }
  };

  __syncthreads();  // gather tables ready (ntot >= NRB >= 1)
  load_tile(0, hregA, xregA, xsclA);
  store_tile(0, hregA, xregA, xsclA);
  __syncthreads();
  for (int t = 0; t < ntot; ++t) {
    const int buf = NB == 2 ? (t & 1) : 0;
    if (t + 1 < ntot) load_tile(t + 1, hregA, xregA, xsclA);
    compute(buf);
    if (NB == 1) __syncthreads();  // every wave is done with the one buffer
    if (t + 1 < ntot) store_tile(NB == 2 ? buf ^ 1 : 0, hregA, xregA, xsclA);
    __syncthreads();
  }

  hx_scatter<T, NRB, WR, RBW, CBW, 8>(a, a.ne, acc, smap, sout, wr, wc, lane, v0);
}

// ---------------------------------------------------------------------------
// Fused HX (FP32) fed by direct global->LDS loads through an S-stage ring
// ---------------------------------------------------------------------------
// Same work items, colouring, MFMA operand mapping and epilogue as hx_kernel;
// only the way K slices reach LDS differs.  global_load_lds_dwordx4 copies 16
// bytes per lane from global memory to LDS address M0 + 16*lane without
// passing through registers, so S-1 slices can be in flight while slice t is
// multiplied, at no register cost.  The copies are issued from inline
// assembly: the compiler does not count them, so it neither waits for them
// nor can its own waits fall short of them (waits are "at most N outstanding"
// in issue order, and uncounted older copies only make them stricter).  The
// kernel waits for its copies itself.  Every wave issues the same number D of
// copies per slice (duplicates of valid copies fill the count), so "this
// wave's copies of slice t have landed" is exactly vmcnt((S-2)*D).
//
// One barrier per slice: wait for slice t, barrier (every wave's copies of t
// have landed and every wave is done with slice t-1), refill the stage that
// held slice t-1 with slice t+S-1, multiply slice t.
//
// LDS layouts (a copy fills 1 KB contiguously):
//   cell-matrix slice: ROWS_MAX rows of 16 floats, four 16-byte units per
//     row; unit q of row r sits at unit position q ^ f(r) with
//     f(r) = ((r >> 1) ^ (r >> 3)) & 3, so the A-operand reads (16 rows, one
//     unit per lane quarter) of any 8 or 16 consecutive lanes hit distinct
//     banks;
//   gathered slice: [k][column], RPO = 256/TN rows per 1 KB copy, the copies
//     spaced 16*RPO bytes apart so that the four k rows of a B-operand read
//     fall in distinct 64-byte bank windows.
// Requirements (checked by the launcher): B a multiple of 4, 16-byte aligned
// rows of src and W (leading dimensions multiples of 4).
__device__ __forceinline__ uint32_t lds_address(const void *p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)p;
}

// 16 bytes per lane from g to LDS address m0 + 16*lane (M0 restored after)
__device__ __forceinline__ void copy16_to_lds(const void *g, uint32_t m0) {
  uint32_t saved;
  asm volatile(
      "s_mov_b32 %0, m0\n\t"
      "s_mov_b32 m0, %2\n\t"
      "s_nop 0\n\t"
      "global_load_lds_dwordx4 %1, off\n\t"
      "s_mov_b32 m0, %0"
      : "=&s"(saved)
      : "v"(g), "s"(m0)
      : "memory");
}

// s_waitcnt vmcnt(N) (expcnt and lgkmcnt not waited for)
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt(0x0F70 | (N & 15) | ((N >> 4) << 14));
}

template <int NW, int TN, int WC, int NRB, int S>
__global__ __launch_bounds__(64 * NW)
__attribute__((amdgpu_waves_per_eu(NW == 8 && S == 2 ? 4 : 1))) void hxp_kernel(HxLaunch<float> a) {
  using T = float;
  using v4 = f32x4;
  constexpr int NT = 64 * NW;
  constexpr int WR = NW / WC;
  constexpr int RBW = (NRB + WR - 1) / WR;
  constexpr int ROWS_MAX = NRB * 16;
  constexpr int CBW = TN / 16 / WC;
  static_assert(CBW >= 1 && TN % (16 * WC) == 0, "bad column split");
  static_assert(TN == 64 || TN == 128, "gathered-slice layout needs TN 64 or 128");
  static_assert(S >= 2, "at least two stages");
  constexpr int NLMAX = 128;
  constexpr int KMAX = ROWS_MAX + NLMAX;
  constexpr int HB = ROWS_MAX * 16 * 4;   // cell-matrix slice bytes (NRB copies)
  constexpr int RPO = 256 / TN;           // gathered rows per copy
  constexpr int LPR = TN / 4;             // lanes per gathered row
  constexpr int XG = 1024 + 16 * RPO;     // bytes per gathered copy incl. spacing
  constexpr int NX = 16 / RPO;            // gathered copies per slice
  constexpr int SB = HB + NX * XG;        // bytes per stage
  constexpr int NOP = NRB + NX;
  constexpr int D = (NOP + NW - 1) / NW;  // copies per wave per slice
  static_assert((S - 2) * D < 64, "vmcnt range");

  __shared__ __align__(16) unsigned char ring[S * SB];
  __shared__ uint32_t smap[ROWS_MAX];
  __shared__ T sout[ROWS_MAX];
  __shared__ __align__(16) int64_t soff[KMAX];
  __shared__ __align__(16) T sscl[KMAX];

  const int wid = blockIdx.x;
  const int local = wid >> 3;
  if (local >= a.per_xcd) return;
  const int item = (wid & 7) * a.per_xcd + local;
  if (item >= a.ncells * a.ntiles) return;
  const int cell = a.cells[item / a.ntiles];
  const int v0 = (item % a.ntiles) * TN;

  const int tid = threadIdx.x;
  __builtin_assume(tid < NT);
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wave_s = __builtin_amdgcn_readfirstlane(wave);
  const int wc = wave % WC;
  const int wr = wave / WC;
  constexpr int nkt = NRB;

  int nl_begin = 0, nslot = 0;
  if (a.nl_off) {
    nl_begin = a.nl_off[cell];
    nslot = a.nl_off[cell + 1] - nl_begin;
  }
  const int ntot = nkt + (nslot >> 4);
  const uint32_t *cmap = a.map + (int64_t)cell * a.ne;
  const T *Hcell = a.Ht + (int64_t)cell * a.ne_pad * a.ne_pad;
  const T *NLcell = a.nl_C ? a.nl_C + (int64_t)nl_begin * a.ne_pad : Hcell;

  // gather tables (padding rows and empty slots read row 0 with scale 0)
  for (int i = tid; i < KMAX; i += NT) {
    int64_t off = 0;
    T sc = T(0);
    if (i < ROWS_MAX) {
      uint32_t w = 0;
      T so = T(1);
      if (i < a.ne) {
        w = cmap[i];
        const uint32_t r = w & 0x7fffffffu;
        off = (int64_t)r * a.ld_src;
        sc = a.s_in ? a.s_in[r] : T(1);
        if (a.s_out) so = a.s_out[r];
      }
      smap[i] = w;
      sout[i] = so;
    } else {
      const int sl = i - ROWS_MAX;
      const int pj = sl < nslot ? a.slot_proj[nl_begin + sl] : -1;
      if (pj >= 0) {
        off = (int64_t)pj * a.ldW;
        sc = T(1);
      }
    }
    soff[i] = off;
    sscl[i] = sc;
  }
  __syncthreads();

  // per-lane parts of the copy addresses
  const uint32_t ring0 = lds_address(ring);
  const int hr = lane >> 2;  // row within a copy's 16 rows (bits 0..3 of the row)
  const int hoff = hr * 16 + (((lane & 3) ^ (((hr >> 1) ^ (hr >> 3)) & 3)) * 4);
  const int xsub = lane / LPR;
  const int xc = v0 + 4 * (lane % LPR);
  const int xcl = xc < a.B - 4 ? xc : a.B - 4;  // columns >= B are never stored

  auto issue = [&](int t, int stage) {
    const bool own = t < nkt;
    const T *tile = own ? Hcell + (int64_t)t * ROWS_MAX * 16
                        : NLcell + (int64_t)(t - nkt) * ROWS_MAX * 16;
    const T *xb = (own ? a.src : a.W) + xcl;
    const uint32_t sbase = ring0 + stage * SB;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int o = (wave_s + NW * j) % NOP;  // wave-uniform
      if (o < NRB) {
        copy16_to_lds(tile + o * 256 + hoff, sbase + o * 1024);
      } else {
        const int x = o - NRB;
        copy16_to_lds(xb + soff[t * 16 + RPO * x + xsub], sbase + HB + x * XG);
      }
    }
  };

  v4 acc[RBW][CBW];
#pragma unroll
  for (int j = 0; j < RBW; ++j)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = v4{0, 0, 0, 0};

  const int arow = lane & 15;
  const int aoff = arow * 16 + (((lane >> 4) ^ (((arow >> 1) ^ (arow >> 3)) & 3)) * 4);
  const int kq = 4 * (lane >> 4);  // first of the lane's four K rows

  auto compute = [&](int stage, int t) {
    const T *Hs = reinterpret_cast<const T *>(ring + stage * SB);
    const T *Xs = reinterpret_cast<const T *>(ring + stage * SB + HB);
    const v4 sc = *reinterpret_cast<const v4 *>(&sscl[t * 16 + kq]);
    v4 b4[CBW];
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int col = (wc * CBW + cb) * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = kq + r;
        b4[cb][r] = Xs[(k / RPO) * (XG / 4) + (k % RPO) * TN + col] * sc[r];
      }
    }
    constexpr int G = RBW < 4 ? RBW : 4;
    const T *hb = Hs + aoff;
    v4 a4[G], a4n[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int rb = wr + WR * g;
      a4[g] = *reinterpret_cast<const v4 *>(hb + (rb < NRB ? rb : 0) * 256);
    }
#pragma unroll
    for (int j0 = 0; j0 < RBW; j0 += G) {
      if (j0 + G < RBW) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const int jn = j0 + G + g;
          const int rb = wr + WR * jn;
          if (jn < RBW) a4n[g] = *reinterpret_cast<const v4 *>(hb + (rb < NRB ? rb : 0) * 256);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const int j = j0 + g;
          if (j < RBW && wr + WR * j < NRB) {
#pragma unroll
            for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = mfma16(a4[g][r], b4[cb][r], acc[j][cb]);
          }
        }
#pragma unroll
      for (int g = 0; g < G; ++g) a4[g] = a4n[g];
    }
  };

#pragma unroll
  for (int s = 0; s < S - 1; ++s) issue(s < ntot ? s : ntot - 1, s);
  for (int t = 0; t < ntot; ++t) {
    wait_vmcnt<(S - 2) * D>();
    __syncthreads();
    const int tn = t + S - 1;
    issue(tn < ntot ? tn : ntot - 1, tn % S);
    compute(t % S, t);
  }
  wait_vmcnt<0>();  // no copy may still target LDS when the workgroup ends

  hx_scatter<T, NRB, WR, RBW, CBW, 8>(a, a.ne, acc, smap, sout, wr, wc, lane, v0);
}

// ---------------------------------------------------------------------------
// Fused HX (FP64) fed by the direct-to-LDS ring
// ---------------------------------------------------------------------------
// The FP64 form of hxp_kernel.  A stage holds half of a 16-wide K slice (K
// rows 8h .. 8h+7 of slice t): per cell-matrix row 8 doubles = 64 bytes, the
// byte geometry of an FP32 slice row, so the copies and the unit swizzle
// q ^ f(r) are the same; a half slice of the cell matrix takes 22.5 KB at
// p = 6, so two stages and the tables fit two workgroups per CU and one
// workgroup's gather and scatter overlap the other's MFMAs.  MFMA (16x16x4
// FP64): lane l supplies A[l%16][k] and B[k][l%16] for k = 2 (l/16) + r in
// the r-th of the two MFMAs of a half slice.  Gathered half slice: 8 K rows
// of TN = 64 doubles, two rows per 1 KB copy, the copies XG bytes apart so
// that the B-operand reads of lane groups 0/1 and 2/3 fall in distinct
// 128-byte bank windows.  Requirements (checked by the launcher): B even,
// 16-byte aligned rows of src and W (leading dimensions even).
template <int NW, int WC, int NRB>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW / 2))) void hxd_kernel(
    HxLaunch<double> a) {
  using T = double;
  using v4 = f64x4;
  constexpr int TN = 64;
  constexpr int S = 2;
  constexpr int NT = 64 * NW;
  constexpr int WR = NW / WC;
  constexpr int RBW = (NRB + WR - 1) / WR;
  constexpr int ROWS_MAX = NRB * 16;
  constexpr int CBW = TN / 16 / WC;
  static_assert(CBW >= 1 && TN % (16 * WC) == 0, "bad column split");
  constexpr int NLMAX = 128;
  constexpr int KMAX = ROWS_MAX + NLMAX;
  constexpr int HB = ROWS_MAX * 64;   // bytes of a half slice of the cell matrix (NRB copies)
  constexpr int XG = 1024 + 128;      // bytes per gathered copy incl. spacing
  constexpr int NX = 4;               // gathered copies per half slice (2 K rows each)
  constexpr int SB = HB + NX * XG;    // bytes per stage
  constexpr int NOP = NRB + NX;
  constexpr int D = (NOP + NW - 1) / NW;  // copies per wave per stage

  __shared__ __align__(16) unsigned char ring[S * SB];
  __shared__ uint32_t smap[ROWS_MAX];
  __shared__ T sout[ROWS_MAX];
  __shared__ __align__(16) int64_t soff[KMAX];
  __shared__ __align__(16) T sscl[KMAX];

  const int wid = blockIdx.x;
  const int local = wid >> 3;
  if (local >= a.per_xcd) return;
  const int item = (wid & 7) * a.per_xcd + local;
  if (item >= a.ncells * a.ntiles) return;
  const int cell = a.cells[item / a.ntiles];
  const int v0 = (item % a.ntiles) * TN;

  const int tid = threadIdx.x;
  __builtin_assume(tid < NT);
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wave_s = __builtin_amdgcn_readfirstlane(wave);
  const int wc = wave % WC;
  const int wr = wave / WC;
  constexpr int nkt = NRB;

  int nl_begin = 0, nslot = 0;
  if (a.nl_off) {
    nl_begin = a.nl_off[cell];
    nslot = a.nl_off[cell + 1] - nl_begin;
  }
  const int ntot = nkt + (nslot >> 4);
  const int nh = 2 * ntot;  // half slices
  const uint32_t *cmap = a.map + (int64_t)cell * a.ne;
  const T *Hcell = a.Ht + (int64_t)cell * a.ne_pad * a.ne_pad;
  const T *NLcell = a.nl_C ? a.nl_C + (int64_t)nl_begin * a.ne_pad : Hcell;

  // gather tables (padding rows and empty slots read row 0 with scale 0)
  for (int i = tid; i < KMAX; i += NT) {
    int64_t off = 0;
    T sc = T(0);
    if (i < ROWS_MAX) {
      uint32_t w = 0;
      T so = T(1);
      if (i < a.ne) {
        w = cmap[i];
        const uint32_t r = w & 0x7fffffffu;
        off = (int64_t)r * a.ld_src;
        sc = a.s_in ? a.s_in[r] : T(1);
        if (a.s_out) so = a.s_out[r];
      }
      smap[i] = w;
      sout[i] = so;
    } else {
      const int sl = i - ROWS_MAX;
      const int pj = sl < nslot ? a.slot_proj[nl_begin + sl] : -1;
      if (pj >= 0) {
        off = (int64_t)pj * a.ldW;
        sc = T(1);
      }
    }
    soff[i] = off;
    sscl[i] = sc;
  }
  __syncthreads();

  // per-lane parts of the copy addresses
  const uint32_t ring0 = lds_address(ring);
  const int hr = lane >> 2;  // row within a copy's 16 rows
  const int hoff = hr * 16 + 2 * ((lane & 3) ^ (((hr >> 1) ^ (hr >> 3)) & 3));
  const int xsub = lane >> 5;  // K row within a gathered copy
  const int xc = v0 + 2 * (lane & 31);
  const int xcl = xc < a.B - 2 ? xc : a.B - 2;  // columns >= B are never stored

  auto issue = [&](int h, int stage) {
    const int t = h >> 1, half = h & 1;
    const bool own = t < nkt;
    const T *tile = (own ? Hcell + (int64_t)t * ROWS_MAX * 16
                         : NLcell + (int64_t)(t - nkt) * ROWS_MAX * 16) + half * 8;
    const T *xb = (own ? a.src : a.W) + xcl;
    const uint32_t sbase = ring0 + stage * SB;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int o = (wave_s + NW * j) % NOP;  // wave-uniform
      if (o < NRB) {
        copy16_to_lds(tile + o * 256 + hoff, sbase + o * 1024);
      } else {
        const int x = o - NRB;
        copy16_to_lds(xb + soff[t * 16 + half * 8 + 2 * x + xsub], sbase + HB + x * XG);
      }
    }
  };

  v4 acc[RBW][CBW];
#pragma unroll
  for (int j = 0; j < RBW; ++j)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = v4{0, 0, 0, 0};

  const int arow = lane & 15;
  const int g = lane >> 4;  // K rows 2g, 2g + 1 of the half slice
  const int aoff = arow * 64 + ((g ^ (((arow >> 1) ^ (arow >> 3)) & 3)) * 16);  // bytes

  auto compute = [&](int stage, int h) {
    const unsigned char *Hs = ring + stage * SB;
    const T *Xs = reinterpret_cast<const T *>(ring + stage * SB + HB + g * XG);
    const int k0 = (h >> 1) * 16 + (h & 1) * 8 + 2 * g;
    const T sc0 = sscl[k0], sc1 = sscl[k0 + 1];
    T b0[CBW], b1[CBW];
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int col = (wc * CBW + cb) * 16 + (lane & 15);
      b0[cb] = Xs[col] * sc0;
      b1[cb] = Xs[TN + col] * sc1;
    }
#pragma unroll
    for (int j = 0; j < RBW; ++j) {
      const int rb = wr + WR * j;
      if (rb < NRB) {
        const double2 a2 = *reinterpret_cast<const double2 *>(Hs + rb * 1024 + aoff);
#pragma unroll
        for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = mfma16(a2.x, b0[cb], acc[j][cb]);
#pragma unroll
        for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = mfma16(a2.y, b1[cb], acc[j][cb]);
      }
    }
  };

  issue(0, 0);
  for (int h = 0; h < nh; ++h) {
    wait_vmcnt<0>();
    __syncthreads();
    if (h + 1 < nh) issue(h + 1, (h + 1) & 1);
    compute(h & 1, h);
  }
  wait_vmcnt<0>();  // no copy may still target LDS when the workgroup ends

  hx_scatter<T, NRB, WR, RBW, CBW, 8>(a, a.ne, acc, smap, sout, wr, wc, lane, v0);
}

// ---------------------------------------------------------------------------
// BF16 form of the fused HX: cell matrices and gathered inputs in BF16, FP32
// accumulation (v_mfma_f32_16x16x16_bf16: one MFMA per 16-wide K slice), FP32
// multivectors.  Same work decomposition, colouring and epilogue as hx_kernel.
// ---------------------------------------------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}

template <int NW, int TN, int WC, int NRB>
__global__ __launch_bounds__(64 * NW) void hx_bf16_kernel(HxLaunch<float> a, const uint16_t *Ht16,
                                                          const uint16_t *nlC16) {
  constexpr int NT = 64 * NW;
  constexpr int WR = NW / WC;
  constexpr int RBW = (NRB + WR - 1) / WR;
  constexpr int RS = 24;  // bf16 elements per LDS row (48 B: 16-B aligned, conflict-free 8-B reads)
  constexpr int ROWS_MAX = NRB * 16;
  constexpr int CBW = TN / 16 / WC;
  constexpr int KPT = kTile * TN / NT;
  static_assert(KPT >= 1 && CBW >= 1 && TN % (16 * WC) == 0, "bad tile");
  constexpr int nchunk = ROWS_MAX * 2;  // 16-byte chunks of a K slice
  constexpr int HCH = (nchunk + NT - 1) / NT;
  constexpr int NLMAX = 128;

  __shared__ __align__(16) uint16_t Hs[2][ROWS_MAX * RS];
  __shared__ __align__(16) uint16_t Xs[2][TN * RS];
  __shared__ uint32_t smap[ROWS_MAX];
  __shared__ float sscale[ROWS_MAX];
  __shared__ int sproj[NLMAX];

  const int wid = blockIdx.x;
  const int local = wid >> 3;
  if (local >= a.per_xcd) return;
  const int item = (wid & 7) * a.per_xcd + local;
  if (item >= a.ncells * a.ntiles) return;
  const int cell = a.cells[item / a.ntiles];
  const int v0 = (item % a.ntiles) * TN;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wc = wave % WC;
  const int wr = wave / WC;
  constexpr int nkt = NRB;

  int nl_begin = 0, nslot = 0;
  if (a.nl_off) {
    nl_begin = a.nl_off[cell];
    nslot = a.nl_off[cell + 1] - nl_begin;
  }
  const int ntot = nkt + (nslot >> 4);
  const uint32_t *cmap = a.map + (int64_t)cell * a.ne;
  const uint16_t *Hcell = Ht16 + (int64_t)cell * ROWS_MAX * ROWS_MAX;
  const uint16_t *NLcell = nlC16 ? nlC16 + (int64_t)nl_begin * ROWS_MAX : Hcell;

  for (int i = tid; i < ROWS_MAX; i += NT) {
    uint32_t w = 0;
    float sc = 0.0f;
    if (i < a.ne) {
      w = cmap[i];
      sc = a.s_in ? a.s_in[w & 0x7fffffffu] : 1.0f;
    }
    smap[i] = w;
    sscale[i] = sc;
  }
  for (int s = tid; s < NLMAX; s += NT) sproj[s] = (s < nslot) ? a.slot_proj[nl_begin + s] : -1;

  const int gv = tid % TN;
  const int gk = (tid / TN) * KPT;
  const bool vok = (v0 + gv) < a.B;
  const int vcl = vok ? v0 + gv : 0;
  const float vmask = vok ? 1.0f : 0.0f;

  f32x4 acc[RBW][CBW];
#pragma unroll
  for (int j = 0; j < RBW; ++j)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = f32x4{0, 0, 0, 0};
  uint4 hreg[HCH];
  float xreg[KPT], xscl[KPT];

  auto load_tile = [&](int t) {
    const uint16_t *tile_src = (t < nkt) ? Hcell + (int64_t)t * ROWS_MAX * 16
                                         : NLcell + (int64_t)(t - nkt) * ROWS_MAX * 16;
#pragma unroll
    for (int j = 0; j < HCH; ++j) {
      const int chunk = (tid + NT * j < nchunk) ? tid + NT * j : nchunk - 1;
      hreg[j] = reinterpret_cast<const uint4 *>(tile_src)[chunk];
    }
    if (t < nkt) {
#pragma unroll
      for (int kk = 0; kk < KPT; ++kk) {
        const int k = t * 16 + gk + kk;
        xreg[kk] = a.src[(int64_t)(smap[k] & 0x7fffffffu) * a.ld_src + vcl];
        xscl[kk] = vmask * sscale[k];
      }
    } else {
#pragma unroll
      for (int kk = 0; kk < KPT; ++kk) {
        const int pj = sproj[(t - nkt) * 16 + gk + kk];
        xreg[kk] = a.W[(int64_t)(pj < 0 ? 0 : pj) * a.ldW + vcl];
        xscl[kk] = pj >= 0 ? vmask : 0.0f;
      }
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int j = 0; j < HCH; ++j) {
      const int chunk = tid + NT * j;
      if (chunk < nchunk)
        *reinterpret_cast<uint4 *>(&Hs[buf][(chunk >> 1) * RS + (chunk & 1) * 8]) = hreg[j];
    }
#pragma unroll
    for (int kk = 0; kk < KPT; ++kk) Xs[buf][gv * RS + gk + kk] = to_bf16(xscl[kk] * xreg[kk]);
  };

  __syncthreads();
  if (ntot > 0) {
    load_tile(0);
    store_tile(0);
  }
  __syncthreads();
  for (int t = 0; t < ntot; ++t) {
    const int buf = t & 1;
    if (t + 1 < ntot) load_tile(t + 1);
    bf16x4 b4[CBW];
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb)
      b4[cb] = *reinterpret_cast<const bf16x4 *>(
          &Xs[buf][((wc * CBW + cb) * 16 + (lane & 15)) * RS + (lane >> 4) * 4]);
#pragma unroll
    for (int j = 0; j < RBW; ++j) {
      const int rb = wr + WR * j;
      if (rb < NRB) {
        const bf16x4 a4 = *reinterpret_cast<const bf16x4 *>(
            &Hs[buf][(rb * 16 + (lane & 15)) * RS + (lane >> 4) * 4]);
#pragma unroll
        for (int cb = 0; cb < CBW; ++cb)
          acc[j][cb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a4, b4[cb], acc[j][cb], 0, 0, 0);
      }
    }
    if (t + 1 < ntot) store_tile(buf ^ 1);
    __syncthreads();
  }

#pragma unroll
  for (int j = 0; j < RBW; ++j) {
    const int rb = wr + WR * j;
    if (rb >= NRB) continue;
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int v = v0 + (wc * CBW + cb) * 16 + (lane & 15);
      if (v >= a.B) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb * 16 + 4 * (lane >> 4) + r;
        if (row >= a.ne) continue;
        const uint32_t w = smap[row];
        const int64_t g = w & 0x7fffffffu;
        const float so = a.s_out ? a.s_out[g] : 1.0f;
        // zero output scale: constrained row, plain cell sums (no combine)
        const bool cons = so == 0.0f;
        const int64_t addr = g * a.ld_dst + v;
        float base;
        if (w >> 31) {
          base = 0.0f;
          if (!cons) {
            if (a.beta_mode == 0)
              base = a.beta * a.dst[addr];
            else if (a.beta_mode == 2)
              base = a.beta * so * a.dst[addr];
            if (a.gamma != 0.0f) base += a.gamma * a.src[g * a.ld_src + v];
          }
        } else {
          base = a.dst[addr];
        }
        a.dst[addr] = base + a.alpha * (cons ? 1.0f : so) * acc[j][cb][r];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Projection partials  P_slot[v] = sum_i C_e[i, slot] * s_in[g_i] src[g_i, v]
// ---------------------------------------------------------------------------
// ST node tiles of 16 per stage (one barrier per stage); accumulators for
// MAXSRB slot blocks of 16 (the largest slot count of any cell, rounded up).
// OWN: the rows are the cell's owned rows (own_map, own_n) instead of its
// nodes, and only their tiles are visited.
template <typename T, int NW, int TN, int MAXSRB, int NRB, int ST, bool OWN = false>
__global__ __launch_bounds__(64 * NW) void proj_kernel(ProjLaunch<T> a) {
  using v4 = typename V4<T>::t;
  constexpr int NT = 64 * NW;
  constexpr int RS = row_stride<T>();
  constexpr int KPT = kTile * TN / NT;
  constexpr int CBW = TN / 16 / NW;
  constexpr int ROWS = NRB * 16;              // padded nodes per cell
  constexpr int ACH = (MAXSRB * 16 * 4 + NT - 1) / NT;
  constexpr int NSTAGE = NRB / ST;
  static_assert(KPT >= 1 && CBW >= 1, "bad projection tile");
  static_assert(NRB % ST == 0, "stage must divide the node tiles");
  __shared__ __align__(16) T As[2][ST][MAXSRB * 16 * RS];
  __shared__ __align__(16) T Xs[2][ST][TN * RS];
  __shared__ uint32_t smap[ROWS];
  __shared__ T sscale[ROWS];

  // XCD-aware: consecutive cells (neighbours in the cell order, sharing face
  // nodes) run on one XCD, so shared rows of the source hit in its L2
  const int wid = blockIdx.x;
  const int local = wid >> 3;
  if (local >= a.per_xcd) return;
  const int item = (wid & 7) * a.per_xcd + local;
  if (item >= a.ncells * a.ntiles) return;
  const int cell = a.cells[item / a.ntiles];
  const int v0 = (item % a.ntiles) * TN;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int s0 = a.nl_off[cell];
  const int nslot = a.nl_off[cell + 1] - s0;
  const int nsrb = nslot >> 4;
  const int nchunk = nslot * 4;
  const T *Acell = a.nl_Ct + (int64_t)s0 * ROWS;
  const uint32_t *cmap = OWN ? a.own_map + (int64_t)cell * ROWS : a.map + (int64_t)cell * a.ne;
  const int nrow = OWN ? a.own_n[cell] : a.ne;
  // stages holding rows (the tiles past the last row of an owned-row cell
  // are zero in nl_Ct and read row 0 with scale 0)
  const int nstage = OWN ? (((nrow + 15) >> 4) + ST - 1) / ST : NSTAGE;

  for (int i = tid; i < ROWS; i += NT) {
    uint32_t w = 0;
    T sc = T(0);
    if (i < nrow) {
      w = cmap[i];
      sc = a.s_in ? a.s_in[w & 0x7fffffffu] : T(1);
    }
    smap[i] = w;
    sscale[i] = sc;
  }

  const int gv = tid % TN;
  const int gk = (tid / TN) * KPT;
  const bool vok = (v0 + gv) < a.B;
  const int vcl = vok ? v0 + gv : 0;
  const T vmask = vok ? T(1) : T(0);

  v4 acc[MAXSRB][CBW];
#pragma unroll
  for (int j = 0; j < MAXSRB; ++j)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = v4{0, 0, 0, 0};
  v4 areg[ST][ACH];
  T xreg[ST][KPT], xscl[ST][KPT];

  auto load_stage = [&](int sg) {
#pragma unroll
    for (int u = 0; u < ST; ++u) {
      const int t = sg * ST + u;
      const T *tile_src = Acell + (int64_t)t * nslot * 16;
#pragma unroll
      for (int j = 0; j < ACH; ++j) {
        const int chunk = (tid + NT * j < nchunk) ? tid + NT * j : nchunk - 1;
        areg[u][j] = reinterpret_cast<const v4 *>(tile_src)[chunk];
      }
#pragma unroll
      for (int kk = 0; kk < KPT; ++kk) {
        const int k = t * 16 + gk + kk;
        xreg[u][kk] = a.src[(int64_t)(smap[k] & 0x7fffffffu) * a.ld_src + vcl];
        xscl[u][kk] = vmask * sscale[k];
      }
    }
  };
  auto store_stage = [&](int buf) {
#pragma unroll
    for (int u = 0; u < ST; ++u) {
#pragma unroll
      for (int j = 0; j < ACH; ++j) {
        const int chunk = tid + NT * j;
        if (chunk < nchunk)
          *reinterpret_cast<v4 *>(&As[buf][u][(chunk >> 2) * RS + (chunk & 3) * 4]) = areg[u][j];
      }
#pragma unroll
      for (int kk = 0; kk < KPT; ++kk) Xs[buf][u][gv * RS + gk + kk] = xscl[u][kk] * xreg[u][kk];
    }
  };

  __syncthreads();
  load_stage(0);
  store_stage(0);
  __syncthreads();
  for (int sg = 0; sg < nstage; ++sg) {
    const int buf = sg & 1;
    if (sg + 1 < nstage) load_stage(sg + 1);
#pragma unroll
    for (int u = 0; u < ST; ++u) {
      v4 b4[CBW];
#pragma unroll
      for (int cb = 0; cb < CBW; ++cb)
        b4[cb] = *reinterpret_cast<const v4 *>(
            &Xs[buf][u][((wave * CBW + cb) * 16 + (lane & 15)) * RS + (lane >> 4) * 4]);
#pragma unroll
      for (int j = 0; j < MAXSRB; ++j) {
        if (j < nsrb) {
          const v4 a4 = *reinterpret_cast<const v4 *>(
              &As[buf][u][(j * 16 + (lane & 15)) * RS + (lane >> 4) * 4]);
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = mfma16(a4[r], b4[cb][r], acc[j][cb]);
        }
      }
    }
    if (sg + 1 < nstage) store_stage(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < MAXSRB; ++j) {
    if (j >= nsrb) continue;
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int v = v0 + (wave * CBW + cb) * 16 + (lane & 15);
      if (v >= a.B) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int srow = j * 16 + acc_row<T>(lane, r);
        a.partial[(int64_t)(s0 + srow) * a.ldP + v] = acc[j][cb][r];
      }
    }
  }
}

template <typename T>
__global__ void proj_reduce_kernel(const int *proj_ptr, const int *proj_slots,
                                   const double *proj_V, int n_proj, const T *partial,
                                   int64_t ldP, T *W, int64_t ldW, int B) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n_proj * B) return;
  const int p = (int)(idx / B);
  const int v = (int)(idx % B);
  T s = T(0);
  int j = proj_ptr[p];
  const int j1 = proj_ptr[p + 1];
  // eight loads in flight, then the adds in slot order (the sum is the same
  // sequence of roundings as one add per load)
  for (; j + 8 <= j1; j += 8) {
    T x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = partial[(int64_t)proj_slots[j + u] * ldP + v];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += x[u];
  }
  for (; j < j1; ++j) s += partial[(int64_t)proj_slots[j] * ldP + v];
  W[(int64_t)p * ldW + v] = T(proj_V[p]) * s;
}

// ---------------------------------------------------------------------------
// Cell Hamiltonian build  H_e = 1/2 K + N^T diag(vJxW_c) N  (+ GGA term)
// (hamMatrixKernelLDA / hamMatrixKernelGGAMemOpt,
//  hamiltonianMatrixCalculatorFlattenedDevice.cc:61-116, :283-456)
// Output written in the padded K-tiled layout [kt][row][16] of the HX kernel.
//
// All cells share the shape functions, so the build is one GEMM of node pairs
// against per-cell quadrature weights:
//   H_e[i][j] = 1/2 K[i][j] + sum_k P_k(i, j) w_k(e)
// LDA: k = q,       P = N_i(q) N_j(q),                          w = vJxW_e(q)
// GGA: k = 4 q + s, s = 0 as LDA; s = 1..3 (d = s - 1):
//                   P = 2 (dN_i/dx_d N_j + N_i dN_j/dx_d),      w = gJxW_e(q, d)
// A workgroup takes a 16 x 16 block (ib <= jb) of node pairs and 64 cells:
// wave w owns nodes i = 16 ib + 4 w + t (t < 4), each a 16 (j) x 64 (cells)
// strip of FP64 16x16x4 MFMA results.  Off-diagonal blocks also write the
// transposed entries, so each pair's quadrature sum is formed once.
// ---------------------------------------------------------------------------
template <bool GGA>
__global__ __launch_bounds__(256, 3) void build_h_kernel(const double *K, int64_t K_stride,
                                                      const double *Nq, const double *G,
                                                      const double *invJ, int jac,
                                                      const double *vJxW, const double *gJxW,
                                                      int C, int ne, int ne_pad, int nq,
                                                      double *Ht64, float *Ht32,
                                                      uint16_t *Ht16) {
  constexpr int QC = GGA ? 16 : 32;            // quadrature points per LDS chunk
  constexpr int CT = 64;                       // cells per workgroup
  constexpr int NS = 18;                       // padded row of 16 nodes
  constexpr int WS = GGA ? CT * 4 + 4 : CT + 2;  // padded row of the cell weights
  constexpr int GQ = GGA ? QC : 1;
  constexpr int OS = 17;                       // padded row of the output staging
  constexpr int SIN = 2 * QC * NS + 2 * GQ * 3 * NS + QC * WS;
  constexpr int SOUT = 16 * 16 * OS;           // 16 cells x 16 x 16 results
  __shared__ double smem[SIN > SOUT ? SIN : SOUT];
  double(*sNi)[NS] = reinterpret_cast<double(*)[NS]>(smem);
  double(*sNj)[NS] = reinterpret_cast<double(*)[NS]>(smem + QC * NS);
  double(*sGi)[3][NS] = reinterpret_cast<double(*)[3][NS]>(smem + 2 * QC * NS);
  double(*sGj)[3][NS] = reinterpret_cast<double(*)[3][NS]>(smem + 2 * QC * NS + GQ * 3 * NS);
  double(*sW)[WS] = reinterpret_cast<double(*)[WS]>(smem + 2 * QC * NS + 2 * GQ * 3 * NS);
  const int nb = ne_pad / 16;
  int p = blockIdx.x, ib = 0;
  while (p >= nb - ib) {
    p -= nb - ib;
    ++ib;
  }
  const int jb = ib + p;
  const int c0 = blockIdx.y * CT;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int n = lane & 15, g = lane >> 4;
  f64x4 acc[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[t][cb] = f64x4{0, 0, 0, 0};

  for (int q0 = 0; q0 < nq; q0 += QC) {
    for (int e = tid; e < QC * 16; e += 256) {
      const int qq = e >> 4, m = e & 15, q = q0 + qq;
      const int i = ib * 16 + m, j = jb * 16 + m;
      const bool iok = q < nq && i < ne, jok = q < nq && j < ne;
      sNi[qq][m] = iok ? Nq[(int64_t)q * ne + i] : 0.0;
      sNj[qq][m] = jok ? Nq[(int64_t)q * ne + j] : 0.0;
      if constexpr (GGA) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          sGi[qq][d][m] = iok ? G[((int64_t)q * ne + i) * 3 + d] : 0.0;
          sGj[qq][d][m] = jok ? G[((int64_t)q * ne + j) * 3 + d] : 0.0;
        }
      }
    }
    // cell weights, read along q (contiguous per cell)
    for (int e = tid; e < CT * QC; e += 256) {
      const int c = e / QC, qq = e % QC, q = q0 + qq, cell = c0 + c;
      const double v = (q < nq && cell < C) ? vJxW[(int64_t)cell * nq + q] : 0.0;
      if constexpr (GGA)
        sW[qq][c * 4] = v;
      else
        sW[qq][c] = v;
    }
    if constexpr (GGA) {
      // with per-cell inverse Jacobians the shape gradients are the reference
      // cell's, and the Jacobian moves onto the weights:
      // sum_d g_d (J grad_ref N)_d = sum_e (sum_d g_d J[d][e]) d_e N_ref
      for (int e = tid; e < CT * QC; e += 256) {
        const int c = e / QC, qq = e % QC, q = q0 + qq, cell = c0 + c;
        const bool ok = q < nq && cell < C;
        const double *gp = gJxW + ((int64_t)(ok ? cell : 0) * nq + (ok ? q : 0)) * 3;
        const double g0 = ok ? gp[0] : 0.0, g1 = ok ? gp[1] : 0.0, g2 = ok ? gp[2] : 0.0;
        double w0 = g0, w1 = g1, w2 = g2;
        if (jac == 3 && ok) {
          const double *J = invJ + (int64_t)cell * 3;
          w0 = g0 * J[0];
          w1 = g1 * J[1];
          w2 = g2 * J[2];
        } else if (jac == 9 && ok) {
          const double *J = invJ + (int64_t)cell * 9;
          w0 = g0 * J[0] + g1 * J[3] + g2 * J[6];
          w1 = g0 * J[1] + g1 * J[4] + g2 * J[7];
          w2 = g0 * J[2] + g1 * J[5] + g2 * J[8];
        }
        sW[qq][c * 4 + 1] = w0;
        sW[qq][c * 4 + 2] = w1;
        sW[qq][c * 4 + 3] = w2;
      }
    }
    __syncthreads();
    constexpr int KS = GGA ? QC : QC / 4;
#pragma unroll 4
    for (int ks = 0; ks < KS; ++ks) {
      double av[4], bv[4];
      if constexpr (GGA) {
        const int q = ks, d = g > 0 ? g - 1 : 0;
        const double nj = sNj[q][n], gj = sGj[q][d][n];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const double ni = sNi[q][4 * w + t], gi = sGi[q][d][4 * w + t];
          av[t] = g == 0 ? ni * nj : 2.0 * fmadd(gi, nj, ni * gj);
        }
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) bv[cb] = sW[q][(cb * 16 + n) * 4 + g];
      } else {
        const int q = 4 * ks + g;
        const double nj = sNj[q][n];
#pragma unroll
        for (int t = 0; t < 4; ++t) av[t] = sNi[q][4 * w + t] * nj;
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) bv[cb] = sW[q][cb * 16 + n];
      }
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) acc[t][cb] = mfma16(av[t], bv[cb], acc[t][cb]);
    }
    __syncthreads();
  }

  // Epilogue through LDS, 16 cells at a time: each cell's 16 x 16 block is one
  // contiguous 2 KB run of the [kt][row][16] layout (kt = jb, rows of ib), and
  // so is its transpose (kt = ib, rows of jb); threads write consecutive elements.
  // The K loop ended with a barrier, so the staging may overwrite its operands.
  const int64_t per = (int64_t)ne_pad * ne_pad;
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        smem[(n * 16 + 4 * w + t) * OS + acc_row<double>(lane, r)] = acc[t][cb][r];
    __syncthreads();
#pragma unroll 2
    for (int e = tid; e < 16 * 256; e += 256) {
      const int c = e >> 8, wi = e & 255;
      const int cell = c0 + cb * 16 + c;
      if (cell >= C) break;  // e grows with c: no later element has a cell
      const int64_t base = (int64_t)cell * per;
      const double *Kc = K + (int64_t)cell * K_stride;
      {
        const int il = wi >> 4, jl = wi & 15, i = ib * 16 + il, j = jb * 16 + jl;
        const double q = smem[(c * 16 + il) * OS + jl];
        const double h = (i < ne && j < ne) ? 0.5 * Kc[(int64_t)i * ne + j] + q : 0.0;
        const int64_t o = base + (int64_t)jb * ne_pad * 16 + (int64_t)i * 16 + jl;
        Ht64[o] = h;
        if (Ht32) Ht32[o] = (float)h;
        if (Ht16) Ht16[o] = to_bf16((float)h);
      }
      if (ib != jb) {
        const int jl = wi >> 4, il = wi & 15, i = ib * 16 + il, j = jb * 16 + jl;
        const double q = smem[(c * 16 + il) * OS + jl];
        const double h = (i < ne && j < ne) ? 0.5 * Kc[(int64_t)j * ne + i] + q : 0.0;
        const int64_t o = base + (int64_t)ib * ne_pad * 16 + (int64_t)j * 16 + il;
        Ht64[o] = h;
        if (Ht32) Ht32[o] = (float)h;
        if (Ht16) Ht16[o] = to_bf16((float)h);
      }
    }
    __syncthreads();
  }
}

__global__ void untile_h_kernel(const double *Ht, int C, int ne, int ne_pad, double *H) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)ne * ne;
  if (idx >= per * C) return;
  const int64_t c = idx / per;
  const int i = (int)((idx % per) / ne), j = (int)(idx % ne);
  H[idx] = Ht[c * ne_pad * ne_pad + (int64_t)(j >> 4) * ne_pad * 16 + (int64_t)i * 16 + (j & 15)];
}

__global__ void tile_h_kernel(const double *H, int C, int ne, int ne_pad, double *Ht64,
                              float *Ht32) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)ne_pad * ne_pad;
  if (idx >= per * C) return;
  const int64_t c = idx / per;
  const int64_t r = idx % per;
  const int kt = (int)(r / (ne_pad * 16));
  const int i = (int)((r / 16) % ne_pad);
  const int j = kt * 16 + (int)(r % 16);
  const double h = (i < ne && j < ne) ? H[c * ne * ne + (int64_t)i * ne + j] : 0.0;
  Ht64[idx] = h;
  if (Ht32) Ht32[idx] = (float)h;
}

// ---------------------------------------------------------------------------
// Elementwise / reduction kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void scale_rows_kernel(T *x, int64_t ld, int64_t M, int B, double a, const double *s) {
  const int64_t n = M * B;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = idx / B;
    const int v = (int)(idx % B);
    x[g * ld + v] = (T)((double)x[g * ld + v] * a * s[g]);
  }
}

template <typename TS, typename TD>
__global__ void copy_block_kernel(const TS *src, int64_t ld_src, TD *dst, int64_t ld_dst,
                                  int64_t M, int B) {
  const int64_t n = M * B;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = idx / B;
    const int v = (int)(idx % B);
    dst[g * ld_dst + v] = (TD)src[g * ld_src + v];
  }
}

template <typename T>
__global__ void axpby_kernel(int64_t n, T a, const T *x, T b, T *y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}

// With rescale (nullable) each element of X is multiplied by rescale[g] after
// it is read (the pass's return to the FE basis folded into this read of X).
__global__ __launch_bounds__(256) void residual_partial_kernel(const double *HX, int64_t ldH,
                                                               double *X, int64_t ldX,
                                                               const double *lam, int64_t M,
                                                               int B, double *partial,
                                                               int nchunks,
                                                               const double *rescale) {
  __shared__ double red[4][64];
  const int tv = threadIdx.x & 63, tg = threadIdx.x >> 6;
  const int v = blockIdx.y * 64 + tv;
  const int64_t rows = (M + nchunks - 1) / nchunks;
  const int64_t g0 = (int64_t)blockIdx.x * rows;
  const int64_t g1 = g0 + rows < M ? g0 + rows : M;
  double s = 0.0;
  if (v < B) {
    const double l = lam[v];
    for (int64_t g = g0 + tg; g < g1; g += 4) {
      const double x = X[g * ldX + v];
      const double r = HX[g * ldH + v] - l * x;
      s += r * r;
      if (rescale) X[g * ldX + v] = x * rescale[g];
    }
  }
  red[tg][tv] = s;
  __syncthreads();
  if (tg == 0 && v < B)
    partial[(int64_t)blockIdx.x * B + v] = red[0][tv] + red[1][tv] + red[2][tv] + red[3][tv];
}

__global__ __launch_bounds__(256) void colstats_partial_kernel(const double *HX, int64_t ldH,
                                                               const double *X, int64_t ldX,
                                                               int64_t M, int B,
                                                               double *partial, int nchunks) {
  __shared__ double red[2][4][64];
  const int tv = threadIdx.x & 63, tg = threadIdx.x >> 6;
  const int v = blockIdx.y * 64 + tv;
  const int64_t rows = (M + nchunks - 1) / nchunks;
  const int64_t g0 = (int64_t)blockIdx.x * rows;
  const int64_t g1 = g0 + rows < M ? g0 + rows : M;
  double xh = 0.0, xx = 0.0;
  if (v < B) {
    for (int64_t g = g0 + tg; g < g1; g += 4) {
      const double x = X[g * ldX + v];
      xh += x * HX[g * ldH + v];
      xx += x * x;
    }
  }
  red[0][tg][tv] = xh;
  red[1][tg][tv] = xx;
  __syncthreads();
  if (tg < 2 && v < B)
    partial[((int64_t)blockIdx.x * 2 + tg) * B + v] =
        red[tg][0][tv] + red[tg][1][tv] + red[tg][2][tv] + red[tg][3][tv];
}

__global__ __launch_bounds__(256) void defect_source_kernel(const double *HX, int64_t ldH,
                                                            const double *X, int64_t ldX,
                                                            const double *theta, double s,
                                                            int64_t M, int B, float *R,
                                                            float *E) {
  const int64_t n = M * B;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = i / B;
    const int v = (int)(i - g * B);
    const double r = HX[g * ldH + v] - X[g * ldX + v] * theta[v];
    R[i] = (float)r;
    E[i] = (float)(s * r);
  }
}

__global__ __launch_bounds__(256) void defect_finish_kernel(double *X, int64_t ldX,
                                                            const double *D, const float *E,
                                                            int64_t ldE, int64_t M, int B) {
  const int64_t n = M * B;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = i / B;
    const int v = (int)(i - g * B);
    X[g * ldX + v] = fma(X[g * ldX + v], D[v], (double)E[g * ldE + v]);
  }
}

__global__ __launch_bounds__(256) void dot_partial_kernel(const double *x, const double *y,
                                                          int64_t n, double *partial,
                                                          int nchunks) {
  __shared__ double red[256];
  const int64_t per = (n + nchunks - 1) / nchunks;
  const int64_t i0 = (int64_t)blockIdx.x * per;
  const int64_t i1 = i0 + per < n ? i0 + per : n;
  double s = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) s += x[i] * y[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// *slot = partial[0] + partial[1] + ... in index order (the order of a host
// loop over the same values, so the sum is the same to the bit)
__global__ void sum_slot_kernel(const double *partial, int n, double *slot) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double s = 0.0;
  for (int k = 0; k < n; ++k) s += partial[k];
  *slot = s;
}

// y = a x + b y with a formed from the device scalar c = *slot: -c (form 0),
// 1 / sqrt(c) (form 1) or -sqrt(c) (form 2); the element update is the
// expression of axpby_kernel
__global__ void axpby_slot_kernel(int64_t n, const double *slot, int form, const double *x,
                                  double b, double *y) {
  const double c = *slot;
  const double a = form == 0 ? -c : (form == 1 ? 1.0 / sqrt(c) : -sqrt(c));
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}

__global__ void sum_partials_kernel(const double *P, int nk, int m, int n, double *C,
                                    int64_t ldc) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t mn = (int64_t)m * n;
  if (idx >= mn) return;
  double s = 0.0;
  for (int k = 0; k < nk; ++k) s += P[k * mn + idx];
  C[(idx / m) * ldc + idx % m] = s;
}

// FP32 split-K partials summed in FP64 (mixed-precision off-diagonal blocks)
__global__ void sum_partials_f32_kernel(const float *P, int nk, int m, int n, double *C,
                                        int64_t ldc) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t mn = (int64_t)m * n;
  if (idx >= mn) return;
  double s = 0.0;
  for (int k = 0; k < nk; ++k) s += (double)P[k * mn + idx];
  C[(idx / m) * ldc + idx % m] = s;
}

// C[i + j*ldc] += A[i + j*lda] (column-major m x n, FP32 into FP64)
__global__ void add_f32_to_f64_kernel(const float *A, int64_t lda, double *C, int64_t ldc, int m,
                                      int64_t n) {
  const int64_t mn = (int64_t)m * n;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < mn;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx % m, j = idx / m;
    C[i + j * ldc] += (double)A[i + j * lda];
  }
}

// B = (float) A for an n x n column-major matrix, with the bs x bs diagonal
// blocks set to zero when zero_diag (their FP64 part is applied separately)
__global__ void cast_blocks_kernel(const double *A, int n, int64_t lda, int bs, int zero_diag,
                                   float *B) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  const bool diag = zero_diag && (i / bs == j / bs);
  B[i + (int64_t)j * n] = diag ? 0.0f : (float)A[i + j * lda];
}

__global__ void zero_upper_kernel(double *A, int n, int64_t lda) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  if (i < j) A[j * lda + i] = 0.0;
}

// x[rows[r], v] = sum_j vals[j] g[cols[j]] x[cols[j], v]  (g nullable: 1)
template <typename T>
__global__ void distribute_kernel(const int64_t *rows, const int64_t *ptr, const int64_t *cols,
                                  const double *vals, int64_t nrows, const T *g, T *X,
                                  int64_t ld, int B) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nrows * B) return;
  const int64_t r = idx / B;
  const int v = (int)(idx % B);
  T s = T(0);
  for (int64_t j = ptr[r]; j < ptr[r + 1]; ++j) {
    const int64_t c = cols[j];
    s += T(vals[j]) * (g ? g[c] * X[c * ld + v] : X[c * ld + v]);
  }
  X[rows[r] * ld + v] = s;
}
// x[m, v] += so[m] sum_j vals[j] x[slave_j, v] over the master-major CSR (so nullable: 1)
template <typename T>
__global__ void slave_to_master_kernel(const int64_t *m_rows, const int64_t *m_ptr,
                                       const int64_t *m_slaves, const double *m_vals, int64_t nm,
                                       const T *so, T *X, int64_t ld, int B) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nm * B) return;
  const int64_t r = idx / B;
  const int v = (int)(idx % B);
  T s = T(0);
  for (int64_t j = m_ptr[r]; j < m_ptr[r + 1]; ++j) s += T(m_vals[j]) * X[m_slaves[j] * ld + v];
  const int64_t m = m_rows[r];
  X[m * ld + v] += (so ? so[m] : T(1)) * s;
}
template <typename T>
__global__ void zero_rows_kernel(const int64_t *rows, int64_t nrows, T *X, int64_t ld, int B) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nrows * B) return;
  X[rows[idx / B] * ld + idx % B] = T(0);
}

// ---------------------------------------------------------------------------
// Products of two multivectors at the quadrature points
//   g[c, q] (+)= JxW[q] * sum_v w_v (N(q) . s a_v|_c) (N(q) . s b_v|_c)
//   a = b = X, w = spin * occupation        -> density rho(q) JxW (densityCalculator.cc:622-689)
//   a = Ybar, b = X, w = scalar              -> dL/dV(q) of Y = scalar * S H(V) S X
// One workgroup per cell.  The interpolation N (n_q x n_e) times the cell's
// gathered rows (n_e x VC columns) is a GEMM on FP64 16x16x4 MFMA: each wave
// owns 2 blocks of 16 quadrature points of a 128-point group and the VC / 16
// column blocks; per slice of 16 nodes the shape values and the scaled
// gathered rows are staged in LDS.  VC = 64: with 128 columns the density
// form (one set of accumulators) holds 2 waves per SIMD instead of 4 and runs
// at 51.4 ms instead of 42.2 ms on 4096 cells x 1024 columns.  The pointwise products are weighted, summed over the
// wave's columns and over the 16 lanes of a row, and added per point into an
// LDS sum in a fixed order, so each g[c, q] is written once (deterministic).
// SAME (a and b the same array, the density) interpolates once.
// ---------------------------------------------------------------------------
template <bool SAME, typename TIn, int VC>
__global__ __launch_bounds__(256, SAME ? 2 : 4) void quad_product_kernel(
    const uint32_t *map, int ne, int nqc, const double *Nq, const double *JxW, const TIn *a,
    int64_t lda, const TIn *b, int64_t ldb, int B, const double *s_in, const double *w,
    int wshift, double wscale, double *g, int accumulate) {
  constexpr int QG = 128;  // quadrature points per group (4 waves x 2 blocks of 16)
  constexpr int CB = VC / 16;  // column blocks
  constexpr int KT = 16;   // nodes per slice
  constexpr int NS = KT + 1, XS = VC + 2;
  __shared__ double sN[QG][NS];
  __shared__ double sXa[KT][XS];
  __shared__ double sXb[SAME ? 1 : KT][XS];
  __shared__ uint32_t smap[384];
  __shared__ double sscale[384];
  __shared__ double gsum[512];
  const int cell = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int n = lane & 15, gl = lane >> 4;
  const uint32_t *cmap = map + (int64_t)cell * ne;
  for (int i = tid; i < ne; i += 256) {
    const uint32_t gg = cmap[i] & 0x7fffffffu;
    smap[i] = gg;
    sscale[i] = s_in ? s_in[gg] : 1.0;
  }
  for (int q = tid; q < nqc; q += 256) gsum[q] = 0.0;
  __syncthreads();
  for (int v0 = 0; v0 < B; v0 += VC) {
    for (int q0 = 0; q0 < nqc; q0 += QG) {
      f64x4 acca[2][CB], accb[2][SAME ? 1 : CB];
#pragma unroll
      for (int qb = 0; qb < 2; ++qb)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
          acca[qb][cb] = f64x4{0, 0, 0, 0};
          if (!SAME) accb[qb][SAME ? 0 : cb] = f64x4{0, 0, 0, 0};
        }
      for (int k0 = 0; k0 < ne; k0 += KT) {
        for (int e = tid; e < QG * KT; e += 256) {
          const int qq = e >> 4, kk = e & 15;
          const int q = q0 + qq, k = k0 + kk;
          sN[qq][kk] = (q < nqc && k < ne) ? Nq[(int64_t)q * ne + k] : 0.0;
        }
        for (int e = tid; e < KT * VC; e += 256) {
          const int kk = e / VC, vv = e % VC;
          const int k = k0 + kk, v = v0 + vv;
          double xa = 0.0, xb = 0.0;
          if (k < ne && v < B) {
            const int64_t gg = smap[k];
            const double sc = sscale[k];
            xa = sc * (double)a[gg * lda + v];
            if (!SAME) xb = sc * (double)b[gg * ldb + v];
          }
          sXa[kk][vv] = xa;
          if (!SAME) sXb[kk][vv] = xb;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int kk = 4 * t + gl;
          double av[2], ba[CB], bb[CB];
#pragma unroll
          for (int qb = 0; qb < 2; ++qb) av[qb] = sN[(wv * 2 + qb) * 16 + n][kk];
#pragma unroll
          for (int cb = 0; cb < CB; ++cb) {
            ba[cb] = sXa[kk][cb * 16 + n];
            if (!SAME) bb[cb] = sXb[kk][cb * 16 + n];
          }
#pragma unroll
          for (int qb = 0; qb < 2; ++qb)
#pragma unroll
            for (int cb = 0; cb < CB; ++cb) {
              acca[qb][cb] = mfma16(av[qb], ba[cb], acca[qb][cb]);
              if (!SAME) accb[qb][SAME ? 0 : cb] = mfma16(av[qb], bb[cb], accb[qb][SAME ? 0 : cb]);
            }
        }
        __syncthreads();
      }
      // row q = q0 + 16 (2 wv + qb) + acc_row, column v = v0 + 16 cb + n
#pragma unroll
      for (int qb = 0; qb < 2; ++qb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          double sacc = 0.0;
#pragma unroll
          for (int cb = 0; cb < CB; ++cb) {
            const int v = v0 + cb * 16 + n;
            const double x = acca[qb][cb][r];
            const double y = SAME ? x : accb[qb][SAME ? 0 : cb][r];
            if (v < B) sacc += (w ? w[v >> wshift] : 1.0) * x * y;
          }
#pragma unroll
          for (int off = 8; off >= 1; off >>= 1) sacc += __shfl_xor(sacc, off);
          const int q = q0 + (wv * 2 + qb) * 16 + acc_row<double>(lane, r);
          if (n == 0 && q < nqc) gsum[q] += sacc;
        }
    }
  }
  __syncthreads();
  for (int q = tid; q < nqc; q += 256) {
    const int64_t o = (int64_t)cell * nqc + q;
    g[o] = (accumulate ? g[o] : 0.0) + (JxW ? JxW[o] : 1.0) * wscale * gsum[q];
  }
}

template <typename TIn>
static hipError_t quad_go(const uint32_t *map, int C, int ne, int nqc, const double *Nq,
                          const double *JxW, const TIn *a, int64_t lda, const TIn *b, int64_t ldb,
                          int B, const double *s_in, const double *w, int wshift, double wscale,
                          double *g, int accumulate, hipStream_t s) {
  if (ne > 384 || nqc > 512) return hipErrorInvalidValue;
  // columns per pass: 64 for one interpolation (density); 32 for two (the
  // V-adjoint), whose accumulators then fit 125 VGPRs at 4 waves per SIMD
  // without spills (72 ms against 77 ms with 64 columns and 37 spilled
  // registers at 3 waves, 884k DOFs x 1024 columns)
  if (a == b && lda == ldb)
    hipLaunchKernelGGL((quad_product_kernel<true, TIn, 64>), dim3(C), dim3(256), 0, s, map, ne, nqc,
                       Nq, JxW, a, lda, b, ldb, B, s_in, w, wshift, wscale, g, accumulate);
  else
    hipLaunchKernelGGL((quad_product_kernel<false, TIn, 32>), dim3(C), dim3(256), 0, s, map, ne, nqc,
                       Nq, JxW, a, lda, b, ldb, B, s_in, w, wshift, wscale, g, accumulate);
  return hipGetLastError();
}

hipError_t launch_quad_product(const uint32_t *map, int C, int ne, int nqc, const double *Nq,
                               const double *JxW, const double *a, int64_t lda, const double *b,
                               int64_t ldb, int B, const double *s_in, const double *w,
                               double wscale, double *g, int accumulate, hipStream_t s,
                               int wshift) {
  return quad_go<double>(map, C, ne, nqc, Nq, JxW, a, lda, b, ldb, B, s_in, w, wshift, wscale, g,
                         accumulate, s);
}

hipError_t launch_quad_product(const uint32_t *map, int C, int ne, int nqc, const double *Nq,
                               const double *JxW, const float *a, int64_t lda, const float *b,
                               int64_t ldb, int B, const double *s_in, const double *w,
                               double wscale, double *g, int accumulate, hipStream_t s,
                               int wshift) {
  return quad_go<float>(map, C, ne, nqc, Nq, JxW, a, lda, b, ldb, B, s_in, w, wshift, wscale, g,
                        accumulate, s);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_for(int64_t n, int block, unsigned cap = 1u << 20) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

template <typename T, int NW, int TN, int WC, int NRB, int NB>
static hipError_t hx_go(HxLaunch<T> a, hipStream_t s) {
  a.ntiles = (a.B + TN - 1) / TN;
  const int items = a.ncells * a.ntiles;
  if (items == 0) return hipSuccess;
  a.per_xcd = (items + 7) / 8;
  hipLaunchKernelGGL((hx_kernel<T, NW, TN, WC, NRB, NB>), dim3(8 * a.per_xcd), dim3(64 * NW), 0,
                     s, a);
  return hipGetLastError();
}

// Row-block count of the padded cell matrix is a compile-time constant so the
// MFMA loop has no guards: (p+1)^3 padded to 16 for p = 1..6.
template <typename T, int NW, int TN, int WC, int NB = 2>
static hipError_t hx_dispatch(const HxLaunch<T> &a, hipStream_t s) {
  switch (a.ne_pad >> 4) {
    case 1: return hx_go<T, NW, TN, WC, 1, NB>(a, s);
    case 2: return hx_go<T, NW, TN, WC, 2, NB>(a, s);
    case 4: return hx_go<T, NW, TN, WC, 4, NB>(a, s);
    case 8: return hx_go<T, NW, TN, WC, 8, NB>(a, s);
    case 14: return hx_go<T, NW, TN, WC, 14, NB>(a, s);
    case 22: return hx_go<T, NW, TN, WC, 22, NB>(a, s);
    default: return hipErrorInvalidValue;
  }
}

template <int NW, int TN, int WC, int NRB, int S>
static hipError_t hxp_go(HxLaunch<float> a, hipStream_t s) {
  a.ntiles = (a.B + TN - 1) / TN;
  const int items = a.ncells * a.ntiles;
  if (items == 0) return hipSuccess;
  a.per_xcd = (items + 7) / 8;
  hipLaunchKernelGGL((hxp_kernel<NW, TN, WC, NRB, S>), dim3(8 * a.per_xcd), dim3(64 * NW), 0,
                     s, a);
  return hipGetLastError();
}

template <int NW, int WC, int NRB>
static hipError_t hxd_go(HxLaunch<double> a, hipStream_t s) {
  a.ntiles = (a.B + 63) / 64;
  const int items = a.ncells * a.ntiles;
  if (items == 0) return hipSuccess;
  a.per_xcd = (items + 7) / 8;
  hipLaunchKernelGGL((hxd_kernel<NW, WC, NRB>), dim3(8 * a.per_xcd), dim3(64 * NW), 0, s, a);
  return hipGetLastError();
}

// Single-column HX (the Lanczos matrix-vector products).  With one column
// the cell matrices stream once and bound the apply; the tile kernels would
// use one column of each 64-wide MFMA tile.  One 4-wave workgroup per cell:
// per K slice, the 16 values of a row are LPR lanes' 16-byte loads, so a wave
// load covers RPW rows (1 KB contiguous) and the workgroup RPI rows; the
// partial sums of a row are reduced over its LPR lanes at the end.  Gather
// tables, nonlocal slices (rows of W) and the first-touch combine are those
// of hx_kernel.  SYM (cell matrices exactly symmetric): slice t of the cell
// matrix is read only for the row blocks rb >= t (52% of the bytes at p = 6);
// a row block below the diagonal also gives its transposed product to the
// rows of block t, summed over a wave's lanes by shuffles and over the waves
// in LDS, in a fixed order (deterministic).
template <typename T, int NRB, bool SYM = false>
__global__ __launch_bounds__(256) void hx1_kernel(HxLaunch<T> a) {
  constexpr int ROWS = NRB * 16;
  constexpr int KMAX = ROWS + 128;
  constexpr int EPL = 16 / sizeof(T);  // elements per 16-byte load
  constexpr int LPR = 16 / EPL;        // lanes per row
  constexpr int RPW = 64 / LPR;        // rows per wave load
  constexpr int RPI = 4 * RPW;         // rows per workgroup load
  constexpr int NG = (ROWS + RPI - 1) / RPI;
  typedef double d2 __attribute__((ext_vector_type(2)));
  using L16 = typename std::conditional<sizeof(T) == 8, d2, f32x4>::type;
  __shared__ uint32_t smap[ROWS];
  __shared__ T sout[ROWS];
  __shared__ T xs[KMAX];
  __shared__ T tpart[SYM ? 4 * ROWS : 1];  // [wave][row]: transposed products

  const int wid = blockIdx.x;
  const int local = wid >> 3;
  if (local >= a.per_xcd) return;
  const int item = (wid & 7) * a.per_xcd + local;
  if (item >= a.ncells) return;
  const int cell = a.cells[item];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  int nl_begin = 0, nslot = 0;
  if (a.nl_off) {
    nl_begin = a.nl_off[cell];
    nslot = a.nl_off[cell + 1] - nl_begin;
  }
  const int ntot = NRB + (nslot >> 4);
  const uint32_t *cmap = a.map + (int64_t)cell * a.ne;
  const T *Hcell = a.Ht + (int64_t)cell * a.ne_pad * a.ne_pad;
  const T *NLcell = a.nl_C ? a.nl_C + (int64_t)nl_begin * a.ne_pad : Hcell;

  for (int i = tid; i < KMAX; i += 256) {
    T x = T(0);
    if (i < ROWS) {
      uint32_t w = 0;
      T so = T(1);
      if (i < a.ne) {
        w = cmap[i];
        const uint32_t r = w & 0x7fffffffu;
        x = (a.s_in ? a.s_in[r] : T(1)) * a.src[(int64_t)r * a.ld_src];
        if (a.s_out) so = a.s_out[r];
      }
      smap[i] = w;
      sout[i] = so;
    } else {
      const int sl = i - ROWS;
      const int pj = sl < nslot ? a.slot_proj[nl_begin + sl] : -1;
      if (pj >= 0) x = a.W[(int64_t)pj * a.ldW];
    }
    xs[i] = x;
  }
  __syncthreads();

  const int lr = lane / LPR, lk = lane % LPR;
  static_assert(16 % RPW == 0 || RPW % 16 == 0, "a wave load stays in one row block");
  T acc[NG];
#pragma unroll
  for (int g = 0; g < NG; ++g) acc[g] = T(0);
  for (int t = 0; t < ntot; ++t) {
    const T *tile = t < NRB ? Hcell + (int64_t)t * ROWS * 16 : NLcell + (int64_t)(t - NRB) * ROWS * 16;
    T xk[EPL];
#pragma unroll
    for (int e = 0; e < EPL; ++e) xk[e] = xs[t * 16 + lk * EPL + e];
    L16 v[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      const int row = g * RPI + wave * RPW + lr;
      const int rc = row < ROWS ? row : 0;
      // SYM: the row blocks above the diagonal of slice t are not read
      // (wave-uniform: a wave load lies in one row block)
      const bool skip = SYM && t < NRB && (g * RPI + wave * RPW) / 16 < t;
      if (skip) {
        v[g] = L16{};
      } else {
        v[g] = *reinterpret_cast<const L16 *>(tile + (int64_t)rc * 16 + lk * EPL);
      }
    }
    // one common first use of every load of the slice, so that all of them
    // are issued before the first wait (left alone, the compiler sinks each
    // load to its own use and keeps one or two in flight)
    static_assert(NG <= 11, "row groups per thread");
    asm volatile("" ::"v"(v[0]), "v"(v[NG > 1 ? 1 : 0]), "v"(v[NG > 2 ? 2 : 0]),
                 "v"(v[NG > 3 ? 3 : 0]), "v"(v[NG > 4 ? 4 : 0]), "v"(v[NG > 5 ? 5 : 0]),
                 "v"(v[NG > 6 ? 6 : 0]), "v"(v[NG > 7 ? 7 : 0]), "v"(v[NG > 8 ? 8 : 0]),
                 "v"(v[NG > 9 ? 9 : 0]), "v"(v[NG > 10 ? 10 : 0]));
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      const T *e = reinterpret_cast<const T *>(&v[g]);
#pragma unroll
      for (int q = 0; q < EPL; ++q) acc[g] = fmadd(e[q], xk[q], acc[g]);
    }
    if constexpr (SYM) {
      if (t < NRB) {
        // transposed products of the row blocks below the diagonal: column
        // t*16 + lk*EPL + q of the slice collects H[row][col] x[row]
        T ts[EPL];
#pragma unroll
        for (int q = 0; q < EPL; ++q) ts[q] = T(0);
#pragma unroll
        for (int g = 0; g < NG; ++g) {
          const int row = g * RPI + wave * RPW + lr;
          if ((g * RPI + wave * RPW) / 16 > t && row < ROWS) {
            const T *e = reinterpret_cast<const T *>(&v[g]);
            const T xr = xs[row];
#pragma unroll
            for (int q = 0; q < EPL; ++q) ts[q] = fmadd(e[q], xr, ts[q]);
          }
        }
#pragma unroll
        for (int q = 0; q < EPL; ++q)
#pragma unroll
          for (int o = LPR; o < 64; o <<= 1) ts[q] += __shfl_xor(ts[q], o);
        if (lr == 0) {
#pragma unroll
          for (int q = 0; q < EPL; ++q) tpart[wave * ROWS + t * 16 + lk * EPL + q] = ts[q];
        }
      }
    }
  }
#pragma unroll
  for (int g = 0; g < NG; ++g)
#pragma unroll
    for (int o = 1; o < LPR; o <<= 1) acc[g] += __shfl_xor(acc[g], o);
  if constexpr (SYM) {
    __syncthreads();  // every wave's transposed products are in LDS
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      const int row = g * RPI + wave * RPW + lr;
      if (row < ROWS)
        acc[g] += ((tpart[row] + tpart[ROWS + row]) + tpart[2 * ROWS + row]) + tpart[3 * ROWS + row];
    }
  }
  if (lk != 0) return;
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const int row = g * RPI + wave * RPW + lr;
    if (row >= a.ne) continue;
    const uint32_t w = smap[row];
    const int64_t gd = w & 0x7fffffffu;
    const bool first = (w >> 31) != 0;
    const T so = sout[row];
    // a zero output scale marks a constrained row: it collects the plain cell
    // sums (no combine) for the slave -> master pass
    const bool cons = so == T(0);
    const T d = a.dst[gd * a.ld_dst];
    const T keep = a.beta_mode == 0 ? a.beta * d : (a.beta_mode == 2 ? (a.beta * so) * d : T(0));
    const T base = first ? (cons ? T(0) : fmadd(a.gamma, a.src[gd * a.ld_src], keep)) : d;
    a.dst[gd * a.ld_dst] = fmadd(a.alpha * (cons ? T(1) : so), acc[g], base);
  }
}

template <typename T, int NRB>
static hipError_t hx1_go(HxLaunch<T> a, hipStream_t s) {
  a.ntiles = 1;
  if (a.ncells == 0) return hipSuccess;
  a.per_xcd = (a.ncells + 7) / 8;
  if (a.h_sym)
    hipLaunchKernelGGL((hx1_kernel<T, NRB, true>), dim3(8 * a.per_xcd), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((hx1_kernel<T, NRB>), dim3(8 * a.per_xcd), dim3(256), 0, s, a);
  return hipGetLastError();
}

template <typename T>
static hipError_t hx1_dispatch(const HxLaunch<T> &a, hipStream_t s) {
  switch (a.ne_pad >> 4) {
    case 1: return hx1_go<T, 1>(a, s);
    case 2: return hx1_go<T, 2>(a, s);
    case 4: return hx1_go<T, 4>(a, s);
    case 8: return hx1_go<T, 8>(a, s);
    case 14: return hx1_go<T, 14>(a, s);
    case 22: return hx1_go<T, 22>(a, s);
    default: return hipErrorInvalidValue;
  }
}

// The direct-to-LDS copies move 16-byte units: B a multiple of 4 and 16-byte
// aligned rows of src and W.  Other shapes take the register-staged kernel.
static bool hxp_fits(const HxLaunch<float> &a) {
  auto al = [](const void *p) { return ((uintptr_t)p & 15) == 0; };
  bool ok = a.B >= 4 && a.B % 4 == 0 && a.ld_src % 4 == 0 && al(a.src) && al(a.Ht);
  if (a.nl_off) ok = ok && a.ldW % 4 == 0 && al(a.W) && al(a.nl_C);
  return ok;
}

static bool hxd_fits(const HxLaunch<double> &a) {
  auto al = [](const void *p) { return ((uintptr_t)p & 15) == 0; };
  bool ok = a.B >= 2 && a.B % 2 == 0 && a.ld_src % 2 == 0 && al(a.src) && al(a.Ht);
  if (a.nl_off) ok = ok && a.ldW % 2 == 0 && al(a.W) && al(a.nl_C);
  return ok;
}

template <int NW, int WC>
static hipError_t hxd_dispatch(const HxLaunch<double> &a, hipStream_t s) {
  if (!hxd_fits(a)) return hx_dispatch<double, 8, 64, 4>(a, s);
  switch (a.ne_pad >> 4) {
    case 1: return hxd_go<NW, WC, 1>(a, s);
    case 2: return hxd_go<NW, WC, 2>(a, s);
    case 4: return hxd_go<NW, WC, 4>(a, s);
    case 8: return hxd_go<NW, WC, 8>(a, s);
    case 14: return hxd_go<NW, WC, 14>(a, s);
    case 22: return hxd_go<NW, WC, 22>(a, s);
    default: return hipErrorInvalidValue;
  }
}

template <int NW, int TN, int WC, int S>
static hipError_t hxp_dispatch(const HxLaunch<float> &a, hipStream_t s) {
  if (!hxp_fits(a)) return hx_dispatch<float, 8, 64, 4>(a, s);
  switch (a.ne_pad >> 4) {
    case 1: return hxp_go<NW, TN, WC, 1, S>(a, s);
    case 2: return hxp_go<NW, TN, WC, 2, S>(a, s);
    case 4: return hxp_go<NW, TN, WC, 4, S>(a, s);
    case 8: return hxp_go<NW, TN, WC, 8, S>(a, s);
    case 14: return hxp_go<NW, TN, WC, 14, S>(a, s);
    case 22: return hxp_go<NW, TN, WC, 22, S>(a, s);
    default: return hipErrorInvalidValue;
  }
}

// Tile variants (waves per workgroup x column tile x column waves):
//   FP32  0: 8 x 128 x 8 (one workgroup per CU, all row blocks per wave)
//         1: 4 x  64 x 4 (two independent workgroups per CU)
//         2: 8 x  64 x 4 (two row groups)
//         3: 16 x 128 x 8 (two row groups, one cell matrix read per 128 columns)
//         4: 8 x  32 x 2 (four row groups)
//         7: 8 x 64 x 2, 8: 8 x 64 x 1, 9: 16 x 128 x 4 (fewer column waves: each
//            A operand feeds more column blocks), 10: 16 x 64 x 4
//   direct-to-LDS ring (stages): 5: 16 x 128 x 8 (4), 6: 8 x 64 x 4 (2, two
//            workgroups per CU), 11: 16 x 64 x 4 (4)
//   FP64  0: 8 x  64 x 4 (two row groups)
//         1: 4 x  32 x 2
//         2: 16 x 64 x 4 (four row groups, 16 waves per CU)
//         3: 8 x 64 x 2, 4: 16 x 64 x 2 (fewer column waves)
//         5: 4 x 32 x 2 with single-buffered slices (two workgroups per CU)
//   direct-to-LDS ring of half slices (two workgroups per CU): 6: 4 x 64 x 2,
//            7: 8 x 64 x 4
int hx_variants(bool fp64) { return fp64 ? 8 : 12; }

template <>
hipError_t launch_hx<float>(const HxLaunch<float> &a, hipStream_t s) {
  if (a.B == 1 && !a.R) return hx1_dispatch<float>(a, s);
  switch (a.variant) {
    case 1: return hx_dispatch<float, 4, 64, 4>(a, s);
    case 2: return hx_dispatch<float, 8, 64, 4>(a, s);
    case 3: return hx_dispatch<float, 16, 128, 8>(a, s);
    case 4: return hx_dispatch<float, 8, 32, 2>(a, s);
    case 5: return hxp_dispatch<16, 128, 8, 4>(a, s);
    case 6: return hxp_dispatch<8, 64, 4, 2>(a, s);
    case 7: return hx_dispatch<float, 8, 64, 2>(a, s);
    case 8: return hx_dispatch<float, 8, 64, 1>(a, s);
    case 9: return hx_dispatch<float, 16, 128, 4>(a, s);
    case 10: return hx_dispatch<float, 16, 64, 4>(a, s);
    case 11: return hxp_dispatch<16, 64, 4, 4>(a, s);
    default: return hx_dispatch<float, 8, 128, 8>(a, s);
  }
}

template <>
hipError_t launch_hx<double>(const HxLaunch<double> &a, hipStream_t s) {
  if (a.B == 1) return hx1_dispatch<double>(a, s);
  switch (a.variant) {
    case 1: return hx_dispatch<double, 4, 32, 2>(a, s);
    case 2: return hx_dispatch<double, 16, 64, 4>(a, s);
    case 3: return hx_dispatch<double, 8, 64, 2>(a, s);
    case 4: return hx_dispatch<double, 16, 64, 2>(a, s);
    case 5: return hx_dispatch<double, 4, 32, 2, 1>(a, s);
    case 6: return hxd_dispatch<4, 2>(a, s);
    case 7: return hxd_dispatch<8, 4>(a, s);
    default: return hx_dispatch<double, 8, 64, 4>(a, s);
  }
}

// ---------------------------------------------------------------------------
// BF16 HX on K = 32 MFMA (v_mfma_f32_16x16x32_bf16) fed by the direct-to-LDS
// ring of hxp_kernel.  A stage holds two 16-wide K slices of the BF16 cell
// matrix (rows of 32 bytes; the two 16-byte units of row r are swapped when
// (r >> 3) & 1, so that the A-operand reads of 16 rows fall in distinct banks)
// and the 32 gathered FP32 rows of the stage ([k][column] groups of RPO rows
// per 1 KB copy, the copies 32 bytes apart so that the four k rows of a
// B-operand read fall in distinct bank windows).  The B operand is read in
// FP32, scaled and rounded to BF16 in registers; one MFMA covers 32 K.  An odd
// count of 16-wide slices ends with a half stage whose second slice reads
// clamped data multiplied by zero.  Same work items, colouring and epilogue
// as hx_kernel; requires an even number of row blocks.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf16x8v pack_bf16x8(const float (&x)[8]) {
  uint32_t w[4];
#pragma unroll
  for (int e = 0; e < 4; ++e)
    w[e] = (uint32_t)to_bf16(x[2 * e]) | ((uint32_t)to_bf16(x[2 * e + 1]) << 16);
  bf16x8v r;
  __builtin_memcpy(&r, w, sizeof(r));
  return r;
}

template <int NW, int TN, int WC, int NRB, int S>
__global__ __launch_bounds__(64 * NW)
__attribute__((amdgpu_waves_per_eu(NW == 8 && S == 2 ? 4 : 1))) void hxb_kernel(
    HxLaunch<float> a, const uint16_t *Ht16, const uint16_t *nlC16) {
  using v4 = f32x4;
  constexpr int NT = 64 * NW;
  constexpr int WR = NW / WC;
  constexpr int RBW = (NRB + WR - 1) / WR;
  constexpr int ROWS_MAX = NRB * 16;
  constexpr int CBW = TN / 16 / WC;
  static_assert((TN == 64 || TN == 128) && CBW >= 1 && TN % (16 * WC) == 0, "bad tile");
  static_assert(NRB % 2 == 0, "even row-block count");
  static_assert(S >= 2, "at least two stages");
  constexpr int NLMAX = 128;
  constexpr int KMAX = ROWS_MAX + NLMAX + 16;  // + the zero half of an odd last stage
  constexpr int HSB = ROWS_MAX * 32;           // bytes of one 16-wide BF16 K slice
  constexpr int HCP = HSB / 1024;              // 1 KB copies per slice
  constexpr int RPO = 256 / TN;                // gathered rows per copy
  constexpr int LPR = TN / 4;                  // lanes per gathered row
  // bytes per gathered copy incl. spacing: the four K groups of a B-operand
  // read (copies 2g or 4g apart) fall in four distinct 64-byte bank windows
  constexpr int XG = 1024 + 8 * RPO;
  constexpr int NXC = 32 / RPO;                // gathered copies per stage
  constexpr int HB = 2 * HSB;
  constexpr int SB = HB + NXC * XG;
  constexpr int NOP = 2 * HCP + NXC;
  constexpr int D = (NOP + NW - 1) / NW;
  static_assert((S - 2) * D < 64, "vmcnt range");

  __shared__ __align__(16) unsigned char ring[S * SB];
  __shared__ uint32_t smap[ROWS_MAX];
  __shared__ float sout[ROWS_MAX];
  __shared__ __align__(16) int64_t soff[KMAX];
  __shared__ __align__(16) float sscl[KMAX];

  const int wid = blockIdx.x;
  const int local = wid >> 3;
  if (local >= a.per_xcd) return;
  const int item = (wid & 7) * a.per_xcd + local;
  if (item >= a.ncells * a.ntiles) return;
  const int cell = a.cells[item / a.ntiles];
  const int v0 = (item % a.ntiles) * TN;

  const int tid = threadIdx.x;
  __builtin_assume(tid < NT);
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wave_s = __builtin_amdgcn_readfirstlane(wave);
  const int wc = wave % WC;
  const int wr = wave / WC;

  int nl_begin = 0, nslot = 0;
  if (a.nl_off) {
    nl_begin = a.nl_off[cell];
    nslot = a.nl_off[cell + 1] - nl_begin;
  }
  const int nt16 = NRB + (nslot >> 4);  // 16-wide K slices
  const int nst = (nt16 + 1) >> 1;      // stages
  const uint32_t *cmap = a.map + (int64_t)cell * a.ne;
  const uint16_t *Hcell = Ht16 + (int64_t)cell * a.ne_pad * a.ne_pad;
  const uint16_t *NLcell = nlC16 ? nlC16 + (int64_t)nl_begin * a.ne_pad : Hcell;

  // gather tables: K row t*16 + kk of slice t (cell nodes, then projector
  // slots); padding rows, empty slots and the zero half read row 0 with scale 0
  for (int i = tid; i < KMAX; i += NT) {
    int64_t off = 0;
    float sc = 0.0f;
    if (i < ROWS_MAX) {
      uint32_t w = 0;
      float so = 1.0f;
      if (i < a.ne) {
        w = cmap[i];
        const uint32_t r = w & 0x7fffffffu;
        off = (int64_t)r * a.ld_src;
        sc = a.s_in ? a.s_in[r] : 1.0f;
        if (a.s_out) so = a.s_out[r];
      }
      smap[i] = w;
      sout[i] = so;
    } else {
      const int sl = i - ROWS_MAX;
      const int pj = sl < nslot ? a.slot_proj[nl_begin + sl] : -1;
      if (pj >= 0) {
        off = (int64_t)pj * a.ldW;
        sc = 1.0f;
      }
    }
    soff[i] = off;
    sscl[i] = sc;
  }
  __syncthreads();

  const uint32_t ring0 = lds_address(ring);
  // H copy: lane L fills LDS unit L of 64 (row L/2, unit L&1) with the global
  // unit (L&1) ^ ((row >> 3) & 1) of that row
  const int hr = lane >> 1;
  const int hoff = hr * 16 + (((lane & 1) ^ ((hr >> 3) & 1)) * 8);
  const int xsub = lane / LPR;
  const int xc = v0 + 4 * (lane % LPR);
  const int xcl = xc < a.B - 4 ? xc : a.B - 4;  // columns >= B are never stored

  auto slice_ptr = [&](int t) -> const uint16_t * {
    return t < NRB ? Hcell + (int64_t)t * ROWS_MAX * 16
                   : NLcell + (int64_t)(t - NRB) * ROWS_MAX * 16;
  };
  auto issue = [&](int st, int stage) {
    const int t0 = 2 * st;
    const int t1 = 2 * st + 1;
    const uint16_t *h0 = slice_ptr(t0);
    const uint16_t *h1 = t1 < nt16 ? slice_ptr(t1) : h0;  // zero half: any finite data
    const float *x0 = (t0 < NRB ? a.src : a.W) + xcl;
    const float *x1 = (t1 < NRB ? a.src : a.W) + xcl;
    const uint32_t sbase = ring0 + stage * SB;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int o = (wave_s + NW * j) % NOP;  // wave-uniform
      // LDS destination of the copy: wave-uniform (M0)
      const uint32_t dst = __builtin_amdgcn_readfirstlane(
          o < HCP ? sbase + o * 1024
                  : (o < 2 * HCP ? sbase + HSB + (o - HCP) * 1024 : sbase + HB + (o - 2 * HCP) * XG));
      if (o < HCP) {
        copy16_to_lds(h0 + o * 512 + hoff, dst);
      } else if (o < 2 * HCP) {
        copy16_to_lds(h1 + (o - HCP) * 512 + hoff, dst);
      } else {
        const int x = o - 2 * HCP;
        const int k = RPO * x + xsub;  // K row within the stage
        const int ti = k < 16 ? t0 * 16 + k : t1 * 16 + (k - 16);
        copy16_to_lds((k < 16 ? x0 : x1) + soff[ti], dst);
      }
    }
  };

  v4 acc[RBW][CBW];
#pragma unroll
  for (int j = 0; j < RBW; ++j)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = v4{0, 0, 0, 0};

  const int arow = lane & 15;
  const int g = lane >> 4;  // K group: k = 8 g .. 8 g + 7
  // A operand: slice g >> 1, unit g & 1 of row 16 rb + arow, swizzled
  const int aoff = (g >> 1) * HSB + arow * 32 + (((g & 1) ^ ((arow >> 3) & 1)) * 16);

  auto compute = [&](int stage, int st) {
    const unsigned char *Hs = ring + stage * SB;
    const float *Xs = reinterpret_cast<const float *>(ring + stage * SB + HB);
    const int tk = (g < 2 ? 2 * st : 2 * st + 1) * 16 + (g & 1) * 8;  // table row of k = 8 g
    float sc[8];
    *reinterpret_cast<v4 *>(&sc[0]) = *reinterpret_cast<const v4 *>(&sscl[tk]);
    *reinterpret_cast<v4 *>(&sc[4]) = *reinterpret_cast<const v4 *>(&sscl[tk + 4]);
    bf16x8v b8[CBW];
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int col = (wc * CBW + cb) * 16 + arow;
      float x[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = 8 * g + e;
        x[e] = Xs[(k / RPO) * (XG / 4) + (k % RPO) * TN + col] * sc[e];
      }
      b8[cb] = pack_bf16x8(x);
    }
#pragma unroll
    for (int j = 0; j < RBW; ++j) {
      const int rb = wr + WR * j;
      if (rb < NRB) {
        const bf16x8v a8 = *reinterpret_cast<const bf16x8v *>(Hs + aoff + rb * 512);
#pragma unroll
        for (int cb = 0; cb < CBW; ++cb)
          acc[j][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a8, b8[cb], acc[j][cb], 0, 0, 0);
      }
    }
  };

#pragma unroll
  for (int s = 0; s < S - 1; ++s) issue(s < nst ? s : nst - 1, s);
  for (int st = 0; st < nst; ++st) {
    wait_vmcnt<(S - 2) * D>();
    __syncthreads();
    const int sn = st + S - 1;
    issue(sn < nst ? sn : nst - 1, sn % S);
    compute(st % S, st);
  }
  wait_vmcnt<0>();  // no copy may still target LDS when the workgroup ends

  hx_scatter<float, NRB, WR, RBW, CBW, 8>(a, a.ne, acc, smap, sout, wr, wc, lane, v0);
}

template <int NRB, int NW, int WC, int S, int TN = 64>
static hipError_t hxb_go(HxLaunch<float> a, const uint16_t *Ht16, const uint16_t *nlC16,
                         hipStream_t s) {
  a.ntiles = (a.B + TN - 1) / TN;
  const int items = a.ncells * a.ntiles;
  if (items == 0) return hipSuccess;
  a.per_xcd = (items + 7) / 8;
  hipLaunchKernelGGL((hxb_kernel<NW, TN, WC, NRB, S>), dim3(8 * a.per_xcd), dim3(64 * NW), 0, s,
                     a, Ht16, nlC16);
  return hipGetLastError();
}

template <int NRB>
static hipError_t hx_bf16_go(HxLaunch<float> a, const uint16_t *Ht16, const uint16_t *nlC16,
                             hipStream_t s) {
  constexpr int NW = 8, TN = 64, WC = 4;
  a.ntiles = (a.B + TN - 1) / TN;
  const int items = a.ncells * a.ntiles;
  if (items == 0) return hipSuccess;
  a.per_xcd = (items + 7) / 8;
  hipLaunchKernelGGL((hx_bf16_kernel<NW, TN, WC, NRB>), dim3(8 * a.per_xcd), dim3(64 * NW), 0, s,
                     a, Ht16, nlC16);
  return hipGetLastError();
}

template <int NW, int WC, int S, int TN = 64>
static hipError_t hxb_dispatch(const HxLaunch<float> &a, const uint16_t *Ht16,
                               const uint16_t *nlC16, hipStream_t s) {
  switch (a.ne_pad >> 4) {
    case 2: return hxb_go<2, NW, WC, S, TN>(a, Ht16, nlC16, s);
    case 4: return hxb_go<4, NW, WC, S, TN>(a, Ht16, nlC16, s);
    case 8: return hxb_go<8, NW, WC, S, TN>(a, Ht16, nlC16, s);
    case 14: return hxb_go<14, NW, WC, S, TN>(a, Ht16, nlC16, s);
    case 22: return hxb_go<22, NW, WC, S, TN>(a, Ht16, nlC16, s);
    default: return hipErrorInvalidValue;
  }
}

int hx_bf16_variants() { return 6; }

hipError_t launch_hx_bf16(const HxLaunch<float> &a, const uint16_t *Ht16, const uint16_t *nlC16,
                          hipStream_t s) {
  // Variants: the K = 32 ring (16-byte aligned rows and blocks, B a multiple
  // of 4, an even number of row blocks) as waves x column waves x stages
  // 0: 8 x 4 x 2 (two workgroups per CU), 2: 8 x 4 x 3, 3: 16 x 4 x 3 (one
  // workgroup per CU); 1, or a layout the ring does not take: the
  // register-staged K = 16 kernel
  auto al = [](const void *p) { return ((uintptr_t)p & 15) == 0; };
  bool ring = a.variant != 1 && a.B >= 4 && a.B % 4 == 0 && a.ld_src % 4 == 0 && al(a.src) &&
              al(Ht16) && ((a.ne_pad >> 4) % 2 == 0) && (a.ne_pad >> 4) >= 2;
  if (a.nl_off) ring = ring && a.ldW % 4 == 0 && al(a.W) && al(nlC16);
  if (ring) {
    switch (a.variant) {
      case 2: return hxb_dispatch<8, 4, 3>(a, Ht16, nlC16, s);
      case 3: return hxb_dispatch<16, 4, 3>(a, Ht16, nlC16, s);
      case 4: return hxb_dispatch<16, 8, 3, 128>(a, Ht16, nlC16, s);
      case 5: return hxb_dispatch<16, 8, 2, 128>(a, Ht16, nlC16, s);
      default: return hxb_dispatch<8, 4, 2>(a, Ht16, nlC16, s);
    }
  }
  switch (a.ne_pad >> 4) {
    case 1: return hx_bf16_go<1>(a, Ht16, nlC16, s);
    case 2: return hx_bf16_go<2>(a, Ht16, nlC16, s);
    case 4: return hx_bf16_go<4>(a, Ht16, nlC16, s);
    case 8: return hx_bf16_go<8>(a, Ht16, nlC16, s);
    case 14: return hx_bf16_go<14>(a, Ht16, nlC16, s);
    case 22: return hx_bf16_go<22>(a, Ht16, nlC16, s);
    default: return hipErrorInvalidValue;
  }
}

__global__ void f64_to_bf16_kernel(const double *x, int64_t n, uint16_t *y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = to_bf16((float)x[i]);
}

hipError_t launch_f64_to_bf16(const double *x, int64_t n, uint16_t *y, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(f64_to_bf16_kernel, dim3(grid_for(n, 256, 65536)), dim3(256), 0, s, x, n, y);
  return hipGetLastError();
}

// FP32 projection P_slot = C_e^T (s X)_e with the K slices copied global ->
// LDS directly (the ring of hxp_kernel): per slice, the cell's slot rows of
// nl_Ct (16 slots = one 1 KB copy, 16-byte units XOR-swizzled as for the cell
// matrices) and the 16 gathered source rows ([k][column] groups).  All waves
// split the columns and hold every slot block of the cell.  The products and
// their order are proj_kernel's (B operand X*s, four 16x16x4 MFMAs per slice),
// so the partials agree bit for bit.  Requirements as hxp_kernel's.  OWN:
// owned rows instead of the cell's nodes, as in proj_kernel.
template <int NW, int TN, int MAXSRB, int NRB, int S, bool OWN = false>
__global__ __launch_bounds__(64 * NW) void projp_kernel(ProjLaunch<float> a) {
  using T = float;
  using v4 = f32x4;
  constexpr int NT = 64 * NW;
  constexpr int CBW = TN / 16 / NW;
  constexpr int ROWS = NRB * 16;
  constexpr int RPO = 256 / TN;
  constexpr int LPR = TN / 4;
  constexpr int XG = 1024 + 16 * RPO;
  constexpr int NX = 16 / RPO;
  constexpr int AB = MAXSRB * 1024;  // slot rows of one slice
  constexpr int SB = AB + NX * XG;
  constexpr int D = (MAXSRB + NX + NW - 1) / NW;
  static_assert(CBW >= 1 && TN % (16 * NW) == 0, "bad column split");
  static_assert(TN == 64 || TN == 128 || TN == 256, "gathered-slice layout needs TN 64..256");
  static_assert(S >= 2 && (S - 2) * D < 64, "ring depth");

  __shared__ __align__(16) unsigned char ring[S * SB];
  __shared__ __align__(16) int64_t soff[ROWS];
  __shared__ __align__(16) T sscl[ROWS];

  // XCD-aware: consecutive cells run on one XCD, the column tiles of a cell
  // next to each other, so the cell's slot rows of nl_Ct are read from HBM
  // once for all its tiles (ordering the tiles outermost instead, to share the
  // face rows of neighbouring cells, measured slower)
  const int wid = blockIdx.x;
  const int local = wid >> 3;
  if (local >= a.per_xcd) return;
  const int item = (wid & 7) * a.per_xcd + local;
  if (item >= a.ncells * a.ntiles) return;
  const int cell = a.cells[item / a.ntiles];
  const int v0 = (item % a.ntiles) * TN;
  const int tid = threadIdx.x;
  __builtin_assume(tid < NT);
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wave_s = __builtin_amdgcn_readfirstlane(wave);
  const int s0 = a.nl_off[cell];
  const int nslot = a.nl_off[cell + 1] - s0;
  const int nsrb = nslot >> 4;
  const int nops = nsrb + NX;
  const T *Acell = a.nl_Ct + (int64_t)s0 * ROWS;
  const uint32_t *cmap = OWN ? a.own_map + (int64_t)cell * ROWS : a.map + (int64_t)cell * a.ne;
  const int nrow = OWN ? a.own_n[cell] : a.ne;
  const int nt = OWN ? (nrow > 0 ? (nrow + 15) >> 4 : 1) : NRB;  // K slices visited

  for (int i = tid; i < ROWS; i += NT) {
    int64_t off = 0;
    T sc = T(0);
    if (i < nrow) {
      const uint32_t r = cmap[i] & 0x7fffffffu;
      off = (int64_t)r * a.ld_src;
      sc = a.s_in ? a.s_in[r] : T(1);
    }
    soff[i] = off;
    sscl[i] = sc;
  }
  __syncthreads();

  const uint32_t ring0 = lds_address(ring);
  const int hr = lane >> 2;
  const int hoff = hr * 16 + (((lane & 3) ^ (((hr >> 1) ^ (hr >> 3)) & 3)) * 4);
  const int xsub = lane / LPR;
  const int xc = v0 + 4 * (lane % LPR);
  const int xcl = xc < a.B - 4 ? xc : a.B - 4;  // columns >= B are never stored
  const T *xb = a.src + xcl;

  auto issue = [&](int t, int stage) {
    const T *tile = Acell + (int64_t)t * nslot * 16;
    const uint32_t sbase = ring0 + stage * SB;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int o = (wave_s + NW * j) % nops;  // wave-uniform
      if (o < nsrb) {
        copy16_to_lds(tile + o * 256 + hoff, sbase + o * 1024);
      } else {
        const int x = o - nsrb;
        copy16_to_lds(xb + soff[t * 16 + RPO * x + xsub], sbase + AB + x * XG);
      }
    }
  };

  v4 acc[MAXSRB][CBW];
#pragma unroll
  for (int j = 0; j < MAXSRB; ++j)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = v4{0, 0, 0, 0};

  const int arow = lane & 15;
  const int aoff = arow * 16 + (((lane >> 4) ^ (((arow >> 1) ^ (arow >> 3)) & 3)) * 4);
  const int kq = 4 * (lane >> 4);

  auto compute = [&](int stage, int t) {
    const T *As = reinterpret_cast<const T *>(ring + stage * SB);
    const T *Xs = reinterpret_cast<const T *>(ring + stage * SB + AB);
    const v4 sc = *reinterpret_cast<const v4 *>(&sscl[t * 16 + kq]);
    v4 b4[CBW];
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int col = (wave * CBW + cb) * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = kq + r;
        b4[cb][r] = sc[r] * Xs[(k / RPO) * (XG / 4) + (k % RPO) * TN + col];
      }
    }
#pragma unroll
    for (int j = 0; j < MAXSRB; ++j) {
      if (j < nsrb) {
        const v4 a4 = *reinterpret_cast<const v4 *>(As + j * 256 + aoff);
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = mfma16(a4[r], b4[cb][r], acc[j][cb]);
      }
    }
  };

#pragma unroll
  for (int st = 0; st < S - 1; ++st) issue(st < nt ? st : nt - 1, st);
  for (int t = 0; t < nt; ++t) {
    wait_vmcnt<(S - 2) * D>();
    __syncthreads();
    const int tn = t + S - 1;
    issue(tn < nt ? tn : nt - 1, tn % S);
    compute(t % S, t);
  }
  wait_vmcnt<0>();  // no copy may still target LDS when the workgroup ends

#pragma unroll
  for (int j = 0; j < MAXSRB; ++j) {
    if (j >= nsrb) continue;
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int v = v0 + (wave * CBW + cb) * 16 + (lane & 15);
      if (v >= a.B) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int srow = j * 16 + acc_row<T>(lane, r);
        a.partial[(int64_t)(s0 + srow) * a.ldP + v] = acc[j][cb][r];
      }
    }
  }
}

// FP64 projection on the direct-to-LDS ring of hxd_kernel.  A stage holds
// half of a 16-wide K slice: the cell's slot rows of nl_Ct (8 doubles = 64
// bytes of each, one 1 KB copy per 16 slots, 16-byte units swizzled as for
// the cell matrices) and the 8 gathered source rows of TN = 128 columns (one
// 1 KB copy per row, the copies XG bytes apart so that the B-operand reads of
// the four lane groups fall in two 128-byte bank windows).  Four waves split
// the columns (32 each) and hold every slot block of the cell.  In the r-th
// of the two MFMAs of a half slice lane l supplies K row 2 (l/16) + r.  OWN:
// owned rows instead of the cell's nodes, as in proj_kernel.  Requirements
// (checked by the launcher): B even, 16-byte aligned src rows (ld_src even)
// and nl_Ct.
template <int MAXSRB, int NRB, int S, bool OWN>
__global__ __launch_bounds__(256) void projd_kernel(ProjLaunch<double> a) {
  using T = double;
  using v4 = f64x4;
  constexpr int NW = 4;
  constexpr int NT = 64 * NW;
  constexpr int TN = 128;
  constexpr int CBW = TN / 16 / NW;
  constexpr int ROWS = NRB * 16;
  constexpr int XG = 1024 + 64;  // bytes per gathered copy incl. spacing
  constexpr int NX = 8;          // gathered copies (K rows) per half slice
  constexpr int AB = MAXSRB * 1024;
  constexpr int SB = AB + NX * XG;
  constexpr int D = (MAXSRB + NX + NW - 1) / NW;  // copies per wave per stage
  static_assert(S >= 2 && (S - 2) * D < 64, "ring depth");

  __shared__ __align__(16) unsigned char ring[S * SB];
  __shared__ __align__(16) int64_t soff[ROWS];
  __shared__ __align__(16) T sscl[ROWS];

  // XCD-aware: consecutive cells run on one XCD, the column tiles of a cell
  // next to each other (the cell's slot rows of nl_Ct stay in its L2)
  const int wid = blockIdx.x;
  const int local = wid >> 3;
  if (local >= a.per_xcd) return;
  const int item = (wid & 7) * a.per_xcd + local;
  if (item >= a.ncells * a.ntiles) return;
  const int cell = a.cells[item / a.ntiles];
  const int v0 = (item % a.ntiles) * TN;
  const int tid = threadIdx.x;
  __builtin_assume(tid < NT);
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wave_s = __builtin_amdgcn_readfirstlane(wave);
  const int s0 = a.nl_off[cell];
  const int nslot = a.nl_off[cell + 1] - s0;
  const int nsrb = nslot >> 4;
  const int nops = nsrb + NX;
  const T *Acell = a.nl_Ct + (int64_t)s0 * ROWS;
  const uint32_t *cmap = OWN ? a.own_map + (int64_t)cell * ROWS : a.map + (int64_t)cell * a.ne;
  const int nrow = OWN ? a.own_n[cell] : a.ne;
  const int nt = OWN ? (nrow > 0 ? (nrow + 15) >> 4 : 1) : NRB;  // K slices visited
  const int nh = 2 * nt;                                         // half slices

  for (int i = tid; i < ROWS; i += NT) {
    int64_t off = 0;
    T sc = T(0);
    if (i < nrow) {
      const uint32_t r = cmap[i] & 0x7fffffffu;
      off = (int64_t)r * a.ld_src;
      sc = a.s_in ? a.s_in[r] : T(1);
    }
    soff[i] = off;
    sscl[i] = sc;
  }
  __syncthreads();

  const uint32_t ring0 = lds_address(ring);
  const int hr = lane >> 2;  // slot row within a copy's 16 rows
  const int hoff = hr * 16 + 2 * ((lane & 3) ^ (((hr >> 1) ^ (hr >> 3)) & 3));
  const int xc = v0 + 2 * lane;
  const int xcl = xc < a.B - 2 ? xc : a.B - 2;  // columns >= B are never stored
  const T *xb = a.src + xcl;

  auto issue = [&](int h, int stage) {
    const int t = h >> 1, half = h & 1;
    const T *tile = Acell + (int64_t)t * nslot * 16 + half * 8;
    const uint32_t sbase = ring0 + stage * SB;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int o = (wave_s + NW * j) % nops;  // wave-uniform
      if (o < nsrb) {
        copy16_to_lds(tile + o * 256 + hoff, sbase + o * 1024);
      } else {
        const int x = o - nsrb;
        copy16_to_lds(xb + soff[t * 16 + half * 8 + x], sbase + AB + x * XG);
      }
    }
  };

  v4 acc[MAXSRB][CBW];
#pragma unroll
  for (int j = 0; j < MAXSRB; ++j)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = v4{0, 0, 0, 0};

  const int arow = lane & 15;
  const int g = lane >> 4;  // K rows 2g, 2g + 1 of the half slice
  const int aoff = arow * 64 + ((g ^ (((arow >> 1) ^ (arow >> 3)) & 3)) * 16);  // bytes
  const int xo0 = 2 * g * XG, xo1 = xo0 + XG;  // copies of K rows 2g, 2g + 1

  auto compute = [&](int stage, int h) {
    const unsigned char *As = ring + stage * SB;
    const T *X0 = reinterpret_cast<const T *>(ring + stage * SB + AB + xo0);
    const T *X1 = reinterpret_cast<const T *>(ring + stage * SB + AB + xo1);
    const int k0 = (h >> 1) * 16 + (h & 1) * 8 + 2 * g;
    const T sc0 = sscl[k0], sc1 = sscl[k0 + 1];
    T b0[CBW], b1[CBW];
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int col = (wave * CBW + cb) * 16 + (lane & 15);
      b0[cb] = X0[col] * sc0;
      b1[cb] = X1[col] * sc1;
    }
#pragma unroll
    for (int j = 0; j < MAXSRB; ++j) {
      if (j < nsrb) {
        const double2 a2 = *reinterpret_cast<const double2 *>(As + j * 1024 + aoff);
#pragma unroll
        for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = mfma16(a2.x, b0[cb], acc[j][cb]);
#pragma unroll
        for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = mfma16(a2.y, b1[cb], acc[j][cb]);
      }
    }
  };

#pragma unroll
  for (int st = 0; st < S - 1; ++st) issue(st < nh ? st : nh - 1, st);
  for (int h = 0; h < nh; ++h) {
    wait_vmcnt<(S - 2) * D>();
    __syncthreads();
    const int hn = h + S - 1;
    issue(hn < nh ? hn : nh - 1, hn % S);
    compute(h % S, h);
  }
  wait_vmcnt<0>();  // no copy may still target LDS when the workgroup ends

#pragma unroll
  for (int j = 0; j < MAXSRB; ++j) {
    if (j >= nsrb) continue;
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int v = v0 + (wave * CBW + cb) * 16 + (lane & 15);
      if (v >= a.B) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int srow = j * 16 + acc_row<T>(lane, r);
        a.partial[(int64_t)(s0 + srow) * a.ldP + v] = acc[j][cb][r];
      }
    }
  }
}

template <int MAXSRB, int NRB, int S, bool OWN>
static hipError_t projd_go(ProjLaunch<double> a, hipStream_t s) {
  a.ntiles = (a.B + 127) / 128;
  const int items = a.ncells * a.ntiles;
  if (items == 0) return hipSuccess;
  a.per_xcd = (items + 7) / 8;
  hipLaunchKernelGGL((projd_kernel<MAXSRB, NRB, S, OWN>), dim3(8 * a.per_xcd), dim3(256), 0, s,
                     a);
  return hipGetLastError();
}

template <int NW, int TN, int MAXSRB, int NRB, int S, bool OWN = false>
static hipError_t projp_go(ProjLaunch<float> a, hipStream_t s) {
  a.ntiles = (a.B + TN - 1) / TN;
  const int items = a.ncells * a.ntiles;
  if (items == 0) return hipSuccess;
  a.per_xcd = (items + 7) / 8;
  hipLaunchKernelGGL((projp_kernel<NW, TN, MAXSRB, NRB, S, OWN>), dim3(8 * a.per_xcd),
                     dim3(64 * NW), 0, s, a);
  return hipGetLastError();
}

template <typename T, int NW, int TN, int MAXSRB, int NRB, int ST, bool OWN = false>
static hipError_t proj_go(ProjLaunch<T> a, hipStream_t s) {
  a.ntiles = (a.B + TN - 1) / TN;
  const int items = a.ncells * a.ntiles;
  if (items == 0) return hipSuccess;
  a.per_xcd = (items + 7) / 8;
  hipLaunchKernelGGL((proj_kernel<T, NW, TN, MAXSRB, NRB, ST, OWN>), dim3(8 * a.per_xcd),
                     dim3(64 * NW), 0, s, a);
  return hipGetLastError();
}

// Projection variants: 0: 4 waves x wide column tile, one node tile per stage,
// accumulators for 128 slots; 1: same tile, two node tiles per stage,
// accumulators sized to the largest cell; 2: 2 waves x half-width tile, two
// node tiles per stage, sized accumulators; 3 (FP32; FP64 and unaligned
// blocks take 0): 8 waves x 256 columns (128 for blocks of at most 256),
// direct-to-LDS ring of 3 slices; 4: the owned-row form (each DOF row read by
// as few cells as cover its projectors) on the direct-to-LDS rings (FP32:
// variant 3's kernel; FP64: projd_kernel, 128 columns, 3 half slices;
// unaligned blocks: variant 5's kernel); 5: the owned-row form on variant 1's
// register-staged kernel.
template <typename T, int NRB, int MAXSRB>
static hipError_t proj_variant(const ProjLaunch<T> &a, hipStream_t s) {
  constexpr int TNW = sizeof(T) == 4 ? 128 : 64;
  constexpr int ST = NRB % 2 == 0 ? 2 : 1;
  if (a.variant >= 4) {
    if (!a.own_map || !a.own_n) return hipErrorInvalidValue;
    auto al = [](const void *p) { return ((uintptr_t)p & 15) == 0; };
    if (a.variant == 4) {
      if constexpr (sizeof(T) == 4) {
        if (a.B >= 4 && a.B % 4 == 0 && a.ld_src % 4 == 0 && al(a.src) && al(a.nl_Ct))
          return a.B > 256 ? projp_go<8, 256, MAXSRB, NRB, 3, true>(a, s)
                           : projp_go<8, 128, MAXSRB, NRB, 3, true>(a, s);
      } else {
        if (a.B >= 2 && a.B % 2 == 0 && a.ld_src % 2 == 0 && al(a.src) && al(a.nl_Ct))
          return projd_go<MAXSRB, NRB, 3, true>(a, s);
      }
    }
    return proj_go<T, 4, TNW, MAXSRB, NRB, ST, true>(a, s);
  }
  if constexpr (sizeof(T) == 4) {
    auto al = [](const void *p) { return ((uintptr_t)p & 15) == 0; };
    if (a.variant == 3 && a.B >= 4 && a.B % 4 == 0 && a.ld_src % 4 == 0 && al(a.src) &&
        al(a.nl_Ct))
      return a.B > 256 ? projp_go<8, 256, MAXSRB, NRB, 3>(a, s)
                       : projp_go<8, 128, MAXSRB, NRB, 3>(a, s);
  }
  switch (a.variant) {
    case 1: return proj_go<T, 4, TNW, MAXSRB, NRB, ST>(a, s);
    case 2: return proj_go<T, 2, TNW / 2, MAXSRB, NRB, ST>(a, s);
    default: return proj_go<T, 4, TNW, 8, NRB, 1>(a, s);
  }
}

template <typename T, int NRB>
static hipError_t proj_fit(const ProjLaunch<T> &a, hipStream_t s) {
  if (a.max_srb <= 2) return proj_variant<T, NRB, 2>(a, s);
  if (a.max_srb <= 4) return proj_variant<T, NRB, 4>(a, s);
  return proj_variant<T, NRB, 8>(a, s);
}

int proj_variants() { return 6; }

template <typename T>
hipError_t launch_proj(const ProjLaunch<T> &a, hipStream_t s) {
  if (a.max_srb < 1 || a.max_srb > 8) return hipErrorInvalidValue;
  switch (a.ne_pad >> 4) {
    case 1: return proj_fit<T, 1>(a, s);
    case 2: return proj_fit<T, 2>(a, s);
    case 4: return proj_fit<T, 4>(a, s);
    case 8: return proj_fit<T, 8>(a, s);
    case 14: return proj_fit<T, 14>(a, s);
    case 22: return proj_fit<T, 22>(a, s);
    default: return hipErrorInvalidValue;
  }
}
template hipError_t launch_proj<float>(const ProjLaunch<float> &, hipStream_t);
template hipError_t launch_proj<double>(const ProjLaunch<double> &, hipStream_t);

template <typename T>
hipError_t launch_proj_reduce(const int *proj_ptr, const int *proj_slots, const double *proj_V,
                              int n_proj, const T *partial, int64_t ldP, T *W, int64_t ldW,
                              int B, hipStream_t s) {
  const int64_t n = (int64_t)n_proj * B;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(proj_reduce_kernel<T>, dim3(grid_for(n, 256, 1u << 30)), dim3(256), 0, s,
                     proj_ptr, proj_slots, proj_V, n_proj, partial, ldP, W, ldW, B);
  return hipGetLastError();
}
template hipError_t launch_proj_reduce<float>(const int *, const int *, const double *, int,
                                              const float *, int64_t, float *, int64_t, int,
                                              hipStream_t);
template hipError_t launch_proj_reduce<double>(const int *, const int *, const double *, int,
                                               const double *, int64_t, double *, int64_t, int,
                                               hipStream_t);

hipError_t launch_build_h(const double *K, const double *Nq, const double *vJxW, const double *G,
                          const double *gJxW, int C, int ne, int ne_pad, int nq, double *Ht64,
                          float *Ht32, uint16_t *Ht16, hipStream_t s, bool K_per_cell,
                          const double *invJ, int jac) {
  if ((gJxW && !G) || ne_pad % 16 != 0 || ne_pad < ne || C <= 0) return hipErrorInvalidValue;
  if (!(jac == 0 || ((jac == 3 || jac == 9) && invJ))) return hipErrorInvalidValue;
  const int nb = ne_pad / 16;
  const dim3 grid(nb * (nb + 1) / 2, (C + 63) / 64);
  if (grid.y > 65535u) return hipErrorInvalidValue;
  const int64_t Ks = K_per_cell ? (int64_t)ne * ne : 0;
  if (gJxW)
    hipLaunchKernelGGL(build_h_kernel<true>, grid, dim3(256), 0, s, K, Ks, Nq, G, invJ, jac, vJxW,
                       gJxW, C, ne, ne_pad, nq, Ht64, Ht32, Ht16);
  else
    hipLaunchKernelGGL(build_h_kernel<false>, grid, dim3(256), 0, s, K, Ks, Nq, G, invJ, jac,
                       vJxW, gJxW, C, ne, ne_pad, nq, Ht64, Ht32, Ht16);
  return hipGetLastError();
}

hipError_t launch_untile_h(const double *Ht, int C, int ne, int ne_pad, double *H, hipStream_t s) {
  const int64_t n = (int64_t)C * ne * ne;
  hipLaunchKernelGGL(untile_h_kernel, dim3(grid_for(n, 256, 1u << 30)), dim3(256), 0, s, Ht, C,
                     ne, ne_pad, H);
  return hipGetLastError();
}

hipError_t launch_tile_h(const double *H, int C, int ne, int ne_pad, double *Ht64, float *Ht32,
                         hipStream_t s) {
  const int64_t n = (int64_t)C * ne_pad * ne_pad;
  hipLaunchKernelGGL(tile_h_kernel, dim3(grid_for(n, 256, 1u << 30)), dim3(256), 0, s, H, C, ne,
                     ne_pad, Ht64, Ht32);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scale_rows(T *x, int64_t ld, int64_t M, int B, double a, const double *s,
                             hipStream_t st) {
  hipLaunchKernelGGL(scale_rows_kernel<T>, dim3(grid_for(M * B, 256, 65536)), dim3(256), 0, st,
                     x, ld, M, B, a, s);
  return hipGetLastError();
}
template hipError_t launch_scale_rows<double>(double *, int64_t, int64_t, int, double,
                                              const double *, hipStream_t);
template hipError_t launch_scale_rows<float>(float *, int64_t, int64_t, int, double,
                                             const double *, hipStream_t);

template <typename TS, typename TD>
hipError_t launch_copy_block(const TS *src, int64_t ld_src, TD *dst, int64_t ld_dst, int64_t M,
                             int B, hipStream_t st) {
  hipLaunchKernelGGL((copy_block_kernel<TS, TD>), dim3(grid_for(M * B, 256, 65536)), dim3(256), 0,
                     st, src, ld_src, dst, ld_dst, M, B);
  return hipGetLastError();
}
template hipError_t launch_copy_block<double, double>(const double *, int64_t, double *, int64_t,
                                                      int64_t, int, hipStream_t);
template hipError_t launch_copy_block<double, float>(const double *, int64_t, float *, int64_t,
                                                     int64_t, int, hipStream_t);
template hipError_t launch_copy_block<float, float>(const float *, int64_t, float *, int64_t,
                                                   int64_t, int, hipStream_t);
template hipError_t launch_copy_block<float, double>(const float *, int64_t, double *, int64_t,
                                                     int64_t, int, hipStream_t);

template <typename T>
hipError_t launch_axpby(int64_t n, T a, const T *x, T b, T *y, hipStream_t st) {
  hipLaunchKernelGGL(axpby_kernel<T>, dim3(grid_for(n, 256, 65536)), dim3(256), 0, st, n, a, x, b,
                     y);
  return hipGetLastError();
}
template hipError_t launch_axpby<double>(int64_t, double, const double *, double, double *,
                                         hipStream_t);
template hipError_t launch_axpby<float>(int64_t, float, const float *, float, float *,
                                        hipStream_t);

hipError_t launch_residual_partial(const double *HX, int64_t ldH, const double *X, int64_t ldX,
                                   const double *lam, int64_t M, int B, double *partial,
                                   int nchunks, hipStream_t st, double *X_rescaled,
                                   const double *rescale) {
  // X_rescaled, when given, is X itself (writable): it is scaled in place
  if ((X_rescaled != nullptr) != (rescale != nullptr) || (X_rescaled && X_rescaled != X))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(residual_partial_kernel, dim3(nchunks, (B + 63) / 64), dim3(256), 0, st, HX,
                     ldH, const_cast<double *>(X), ldX, lam, M, B, partial, nchunks, rescale);
  return hipGetLastError();
}

hipError_t launch_colstats_partial(const double *HX, int64_t ldH, const double *X, int64_t ldX,
                                   int64_t M, int B, double *partial, int nchunks,
                                   hipStream_t st) {
  hipLaunchKernelGGL(colstats_partial_kernel, dim3(nchunks, (B + 63) / 64), dim3(256), 0, st, HX,
                     ldH, X, ldX, M, B, partial, nchunks);
  return hipGetLastError();
}

hipError_t launch_defect_source(const double *HX, int64_t ldH, const double *X, int64_t ldX,
                                const double *theta, double s, int64_t M, int B, float *R,
                                float *E, hipStream_t st) {
  if (M * B == 0) return hipSuccess;
  hipLaunchKernelGGL(defect_source_kernel, dim3(grid_for(M * B, 256, 65536)), dim3(256), 0, st,
                     HX, ldH, X, ldX, theta, s, M, B, R, E);
  return hipGetLastError();
}

hipError_t launch_defect_finish(double *X, int64_t ldX, const double *D, const float *E,
                                int64_t ldE, int64_t M, int B, hipStream_t st) {
  if (M * B == 0) return hipSuccess;
  hipLaunchKernelGGL(defect_finish_kernel, dim3(grid_for(M * B, 256, 65536)), dim3(256), 0, st, X,
                     ldX, D, E, ldE, M, B);
  return hipGetLastError();
}

hipError_t launch_dot_partial(const double *x, const double *y, int64_t n, double *partial,
                              int nchunks, hipStream_t st) {
  hipLaunchKernelGGL(dot_partial_kernel, dim3(nchunks), dim3(256), 0, st, x, y, n, partial,
                     nchunks);
  return hipGetLastError();
}

hipError_t launch_sum_slot(const double *partial, int n, double *slot, hipStream_t st) {
  hipLaunchKernelGGL(sum_slot_kernel, dim3(1), dim3(64), 0, st, partial, n, slot);
  return hipGetLastError();
}

hipError_t launch_axpby_slot(int64_t n, const double *slot, int form, const double *x, double b,
                             double *y, hipStream_t st) {
  if (form < 0 || form > 2) return hipErrorInvalidValue;
  hipLaunchKernelGGL(axpby_slot_kernel, dim3(grid_for(n, 256, 65536)), dim3(256), 0, st, n, slot,
                     form, x, b, y);
  return hipGetLastError();
}

hipError_t launch_sum_partials(const double *P, int nk, int m, int n, double *C, int64_t ldc,
                               hipStream_t st) {
  const int64_t mn = (int64_t)m * n;
  hipLaunchKernelGGL(sum_partials_kernel, dim3(grid_for(mn, 256, 1u << 30)), dim3(256), 0, st, P,
                     nk, m, n, C, ldc);
  return hipGetLastError();
}

hipError_t launch_sum_partials_f32(const float *P, int nk, int m, int n, double *C, int64_t ldc,
                                   hipStream_t st) {
  const int64_t mn = (int64_t)m * n;
  if (mn == 0) return hipSuccess;
  hipLaunchKernelGGL(sum_partials_f32_kernel, dim3(grid_for(mn, 256, 1u << 30)), dim3(256), 0, st,
                     P, nk, m, n, C, ldc);
  return hipGetLastError();
}

hipError_t launch_add_f32_to_f64(const float *A, int64_t lda, double *C, int64_t ldc, int m,
                                 int64_t n, hipStream_t st) {
  const int64_t mn = (int64_t)m * n;
  if (mn == 0) return hipSuccess;
  hipLaunchKernelGGL(add_f32_to_f64_kernel, dim3(grid_for(mn, 256, 65536)), dim3(256), 0, st, A,
                     lda, C, ldc, m, n);
  return hipGetLastError();
}

hipError_t launch_cast_blocks(const double *A, int n, int64_t lda, int bs, int zero_diag,
                              float *B, hipStream_t st) {
  const int64_t nn = (int64_t)n * n;
  hipLaunchKernelGGL(cast_blocks_kernel, dim3(grid_for(nn, 256, 1u << 30)), dim3(256), 0, st, A,
                     n, lda, bs, zero_diag, B);
  return hipGetLastError();
}

hipError_t launch_zero_upper(double *A, int n, int64_t lda, hipStream_t st) {
  const int64_t nn = (int64_t)n * n;
  hipLaunchKernelGGL(zero_upper_kernel, dim3(grid_for(nn, 256, 1u << 30)), dim3(256), 0, st, A, n,
                     lda);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_distribute_scaled(const int64_t *rows, const int64_t *ptr, const int64_t *cols,
                                    const double *vals, int64_t nrows, const T *g, T *X,
                                    int64_t ld, int B, hipStream_t st) {
  if (nrows == 0) return hipSuccess;
  hipLaunchKernelGGL(distribute_kernel<T>, dim3(grid_for(nrows * B, 256, 1u << 30)), dim3(256),
                     0, st, rows, ptr, cols, vals, nrows, g, X, ld, B);
  return hipGetLastError();
}
template hipError_t launch_distribute_scaled<double>(const int64_t *, const int64_t *,
                                                     const int64_t *, const double *, int64_t,
                                                     const double *, double *, int64_t, int,
                                                     hipStream_t);
template hipError_t launch_distribute_scaled<float>(const int64_t *, const int64_t *,
                                                    const int64_t *, const double *, int64_t,
                                                    const float *, float *, int64_t, int,
                                                    hipStream_t);

template <typename T>
hipError_t launch_slave_to_master_scaled(const int64_t *m_rows, const int64_t *m_ptr,
                                         const int64_t *m_slaves, const double *m_vals,
                                         int64_t nm, const T *so, T *X, int64_t ld, int B,
                                         hipStream_t st) {
  if (nm == 0) return hipSuccess;
  hipLaunchKernelGGL(slave_to_master_kernel<T>, dim3(grid_for(nm * B, 256, 1u << 30)), dim3(256),
                     0, st, m_rows, m_ptr, m_slaves, m_vals, nm, so, X, ld, B);
  return hipGetLastError();
}
template hipError_t launch_slave_to_master_scaled<double>(const int64_t *, const int64_t *,
                                                          const int64_t *, const double *,
                                                          int64_t, const double *, double *,
                                                          int64_t, int, hipStream_t);
template hipError_t launch_slave_to_master_scaled<float>(const int64_t *, const int64_t *,
                                                         const int64_t *, const double *,
                                                         int64_t, const float *, float *,
                                                         int64_t, int, hipStream_t);

template <typename T>
hipError_t launch_zero_rows_t(const int64_t *rows, int64_t nrows, T *X, int64_t ld, int B,
                              hipStream_t st) {
  if (nrows == 0) return hipSuccess;
  hipLaunchKernelGGL(zero_rows_kernel<T>, dim3(grid_for(nrows * B, 256, 1u << 30)), dim3(256), 0,
                     st, rows, nrows, X, ld, B);
  return hipGetLastError();
}
template hipError_t launch_zero_rows_t<double>(const int64_t *, int64_t, double *, int64_t, int,
                                               hipStream_t);
template hipError_t launch_zero_rows_t<float>(const int64_t *, int64_t, float *, int64_t, int,
                                              hipStream_t);

hipError_t launch_distribute(const int64_t *rows, const int64_t *ptr, const int64_t *cols,
                             const double *vals, int64_t nrows, double *X, int64_t ld, int B,
                             hipStream_t st) {
  return launch_distribute_scaled<double>(rows, ptr, cols, vals, nrows, nullptr, X, ld, B, st);
}

hipError_t launch_slave_to_master(const int64_t *m_rows, const int64_t *m_ptr,
                                  const int64_t *m_slaves, const double *m_vals, int64_t nm,
                                  double *X, int64_t ld, int B, hipStream_t st) {
  return launch_slave_to_master_scaled<double>(m_rows, m_ptr, m_slaves, m_vals, nm, nullptr, X,
                                               ld, B, st);
}

hipError_t launch_zero_rows(const int64_t *rows, int64_t nrows, double *X, int64_t ld, int B,
                            hipStream_t st) {
  return launch_zero_rows_t<double>(rows, nrows, X, ld, B, st);
}
}  // namespace chfsi

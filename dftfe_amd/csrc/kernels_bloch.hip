// gfx950 kernels of the complex (Bloch k-point) fused HX, batched over
// k-points: one launch per colour class covers every (k-point, cell, column
// tile) work item of the batch, so many k-points with uneven band counts fill
// the GPU together instead of one small launch each.
//
// Reference structure being replaced: the complex cell GEMM
// (src/dftOperator/matrixVectorProductImplementationsDevice.cc:27-117 with
// dataTypes::number = complex), the complex nonlocal term
// (computeNonLocalHamiltonianTimesXMemoryOptBatchGEMMDevice.cc:27-283) and
// the per-k-point loop of the eigensolve (src/dft/kohnShamEigenSolve.cc:335-366),
// which filters one k-point after the other.

#include "bloch.hpp"
#include "kernels.hpp"
#include "mfma.hpp"
#include "scatter.hpp"

namespace chfsi {

// Batch entry owning work item `item` (kbegin is ascending; nk is small and
// the search is uniform across the workgroup).
__device__ __forceinline__ int batch_entry(const int *kbegin, int nk, int item) {
  int k = 0;
  while (k + 1 < nk && item >= kbegin[k + 1]) ++k;
  return k;
}

// ---------------------------------------------------------------------------
// Fused complex HX over one colour class, all batch entries
// ---------------------------------------------------------------------------
// Same decomposition as the real kernel (work item = (entry, cell, column
// tile), XCD-sliced 1-D grid, 16x16x4 MFMA, first-touch combine in the
// epilogue).  The K loop runs over tile pairs: tile 2s is the real part of
// slice s against the gathered slice, tile 2s+1 the imaginary part against the
// same slice read with swapped (re, im) columns and the real column negated.
template <typename T, int NW, int TN, int WC, int NRB>
__global__ __launch_bounds__(64 * NW) void zhx_kernel(ZHxLaunch<T> a) {
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
  constexpr int NLMAX = 128;  // projector slots per cell

  __shared__ __align__(16) T Hs[2][ROWS_MAX * RS];
  __shared__ __align__(16) T Xs[2][TN * RS];
  __shared__ uint32_t smap[ROWS_MAX];
  __shared__ T sscale[ROWS_MAX];
  __shared__ T sout[ROWS_MAX];
  __shared__ int sproj[NLMAX];

  const int wid = blockIdx.x;
  const int local = wid >> 3;
  if (local >= a.per_xcd) return;
  const int item = (wid & 7) * a.per_xcd + local;
  if (item >= a.kbegin[a.nk]) return;
  const int k = batch_entry(a.kbegin, a.nk, item);
  const ZHxItem<T> it = a.items[k];
  const int ntiles = (it.B + TN - 1) / TN;
  const int rel = item - a.kbegin[k];
  const int cell = a.cells[rel / ntiles];
  const int v0 = (rel % ntiles) * TN;

  const int tid = threadIdx.x;
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
  const int ntot = 2 * (nkt + (nslot >> 4));

  const uint32_t *cmap = a.map + (int64_t)cell * a.ne;
  const T *Hcell = it.Ht + (int64_t)cell * 2 * a.ne_pad * a.ne_pad;
  const T *NLcell = a.nl_off ? it.nlC + (int64_t)nl_begin * 2 * a.ne_pad : Hcell;

  for (int i = tid; i < ROWS_MAX; i += NT) {
    uint32_t w = 0;
    T sc = T(0), so = T(1);
    if (i < a.ne) {
      w = cmap[i];
      sc = a.s_in ? a.s_in[w & 0x7fffffffu] : T(1);
      if (a.s_out) so = a.s_out[w & 0x7fffffffu];
    }
    smap[i] = w;
    sscale[i] = sc;
    sout[i] = so;
  }
  for (int s = tid; s < NLMAX; s += NT) sproj[s] = (s < nslot) ? a.slot_proj[nl_begin + s] : -1;

  const int gv = tid % TN;
  const int gk = (tid / TN) * KPT;
  const bool vok = (v0 + gv) < it.B;
  const int vcl = vok ? v0 + gv : 0;
  constexpr int nchunk = ROWS_MAX * 4;

  v4 acc[RBW][CBW];
#pragma unroll
  for (int j = 0; j < RBW; ++j)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = v4{0, 0, 0, 0};

  v4 hreg[HCH];
  T xreg[KPT], xscl[KPT];
  const T vmask = vok ? T(1) : T(0);
  // tile t: cell-matrix tile t (both parts interleaved); gathered slice t/2,
  // loaded with the even tile only
  auto load_tile = [&](int t) {
    const T *tile_src = (t < 2 * nkt) ? Hcell + (int64_t)t * ROWS_MAX * 16
                                      : NLcell + (int64_t)(t - 2 * nkt) * ROWS_MAX * 16;
#pragma unroll
    for (int j = 0; j < HCH; ++j) {
      const int chunk = (tid + NT * j < nchunk) ? tid + NT * j : nchunk - 1;
      hreg[j] = reinterpret_cast<const v4 *>(tile_src)[chunk];
    }
    if (t & 1) return;
    const int sl = t >> 1;
    if (sl < nkt) {
#pragma unroll
      for (int kk = 0; kk < KPT; ++kk) {
        const int kr = sl * 16 + gk + kk;
        const uint32_t g = smap[kr] & 0x7fffffffu;
        xreg[kk] = it.src[(int64_t)g * it.ld_src + vcl];
        xscl[kk] = vmask * sscale[kr];
      }
    } else {
#pragma unroll
      for (int kk = 0; kk < KPT; ++kk) {
        const int pj = sproj[(sl - nkt) * 16 + gk + kk];
        xreg[kk] = it.W[(int64_t)(pj < 0 ? 0 : pj) * it.ldW + vcl];
        xscl[kk] = pj >= 0 ? vmask : T(0);
      }
    }
  };
  auto store_tile = [&](int t) {
    const int hb = t & 1;
#pragma unroll
    for (int j = 0; j < HCH; ++j) {
      const int chunk = tid + NT * j;
      if (chunk < nchunk)
        *reinterpret_cast<v4 *>(&Hs[hb][(chunk >> 2) * RS + (chunk & 3) * 4]) = hreg[j];
    }
    if (t & 1) return;
    const int xb = (t >> 1) & 1;
#pragma unroll
    for (int kk = 0; kk < KPT; ++kk) Xs[xb][gv * RS + gk + kk] = xscl[kk] * xreg[kk];
  };

  __syncthreads();
  if (ntot > 0) {
    load_tile(0);
    store_tile(0);
  }
  __syncthreads();
  // B operand of an imaginary tile: column pair swapped, real column negated
  const T sgn = (lane & 1) ? T(1) : T(-1);
  for (int t = 0; t < ntot; ++t) {
    const int hb = t & 1;
    const int xb = (t >> 1) & 1;
    const int flip = t & 1;
    const T mul = flip ? sgn : T(1);
    if (t + 1 < ntot) load_tile(t + 1);
    v4 b4[CBW];
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int col = ((wc * CBW + cb) * 16 + (lane & 15)) ^ flip;
      const v4 x = *reinterpret_cast<const v4 *>(&Xs[xb][col * RS + (lane >> 4) * 4]);
      b4[cb] = x * mul;
    }
    constexpr int GMAX = sizeof(T) == 8 ? 2 : 4;
    constexpr int G = RBW < GMAX ? RBW : GMAX;
    const T *hbp = &Hs[hb][(lane & 15) * RS + (lane >> 4) * 4];
    v4 a4[G], a4n[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int rb = wr + WR * g;
      a4[g] = *reinterpret_cast<const v4 *>(hbp + (rb < nrb ? rb : 0) * 16 * RS);
    }
#pragma unroll
    for (int j0 = 0; j0 < RBW; j0 += G) {
      if (j0 + G < RBW) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const int jn = j0 + G + g;
          const int rb = wr + WR * jn;
          if (jn < RBW) a4n[g] = *reinterpret_cast<const v4 *>(hbp + (rb < nrb ? rb : 0) * 16 * RS);
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
    if (t + 1 < ntot) store_tile(t + 1);
    __syncthreads();
  }

  // epilogue: scaled scatter with the first-touch combine (real scalars)
  hx_scatter<T, NRB, WR, RBW, CBW, 8>(it, a.ne, acc, smap, sout, wr, wc, lane, v0);
}

// ---------------------------------------------------------------------------
// Complex projection partials  P_slot[v] = sum_i conj(C_e[i, slot]) (s_in src)[g_i, v]
// ---------------------------------------------------------------------------
// Tile 2t: Re C against node slice t; tile 2t+1: Im C against the same slice
// with the column pair swapped and the imaginary column negated.
// OWN: the rows are the cell's owned DOF rows (own_map, own_n) with the
// projector values assembled over the cell nodes, and only their tiles are
// visited (as proj_kernel's owned-row form).
template <typename T, int NW, int TN, int MAXSRB, int NRB, bool OWN = false>
__global__ __launch_bounds__(64 * NW) void zproj_kernel(ZProjLaunch<T> a) {
  using v4 = typename V4<T>::t;
  constexpr int NT = 64 * NW;
  constexpr int RS = row_stride<T>();
  constexpr int KPT = kTile * TN / NT;
  constexpr int CBW = TN / 16 / NW;
  constexpr int ROWS = NRB * 16;
  constexpr int ACH = (MAXSRB * 16 * 4 + NT - 1) / NT;
  static_assert(KPT >= 1 && CBW >= 1, "bad projection tile");
  __shared__ __align__(16) T As[2][MAXSRB * 16 * RS];
  __shared__ __align__(16) T Xs[2][TN * RS];
  __shared__ uint32_t smap[ROWS];
  __shared__ T sscale[ROWS];

  const int wid = blockIdx.x;
  const int local = wid >> 3;
  if (local >= a.per_xcd) return;
  const int item = (wid & 7) * a.per_xcd + local;
  if (item >= a.kbegin[a.nk]) return;
  const int k = batch_entry(a.kbegin, a.nk, item);
  const ZHxItem<T> it = a.items[k];
  const int ntiles = (it.B + TN - 1) / TN;
  const int rel = item - a.kbegin[k];
  const int cell = a.cells[rel / ntiles];
  const int v0 = (rel % ntiles) * TN;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int s0 = a.nl_off[cell];
  const int nslot = a.nl_off[cell + 1] - s0;
  const int nsrb = nslot >> 4;
  const int nchunk = nslot * 4;
  const T *Acell = it.nlCt + (int64_t)s0 * ROWS * 2;
  const uint32_t *cmap = OWN ? a.own_map + (int64_t)cell * ROWS : a.map + (int64_t)cell * a.ne;
  const int nrow = OWN ? a.own_n[cell] : a.ne;
  // (Re, Im) tile pairs holding rows; the tiles past an owned-row cell's last
  // row are zero and read row 0 with scale 0
  const int ntot = OWN ? 2 * (nrow > 0 ? (nrow + 15) >> 4 : 1) : 2 * NRB;

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
  const bool vok = (v0 + gv) < it.B;
  const int vcl = vok ? v0 + gv : 0;
  const T vmask = vok ? T(1) : T(0);

  v4 acc[MAXSRB][CBW];
#pragma unroll
  for (int j = 0; j < MAXSRB; ++j)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = v4{0, 0, 0, 0};
  v4 areg[ACH];
  T xreg[KPT], xscl[KPT];

  auto load_tile = [&](int t) {
    const T *tile_src = Acell + (int64_t)t * nslot * 16;
#pragma unroll
    for (int j = 0; j < ACH; ++j) {
      const int chunk = (tid + NT * j < nchunk) ? tid + NT * j : nchunk - 1;
      areg[j] = reinterpret_cast<const v4 *>(tile_src)[chunk];
    }
    if (t & 1) return;
#pragma unroll
    for (int kk = 0; kk < KPT; ++kk) {
      const int kr = (t >> 1) * 16 + gk + kk;
      xreg[kk] = it.src[(int64_t)(smap[kr] & 0x7fffffffu) * it.ld_src + vcl];
      xscl[kk] = vmask * sscale[kr];
    }
  };
  auto store_tile = [&](int t) {
    const int ab = t & 1;
#pragma unroll
    for (int j = 0; j < ACH; ++j) {
      const int chunk = tid + NT * j;
      if (chunk < nchunk)
        *reinterpret_cast<v4 *>(&As[ab][(chunk >> 2) * RS + (chunk & 3) * 4]) = areg[j];
    }
    if (t & 1) return;
    const int xb = (t >> 1) & 1;
#pragma unroll
    for (int kk = 0; kk < KPT; ++kk) Xs[xb][gv * RS + gk + kk] = xscl[kk] * xreg[kk];
  };

  __syncthreads();
  load_tile(0);
  store_tile(0);
  __syncthreads();
  const T sgn = (lane & 1) ? T(-1) : T(1);
  for (int t = 0; t < ntot; ++t) {
    const int ab = t & 1;
    const int xb = (t >> 1) & 1;
    const int flip = t & 1;
    const T mul = flip ? sgn : T(1);
    if (t + 1 < ntot) load_tile(t + 1);
    v4 b4[CBW];
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int col = ((wave * CBW + cb) * 16 + (lane & 15)) ^ flip;
      const v4 x = *reinterpret_cast<const v4 *>(&Xs[xb][col * RS + (lane >> 4) * 4]);
      b4[cb] = x * mul;
    }
#pragma unroll
    for (int j = 0; j < MAXSRB; ++j) {
      if (j < nsrb) {
        const v4 a4 =
            *reinterpret_cast<const v4 *>(&As[ab][(j * 16 + (lane & 15)) * RS + (lane >> 4) * 4]);
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int cb = 0; cb < CBW; ++cb) acc[j][cb] = mfma16(a4[r], b4[cb][r], acc[j][cb]);
      }
    }
    if (t + 1 < ntot) store_tile(t + 1);
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < MAXSRB; ++j) {
    if (j >= nsrb) continue;
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int v = v0 + (wave * CBW + cb) * 16 + (lane & 15);
      if (v >= it.B) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int srow = j * 16 + acc_row<T>(lane, r);
        it.partial[(int64_t)(s0 + srow) * it.ldP + v] = acc[j][cb][r];
      }
    }
  }
}

// W[p][v] = V_p sum over the projector's slots of the partials (all entries)
template <typename T>
__global__ void zproj_reduce_kernel(const ZHxItem<T> *items, const int *proj_ptr,
                                    const int *proj_slots, const double *proj_V, int n_proj) {
  const ZHxItem<T> it = items[blockIdx.y];
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n_proj * it.B) return;
  const int p = (int)(idx / it.B);
  const int v = (int)(idx % it.B);
  T s = T(0);
  for (int j = proj_ptr[p]; j < proj_ptr[p + 1]; ++j)
    s += it.partial[(int64_t)proj_slots[j] * it.ldP + v];
  it.W[(int64_t)p * it.ldW + v] = T(proj_V[p]) * s;
}

// Complex cell matrices of one k-point in the paired tile layout.
__global__ void build_hk_kernel(const double *Ht64, const double *Mc, const double *Gd, int C,
                                int ne, int ne_pad, double k0, double k1, double k2,
                                double *Hk64, float *Hk32) {
  const int64_t per = 2 * (int64_t)ne_pad * ne_pad;
  const int64_t n = per * C;
  const double half_k2 = 0.5 * (k0 * k0 + k1 * k1 + k2 * k2);
  const int64_t ne2 = (int64_t)ne * ne;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = idx / per;
    const int64_t r = idx % per;
    const int t2 = (int)(r / ((int64_t)ne_pad * 16));
    const int i = (int)((r / 16) % ne_pad);
    const int jj = (int)(r % 16);
    const int t = t2 >> 1;
    const int j = t * 16 + jj;
    double v = 0.0;
    if (i < ne && j < ne) {
      const int64_t ij = (int64_t)i * ne + j;
      if ((t2 & 1) == 0)
        v = Ht64[c * ne_pad * ne_pad + (int64_t)t * ne_pad * 16 + (int64_t)i * 16 + jj] +
            half_k2 * Mc[ij];
      else
        v = -(k0 * Gd[ij] + k1 * Gd[ne2 + ij] + k2 * Gd[2 * ne2 + ij]);
    }
    Hk64[idx] = v;
    if (Hk32) Hk32[idx] = (float)v;
  }
}

template <typename T>
__global__ void fill_zero_kernel(T *x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = T(0);
}

__global__ void zero_upper_complex_kernel(double *A, int n, int64_t lda) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  if (i < j) {
    A[2 * (i + j * lda)] = 0.0;
    A[2 * (i + j * lda) + 1] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline unsigned zgrid(int64_t n, int block, unsigned cap) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

template <typename T, int NW, int TN, int WC, int NRB>
static hipError_t zhx_go(ZHxLaunch<T> a, const ZHxItem<T> *h, hipStream_t s) {
  int total = 0;
  for (int k = 0; k < a.nk; ++k) {
    a.kbegin[k] = total;
    total += a.ncells * ((h[k].B + TN - 1) / TN);
  }
  a.kbegin[a.nk] = total;
  if (total == 0) return hipSuccess;
  a.per_xcd = (total + 7) / 8;
  hipLaunchKernelGGL((zhx_kernel<T, NW, TN, WC, NRB>), dim3(8 * a.per_xcd), dim3(64 * NW), 0, s,
                     a);
  return hipGetLastError();
}

template <typename T, int NW, int TN, int WC>
static hipError_t zhx_dispatch(const ZHxLaunch<T> &a, const ZHxItem<T> *h, hipStream_t s) {
  switch (a.ne_pad >> 4) {
    case 1: return zhx_go<T, NW, TN, WC, 1>(a, h, s);
    case 2: return zhx_go<T, NW, TN, WC, 2>(a, h, s);
    case 4: return zhx_go<T, NW, TN, WC, 4>(a, h, s);
    case 8: return zhx_go<T, NW, TN, WC, 8>(a, h, s);
    case 14: return zhx_go<T, NW, TN, WC, 14>(a, h, s);
    case 22: return zhx_go<T, NW, TN, WC, 22>(a, h, s);
    default: return hipErrorInvalidValue;
  }
}

// Tile variants (waves x real-column tile x column waves), as for the real
// kernel: FP32 0: 8x128x8, 1: 4x64x4, 2: 8x64x4, 3: 16x128x8, 4: 8x32x2;
// FP64 0: 8x64x4, 1: 4x32x2, 2: 16x64x4.
int zhx_variants(bool fp64) { return fp64 ? 3 : 5; }
int zhx_tile_width(bool fp64, int v) {
  if (fp64) return v == 1 ? 32 : 64;
  switch (v) {
    case 1: case 2: return 64;
    case 4: return 32;
    default: return 128;
  }
}

template <>
hipError_t launch_zhx<float>(ZHxLaunch<float> a, const ZHxItem<float> *h, int variant,
                             hipStream_t s) {
  if (a.nk < 1 || a.nk > kMaxBatch) return hipErrorInvalidValue;
  switch (variant) {
    case 1: return zhx_dispatch<float, 4, 64, 4>(a, h, s);
    case 2: return zhx_dispatch<float, 8, 64, 4>(a, h, s);
    case 3: return zhx_dispatch<float, 16, 128, 8>(a, h, s);
    case 4: return zhx_dispatch<float, 8, 32, 2>(a, h, s);
    default: return zhx_dispatch<float, 8, 128, 8>(a, h, s);
  }
}

template <>
hipError_t launch_zhx<double>(ZHxLaunch<double> a, const ZHxItem<double> *h, int variant,
                              hipStream_t s) {
  if (a.nk < 1 || a.nk > kMaxBatch) return hipErrorInvalidValue;
  switch (variant) {
    case 1: return zhx_dispatch<double, 4, 32, 2>(a, h, s);
    case 2: return zhx_dispatch<double, 16, 64, 4>(a, h, s);
    default: return zhx_dispatch<double, 8, 64, 4>(a, h, s);
  }
}

template <typename T, int NRB, int MAXSRB, bool OWN>
static hipError_t zproj_go(ZProjLaunch<T> a, const ZHxItem<T> *h, hipStream_t s) {
  constexpr int TN = sizeof(T) == 4 ? 128 : 64;
  int total = 0;
  for (int k = 0; k < a.nk; ++k) {
    a.kbegin[k] = total;
    total += a.ncells * ((h[k].B + TN - 1) / TN);
  }
  a.kbegin[a.nk] = total;
  if (total == 0) return hipSuccess;
  a.per_xcd = (total + 7) / 8;
  hipLaunchKernelGGL((zproj_kernel<T, 4, TN, MAXSRB, NRB, OWN>), dim3(8 * a.per_xcd), dim3(256), 0,
                     s, a);
  return hipGetLastError();
}

// accumulators sized to the largest slot count of any cell
template <typename T, int NRB, bool OWN>
static hipError_t zproj_msrb(ZProjLaunch<T> a, const ZHxItem<T> *h, hipStream_t s) {
  if (a.max_srb <= 2) return zproj_go<T, NRB, 2, OWN>(a, h, s);
  if (a.max_srb <= 4) return zproj_go<T, NRB, 4, OWN>(a, h, s);
  return zproj_go<T, NRB, 8, OWN>(a, h, s);
}

template <typename T, int NRB>
static hipError_t zproj_fit(ZProjLaunch<T> a, const ZHxItem<T> *h, hipStream_t s) {
  if (a.own_map && a.own_n) return zproj_msrb<T, NRB, true>(a, h, s);
  return zproj_msrb<T, NRB, false>(a, h, s);
}

template <typename T>
hipError_t launch_zproj(ZProjLaunch<T> a, const ZHxItem<T> *h, hipStream_t s) {
  if (a.nk < 1 || a.nk > kMaxBatch || a.max_srb < 1 || a.max_srb > 8) return hipErrorInvalidValue;
  switch (a.ne_pad >> 4) {
    case 1: return zproj_fit<T, 1>(a, h, s);
    case 2: return zproj_fit<T, 2>(a, h, s);
    case 4: return zproj_fit<T, 4>(a, h, s);
    case 8: return zproj_fit<T, 8>(a, h, s);
    case 14: return zproj_fit<T, 14>(a, h, s);
    case 22: return zproj_fit<T, 22>(a, h, s);
    default: return hipErrorInvalidValue;
  }
}
template hipError_t launch_zproj<float>(ZProjLaunch<float>, const ZHxItem<float> *, hipStream_t);
template hipError_t launch_zproj<double>(ZProjLaunch<double>, const ZHxItem<double> *,
                                         hipStream_t);

template <typename T>
hipError_t launch_zproj_reduce(const ZHxItem<T> *items, int nk, int maxB, const int *proj_ptr,
                               const int *proj_slots, const double *proj_V, int n_proj,
                               hipStream_t s) {
  const int64_t n = (int64_t)n_proj * maxB;
  if (n == 0 || nk == 0) return hipSuccess;
  hipLaunchKernelGGL(zproj_reduce_kernel<T>, dim3(zgrid(n, 256, 1u << 30), nk), dim3(256), 0, s,
                     items, proj_ptr, proj_slots, proj_V, n_proj);
  return hipGetLastError();
}
template hipError_t launch_zproj_reduce<float>(const ZHxItem<float> *, int, int, const int *,
                                               const int *, const double *, int, hipStream_t);
template hipError_t launch_zproj_reduce<double>(const ZHxItem<double> *, int, int, const int *,
                                                const int *, const double *, int, hipStream_t);

hipError_t launch_build_hk(const double *Ht64, const double *Mc, const double *Gd, int C, int ne,
                           int ne_pad, double k0, double k1, double k2, double *Hk64, float *Hk32,
                           hipStream_t s) {
  const int64_t n = 2 * (int64_t)ne_pad * ne_pad * C;
  hipLaunchKernelGGL(build_hk_kernel, dim3(zgrid(n, 256, 65536)), dim3(256), 0, s, Ht64, Mc, Gd,
                     C, ne, ne_pad, k0, k1, k2, Hk64, Hk32);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fill_zero(T *x, int64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_zero_kernel<T>, dim3(zgrid(n, 256, 65536)), dim3(256), 0, s, x, n);
  return hipGetLastError();
}
template hipError_t launch_fill_zero<double>(double *, int64_t, hipStream_t);
template hipError_t launch_fill_zero<float>(float *, int64_t, hipStream_t);

hipError_t launch_zero_upper_complex(double *A, int n, int64_t lda, hipStream_t st) {
  const int64_t nn = (int64_t)n * n;
  hipLaunchKernelGGL(zero_upper_complex_kernel, dim3(zgrid(nn, 256, 1u << 30)), dim3(256), 0, st,
                     A, n, lda);
  return hipGetLastError();
}

}  // namespace chfsi

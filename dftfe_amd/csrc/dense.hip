// Dense FP64 steps of CholGS and Rayleigh-Ritz on the tall node-major block
// X (M x N, M ~ 10^6, N ~ 10^3) as FP64 MFMA kernels:
//   * gram:   C = A^T B over all M rows (the overlap X^T X and the projected
//             Hamiltonian X^T (H~X)), lower-triangle tiles only when asked,
//             split over M into chunks whose partial tiles are summed in a
//             fixed order (deterministic);
//   * rotate: Y = X R over row panels (R = Q, the Ritz rotation, or a column
//             block of R = L^{-T}, the Cholesky-Gram-Schmidt one, with the
//             inner extent cut at the block's last row: the blocks below the
//             diagonal are zero).
// Reference steps replaced: fillParallelOverlapMatScalapack
// (src/linAlg/linearAlgebraOperationsDevice.cc:3078-3240), XtHX
// (src/dftOperator/kohnShamDFTOperatorDevice.cc:4000-4157),
// subspaceRotationScalapack (linearAlgebraOperationsDevice.cc:1832-2242) and
// the rotations of pseudoGSDevice.cc / rayleighRitzDevice.cc, which call
// hipBLAS dgemm per block.

#include <stdint.h>

#include "dense.hpp"
#include "mfma.hpp"

namespace chfsi {

namespace {

constexpr int TT = 128;   // output tile edge
constexpr int KS = 16;    // K rows per LDS slab
constexpr int NT = 256;   // 4 waves, each a 64 x 64 quarter of the tile
constexpr int PADW = 16;  // LDS row padding of [k][128] slabs (doubles)

// Work items are spread over the 8 XCDs so that consecutive items (the tiles
// of one K chunk in gram, the column tiles of one row panel in rotate) share
// an XCD and its L2.
__device__ __forceinline__ int xcd_item(int per_xcd) {
  const int wid = blockIdx.x;
  return (wid & 7) * per_xcd + (wid >> 3);
}

}  // namespace

// C_tile(I, J) = sum_{k in chunk} A[k, I + i] B[k, J + j], written to
// P[item][i + 128 j] (item = chunk * ntiles + tile)
__global__ __launch_bounds__(NT, 2) void gram_kernel(const double *A, int64_t lda, const double *B,
                                                     int64_t ldb, int m, int n, int64_t K,
                                                     int64_t kchunk, const int2 *tiles,
                                                     int ntiles, int per_xcd, int nitems,
                                                     double *P) {
  __shared__ __align__(16) double As[2][KS][TT + PADW];
  __shared__ __align__(16) double Bs[2][KS][TT + PADW];
  const int item = xcd_item(per_xcd);
  if ((blockIdx.x >> 3) >= per_xcd || item >= nitems) return;
  const int chunk = item / ntiles;
  const int2 tl = tiles[item % ntiles];
  const int I = tl.x * TT, J = tl.y * TT;
  const int64_t kb = (int64_t)chunk * kchunk;
  const int64_t ke = kb + kchunk < K ? kb + kchunk : K;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wi = w & 1, wj = w >> 1;

  // staging: thread t loads row t/16 of the slab, 8 consecutive columns
  const int lr = tid >> 4, lc = (tid & 15) * 8;
  double ra[8], rb[8];
  // 16-byte loads when the row pieces are aligned and inside the matrix
  const bool avec = (lda % 2 == 0) && (((uintptr_t)A & 15) == 0) && I + lc + 8 <= m;
  const bool bvec = (ldb % 2 == 0) && (((uintptr_t)B & 15) == 0) && J + lc + 8 <= n;
  auto load = [&](int64_t k0) {
    const int64_t k = k0 + lr;
    const bool kok = k < ke;
    const int64_t kc = kok ? k : kb;
    const double z = kok ? 1.0 : 0.0;
    if (avec) {
      const double2 *p = reinterpret_cast<const double2 *>(A + kc * lda + I + lc);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const double2 v = p[e];
        ra[2 * e] = z * v.x;
        ra[2 * e + 1] = z * v.y;
      }
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int i = I + lc + e;
        const double a = A[kc * lda + (i < m ? i : 0)];
        ra[e] = (kok && i < m) ? a : 0.0;
      }
    }
    if (bvec) {
      const double2 *p = reinterpret_cast<const double2 *>(B + kc * ldb + J + lc);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const double2 v = p[e];
        rb[2 * e] = z * v.x;
        rb[2 * e + 1] = z * v.y;
      }
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int j = J + lc + e;
        const double b = B[kc * ldb + (j < n ? j : 0)];
        rb[e] = (kok && j < n) ? b : 0.0;
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
      *reinterpret_cast<double2 *>(&As[buf][lr][lc + e]) = double2{ra[e], ra[e + 1]};
      *reinterpret_cast<double2 *>(&Bs[buf][lr][lc + e]) = double2{rb[e], rb[e + 1]};
    }
  };

  f64x4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f64x4{0, 0, 0, 0};

  const int nslab = (int)((ke - kb + KS - 1) / KS);
  if (nslab > 0) {
    load(kb);
    store(0);
  }
  __syncthreads();
  for (int s = 0; s < nslab; ++s) {
    const int buf = s & 1;
    if (s + 1 < nslab) load(kb + (int64_t)(s + 1) * KS);
#pragma unroll
    for (int kk = 0; kk < KS / 4; ++kk) {
      const int k = kk * 4 + (lane >> 4);
      double av[4], bv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        av[q] = As[buf][k][wi * 64 + q * 16 + (lane & 15)];
        bv[q] = Bs[buf][k][wj * 64 + q * 16 + (lane & 15)];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = mfma16(av[a], bv[b], acc[a][b]);
    }
    if (s + 1 < nslab) store(buf ^ 1);
    __syncthreads();
  }
  double *Pt = P + (int64_t)item * TT * TT;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = wi * 64 + a * 16 + acc_row<double>(lane, r);
        const int j = wj * 64 + b * 16 + (lane & 15);
        Pt[i + (int64_t)j * TT] = acc[a][b][r];
      }
}

// C[I + i + (J + j) ldc] = sum over chunks of the partial tiles (chunk order)
__global__ void gram_reduce_kernel(const double *P, const int2 *tiles, int ntiles, int nchunks,
                                   int m, int n, double *C, int64_t ldc) {
  const int t = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= TT * TT) return;
  const int i = e % TT, j = e / TT;
  const int gi = tiles[t].x * TT + i, gj = tiles[t].y * TT + j;
  if (gi >= m || gj >= n) return;
  double s = 0.0;
  for (int c = 0; c < nchunks; ++c) s += P[((int64_t)c * ntiles + t) * TT * TT + e];
  C[gi + (int64_t)gj * ldc] = s;
}

// Y[k, J + j] = sum_{i < K} X[k, i] R[i, J + j] for the rows [r0, r0 + rows)
// of the panel and J + j < N; R[i][j] = Rm[i + j ldr] (TR = 0) or
// Rm[j + i ldr] (TR = 1).  LDS slabs: X as [128 rows][KS + 1] (odd stride:
// the 16 rows of an A-operand read fall in distinct banks); R as
// [128 columns][KS + 1] when its columns are contiguous in memory (TR = 0),
// else [KS rows][128 + PADW].
template <int TR>
__global__ __launch_bounds__(NT, 2) void rotate_kernel(const double *X, int64_t ldx,
                                                       const double *Rm, int64_t ldr, int K,
                                                       int N, int64_t rows, double *Y,
                                                       int64_t ldy, int ntj, int per_xcd,
                                                       int nitems) {
  constexpr int XS = KS + 1;
  constexpr int RSZ = TR ? KS * (TT + PADW) : TT * XS;
  constexpr int HK = KS / 2;  // doubles per thread and row of a slab
  __shared__ __align__(16) double Xs[2][TT * XS];
  __shared__ __align__(16) double Rs[2][RSZ];
  const int item = xcd_item(per_xcd);
  if ((blockIdx.x >> 3) >= per_xcd || item >= nitems) return;
  const int64_t R0 = (int64_t)(item / ntj) * TT;
  const int J = (item % ntj) * TT;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wi = w & 1, wj = w >> 1;
  const int nslab = (K + KS - 1) / KS;

  // X slab: thread t takes row t/2, HK consecutive columns from HK*(t%2)
  const int xr = tid >> 1, xc = (tid & 1) * HK;
  double rx[HK], rr[HK];
  // 16-byte loads when rows are aligned and every slab lies inside the matrix
  const bool vec = (ldx % 2 == 0) && (ldr % 2 == 0) && (((uintptr_t)X & 15) == 0) &&
                   (((uintptr_t)Rm & 15) == 0) && (K % KS == 0);
  auto load = [&](int i0) {
    const int64_t k = R0 + xr;
    const bool kok = k < rows;
    const double *xp = X + (kok ? k : 0) * ldx;
    if (vec) {
      const double z = kok ? 1.0 : 0.0;
      const double2 *p = reinterpret_cast<const double2 *>(xp + i0 + xc);
#pragma unroll
      for (int e = 0; e < HK / 2; ++e) {
        const double2 v = p[e];
        rx[2 * e] = z * v.x;
        rx[2 * e + 1] = z * v.y;
      }
    } else {
#pragma unroll
      for (int e = 0; e < HK; ++e) {
        const int i = i0 + xc + e;
        const double v = xp[i < K ? i : 0];
        rx[e] = (kok && i < K) ? v : 0.0;
      }
    }
    if (TR) {
      // (i, j) at Rm[j + i ldr]: thread t takes row i0 + t/16, 8 consecutive j
      const int ii = i0 + (tid >> 4), jj = J + (tid & 15) * 8;
      if (vec && jj + 8 <= N) {
        const double z = ii < K ? 1.0 : 0.0;
        const double2 *p = reinterpret_cast<const double2 *>(Rm + (int64_t)(ii < K ? ii : 0) * ldr + jj);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const double2 v = p[e];
          rr[2 * e] = z * v.x;
          rr[2 * e + 1] = z * v.y;
        }
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int j = jj + e;
          const double v = Rm[(int64_t)(ii < K ? ii : 0) * ldr + (j < N ? j : 0)];
          rr[e] = (ii < K && j < N) ? v : 0.0;
        }
      }
    } else {
      // (i, j) at Rm[i + j ldr]: thread t takes column J + t/2, HK consecutive i
      const int j = J + (tid >> 1), ib = i0 + (tid & 1) * HK;
      const double *rp = Rm + (int64_t)(j < N ? j : 0) * ldr;
      if (vec && j < N) {
        const double2 *p = reinterpret_cast<const double2 *>(rp + ib);
#pragma unroll
        for (int e = 0; e < HK / 2; ++e) {
          const double2 v = p[e];
          rr[2 * e] = v.x;
          rr[2 * e + 1] = v.y;
        }
      } else {
#pragma unroll
        for (int e = 0; e < HK; ++e) {
          const int i = ib + e;
          const double v = rp[i < K ? i : 0];
          rr[e] = (i < K && j < N) ? v : 0.0;
        }
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < HK; ++e) Xs[buf][xr * XS + xc + e] = rx[e];
    if (TR) {
      const int ii = tid >> 4, jj = (tid & 15) * 8;
#pragma unroll
      for (int e = 0; e < 8; e += 2)
        *reinterpret_cast<double2 *>(&Rs[buf][ii * (TT + PADW) + jj + e]) = double2{rr[e], rr[e + 1]};
    } else {
      const int jl = tid >> 1, ib = (tid & 1) * HK;
#pragma unroll
      for (int e = 0; e < HK; ++e) Rs[buf][jl * XS + ib + e] = rr[e];
    }
  };

  f64x4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f64x4{0, 0, 0, 0};
  load(0);
  store(0);
  __syncthreads();
  for (int s = 0; s < nslab; ++s) {
    const int buf = s & 1;
    if (s + 1 < nslab) load((s + 1) * KS);
#pragma unroll
    for (int kk = 0; kk < KS / 4; ++kk) {
      const int k = kk * 4 + (lane >> 4);
      double av[4], bv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = q * 16 + (lane & 15);
        av[q] = Xs[buf][(wi * 64 + c) * XS + k];
        bv[q] = TR ? Rs[buf][k * (TT + PADW) + wj * 64 + c] : Rs[buf][(wj * 64 + c) * XS + k];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = mfma16(av[a], bv[b], acc[a][b]);
    }
    if (s + 1 < nslab) store(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t k = R0 + wi * 64 + a * 16 + acc_row<double>(lane, r);
      if (k >= rows) continue;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int j = J + wj * 64 + b * 16 + (lane & 15);
        if (j < N) Y[k * ldy + j] = acc[a][b][r];
      }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static int64_t gram_chunk_rows(int64_t K, int ntiles) {
  // at most one wave of work items over the chip's 512 workgroup slots (256
  // CUs, two workgroups each): a second, partial wave would double the time.
  // Chunks of at least 4096 rows, at most 256 chunks.
  int64_t nch = 512 / ntiles;
  nch = nch < 1 ? 1 : (nch > 256 ? 256 : nch);
  int64_t kc = (K + nch - 1) / nch;
  if (kc < 4096) kc = 4096;
  return (kc + KS - 1) / KS * KS;
}

static std::vector<int2> gram_tiles(int m, int n, int diag_off) {
  std::vector<int2> t;
  const int tm = (m + TT - 1) / TT, tn = (n + TT - 1) / TT;
  for (int J = 0; J < tn; ++J)
    for (int I = 0; I < tm; ++I)
      if (diag_off < 0 || (I + 1) * TT - 1 >= J * TT + diag_off) t.push_back(int2{I, J});
  return t;
}

size_t gram_partial_bytes(int64_t K, int m, int n, int diag_off) {
  const int nt = (int)gram_tiles(m, n, diag_off).size();
  if (nt == 0) return 0;
  const int64_t kc = gram_chunk_rows(K, nt);
  const int64_t nch = (K + kc - 1) / kc;
  return sizeof(double) * (size_t)nch * nt * TT * TT;
}

hipError_t launch_gram(const double *A, int64_t lda, const double *B, int64_t ldb, int m, int n,
                       int64_t K, int diag_off, double *C, int64_t ldc, double *P,
                       int2 *tiles_dev, hipStream_t s) {
  const std::vector<int2> t = gram_tiles(m, n, diag_off);
  const int nt = (int)t.size();
  if (nt == 0 || K <= 0) return hipSuccess;
  if (nt > kGramMaxTiles) return hipErrorInvalidValue;
  hipError_t e = hipMemcpyAsync(tiles_dev, t.data(), sizeof(int2) * nt, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  const int64_t kc = gram_chunk_rows(K, nt);
  const int nch = (int)((K + kc - 1) / kc);
  const int nitems = nch * nt;
  const int per_xcd = (nitems + 7) / 8;
  hipLaunchKernelGGL(gram_kernel, dim3(8 * per_xcd), dim3(NT), 0, s, A, lda, B, ldb, m, n, K, kc,
                     tiles_dev, nt, per_xcd, nitems, P);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(gram_reduce_kernel, dim3(TT * TT / 256, nt), dim3(256), 0, s, P, tiles_dev, nt,
                     nch, m, n, C, ldc);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  // the host table is read by the copy above before this call returns
  return hipStreamSynchronize(s);
}

hipError_t launch_rotate(const double *X, int64_t ldx, const double *Rm, int64_t ldr, int trans,
                         int K, int N, int64_t rows, double *Y, int64_t ldy, hipStream_t s) {
  if (rows <= 0 || N <= 0 || K <= 0) return hipSuccess;
  const int ntj = (N + TT - 1) / TT;
  const int64_t nti = (rows + TT - 1) / TT;
  const int64_t nitems64 = nti * ntj;
  if (nitems64 > (int64_t(1) << 30)) return hipErrorInvalidValue;
  const int nitems = (int)nitems64;
  const int per_xcd = (nitems + 7) / 8;
  if (trans)
    hipLaunchKernelGGL(rotate_kernel<1>, dim3(8 * per_xcd), dim3(NT), 0, s, X, ldx, Rm, ldr, K,
                       N, rows, Y, ldy, ntj, per_xcd, nitems);
  else
    hipLaunchKernelGGL(rotate_kernel<0>, dim3(8 * per_xcd), dim3(NT), 0, s, X, ldx, Rm, ldr, K,
                       N, rows, Y, ldy, ntj, per_xcd, nitems);
  return hipGetLastError();
}

}  // namespace chfsi

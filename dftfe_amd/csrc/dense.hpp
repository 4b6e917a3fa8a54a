// FP64 MFMA kernels of the dense CholGS / Rayleigh-Ritz steps (dense.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace chfsi {

// largest number of 128 x 128 output tiles of one gram call (N <= 4096)
constexpr int kGramMaxTiles = 1024;

// C[i + j ldc] = sum_{k < K} A[k lda + i] B[k ldb + j] for i < m, j < n
// (A^T B of node-major blocks), split over K with partial tiles in P summed in
// a fixed order.  diag_off >= 0: only the tiles that reach the lower triangle
// of C, i >= j + diag_off (the diagonal tiles are written whole); -1: all.
// P holds gram_partial_bytes(K, m, n, diag_off) bytes and tiles_dev
// kGramMaxTiles entries; synchronises the stream (the tile table is copied
// from the host).
hipError_t launch_gram(const double *A, int64_t lda, const double *B, int64_t ldb, int m, int n,
                       int64_t K, int diag_off, double *C, int64_t ldc, double *P,
                       int2 *tiles_dev, hipStream_t s);
size_t gram_partial_bytes(int64_t K, int m, int n, int diag_off);

// Y[k ldy + j] = sum_{i < K} X[k ldx + i] R(i, j) for k < rows, j < N, with
// R(i, j) = Rm[i + j ldr] (trans = 0) or Rm[j + i ldr] (trans = 1).  Y must
// not overlap X.
hipError_t launch_rotate(const double *X, int64_t ldx, const double *Rm, int64_t ldr, int trans,
                         int K, int N, int64_t rows, double *Y, int64_t ldy, hipStream_t s);

}  // namespace chfsi

// Split-BF16 (Ozaki-style) products on gfx950 matrix cores.
//
// Every FP64 row x (a final eigenvector record, a row of mu, a projected
// vector) is cut along K into chunks of C elements (C = 256 for S <= 3,
// 128 for S = 4).  Each chunk is scaled by 2^-e with e the exponent of its
// largest magnitude (block floating point), then cut into S slices
// x 2^-e = sum_i n_i 2^{-8(i+1)},  |n_0| <= 256, |n_i| <= 128 (i >= 1): each
// slice has at most 8 significant bits and is exact in bf16.  Slice products
// n_i m_j are exact in FP32; products of one level L = i + j are summed in
// FP32 over one chunk, which is short enough that every partial sum stays
// below 2^24 units of the level's grid, so the chunk sums are exact.  Each
// chunk sum is scaled by 2^{e_A + e_B} and accumulated in FP64.  The only
// error is the truncation of the slicing, below 2^{-8S} of the chunk
// maximum.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gain {

inline int slice_chunk(int S) { return S <= 3 ? 256 : 128; }

inline int64_t slice_kpad(int64_t K) {
  return (K + 255) / 256 * 256;  // a multiple of every chunk length
}

inline int64_t slice_nchunk(int S, int64_t K) {
  return slice_kpad(K) / slice_chunk(S);
}

// rows x K FP64 (row stride ld, row r at src + r*ld) -> S bf16 planes
// [S][rows][Kp] (Kp = slice_kpad(K), zero padded) and chunk exponents
// e[rows][Kp / slice_chunk(S)].
void slice_rows_bf16(int S, const double* src, int64_t ld, int rows, int K,
                     uint16_t* planes, int* expo, hipStream_t s);

// mu [3][vI][vF] (v_F fastest) -> [3][vF][vI] FP64 (one-off, for slicing the
// dipole along v_I).
void transpose_mu(const double* mu, int nc, double* muT, hipStream_t s);

// C[z][n*ldc + m] = sum_chunks 2^{eA[m][c] + eB[n][c]} sum_{L<S}
//                   sum_{i+j=L} A_i B_j
// with A planes [S][z][M][Kp] (z stride a_z inside a plane), A exponents
// [z][M][nch], B planes [S][N][Kp], B exponents [N][nch].
struct SlicedGemmArgs {
  int S;
  int M, N;
  int64_t Kp;
  const uint16_t* A;
  int64_t a_plane;  // elements between A slice planes
  int64_t a_z;      // elements between z slices inside a plane
  const int* eA;    // [z][M][nch]
  int64_t ea_z;
  int nch;          // chunks per row = Kp / slice_chunk(S)
  const uint16_t* B;
  int64_t b_plane;
  const int* eB;
  double* C;
  int64_t ldc;
  int64_t c_z;
  int nz;
};
void gemm_sliced_bf16(const SlicedGemmArgs& g, hipStream_t s);

}  // namespace gain

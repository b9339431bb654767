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

// ---------------------------------------------------------------------------
// Split-INT8 (Ozaki-style on the integer matrix cores).
//
// Every FP64 row x is scaled by 2^-e, e the exponent of the row's largest
// magnitude (|x| 2^-e < 1), and cut into S slices
//   x = 2^e sum_{i<S} n_i 2^{-6-7i},   n_i = rint(remainder * 2^{6+7i}),
// so |n_i| <= 64 (an int8) and |x - slices| <= 2^{e-7S}.  Slice products
// n_i m_j are exact int32; the products of one level L = i + j are summed
// by v_mfma_i32_32x32x32_i8 in int32 over a K range, exact while
// S * range * 64 * 64 < 2^31 (i8_max_krange).  Levels L >= S are dropped;
// the S level sums are combined in FP64 and scaled by 2^{e_A + e_B - 12}.
//
// Planes are stored tiled: plane i of an R x K matrix is [R/32][Kp/32] blocks
// of 32 rows x 32 bytes (1 KB), so one MFMA operand fragment is one
// contiguous 1 KB block.  Rows are padded to i8_rows_pad(R) and K to
// i8_kpad(K) with zeros, so the GEMM never bounds-checks its loads.
inline int64_t i8_kpad(int64_t K) { return (K + 63) / 64 * 64; }
inline int64_t i8_rows_pad(int64_t R) { return (R + 31) / 32 * 32 + 256; }
inline int64_t i8_max_krange(int S) {
  return ((1ll << 31) - 1) / (4096ll * S) / 32 * 32;
}

// rows x K FP64 (row r at src + r*ld) -> S tiled int8 planes of
// i8_rows_pad(rows) x i8_kpad(K) (plane stride `plane` bytes) and one
// exponent per padded row.
void slice_rows_i8(int S, const double* src, int64_t ld, int rows, int K,
                   int8_t* planes, int64_t plane, int* expo, hipStream_t s);

// C[z][ks][n*ldc + m] = 2^{eA + eB - 12} sum_{L<S} 2^{-7L} sum_{i+j=L} A_i B_j
// over the k blocks of split ks, for output rows m in [0, M) (A rows
// row0 + m) and columns n in [0, N) (B rows n).
struct I8GemmArgs {
  int S;
  int M, N;
  int row0;          // A row of output row 0
  int64_t Kp;        // i8_kpad(K)
  const int8_t* A;   // tiled planes, Kp bytes per row
  int64_t a_plane;   // bytes between A slice planes
  int64_t a_z;       // bytes between z matrices inside a plane
  const int* eA;     // exponent per A row; z matrices ea_z apart
  int64_t ea_z;
  const int8_t* B;
  int64_t b_plane;
  const int* eB;
  double* C;
  int64_t ldc, c_z, c_ks;
  int nz = 1;
  int ksplit = 1;
  int cfg = 0;
  int beta = 0;      // CRT form: integer bits per operand element
  int8_t* R = nullptr;  // CRT form: residue buffer, gemm_crt_res_bytes(g)
  // CRT form, optional: nonzero k-block range of the A operand.
  // a_kr[z * akr_z + rb] = [x, y) holds every k block (32 columns) of A row
  // block rb (32 rows) of matrix z that has a nonzero element in some plane
  // (x >= y: the row block is zero).  A tile multiplies only the k blocks
  // inside the union of its row blocks' ranges; the products it skips are
  // exact zeros, so the residues are those of the full product.
  const int2* a_kr = nullptr;
  int64_t akr_z = 0;
  int akr_n = 0;     // row blocks per matrix covered by a_kr
};
constexpr int kI8Configs = 12;
// whether configuration cfg supports S slices
bool i8_config_ok(int cfg, int S);
const char* i8_config_name(int cfg);
void gemm_sliced_i8(const I8GemmArgs& g, hipStream_t s);
int64_t gemm_i8_grid_items(const I8GemmArgs& g);

// ---------------------------------------------------------------------------
// Split-INT8, CRT form (Ozaki scheme with integer moduli).
//
// Every FP64 row x is scaled by 2^{beta-1-e} (e the exponent of a bound of
// the row's 2-norm, ||x||_2 < 2^e, or, where that width would pass the
// 50-bit cap, of its maximum) and rounded entry by entry to an integer row
// X, |X_k| <= 2^{beta-1}; X is stored as its residues modulo
// nm pairwise coprime odd moduli m_i <= 255, each in [-(m_i-1)/2,
// (m_i-1)/2] (an int8).  For every modulus the int8 matrix cores form the
// exact int32 product of the residue matrices (exact while
// K * 127^2 < 2^31); it is reduced modulo m_i, and the CRT reconstruction
// of the nm residues gives the exact integer product sum_k X_A X_B, since
// ||X||_2 < 2^{beta-1} + sqrt(K)/2 and beta is chosen so that
// (2^{beta-1} + sqrt(K)/2)^2 < prod m_i / 2 (crt_beta; by Cauchy-Schwarz,
// which bounds the products by the row norms instead of K times the row
// maxima: about log2(K)/2 more bits per operand at the same moduli; with
// the maximum scaling 2 K 4^{beta-1} < prod m_i).  The only error is the
// rounding of the operands to integers, 2^{-beta} of the row norms (or
// maxima) per element; the cost is nm int8 products per FMA instead of
// S(S+1)/2 for the sliced form.
constexpr int kCrtMaxModuli = 14;
struct CrtConsts {
  int nm;
  int m[kCrtMaxModuli];
  float minv[kCrtMaxModuli];   // 1 / m_i
  double minv_d[kCrtMaxModuli];
  // Reconstruction by pairs: the residues modulo m_{2k} and m_{2k+1} are
  // joined into one modulo P_k = m_{2k} m_{2k+1} (P_k = m_{2k} for the last
  // of an odd count), r = r_a + m_a ((r_b - r_a) pinv_k mod m_b) with
  // pinv_k = m_a^{-1} mod m_b, exact in FP32 (|.| < 2^16); then Garner over
  // the np pair moduli in FP64: digit k is
  // v_k = (r_k pgi_k + sum_{j<k} v_j pnw_kj) mod P_k with
  // pgi_k = (P_0 ... P_{k-1})^{-1} mod P_k and
  // pnw_kj = -(P_0 ... P_{j-1}) pgi_k mod P_k, all symmetric residues, so
  // every product is below 2^30 and every sum below 2^33 (exact).
  // Explicit form (at most 12 moduli): X = sum_i r_i C_i - q M with
  // C_i = (M / m_i) ((M / m_i)^{-1} mod m_i) and q = rint(sum_i r_i ef_i),
  // ef_i = fl(C_i / M); C_i and M are split at 2^es into a high part (a
  // multiple of 2^es below 2^(es + 40): every partial sum of the high parts
  // is exact in FP64) and a low part (below 2^es, summed with rounding).
  double ef[kCrtMaxModuli], ech[kCrtMaxModuli], ecl[kCrtMaxModuli];
  double emh, eml;
  static constexpr int kMaxPairs = (kCrtMaxModuli + 1) / 2;
  int np;
  float pinv[kMaxPairs];
  double P[kMaxPairs], Pinv[kMaxPairs], pgi[kMaxPairs];
  double pnw[kMaxPairs][kMaxPairs];
};
const CrtConsts& crt_consts(int nm);
// The reconstruction of the CRT-form kernels run on the host: words
// [n4][nm] (byte q of word i = symmetric residue of output q modulo m_i),
// out [n4][4] = the integer products in FP64 (unscaled).
void crt_reconstruct_host(int nm, const uint32_t* words, int64_t n4,
                          double* out);
int crt_modulus(int nm, int i);
// Operand width of a product of length K: the largest beta <= 50 with
// (2^{beta-1} + sqrt(K)/2)^2 < prod_{i<nm} m_i (1/2 - 3.5e-7), rows
// scaled by their 2-norm (crt_norm_scaled true), or where that passes 50,
// the largest with 2 K 4^{beta-1} < prod m_i 2^{-1e-6}, rows scaled by
// their maximum.
int crt_beta(int nm, int64_t K);
bool crt_norm_scaled(int nm, int64_t K);
struct ScatterArgs;
// The half line strength's X (the scatter of `sa`: sa.nvec vectors x ncol
// columns = rows rows of nc) cut straight into nm residue planes and row
// exponents, as launch_scatter_expand followed by slice_rows_crt, without
// the FP64 X; desc: sa.nslot entries of scratch.  Fits when nc <= 4096 and
// the irrep's repres has fewer than 2^31 entries.  `single`: every slot
// holds at most one primitive and every case one term (the branch-free
// form).
bool scatter_slice_crt_fits(int nm, int nc, int64_t repres_len);
void scatter_slice_rows_crt(int nm, int beta, const ScatterArgs& sa, int ncol,
                            int nc, int rows, int8_t* planes, int64_t plane,
                            int* expo, int4* desc, bool single,
                            hipStream_t s);
// rows x K FP64 -> nm tiled int8 residue planes (layout and padding as
// slice_rows_i8) and one exponent per padded row (||row||_2 < 2^e or
// max|row| < 2^e, as crt_norm_scaled(nm, Kp) says); beta must be
// crt_beta(nm, Kp), Kp = K padded to 64.
void slice_rows_crt(int nm, int beta, const double* src, int64_t ld,
                    int rows, int K, int8_t* planes, int64_t plane,
                    int* expo, hipStream_t s);
constexpr int kCrtConfigs = 14;
// Tile configuration used when nothing is tuned or forced: 256 x 256
// workgroup tiles through the 4-stage LDS pipeline, the autotuner's pick for
// both CRT shapes of the J=20 -> 21 workload.
constexpr int kCrtDefaultConfig = 5;
const char* crt_config_name(int cfg);
// C[z][n*ldc + m] = 2^{eA + eB - 2(beta-1)} * (exact integer product of the
// operands), g.S = number of moduli, g.beta as sliced, no split-K: one
// int8 GEMM per (operand matrix, modulus) writing int8 residues to g.R,
// then the reconstruction kernel.
void gemm_crt_i8(const I8GemmArgs& g, hipStream_t s);
// The two launches of gemm_crt_i8: the per-modulus int8 products with their
// residue epilogue, then the reconstruction into FP64 (timed apart).
void gemm_crt_res(const I8GemmArgs& g, hipStream_t s);
void gemm_crt_combine(const I8GemmArgs& g, hipStream_t s);
int64_t gemm_crt_res_bytes(const I8GemmArgs& g);
// Half line strength of K-block bases straight from the residues of the
// CRT-form product Y_alpha = mu_alpha X (g as given to gemm_crt_res, with
// nz = 3, row0 = 0, residues in g.R): the reconstruction of
// gemm_crt_combine fused with the gather of launch_gather_h (same
// arithmetic in the same order, so H is identical), Y never stored.
// Final primitives are reached through the final basis's (column, v) slot
// CSR (slot = (2 K + tau) * nc + v).
struct CrtGatherArgs {
  I8GemmArgs g;
  int nc, nvec, ncolI, jI, jF, jmax;
  const double* threej;
  const int* slot_ptr;
  const int* slot_prim;
  const int* slot_one;  // the slot's primitive if it has exactly one, -1
                        // if none, -2 if several (listed by the CSR)
  double* H;
  int64_t hstride;
};
void launch_crt_gather_h(const CrtGatherArgs& a, hipStream_t s);
// Nonzero k-block range of every row block of nmat tiled int8 matrices
// (nrb row blocks of nkb k blocks each, matrices mat_stride bytes apart
// inside a plane, nplanes planes `plane` bytes apart): out[m * nrb + rb] =
// [first, last + 1) over the k blocks with a nonzero byte in any plane,
// (0, 0) for a zero row block.
void i8_block_kranges(const int8_t* planes, int64_t plane, int nplanes,
                      int64_t mat_stride, int nmat, int nrb, int nkb,
                      int2* out, hipStream_t s);
// k blocks the product of a CRT-form GEMM with configuration cfg multiplies
// for the ranges kr (host copy, nz matrices of nrb row blocks), as a
// fraction of the dense count; 1 without ranges.
double crt_kblock_fraction(const I8GemmArgs& g, const int2* kr_host);
// Work-items of the per-modulus GEMM launch.
int64_t gemm_crt_grid_items(const I8GemmArgs& g);

}  // namespace gain

// Device kernels of the plain FP64 path and their host launchers.
//
// Data layout in HBM (one context, one GPU):
//   mu      [alpha][vI][vF]            3*nc*nc doubles (reference piece layout)
//   X       [vec][col][vI]             scattered initial vectors, col = 2K+tau
//   Ypart   [ks][alpha][vec*ncol][vF]  mu_alpha * X, split-K partials
//   H       [vec][r']                  half line strengths (primitive basis F)
//   G       [vec][t]                   P^T h in the symmetrised basis of F
//   C       [F][ld]                    final eigenvector records (rows)
//   partial [ks][F-F0][vec]            GEMV split-K partials
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace gain {

// Truncated constant of the reference (defines.h:23, dipole_GPU.cu:24).
constexpr double kSqrt2Ref = 0.70710678118;
constexpr double kACoef = 3.13618923e-7;   // defines.h:22
constexpr double kCMCoef = 1.327209365e-12;  // defines.h:24
constexpr double kPlanck = 6.62606896e-27;   // defines.h:19
constexpr double kVellgt = 2.99792458e+10;   // defines.h:20
constexpr double kBoltz = 1.380658e-16;      // defines.h:21

struct ScatterArgs {
  int nslot;  // ncol * nc
  int nprim;
  const int* slot_ptr;   // [nslot+1]
  const int* slot_prim;  // primitives grouped by slot
  const int* kval;
  const int* tau;
  const int* icontr;     // [2*nprim]
  const int* ijt;        // ijterms + gamma*maxsymcoeffs
  const int* N;          // N of this irrep
  const double* repres;  // repres of this irrep
  int sdeg, ntot;
  const double* cbase;   // initial records
  const int64_t* c_off;  // [nvec] offset of each vector's record in cbase
  const int* ideg;       // [nvec] component of each vector
  int nvec;
  double* X;
  int64_t xstride;       // ncol * nc
};
constexpr int kScatterNV = 4;  // vectors per scatter thread
void launch_scatter_expand(const ScatterArgs& a, hipStream_t s);

// Ypart[ks][alpha][j][vF] = sum_{vI in chunk ks} mu[alpha][vI][vF] X[j][vI]
void launch_mu_gemm_valu(const double* mu, int nc, const double* X,
                         int ncolTot, double* Ypart, int ksplit,
                         hipStream_t s);

struct GatherArgs {
  int nprimF, nc, jI, jF, jmax, ncolI, nvec, ksplit;
  const int* kvalF;
  const int* tauF;
  const int* vibF;
  const double* threej;
  const double* Y;
  int64_t ys_alpha;  // ncolTot * nc
  int64_t ys_ks;     // 3 * ncolTot * nc
  double* H;
  int64_t hstride;
};
void launch_gather_h(const GatherArgs& a, hipStream_t s);

// Gather of the symmetric-top (ROTSYM) half line strength from Y_n = mu_n X
// (dipole_GPU.cu:641-723 regrouped):
//   h[r'] = sum_n sum_{(l_I, d_I)} W[n, l_I, l_F(r'), d_I, d_F(r')]
//           Y_n[(l_I, d_I)][v(r')]
// formed once per final slot (l_F, d_F, v_F) and written to every primitive
// of the slot (final-basis slot CSR).  X columns are l * md + d.
struct GatherRotArgs {
  int nc, ncolI, ncolF, nvec, ksplit;
  int nlI, nlF, md;
  const double* W;       // [3][nlI][nlF][md][md], first index fastest
  const double* Y;
  int64_t ys_alpha;      // ncolTot * nc
  int64_t ys_ks;         // 3 * ncolTot * nc
  const int* slot_ptrF;  // [ncolF * nc + 1]
  const int* slot_primF;
  double* H;
  int64_t hstride;
};
void launch_gather_rotsym(const GatherRotArgs& a, hipStream_t s);

struct ExpandTArgs {
  int ntot, nprim, sdeg;
  const int* tcase;      // [ntot] case of each symmetrised function
  const int* case_ptr;   // [ncase+1]
  const int* case_prim;  // primitives of each case
  const int* icontr;     // lambda = icontr[r + nprim]
  const double* repres;
  const double* H;
  int64_t hstride;
  const int* h_of_vec;   // [nvec] which H vector feeds output vector v
  const int* ideg_of_vec;  // [nvec] final-side component
  int nvec;
  double* G;
  int64_t gstride;
};
void launch_expand_T(const ExpandTArgs& a, hipStream_t s);

struct GemvArgs {
  const double* C;
  int64_t ld;
  int ntot;
  int F0, nF;
  const double* G;
  int64_t gstride;
  int nvec;
  double* partial;  // [ksplit][nF][nvec]
  int ksplit;
  int rows;         // rows per wave: 2, 4 or 8
};
void launch_gemv(const GemvArgs& a, hipStream_t s);
int gemv_kchunk(int ntot, int ksplit);
// Work-items of the launch launch_gemv makes for a.
int64_t gemv_grid_items(const GemvArgs& a);

struct LinesArgs {
  int nI, nF, Flo, Fhi, maxdeg;
  int nvec_per_i;             // vectors per initial state
  const int* slot_of_vec;     // [nvec_per_i] -> dI + dF*maxdeg
  const double* src;          // projected line strengths
  int ksplit;
  int64_t src_ks, src_f, src_v;  // strides (F index is F - Flo)
  const double* eI;           // [nI]
  const int* levelI;          // [nI]
  const uint8_t* okI;         // [nI] lower-state filter
  const double* eF;           // [nF]
  int same_block;
  int block_ok;
  int jI, jF, ndegI, ndegF, reduced;
  double gnsF, zpe, threshold, beta, partition;
  double fw0, fw1, eu0, eu1;
  int isym_do_F;
  // outputs (device, may be null)
  double* ls;
  double* S;
  double* A;
  double* I;
  uint8_t* flags;
};
void launch_lines(const LinesArgs& a, hipStream_t s);

// C[n*ldc + m] = sum_k A[m*lda + k] * B[n*ldb + k]  (FP64, vector ALUs)
void launch_gemm_tn_valu(int M, int N, int K, const double* A, int64_t lda,
                         const double* B, int64_t ldb, double* C,
                         int64_t ldc, hipStream_t s);

}  // namespace gain

namespace gain {

// ---- adjoint (backward of the batched path) ------------------------------
// dLS[v][F-Flo] = dS[i][F] * dS/dls for every pair that passes the filters
// and whose S is formed (nu >= 1e-40; below it S is 0 whatever ls is)
// (S = sum ls^2 / ndegI, or sum ls^2 in the reduced degenerate case).
struct DlsArgs {
  int nI, nF, Flo, nR, nvpi;
  const double* dS;       // [nI][nF]
  const uint8_t* flags;   // [nI][nF]
  const double* eI;       // [nI] initial energies
  const double* eF;       // [nF] final energies
  const double* ls;       // [nvec][nR] forward projection
  double scale;           // 2 / ndegI (or 2)
  double* dls;            // [nvec][nR]
  // optional: the forward's per-pair line strengths (gain_lines.ls,
  // [nI][nF][md2]) read in place of `ls`, vector v of initial i at slot
  // slot_of_vec[v - i * nvpi]
  const double* ls_pair = nullptr;
  const int* slot_of_vec = nullptr;
  int md2 = 0;
};
void launch_dls(const DlsArgs& a, hipStream_t s);

// dH[hv][r] = sum_{v in list(hv)} (P_{dF(v)} dG_v)[r]   (expansion, batched)
struct ExpandAccArgs {
  int nprim, ntot, sdeg;
  const int* icontr;
  const int* ijt;         // ijterms + gamma * maxsymcoeffs
  const int* N;
  const double* repres;
  const double* dG;
  int64_t gstride;
  const int* vptr;        // [nH+1] CSR of projection vectors per H vector
  const int* vidx;        // vector index
  const int* vdeg;        // final-side component of each listed vector
  int nH;
  double* dH;
  int64_t hstride;
};
void launch_expand_acc(const ExpandAccArgs& a, hipStream_t s);

// Transpose of the gather: dY[alpha][b*ncolI + 2K_I+tau_I][v] from dH over
// the final primitives with vibrational index v (final-basis slot CSR).
struct GatherTArgs {
  int nc, jI, jF, jmax, ncolI, ncolF, nvec;
  const int* slot_ptrF;   // [(2(J_F+1)) * nc + 1]
  const int* slot_primF;
  const double* threej;
  const double* dH;
  int64_t hstride;
  double* dY;             // [3][nvec*ncolI][nc]
};
void launch_gather_T(const GatherTArgs& a, hipStream_t s);

// out[c][r] = in[r][c] for an rows x cols FP64 matrix
void launch_transpose(const double* in, int rows, int cols, double* out,
                      hipStream_t s);
// y[i] (+)= x[i]
void launch_axpy1(const double* x, double* y, int64_t n, int accumulate,
                  hipStream_t s);
// mu[i] = sum_p c[p] M[p][i]
void launch_combine(int P, int64_t n, const double* c, const double* M,
                    double* mu, hipStream_t s);


// Grouped launches for mixed-J sweeps: one kernel runs a table of argument
// blocks (device memory); entry keep[t] of the host vector owns workgroups
// [start[t], start[t+1]).  The table builders skip empty entries.
struct GroupedTable {
  std::vector<int> start;  // size = keep.size() + 1; back() = workgroups
  std::vector<int> keep;   // indices into the argument vector
};
enum {
  kGroupedScatter = 0,
  kGroupedGather = 1,
  kGroupedExpandT = 2,
  kGroupedLines = 3,
  kGroupedGemv4 = 4,  // GEMV, 4 rows per wave, up to 4 vectors
  kGroupedGemv8 = 5   // GEMV, 8 rows per wave, up to 9 vectors
};
GroupedTable grouped_scatter_table(const std::vector<ScatterArgs>& v);
GroupedTable grouped_gather_table(const std::vector<GatherArgs>& v);
GroupedTable grouped_expand_T_table(const std::vector<ExpandTArgs>& v);
GroupedTable grouped_lines_table(const std::vector<LinesArgs>& v);
// R = rows per wave of the GEMV kind the table is for (4 or 8); requires
// 16-byte aligned records and vectors with even strides.
GroupedTable grouped_gemv_table(const std::vector<GemvArgs>& v, int R);
void launch_grouped(int kind, const void* d_args, const int* d_start, int n,
                    int nblocks, hipStream_t s);

}  // namespace gain

/*
 * CPU oracle for the dipole line-strength path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain C restatement of GAIN-MPI's half-linestrength / projection /
 * S->A->intensity path, used exclusively as the checker by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  Nothing in
 * the product library links or calls it.
 *
 * Every function cites the reference file:line it restates.  Sums use
 * double-double (compensated) accumulation so the oracle is more accurate
 * than either the reference or the HIP engine.
 *
 * Pinning (see tests/test_oracle.py):
 *   orc_three_j            exact sympy values (tests/golden/threej_*.json),
 *                          orthogonality sums, j2=1 closed forms
 *   orc_rot_coeff          Fortran textbook form (trove_wrappers.f90:329-381)
 *                          restated independently in the test
 *   orc_expand             identity / Hadamard isometry, dense brute force
 *   orc_half_ls*           dense-operator brute force (numpy, textbook form)
 *   orc_dot                math.fsum (exactly rounded sum)
 *   filters / S,A,I        hand-enumerated cases and literal closed forms
 */
#ifndef GAIN_ORACLE_H
#define GAIN_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Reference constants, defines.h:17-24 (SQRT2 is the truncated literal). */
#define ORC_PLANCK 6.62606896e-27
#define ORC_VELLGT 2.99792458e+10
#define ORC_BOLTZ 1.380658e-16
#define ORC_ACOEF 3.13618923e-7
#define ORC_SQRT2 0.70710678118
#define ORC_CMCOEF 1.327209365e-12

double orc_three_j(int j1, int j2, int j3, int m1, int m2, int m3);
int orc_threej_index(int jmax, int jI, int kI, int jF, int kF);
void orc_threej_table(int jmax, double* tab);
double orc_rot_coeff(int jmax, const double* tab, int jI, int jF, int kI,
                     int tauI, int kF, int tauF, int* alpha);

void orc_expand(int dimen, int igamma, int maxcoeff, int ideg, int sdeg,
                int ntot, const int* ijterms, const int* icontr, const int* N,
                const double* repres, const double* vin, double* vout);

void orc_half_ls_kblock(int dimenF, int dimenI, int jI, int jF, int kF,
                        const int* tauI, const int* tauF, const int* vibI,
                        const int* vibF, int startF_idx, int startI_idx,
                        int startFblock, int endFblock, int ncontrF,
                        const int* kblockI, const double* dipole,
                        int dip_stride, const double* x, const double* threej,
                        int jmax, double* half_ls);

void orc_half_ls(int jI, int dimenI, const int* kstartI, const int* kblockI,
                 const int* tauI, const int* vibI, int jF, int dimenF,
                 const int* kstartF, const int* kblockF, const int* tauF,
                 const int* vibF, const double* dipole, int nc, int startF,
                 int endF, const double* threej, int jmax, const double* x,
                 double* h);

void orc_half_ls_rotsym(int dimenF, int dimenI, int nlevelI, int nlevelF,
                        int maxdeg, const int* degI, const int* degF,
                        const int* levelI, const int* levelF,
                        const int* vibI, const int* vibF, int startFblock,
                        int endFblock, int ncontrF, const double* dipole,
                        int dip_stride, const double* wigner,
                        const double* x, double* half_ls);

double orc_dot(int64_t n, const double* x, const double* y);
/* 1: plain FP64 accumulations instead of double-double (CPU timing only). */
void orc_set_plain_fp64(int plain);

/* Intensity / filter parameters (Input.h getters used by States/Output). */
typedef struct {
  int nrepres;
  int maxdeg;
  int maxJ;
  int reduced;
  int symmetry_type;
  int sym_degen[16];
  int isym_do[16];
  int isym_pairs[16];
  int igamma_pair[16];
  double gns[16];
  double erange_lower[2];
  double erange_upper[2];
  double freq_window[2];
  double zpe;
  double temperature;
  double partition;
  double threshold;
} orc_params;

int orc_gamma_pairs(orc_params* p);
int orc_filter_lower(const orc_params* p, double energy, int igamma);
int orc_filter_upper(const orc_params* p, double energy, int igamma);
int orc_filter_intensity(const orc_params* p, int jI, int gI, double eI,
                         int jF, int gF, double eF);
int orc_degeneracy_filter(const orc_params* p, int gI, int gF, int dI, int dF);

/* S->A->I for one pair.  Returns 1 if the line is emitted, 0 if dropped
 * (nu < 1e-40 or A < threshold).  S is the degeneracy-averaged strength. */
int orc_line(const orc_params* p, int jI, int gI, int ndegI, double eI,
             int jF, int gF, int ndegF, double eF, const double* ls,
             double* S, double* A, double* I);
int orc_format_line(const orc_params* p, char* buf, int buflen, int jI,
                    int gI, int ndegI, int ilevelI, double eI, int jF, int gF,
                    int ndegF, int ilevelF, double eF, const double* ls,
                    int with_intens, int with_ls);

/* Per-J basis + inflation description (TroveBasisSet / GpuInflation). */
typedef struct {
  int J;
  int nprim;
  const int* kstart;     /* J+1 */
  const int* kblock;     /* J+1 */
  const int* tau;        /* nprim, ktau & 1 */
  const int* vib;        /* nprim */
  int maxsymcoeffs;
  const int* ijterms;    /* maxsymcoeffs * nrepres */
  const int* icontr;     /* 2 * nprim */
  int matsize;
  const int* N[16];      /* per gamma: maxsymcoeffs */
  const double* repres[16];
  int ntot[16];
  int sdeg[16];
  /* symmetric-top (ROTSYM) basis: level and degenerate component per
   * primitive instead of K blocks, and the rotational tables to J - 1, J,
   * J + 1 (index J_F - J + 1; NULL when absent), each
   * 3 x nlevels x nlevels(J_F) x rmd x rmd, first index fastest */
  int rotsym;
  int nlevels;
  int rmd;
  const int* level;
  const int* deg;
  const double* wig[3];
} orc_basis;

typedef struct {
  int n;                  /* number of states */
  const int* jind;
  const int* jval;
  const int* igamma;
  const int* ndeg;
  const int* ilevel;
  const int* rec_len;
  const double* energy;
  const double* vectors;  /* concatenated eigenvector records */
  const int64_t* offset;  /* start of each state's record in vectors */
} orc_states;

typedef struct {
  int iI, iF;
  int emitted;
  double S, A, I;
  double ls[9];
} orc_transition;

/* The main.cpp:238-338 loop contract over a list of initial states.
 * Returns the number of transitions written (<= cap), or -1 on error. */
int64_t orc_run(const orc_params* p, const orc_basis* bases, int nJ,
                const orc_states* st, const double* dipole, int nc,
                int jmax, const int* initial, int ninitial,
                orc_transition* out, int64_t cap, int nthreads);

#ifdef __cplusplus
}
#endif
#endif

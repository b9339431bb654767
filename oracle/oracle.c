/*
 * CPU oracle for the dipole line-strength path — TEST INFRASTRUCTURE ONLY.
 * See oracle.h for scope, citations and how each function is pinned.
 * Build with -ffp-contract=off: the compensated sums below depend on
 * exactly rounded adds and an explicit fma() for the product error.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- double-double accumulator ------------------------------------- */
typedef struct {
  double hi, lo;
} dd_t;

static inline void dd_add(dd_t* a, double b) {
  double s = a->hi + b;
  double bb = s - a->hi;
  double err = (a->hi - (s - bb)) + (b - bb);
  a->hi = s;
  a->lo += err;
}

/* Arithmetic of the accumulations: 0 = double-double (the checker), 1 =
 * plain FP64 multiply-adds, the arithmetic of the reference's own kernels
 * and of cuBLAS (used only to time a CPU baseline). */
static int g_plain_fp64 = 0;

void orc_set_plain_fp64(int plain) { g_plain_fp64 = plain != 0; }

static inline void dd_add_prod(dd_t* a, double x, double y) {
  if (g_plain_fp64) {
    a->hi += x * y;
    return;
  }
  double p = x * y;
  double e = fma(x, y, -p);
  dd_add(a, p);
  a->lo += e;
}

static inline double dd_val(dd_t a) { return a.hi + a.lo; }

static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

/* ---- Wigner 3j: Racah formula (fort_func.f90:7-104) ----------------- */
static long double lfact(int n) {
  long double f = 1.0L;
  for (int i = 2; i <= n; i++) f *= (long double)i;
  return f;
}

double orc_three_j(int j1, int j2, int j3, int m1, int m2, int m3) {
  /* selection rules, fort_func.f90:29-33 */
  if (j3 > j1 + j2) return 0.0;
  if (j3 < abs(j1 - j2)) return 0.0;
  if (j1 < 0 || j2 < 0 || j3 < 0) return 0.0;
  if (j1 < abs(m1) || j2 < abs(m2) || j3 < abs(m3)) return 0.0;
  if (m1 + m2 + m3 != 0) return 0.0;
  long double delta = lfact(j1 + j2 - j3) * lfact(j1 - j2 + j3) *
                      lfact(-j1 + j2 + j3) / lfact(j1 + j2 + j3 + 1);
  long double pre = sqrtl(delta * lfact(j1 + m1) * lfact(j1 - m1) *
                          lfact(j2 + m2) * lfact(j2 - m2) * lfact(j3 + m3) *
                          lfact(j3 - m3));
  int kmin = imax(0, imax(j2 - j3 - m1, j1 - j3 + m2));
  int kmax = imin(j1 + j2 - j3, imin(j1 - m1, j2 + m2));
  long double s = 0.0L;
  for (int k = kmin; k <= kmax; k++) {
    long double den = lfact(k) * lfact(j3 - j2 + k + m1) *
                      lfact(j3 - j1 + k - m2) * lfact(j1 + j2 - j3 - k) *
                      lfact(j1 - k - m1) * lfact(j2 - k + m2);
    long double t = 1.0L / den;
    s += (k & 1) ? -t : t;
  }
  int ph = j1 - j2 - m3; /* fort_func.f90:94-97 */
  long double sign = (((ph % 2) + 2) % 2 == 0) ? 1.0L : -1.0L;
  return (double)(sign * pre * s);
}

/* 3j table index, GpuManager.cpp:180 and dipole_GPU.cu:445 */
int orc_threej_index(int jmax, int jI, int kI, int jF, int kF) {
  int n = jmax + 1;
  return jI + kI * n + (jF - jI + 1) * n * n + (kF - kI + 1) * n * n * 3;
}

/* GpuManager.cpp:164-184 */
void orc_threej_table(int jmax, double* tab) {
  int n = jmax + 1;
  memset(tab, 0, sizeof(double) * (size_t)n * n * 9);
  for (int jI = 0; jI <= jmax; jI++)
    for (int jF = imax(jI - 1, 0); jF <= imin(jI + 1, jmax); jF++)
      for (int kI = 0; kI <= jI; kI++)
        for (int kF = imax(kI - 1, 0); kF <= imin(kI + 1, jF); kF++)
          tab[orc_threej_index(jmax, jI, kI, jF, kF)] =
              orc_three_j(jI, 1, jF, kI, kF - kI, -kF);
}

/* Rotational factor of one (K_I,tau_I)->(K_F,tau_F) element, following the
 * branch-free arithmetic of dipole_GPU.cu:430-495 (sign of the initial side
 * :437-438, sqrt factor :439,464-466, component :490, multiplier :492, final
 * sign :511).  Returns the scalar multiplying mu_alpha[vF, vI] * x[r]. */
double orc_rot_coeff(int jmax, const double* tab, int jI, int jF, int kI,
                     int tauI, int kF, int tauF, int* alpha) {
  int dk = kF - kI;
  if (abs(dk) > 1) {
    *alpha = -1;
    return 0.0;
  }
  double f3j = tab[orc_threej_index(jmax, jI, kI, jF, kF)];
  double s = (dk != 0 && kI * kF != 0) ? ORC_SQRT2 : 1.0;
  int a, m;
  if (dk == 0) {
    a = 2;
    m = tauF - tauI;
  } else if (tauF != tauI) {
    a = 0;
    m = (tauF - tauI) * dk;
  } else {
    a = 1;
    m = -1;
  }
  int sI = ((((kI % 3) * tauI) + kI) & 1) ? -1 : 1;
  int sF = (((kF % 3) * tauF) & 1) ? -1 : 1;
  *alpha = a;
  return f3j * s * (double)(m * sI * sF);
}

/* Symmetrised -> primitive expansion, dipole_GPU.cu:859-889 */
void orc_expand(int dimen, int igamma, int maxcoeff, int ideg, int sdeg,
                int ntot, const int* ijterms, const int* icontr, const int* N,
                const double* repres, const double* vin, double* vout) {
  for (int r = 0; r < dimen; r++) {
    int irow = icontr[r];
    int ib = icontr[r + dimen];
    int iterm = ijterms[irow + igamma * maxcoeff];
    int nelem = N[irow];
    dd_t acc = {0.0, 0.0};
    for (int i = 0; i < nelem; i++) {
      int t = iterm + i;
      dd_add_prod(&acc, vin[t],
                  repres[(size_t)t + (size_t)ideg * ntot +
                         (size_t)ib * sdeg * ntot]);
    }
    vout[r] = dd_val(acc);
  }
}

/* One K_F block of the half line strength, restating the live kernel
 * device_compute_1st_half_ls_flipped_dipole_shared_block_nontrove
 * (dipole_GPU.cu:392-515).  Entries outside [startFblock,endFblock) are left
 * untouched (Q8: the kernel assigns only valid entries). */
void orc_half_ls_kblock(int dimenF, int dimenI, int jI, int jF, int kF,
                        const int* tauI, const int* tauF, const int* vibI,
                        const int* vibF, int startF_idx, int startI_idx,
                        int startFblock, int endFblock, int ncontrF,
                        const int* kblockI, const double* dipole,
                        int dip_stride, const double* x, const double* threej,
                        int jmax, double* half_ls) {
  size_t stride = (size_t)ncontrF * (size_t)dip_stride; /* :407 */
#pragma omp parallel for schedule(dynamic, 64)
  for (int t = 0; t < dimenF; t++) {
    int irootF = startF_idx + t;
    int icontrF = vibF[irootF];
    int tF = tauF[irootF] & 1;
    if (!(icontrF >= startFblock && icontrF < endFblock)) continue; /* :427 */
    icontrF -= startFblock;
    int sigmaF = (kF % 3) * tF;
    dd_t acc = {0.0, 0.0};
    int sI = startI_idx;
    for (int kI = imax(kF - 1, 0); kI < kF + 2 && kI < jI + 1; kI++) {
      int kbs = kblockI[kI];
      double sq = (kI * kF != 0) ? ORC_SQRT2 : 1.0;
      double f3j = threej[orc_threej_index(jmax, jI, kI, jF, kF)];
      for (int j = 0; j < kbs; j++) {
        int irootI = sI + j;
        int tI = tauI[irootI];
        double ls = x[irootI];
        if (kF != kI) ls *= sq;
        ls *= f3j;
        if (ls == 0.0) continue; /* :477 */
        int sigmaI = (kI % 3) * tI;
        double signI = (((sigmaI + kI) & 1) ? -1.0 : 1.0);
        int eq = (kI == kF);
        int tneq = (tF != tI);
        int dip_idx = eq ? 2 : (tneq ? 0 : 1);
        int m = eq ? (tF - tI) : (tneq ? (tF - tI) * (kF - kI) : -1);
        double a = ls * (double)m * signI;
        double mu = dipole[(size_t)icontrF +
                           (size_t)vibI[irootI] * (size_t)ncontrF +
                           (size_t)dip_idx * stride];
        dd_add_prod(&acc, a, mu);
      }
      sI += kbs; /* :502-503 */
      if (sI >= dimenI) break;
    }
    double sF = (sigmaF & 1) ? -1.0 : 1.0;
    half_ls[irootF] = sF * dd_val(acc);
  }
}

/* Whole half line strength for one (J_I -> J_F): zero the vector
 * (GpuManager.cpp:761-765), then one K_F block at a time
 * (GpuManager::ExecuteKBlockHalfLs, GpuManager.cpp:645-732). */
void orc_half_ls(int jI, int dimenI, const int* kstartI, const int* kblockI,
                 const int* tauI, const int* vibI, int jF, int dimenF,
                 const int* kstartF, const int* kblockF, const int* tauF,
                 const int* vibF, const double* dipole, int nc, int startF,
                 int endF, const double* threej, int jmax, const double* x,
                 double* h) {
  (void)kstartI;
  memset(h, 0, sizeof(double) * (size_t)dimenF);
  for (int kF = 0; kF <= jF; kF++) {
    orc_half_ls_kblock(kblockF[kF], dimenI, jI, jF, kF, tauI, tauF, vibI,
                       vibF, kstartF[kF], kstartI[imax(kF - 1, 0)], startF,
                       endF, endF - startF, kblockI, dipole, nc, x, threej,
                       jmax, h);
  }
}

/* Half line strength in the ROTSYM (Euler) basis, restating
 * device_compute_1st_half_ls_flipped_dipole_shared_rotsym_nontrove
 * (dipole_GPU.cu:641-723): for every final primitive whose vibrational
 * index lies in [startFblock, endFblock),
 *   h[rF] = sum_rI x[rI] sum_n W[n, lvI(rI), lvF(rF), dgI(rI), dgF(rF)]
 *           mu_n[vF - startFblock + vI ncontrF]
 * with W indexed n + 3 lvI + 3 nlevelI lvF + 3 nlevelI nlevelF dgI
 * + 3 nlevelI nlevelF maxdeg dgF (:693) and the dipole components
 * ncontrF * dip_stride apart (:661).  Other entries are left untouched. */
void orc_half_ls_rotsym(int dimenF, int dimenI, int nlevelI, int nlevelF,
                        int maxdeg, const int* degI, const int* degF,
                        const int* levelI, const int* levelF,
                        const int* vibI, const int* vibF, int startFblock,
                        int endFblock, int ncontrF, const double* dipole,
                        int dip_stride, const double* wigner,
                        const double* x, double* half_ls) {
  const size_t comp = (size_t)ncontrF * (size_t)dip_stride;
  const size_t sF = (size_t)3 * nlevelI, sdI = sF * (size_t)nlevelF,
               sdF = sdI * (size_t)maxdeg;
#pragma omp parallel for schedule(dynamic, 64)
  for (int rF = 0; rF < dimenF; rF++) {
    int vF = vibF[rF];
    if (!(vF >= startFblock && vF < endFblock)) continue; /* :673 */
    vF -= startFblock;
    dd_t acc = {0.0, 0.0};
    for (int rI = 0; rI < dimenI; rI++) {
      const size_t w = (size_t)3 * levelI[rI] + sF * levelF[rF] +
                       sdI * degI[rI] + sdF * degF[rF];
      const size_t d = (size_t)vF + (size_t)vibI[rI] * (size_t)ncontrF;
      dd_t t = {0.0, 0.0};
      for (int n = 0; n < 3; n++)
        dd_add_prod(&t, wigner[w + n], dipole[d + n * comp]);
      dd_add_prod(&acc, x[rI], dd_val(t));
    }
    half_ls[rF] = dd_val(acc);
  }
}

/* cublasDdot replacement (GpuManager.cpp:741), compensated. */
double orc_dot(int64_t n, const double* x, const double* y) {
  if (g_plain_fp64) { /* eight independent FP64 partial sums */
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t i = 0;
    for (; i + 8 <= n; i += 8)
      for (int q = 0; q < 8; q++) s[q] += x[i + q] * y[i + q];
    for (; i < n; i++) s[0] += x[i] * y[i];
    return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
  }
  dd_t acc = {0.0, 0.0};
  for (int64_t i = 0; i < n; i++) dd_add_prod(&acc, x[i], y[i]);
  return dd_val(acc);
}

/* ---- filters: States.cpp:41-116, TroveInput.cpp:202-236 ------------- */
int orc_gamma_pairs(orc_params* p) {
  for (int gI = 0; gI < p->nrepres; gI++) {
    int ng = 0;
    p->igamma_pair[gI] = gI;
    for (int gF = 0; gF < p->nrepres; gF++) {
      if (gI != gF && p->isym_pairs[gI] == p->isym_pairs[gF]) {
        p->igamma_pair[gI] = gF;
        if (++ng > 1) return -1; /* TroveInput.cpp:218-222 */
      }
    }
    if (p->gns[gI] != p->gns[p->igamma_pair[gI]]) return -2; /* :226-229 */
  }
  return 0;
}

int orc_filter_lower(const orc_params* p, double energy, int igamma) {
  return (energy - p->zpe) <= p->erange_lower[1] &&
         (energy - p->zpe) >= p->erange_lower[0] && p->isym_do[igamma];
}

int orc_filter_upper(const orc_params* p, double energy, int igamma) {
  return (energy - p->zpe) <= p->erange_upper[1] &&
         (energy - p->zpe) >= p->erange_upper[0] && p->isym_do[igamma];
}

int orc_filter_intensity(const orc_params* p, int jI, int gI, double eI,
                         int jF, int gF, double eF) {
  double nu = eF - eI;
  return p->gns[gI] > 0.0 && p->gns[gF] > 0.0 && nu >= p->freq_window[0] &&
         nu <= p->freq_window[1] && orc_filter_lower(p, eI, gI) &&
         orc_filter_upper(p, eF, gF) && (jF != p->maxJ || jI != p->maxJ) &&
         p->isym_pairs[gI] == p->isym_pairs[gF] && p->igamma_pair[gI] == gF &&
         abs(jF - jI) <= 1 && jI + jF >= 1;
}

int orc_degeneracy_filter(const orc_params* p, int gI, int gF, int dI,
                          int dF) {
  if (p->sym_degen[gI] == 1 || p->sym_degen[gF] == 1) return 1;
  if (!p->reduced) return 1;
  if (p->symmetry_type == 1) return 1;
  if (p->symmetry_type == 2) return (dI == 1 && dF == 0);
  if (p->symmetry_type == 3) {
    if (gI == 3 && gF == 4 && dI == 1 && dF == 0) return 1;
    if (gI == 4 && gF == 3 && dI == 2 && dF == 0) return 1;
    if (gI == 2 && gF == 2 && dI == 1 && dF == 0) return 1;
  }
  return 0;
}

/* ---- S -> A -> intensity: Output.cpp:27,52-112 ---------------------- */
int orc_line(const orc_params* p, int jI, int gI, int ndegI, double eI,
             int jF, int gF, int ndegF, double eF, const double* ls,
             double* S, double* A, double* I) {
  double nu = eF - eI;
  double energyI = eI - p->zpe;
  *S = *A = *I = 0.0;
  if (nu < 1e-40) return 0;
  double s = 0.0;
  for (int dF = 0; dF < ndegF; dF++)
    for (int dI = 0; dI < ndegI; dI++) {
      if (!orc_degeneracy_filter(p, gI, gF, dI, dF)) continue;
      double v = ls[dI + dF * p->maxdeg];
      s += v * v;
    }
  s /= (double)ndegI;
  if (p->reduced && ndegF != 1 && ndegI != 1) s *= (double)ndegI;
  double a = ORC_ACOEF * (double)(2 * jI + 1) * s * fabs(nu) * fabs(nu) *
             fabs(nu);
  *S = s;
  *A = a;
  if (a < p->threshold) return 0;
  double beta = ORC_PLANCK * ORC_VELLGT / (ORC_BOLTZ * p->temperature);
  *I = ORC_CMCOEF * a * p->gns[gF] * (2.0 * (double)jF + 1.0) *
       exp(-beta * energyI) * (1.0 - exp(-beta * nu)) /
       (nu * nu * p->partition);
  return 1;
}

int orc_format_line(const orc_params* p, char* buf, int buflen, int jI,
                    int gI, int ndegI, int ilevelI, double eI, int jF, int gF,
                    int ndegF, int ilevelF, double eF, const double* ls,
                    int with_intens, int with_ls) {
  double S, A, I;
  if (!orc_line(p, jI, gI, ndegI, eI, jF, gF, ndegF, eF, ls, &S, &A, &I)) {
    if (buflen > 0) buf[0] = '\0';
    return 0;
  }
  char lsbuf[300];
  lsbuf[0] = '\0';
  if (with_ls) {
    for (int dF = 0; dF < ndegF; dF++)
      for (int dI = 0; dI < ndegI; dI++) {
        if (!orc_degeneracy_filter(p, gI, gF, dI, dF)) continue;
        size_t l = strlen(lsbuf);
        snprintf(lsbuf + l, sizeof(lsbuf) - l, " %16.9E",
                 ls[dI + dF * p->maxdeg]);
      }
  }
  double nu = eF - eI;
  int n = snprintf(buf, buflen, "%12.6f %8d %4d %4d <- %8d %4d %4d %16.9E ",
                   nu, ilevelF + 1, jF, gF + 1, ilevelI + 1, jI, gI + 1, A);
  if (with_intens) n += snprintf(buf + n, buflen - n, "%16.9E ", I);
  if (with_ls) n += snprintf(buf + n, buflen - n, "[ %s ] ", lsbuf);
  n += snprintf(buf + n, buflen - n, " || \n");
  return n;
}

/* ---- main.cpp:238-338 loop contract --------------------------------- */
static int any_from_j(const orc_params* p, const orc_states* st, int I,
                      int J) {
  /* States::FilterAnyTransitionsFromJ, States.cpp:77-92 */
  int jI = st->jval[I];
  if (abs(jI - J) > 1 || jI + J == 0) return 0;
  for (int F = 0; F < st->n; F++) {
    if (st->jval[F] != J) continue;
    if (orc_filter_intensity(p, jI, st->igamma[I], st->energy[I], J,
                             st->igamma[F], st->energy[F]))
      return 1;
  }
  return 0;
}

int64_t orc_run(const orc_params* p, const orc_basis* bases, int nJ,
                const orc_states* st, const double* dipole, int nc,
                int jmax, const int* initial, int ninitial,
                orc_transition* out, int64_t cap, int nthreads) {
  if (p->maxdeg > 3 || nJ <= 0) return -1;
  double* tab =
      (double*)malloc(sizeof(double) * (size_t)(jmax + 1) * (jmax + 1) * 9);
  orc_threej_table(jmax, tab);
  int dimmax = 0;
  for (int j = 0; j < nJ; j++) dimmax = imax(dimmax, bases[j].nprim);
  double* xprim = (double*)malloc(sizeof(double) * (size_t)dimmax);
  double* h = (double*)calloc((size_t)nJ * 3 * dimmax, sizeof(double));
  int* flist = (int*)malloc(sizeof(int) * (size_t)st->n);
  int64_t count = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  for (int ii = 0; ii < ninitial; ii++) {
    int I = initial[ii];
    if (!orc_filter_lower(p, st->energy[I], st->igamma[I])) continue;
    int jI = st->jval[I], gI = st->igamma[I], ndegI = st->ndeg[I];
    int indI = st->jind[I];
    int gFpair = p->igamma_pair[gI];
    const orc_basis* bI = &bases[indI];
    const double* cI = st->vectors + st->offset[I];
    memset(h, 0, sizeof(double) * (size_t)nJ * 3 * dimmax);
    for (int indF = 0; indF < nJ; indF++) {
      if (!any_from_j(p, st, I, bases[indF].J)) continue;
      const orc_basis* bF = &bases[indF];
      for (int dI = 0; dI < ndegI; dI++) {
        if (!orc_degeneracy_filter(p, gI, gFpair, dI, 0)) continue;
        orc_expand(bI->nprim, gI, bI->maxsymcoeffs, dI, bI->sdeg[gI],
                   bI->ntot[gI], bI->ijterms, bI->icontr, bI->N[gI],
                   bI->repres[gI], cI, xprim);
        if (bI->rotsym) {
          /* GpuManager.cpp:612-618 -> ExecuteRotSymHalfLs (:547-600) */
          const int dJ = bF->J - bI->J + 1;
          const double* w = (dJ >= 0 && dJ <= 2) ? bI->wig[dJ] : NULL;
          if (!w || !bF->rotsym || bF->rmd != bI->rmd) {
            count = -1;
            goto done;
          }
          orc_half_ls_rotsym(bF->nprim, bI->nprim, bI->nlevels, bF->nlevels,
                             bI->rmd, bI->deg, bF->deg, bI->level, bF->level,
                             bI->vib, bF->vib, 0, nc, nc, dipole, nc, w,
                             xprim, h + ((size_t)indF * 3 + dI) * dimmax);
        } else {
          orc_half_ls(jI, bI->nprim, bI->kstart, bI->kblock, bI->tau,
                      bI->vib, bF->J, bF->nprim, bF->kstart, bF->kblock,
                      bF->tau, bF->vib, dipole, nc, 0, nc, tab, jmax, xprim,
                      h + ((size_t)indF * 3 + dI) * dimmax);
        }
      }
    }
    int nf = 0;
    for (int F = 0; F < st->n; F++) {
      if (F == I) continue; /* main.cpp:294-295 */
      if (!orc_filter_intensity(p, jI, gI, st->energy[I], st->jval[F],
                                st->igamma[F], st->energy[F]))
        continue;
      flist[nf++] = F;
    }
    if (count + nf > cap) {
      count = -1;
      break;
    }
#pragma omp parallel
    {
      double* xf = (double*)malloc(sizeof(double) * (size_t)dimmax);
#pragma omp for schedule(dynamic, 16)
      for (int q = 0; q < nf; q++) {
        int F = flist[q];
        int indF = st->jind[F], gF = st->igamma[F], ndegF = st->ndeg[F];
        const orc_basis* bF = &bases[indF];
        const double* cF = st->vectors + st->offset[F];
        orc_transition* tr = &out[count + q];
        memset(tr, 0, sizeof(*tr));
        tr->iI = I;
        tr->iF = F;
        for (int dF = 0; dF < ndegF; dF++) {
          int expanded = 0;
          for (int dI = 0; dI < ndegI; dI++) {
            if (!orc_degeneracy_filter(p, gI, gF, dI, dF)) continue;
            if (!expanded) {
              orc_expand(bF->nprim, gF, bF->maxsymcoeffs, dF, bF->sdeg[gF],
                         bF->ntot[gF], bF->ijterms, bF->icontr, bF->N[gF],
                         bF->repres[gF], cF, xf);
              expanded = 1;
            }
            tr->ls[dI + dF * p->maxdeg] =
                orc_dot(bF->nprim, xf, h + ((size_t)indF * 3 + dI) * dimmax);
          }
        }
        tr->emitted = orc_line(p, jI, gI, ndegI, st->energy[I], st->jval[F],
                               gF, ndegF, st->energy[F], tr->ls, &tr->S,
                               &tr->A, &tr->I);
      }
      free(xf);
    }
    count += nf;
  }
done:
  free(flist);
  free(h);
  free(xprim);
  free(tab);
  return count;
}

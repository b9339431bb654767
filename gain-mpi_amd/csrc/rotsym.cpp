// Symmetrised rotational functions of a rigid symmetric top and the
// contracted rotational dipole factors of the symmetric-top (ROTSYM) basis:
// the Wigner table that compute_gpu_half_linestrength_rotsym_
// (dipole_GPU.cu:641-723) consumes and that the reference builds in Fortran
// (c_wignerGen, trove_wrappers.f90:170-414; MLrotsymmetry_generate
// :416-589; MLsymmetrization_rot :591-914; group tables symmetry.f90:58-745;
// Wigner D function fort_func.f90:150-256).  Host code, long double.
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "engine_internal.hpp"

namespace gain {
namespace {

typedef long double ld;
typedef std::complex<ld> cld;

const ld kPi = 3.141592653589793238462643383279502884L;
// model epsilon of the reference's working precision (accuracy.f90: small_)
const ld kSmall = 2.220446049250313080847263336181640625e-16L;

// ---------------------------------------------------------------- groups
// Molecular symmetry groups with their character tables, the Euler angles
// of the rotation equivalent to each operation and the matrices of the
// degenerate irreducible representations (symmetry.f90:58-540).
// Operations are numbered class by class.
struct Group {
  int nrep = 0, noper = 0, nclass = 0, maxdeg = 1;
  std::vector<int> degen, nelem;
  std::vector<std::vector<ld>> chars;         // [irrep][class]
  std::vector<std::array<ld, 3>> euler;       // [operation] (phi, theta, chi)
  std::vector<std::vector<std::vector<ld>>> rep;  // [irrep][op] d x d, row major
  std::vector<int> class_of;                  // [operation]
};

// 2 x 2 representation matrices are given by their rows.
typedef std::array<ld, 4> M2;

void finish(Group& g, const std::vector<std::vector<M2>>& two) {
  g.class_of.clear();
  for (int c = 0; c < g.nclass; ++c)
    for (int e = 0; e < g.nelem[c]; ++e) g.class_of.push_back(c);
  g.maxdeg = 1;
  g.rep.assign(g.nrep, {});
  int t = 0;
  for (int r = 0; r < g.nrep; ++r) {
    g.maxdeg = std::max(g.maxdeg, g.degen[r]);
    g.rep[r].resize(g.noper);
    for (int o = 0; o < g.noper; ++o) {
      if (g.degen[r] == 1) {
        g.rep[r][o] = {g.chars[r][g.class_of[o]]};
      } else {
        const M2& m = two[t][o];
        g.rep[r][o] = {m[0], m[1], m[2], m[3]};
      }
    }
    if (g.degen[r] == 2) ++t;
  }
}

bool make_group(const std::string& name, Group& g) {
  const ld o = 0.0L, pi = kPi;
  if (name == "C(M)" || name == "C") {  // symmetry.f90:60-90
    g.nrep = 1, g.noper = 1, g.nclass = 1;
    g.degen = {1}, g.nelem = {1}, g.chars = {{1}};
    g.euler = {{o, o, o}};
    finish(g, {});
    return true;
  }
  if (name == "CS(M)" || name == "CS") {  // symmetry.f90:92-108
    g.nrep = 2, g.noper = 2, g.nclass = 2;
    g.degen = {1, 1}, g.nelem = {1, 1};
    g.chars = {{1, 1}, {1, -1}};
    g.euler = {{o, o, o}, {o, o, o}};  // no angles are assigned
    finish(g, {});
    return true;
  }
  if (name == "C2V(M)" || name == "C2V") {  // symmetry.f90:110-138
    g.nrep = 4, g.noper = 4, g.nclass = 4;
    g.degen = {1, 1, 1, 1}, g.nelem = {1, 1, 1, 1};
    g.chars = {{1, 1, 1, 1}, {1, 1, -1, -1}, {1, -1, -1, 1}, {1, -1, 1, -1}};
    g.euler = {{o, o, o}, {pi, o, o}, {o, pi, o}, {0.5L * pi, pi, 1.5L * pi}};
    finish(g, {});
    return true;
  }
  // Abelian groups of order 4 and 8: one operation per class; their
  // equivalent rotations are the identity and the twofold rotations about
  // z (pi, 0, 0), y (0, pi, 0) and x (pi/2, pi, 3pi/2).
  const std::array<ld, 3> E0 = {o, o, o}, Rz = {pi, o, o}, Ry = {o, pi, o},
                          Rx = {0.5L * pi, pi, 1.5L * pi};
  auto abelian = [&](std::vector<std::vector<ld>> chars,
                     std::vector<std::array<ld, 3>> euler) {
    const int n = (int)chars.size();
    g.nrep = g.noper = g.nclass = n;
    g.degen.assign(n, 1), g.nelem.assign(n, 1);
    g.chars = std::move(chars);
    g.euler = std::move(euler);
    finish(g, {});
    return true;
  };
  const std::vector<std::vector<ld>> klein = {
      {1, 1, 1, 1}, {1, 1, -1, -1}, {1, -1, -1, 1}, {1, -1, 1, -1}};
  // D_nh(M) (symmetry.f90:746-1006 for odd n) at the order the reference
  // runs it: nothing assigns the order sym%N, which keeps its default 1
  // (symmetry.f90:33), so the group is E, sigma_v, i, C'_2 with the
  // characters of A1', A2', A1", A2" and no Euler angles assigned (the
  // C_n / S_n loops that set them are empty for n = 1).
  if (name == "DNH(M)" || name == "DNH")
    return abelian({{1, 1, 1, 1}, {1, -1, 1, -1}, {1, 1, -1, -1},
                    {1, -1, -1, 1}},
                   {E0, E0, E0, E0});
  if (name == "C2H(M)" || name == "C2H")  // symmetry.f90:140-170
    return abelian(klein, {E0, Ry, Rx, Rz});
  if (name == "CS(EM)")  // symmetry.f90:171-202
    return abelian({{1, 1, 1, 1}, {1, 1, -1, -1}, {1, -1, 1, -1}, {1, -1, -1, 1}},
                   {E0, Rz, Rx, Ry});
  if (name == "G4(M)" || name == "G4")  // symmetry.f90:203-230
    return abelian(klein, {E0, Rx, Ry, Rz});
  if (name == "G4(EM)")  // symmetry.f90:231-282
    return abelian({{1, 1, 1, 1, 1, 1, 1, 1},
                    {1, 1, -1, -1, 1, 1, -1, -1},
                    {1, -1, -1, 1, 1, -1, -1, 1},
                    {1, -1, 1, -1, 1, -1, 1, -1},
                    {1, 1, -1, -1, -1, -1, 1, 1},
                    {1, 1, 1, 1, -1, -1, -1, -1},
                    {1, -1, 1, -1, -1, 1, -1, 1},
                    {1, -1, -1, 1, -1, 1, 1, -1}},
                   {E0, Ry, Rx, Rz, Rz, Rx, Ry, E0});
  if (name == "D2H(M)")  // symmetry.f90:283-323
    return abelian({{1, 1, 1, 1, 1, 1, 1, 1},
                    {1, 1, 1, 1, -1, -1, -1, -1},
                    {1, 1, -1, -1, -1, -1, 1, 1},
                    {1, 1, -1, -1, 1, 1, -1, -1},
                    {1, -1, 1, -1, -1, 1, -1, 1},
                    {1, -1, 1, -1, 1, -1, 1, -1},
                    {1, -1, -1, 1, 1, -1, -1, 1},
                    {1, -1, -1, 1, -1, 1, 1, -1}},
                   {E0, Rz, Ry, Rx, Rx, Ry, Rz, E0});
  if (name == "D2H")  // symmetry.f90:324-364
    return abelian({{1, 1, 1, 1, 1, 1, 1, 1},
                    {1, 1, 1, 1, -1, -1, -1, -1},
                    {1, 1, -1, -1, 1, 1, -1, -1},
                    {1, 1, -1, -1, -1, -1, 1, 1},
                    {1, -1, 1, -1, 1, -1, 1, -1},
                    {1, -1, 1, -1, -1, 1, -1, 1},
                    {1, -1, -1, 1, 1, -1, -1, 1},
                    {1, -1, -1, 1, -1, 1, 1, -1}},
                   {E0, Rz, Ry, Rx, Rz, E0, Rx, Ry});
  const ld a = 0.5L, b = 0.5L * std::sqrt(3.0L);
  if (name == "C3V(M)" || name == "C3V") {  // symmetry.f90:365-415
    g.nrep = 3, g.noper = 6, g.nclass = 3;
    g.degen = {1, 1, 2}, g.nelem = {1, 2, 3};
    g.chars = {{1, 1, 1}, {1, 1, -1}, {2, -1, 0}};
    const ld p3 = 2.0L / 3.0L * pi, p4 = 4.0L / 3.0L * pi;
    g.euler = {{o, o, o},  {p4, o, o},  {p3, o, o},
               {o, pi, o}, {p3, pi, o}, {p4, pi, o}};
    // E, rows of the matrices (symmetry.f90:397-413)
    finish(g, {{M2{1, 0, 0, 1}, M2{-a, b, -b, -a}, M2{-a, -b, b, -a},
                M2{1, 0, 0, -1}, M2{-a, b, b, a}, M2{-a, -b, -b, a}}});
    return true;
  }
  if (name == "D3H(M)" || name == "D3H") {  // symmetry.f90:417-538
    g.nrep = 6, g.noper = 12, g.nclass = 6;
    g.degen = {1, 1, 2, 1, 1, 2}, g.nelem = {1, 2, 3, 1, 2, 3};
    g.chars = {{1, 1, 1, 1, 1, 1},   {1, 1, -1, 1, 1, -1},
               {2, -1, 0, 2, -1, 0}, {1, 1, 1, -1, -1, -1},
               {1, 1, -1, -1, -1, 1}, {2, -1, 0, -2, 1, 0}};
    const ld p3 = pi / 3.0L, p23 = 2.0L / 3.0L * pi, p43 = 4.0L / 3.0L * pi;
    g.euler = {{o, o, o},    {p23, o, o},  {p43, o, o},  {pi, pi, o},
               {p3, pi, o},  {-p3, pi, o}, {pi, o, o},   {-p3, o, o},
               {p3, o, o},   {o, pi, o},   {p43, pi, o}, {p23, pi, o}};
    // E' (symmetry.f90:453-487) and E" (:489-523), rows of the matrices
    const std::vector<M2> e1 = {
        M2{1, 0, 0, 1},    M2{-a, -b, b, -a}, M2{-a, b, -b, -a},
        M2{1, 0, 0, -1},   M2{-a, -b, -b, a}, M2{-a, b, b, a},
        M2{1, 0, 0, 1},    M2{-a, -b, b, -a}, M2{-a, b, -b, -a},
        M2{1, 0, 0, -1},   M2{-a, -b, -b, a}, M2{-a, b, b, a}};
    const std::vector<M2> e2 = {
        M2{1, 0, 0, 1},   M2{-a, -b, b, -a}, M2{-a, b, -b, -a},
        M2{-1, 0, 0, 1},  M2{a, b, b, -a},   M2{a, -b, -b, -a},
        M2{-1, 0, 0, -1}, M2{a, b, -b, a},   M2{a, -b, b, a},
        M2{1, 0, 0, -1},  M2{-a, -b, -b, a}, M2{-a, b, b, a}};
    finish(g, {e1, e2});
    return true;
  }
  if (name == "TD(M)" || name == "TD") {  // symmetry.f90:539-745
    g.nrep = 5, g.noper = 24, g.nclass = 5;
    g.degen = {1, 1, 2, 3, 3}, g.nelem = {1, 8, 3, 6, 6};
    g.chars = {{1, 1, 1, 1, 1}, {1, 1, 1, -1, -1}, {2, -1, 2, 0, 0},
               {3, 0, -1, 1, -1}, {3, 0, -1, -1, 1}};
    const ld p2 = 0.5L * pi, p3 = 1.5L * pi;
    g.euler = {{o, o, o},    {p3, p2, o},  {pi, p2, p3}, {o, p2, p3},
               {p3, p2, pi}, {p2, p2, o},  {pi, p2, p2}, {o, p2, p2},
               {p2, p2, pi}, {p2, pi, p3}, {o, pi, o},   {pi, o, o},
               {o, p3, o},   {o, p2, o},   {p2, o, o},   {p3, o, o},
               {p2, p3, p3}, {p2, p2, p3}, {p2, p2, p2}, {o, p2, pi},
               {o, pi, p2},  {pi, pi, p2}, {pi, p2, o},  {p3, p2, p3}};
    g.class_of.clear();
    for (int c = 0; c < g.nclass; ++c)
      for (int e = 0; e < g.nelem[c]; ++e) g.class_of.push_back(c);
    g.maxdeg = 3;
    g.rep.assign(g.nrep, std::vector<std::vector<ld>>(g.noper));
    for (int r = 0; r < 2; ++r)
      for (int op = 0; op < g.noper; ++op)
        g.rep[r][op] = {g.chars[r][g.class_of[op]]};
    // degenerate irreps from the generators (123) (operation 2) and (14)*
    // (operation 21), rows of the matrices, and the products of
    // symmetry.f90:714-737
    typedef std::vector<ld> Mx;
    auto mul = [](const Mx& x, const Mx& y, int d) {
      Mx z((size_t)d * d, 0.0L);
      for (int i = 0; i < d; ++i)
        for (int k = 0; k < d; ++k)
          for (int j = 0; j < d; ++j) z[i * d + j] += x[i * d + k] * y[k * d + j];
      return z;
    };
    const Mx gen2[3] = {{-a, b, -b, -a},
                        {0, 1, 0, 0, 0, -1, -1, 0, 0},
                        {0, -1, 0, 0, 0, -1, 1, 0, 0}};
    const Mx gen21[3] = {{-a, b, b, a},
                         {0, 1, 0, 1, 0, 0, 0, 0, -1},
                         {0, 0, -1, 0, 1, 0, -1, 0, 0}};
    for (int r = 2; r < 5; ++r) {
      const int d = g.degen[r];
      std::vector<Mx>& m = g.rep[r];
      m[0].assign((size_t)d * d, 0.0L);
      for (int i = 0; i < d; ++i) m[0][i * d + i] = 1.0L;
      m[1] = gen2[r - 2];
      m[20] = gen21[r - 2];
      // 1-based operation numbers: (target, left, right)
      static const int seq[][3] = {
          {3, 2, 2},    {13, 21, 2},  {8, 13, 21},  {5, 8, 3},
          {4, 5, 5},    {6, 4, 3},    {7, 6, 6},    {9, 8, 8},
          {10, 7, 2},   {11, 8, 2},   {12, 4, 2},   {14, 3, 21},
          {15, 11, 21}, {16, 10, 21}, {17, 9, 21},  {18, 2, 21},
          {19, 5, 21},  {20, 7, 21},  {22, 12, 21}, {23, 4, 21},
          {24, 6, 21}};
      for (const auto& q : seq) m[q[0] - 1] = mul(m[q[1] - 1], m[q[2] - 1], d);
    }
    return true;
  }
  return false;
}

// ------------------------------------------------------ Wigner D function
ld faclog(int k) {
  ld v = 0.0L;
  for (int i = 2; i <= k; ++i) v += std::log((ld)i);
  return v;
}

// Wigner small-d in the reference's convention (fort_func.f90:150-227):
// (-1)^(j-n) sum_s (-1)^s sqrt((j-m)!(j+m)!(j-n)!(j+n)!) /
//   ((j-m-s)!(j-n-s)! s! (m+n+s)!) cos(t/2)^(2s+m+n) sin(t/2)^(2j-2s-m-n)
ld small_d(int j, int n, int m, ld t) {
  const ld c = std::cos(0.5L * t), s = std::sin(0.5L * t);
  const ld g2 =
      0.5L * (faclog(j - m) + faclog(j + m) + faclog(j - n) + faclog(j + n));
  ld sum = 0.0L;
  for (int k = 0; k <= std::min(j - m, j - n); ++k) {
    if (j - m - k < 0 || j - n - k < 0 || m + n + k < 0) continue;
    const ld g1 =
        faclog(j - m - k) + faclog(j - n - k) + faclog(k) + faclog(m + n + k);
    const ld cs = std::pow(c, 2 * k + m + n) *
                  std::pow(s, 2 * j - 2 * k - m - n) * ((k & 1) ? -1.0L : 1.0L);
    const ld h = std::fabs(cs);
    if (h > 0.0L) sum += std::exp(g2 - g1 + std::log(h)) * (cs < 0 ? -1 : 1);
  }
  return ((j - n) & 1) ? -sum : sum;
}

// Conjugated Wigner D (fort_func.f90:232-256):
// (-1)^(n-m) d(j, -n, -m, theta) exp(i (n phi + m chi))
cld big_d_conj(int j, int n, int m, ld phi, ld theta, ld chi) {
  const ld pre = (((n - m) & 1) ? -1.0L : 1.0L) * small_d(j, -n, -m, theta);
  const ld x = (ld)n * phi + (ld)m * chi;
  return cld(std::cos(x) * pre, std::sin(x) * pre);
}

typedef std::vector<ld> Mat;  // square, row major

// --------------------------------------------------------- symmetrisation
struct RotSym {
  int dim = 0, nlevels = 0;
  std::vector<int> ntotal;        // levels per irrep
  Mat vec;                        // [i][root]: function root in the Wang basis
  std::vector<int> root_of;       // [level][deg] -> root, -1 if unused
  std::vector<int> level_irrep;   // [level]
  std::vector<Mat> tmat;          // [op] real part of R^H D R (Wang basis)
};

// Irreducible components of the rotation matrices `tm` by the projection
// operators of each irrep, Gram-Schmidt against the components already
// found, the partners of a degenerate irrep generated by the off-diagonal
// projectors, and the reference's acceptance tests (transformation as the
// irrep, not a repeat): MLsymmetrization_rot, trove_wrappers.f90:591-914.
void symmetrize(const Group& g, int dim, const std::vector<Mat>& tm,
                std::vector<int>& nirr,
                std::vector<std::vector<std::vector<ld>>>& irr) {
  auto T = [&](int op, int r, int c) { return tm[op][r * dim + c]; };
  auto proj = [&](int s, int op, int i, int k) {
    return g.rep[s][op][i * g.degen[s] + k] * (ld)g.degen[s] / (ld)g.noper;
  };
  // characters of the rotation representation, one operation per class
  std::vector<ld> chi(g.nclass, 0.0L);
  for (int c = 0, op = 0; c < g.nclass; op += g.nelem[c], ++c)
    for (int i = 0; i < dim; ++i) chi[c] += T(op, i, i);
  nirr.assign(g.nrep, 0);
  bool any = false;
  for (int s = 0; s < g.nrep; ++s) {
    ld n = 0.0L;
    for (int c = 0; c < g.nclass; ++c) n += (ld)g.nelem[c] * g.chars[s][c] * chi[c];
    nirr[s] = (int)std::lround(n / (ld)g.noper);
    any = any || nirr[s] != 0;
  }
  if (!any) fail(GAIN_ERR_ARG, "rotational symmetrisation: no irreps");
  irr.assign(g.nrep, {});
  const ld tol_rep = 100.0L * std::sqrt(kSmall), tol_same = std::sqrt(kSmall);
  for (int s = 0; s < g.nrep; ++s) {
    if (nirr[s] == 0) continue;
    const int nd = g.degen[s];
    int found = 0;
    for (int tr = 0; tr < dim && found < nirr[s]; ++tr) {
      std::vector<std::vector<ld>> tv(nd, std::vector<ld>(dim, 0.0L));
      for (int d = 0; d < nd; ++d)
        for (int e = 0; e < dim; ++e) {
          ld v = 0.0L;
          for (int op = 0; op < g.noper; ++op) v += proj(s, op, d, d) * T(op, tr, e);
          tv[d][e] = v;
        }
      for (int it = 0; it < nd && found < nirr[s]; ++it) {
        std::vector<std::vector<ld>> v(nd, std::vector<ld>(dim, 0.0L));
        v[0] = tv[it];
        auto norm2 = [&](const std::vector<ld>& x) {
          ld f = 0.0L;
          for (ld y : x) f += y * y;
          return f;
        };
        ld f = norm2(v[0]);
        if (!(f > kSmall)) continue;
        for (ld& y : v[0]) y /= std::sqrt(f);
        for (int q = 0; q < found; ++q)
          for (int d = 0; d < nd; ++d) {
            const std::vector<ld>& w = irr[s][q * nd + d];
            ld p = 0.0L;
            for (int e = 0; e < dim; ++e) p += w[e] * v[0][e];
            for (int e = 0; e < dim; ++e) v[0][e] -= w[e] * p;
          }
        f = norm2(v[0]);
        if (!(f > kSmall)) continue;
        for (ld& y : v[0]) y /= std::sqrt(f);
        // partners of a degenerate irrep
        for (int d = 1; d < nd; ++d)
          for (int op = 0; op < g.noper; ++op) {
            const ld pj = proj(s, op, 0, d);
            for (int e = 0; e < dim; ++e) {
              const ld ve = v[0][e];
              if (ve == 0.0L) continue;
              for (int c = 0; c < dim; ++c) v[d][c] += pj * T(op, e, c) * ve;
            }
          }
        bool ok = true;
        for (int d = 1; d < nd && ok; ++d) {
          f = norm2(v[d]);
          if (!(f > kSmall)) {
            ok = false;
            break;
          }
          for (ld& y : v[d]) y /= std::sqrt(f);
          for (int k = 0; k < d; ++k) {
            ld p = 0.0L;
            for (int e = 0; e < dim; ++e) p += v[k][e] * v[d][e];
            for (int e = 0; e < dim; ++e) v[d][e] -= v[k][e] * p;
          }
        }
        if (!ok) continue;
        // transforms as the irrep under every operation
        for (int op = 0; op < g.noper && ok; ++op)
          for (int d = 0; d < nd && ok; ++d)
            for (int c = 0; c < dim && ok; ++c) {
              ld ft = 0.0L, gt = 0.0L;
              for (int e = 0; e < dim; ++e) ft += v[d][e] * T(op, e, c);
              for (int k = 0; k < nd; ++k)
                gt += g.rep[s][op][d * nd + k] * v[k][c];
              if (std::fabs(ft - gt) > tol_rep) ok = false;
            }
        if (!ok) continue;
        // not a repeat of a component already found (the reference bounds
        // the search by the count found for this irrep)
        for (int s2 = 0; s2 <= s && ok; ++s2)
          for (int q = 0; q < std::min(nirr[s2], found) && ok; ++q)
            for (int k = 0; k < nd && ok; ++k)
              for (int d = 0; d < g.degen[s2] && ok; ++d) {
                if (q * g.degen[s2] + d >= (int)irr[s2].size()) continue;
                const std::vector<ld>& w = irr[s2][q * g.degen[s2] + d];
                bool same = true;
                for (int e = 0; e < dim && same; ++e)
                  same = std::fabs(v[k][e] - w[e]) < tol_same;
                if (same) ok = false;
              }
        if (!ok) continue;
        ++found;
        for (int d = 0; d < nd; ++d) irr[s].push_back(v[d]);
      }
    }
    if (found != nirr[s])
      fail(GAIN_ERR_ARG,
           "rotational symmetrisation: %d of %d components of irrep %d found",
           found, nirr[s], s);
  }
}

// Symmetrised rotational functions of one J in the Wang basis
// (MLrotsymmetry_generate, trove_wrappers.f90:416-589).  Wang functions
// are ordered k = 0 (one function), then k = 1..J with tau = 0, 1.
RotSym rot_functions(const Group& g, int J) {
  const int n = 2 * J + 1;
  RotSym out;
  out.dim = n;
  // Wang transform R (columns: Wang functions; rows: |J, k> with the
  // signed k ordering of the operation matrices below)
  std::vector<cld> R((size_t)n * n, cld(0, 0));
  auto r = [&](int i, int c) -> cld& { return R[(size_t)i * n + c]; };
  const ld h = 1.0L / std::sqrt(2.0L);
  r(0, 0) = (J % 2 == 0) ? cld(1, 0) : cld(0, -1);
  for (int k = 1, i = 1; k <= J; ++k, i += 2) {
    const int sg = k % 3;
    const ld p = (sg & 1) ? -1.0L : 1.0L;            // (-1)^sigma
    const ld q = ((J + k) & 1) ? -1.0L : 1.0L;       // (-1)^(J+k)
    r(i, i) = cld(h, 0);
    r(i, i + 1) = cld(0, -p * h);
    r(i + 1, i) = cld(q * h, 0);
    r(i + 1, i + 1) = cld(0, q * p * h);
  }
  // signed projection of each function: n = k (-1)^tau
  std::vector<int> nk;
  for (int k = 0; k <= J; ++k)
    for (int tau = 0; tau <= std::min(1, k); ++tau)
      nk.push_back(tau ? -k : k);
  out.tmat.resize(g.noper);
  for (int op = 0; op < g.noper; ++op) {
    std::vector<cld> T((size_t)n * n), TR((size_t)n * n, cld(0, 0));
    for (int i1 = 0; i1 < n; ++i1)
      for (int i2 = 0; i2 < n; ++i2)
        T[(size_t)i2 * n + i1] = big_d_conj(J, nk[i2], nk[i1], g.euler[op][0],
                                            g.euler[op][1], g.euler[op][2]);
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < n; ++k) {
        const cld t = T[(size_t)i * n + k];
        if (t == cld(0, 0)) continue;
        for (int c = 0; c < n; ++c) TR[(size_t)i * n + c] += t * R[(size_t)k * n + c];
      }
    Mat& q = out.tmat[op];
    q.assign((size_t)n * n, 0.0L);
    for (int a = 0; a < n; ++a)
      for (int c = 0; c < n; ++c) {
        cld s(0, 0);
        for (int k = 0; k < n; ++k) s += std::conj(R[(size_t)k * n + a]) * TR[(size_t)k * n + c];
        q[(size_t)a * n + c] = s.real();
      }
  }
  std::vector<int> nirr;
  std::vector<std::vector<std::vector<ld>>> irr;
  symmetrize(g, n, out.tmat, nirr, irr);
  out.ntotal = nirr;
  out.vec.assign((size_t)n * n, 0.0L);
  int root = 0;
  for (int s = 0; s < g.nrep; ++s)
    for (int q = 0; q < nirr[s]; ++q) {
      out.level_irrep.push_back(s);
      for (int d = 0; d < g.maxdeg; ++d) {
        if (d < g.degen[s]) {
          for (int e = 0; e < n; ++e)
            out.vec[(size_t)e * n + root] = irr[s][q * g.degen[s] + d][e];
          out.root_of.push_back(root++);
        } else {
          out.root_of.push_back(-1);
        }
      }
    }
  out.nlevels = (int)out.level_irrep.size();
  if (root != n)
    fail(GAIN_ERR_ARG, "rotational symmetrisation: %d functions for J = %d",
         root, J);
  return out;
}

const Group& group_of(const char* name) {
  GAIN_CHECK_ARG(name, "null group name");
  std::string s(name);
  for (char& ch : s) ch = (char)std::toupper((unsigned char)ch);
  static thread_local Group g;
  static thread_local std::string last;
  if (s != last) {
    Group t;
    // D_infinity_h(M) (symmetry.f90:1444-1543) leaves the matrices of its
    // 2-dimensional irreps unset and gives sigma_v and C'_inf the identity
    // as equivalent rotation, so its operations do not form a
    // representation on the rotational functions
    if (s == "DINFTYH(M)")
      fail(GAIN_ERR_ARG,
           "symmetry group DINFTYH(M): the reference's table does not define "
           "the matrices of its degenerate irreps or the equivalent "
           "rotations of sigma_v and C'_inf, so no symmetrised rotational "
           "functions exist for it");
    if (!make_group(s, t))
      fail(GAIN_ERR_ARG,
           "symmetry group \"%s\": C(M), CS(M), C2V(M), C2H(M), CS(EM), "
           "G4(M), G4(EM), D2H(M), D2H, C3V(M), D3H(M), TD(M), DNH(M)",
           name);
    g = t;
    last = s;
  }
  return g;
}

}  // namespace
}  // namespace gain

using namespace gain;

extern "C" int gain_rotsym_functions(const char* group, int J, int* nlevels,
                                     int* maxdeg, int* level_irrep,
                                     int* root_of, double* vectors,
                                     double* operations) {
  return guarded([&] {
    GAIN_CHECK_ARG(J >= 0 && J <= 200 && nlevels && maxdeg,
                   "rotsym_functions: bad arguments");
    const Group& g = group_of(group);
    const RotSym rs = rot_functions(g, J);
    *nlevels = rs.nlevels;
    *maxdeg = g.maxdeg;
    const int n = rs.dim;
    if (level_irrep)
      for (int l = 0; l < rs.nlevels; ++l) level_irrep[l] = rs.level_irrep[l];
    if (root_of)
      for (size_t i = 0; i < rs.root_of.size(); ++i) root_of[i] = rs.root_of[i];
    if (vectors)
      for (size_t i = 0; i < (size_t)n * n; ++i) vectors[i] = (double)rs.vec[i];
    if (operations)
      for (int op = 0; op < g.noper; ++op)
        for (size_t i = 0; i < (size_t)n * n; ++i)
          operations[(size_t)op * n * n + i] = (double)rs.tmat[op][i];
  });
}

extern "C" int gain_rotsym_wigner(const char* group, int jI, int jF,
                                  int* nlevelsI, int* nlevelsF, int* maxdeg,
                                  double* rot) {
  return guarded([&] {
    GAIN_CHECK_ARG(nlevelsI && nlevelsF && maxdeg,
                   "rotsym_wigner: null size output");
    GAIN_CHECK_ARG(jI >= 0 && jF >= 0 && std::abs(jF - jI) <= 1 &&
                       jI <= 200 && jF <= 200,
                   "rotsym_wigner: need 0 <= J <= 200 and |jF - jI| <= 1");
    const Group& g = group_of(group);
    const RotSym fi = rot_functions(g, jI), ff = rot_functions(g, jF);
    *nlevelsI = fi.nlevels;
    *nlevelsF = ff.nlevels;
    *maxdeg = g.maxdeg;
    if (!rot) return;
    const int nI = fi.dim, nF = ff.dim, md = g.maxdeg;
    // K-block rotational factor between Wang functions, in the reference's
    // signs (trove_wrappers.f90:327-381; 1/sqrt(2) exact here, unlike the
    // truncated constant of the K-block kernel)
    std::vector<int> kI(nI), tI(nI), kF(nF), tF(nF);
    kI[0] = 0, tI[0] = jI % 2;
    for (int k = 1, i = 1; k <= jI; ++k, i += 2) kI[i] = kI[i + 1] = k, tI[i] = 0, tI[i + 1] = 1;
    kF[0] = 0, tF[0] = jF % 2;
    for (int k = 1, i = 1; k <= jF; ++k, i += 2) kF[i] = kF[i + 1] = k, tF[i] = 0, tF[i + 1] = 1;
    std::vector<Mat> mat(3, Mat((size_t)nI * nF, 0.0L));
    for (int a = 0; a < nI; ++a) {
      const int sI = (kI[a] % 3) * tI[a];
      const ld fI = ((sI + kI[a]) & 1) ? -1.0L : 1.0L;
      for (int b = 0; b < nF; ++b) {
        const int sF = (kF[b] % 3) * tF[b];
        const ld fIF = fI * ((sF & 1) ? -1.0L : 1.0L);
        if (std::abs(kI[a] - kF[b]) > 1) continue;
        const ld f3j = three_j(jI, 1, jF, kI[a], kF[b] - kI[a], -kF[b]);
        if (std::fabs(f3j) < kSmall) continue;
        const ld inv2 = (kI[a] * kF[b] != 0) ? 1.0L / std::sqrt(2.0L) : 1.0L;
        if (kF[b] == kI[a])
          mat[2][(size_t)a * nF + b] = (ld)(tF[b] - tI[a]) * f3j * fIF;
        else if (tF[b] != tI[a])
          mat[0][(size_t)a * nF + b] =
              (ld)((kF[b] - kI[a]) * (tF[b] - tI[a])) * f3j * fIF * inv2;
        else
          mat[1][(size_t)a * nF + b] = -f3j * fIF * inv2;
      }
    }
    const size_t total = (size_t)3 * fi.nlevels * ff.nlevels * md * md;
    std::memset(rot, 0, total * sizeof(double));
    // E_I^T mat_n E_F, stored as rot(n, levelI, levelF, degI, degF)
    // (first index fastest: the ROTSYM kernel's Wigner index)
    Mat tmp((size_t)nI * nF);
    for (int c = 0; c < 3; ++c) {
      for (int a = 0; a < nI; ++a)
        for (int q = 0; q < nF; ++q) {
          ld s = 0.0L;
          for (int b = 0; b < nF; ++b) s += mat[c][(size_t)a * nF + b] * ff.vec[(size_t)b * nF + q];
          tmp[(size_t)a * nF + q] = s;
        }
      for (int lI = 0; lI < fi.nlevels; ++lI)
        for (int dI = 0; dI < md; ++dI) {
          const int p = fi.root_of[(size_t)lI * md + dI];
          if (p < 0) continue;
          for (int lF = 0; lF < ff.nlevels; ++lF)
            for (int dF = 0; dF < md; ++dF) {
              const int q = ff.root_of[(size_t)lF * md + dF];
              if (q < 0) continue;
              ld s = 0.0L;
              for (int a = 0; a < nI; ++a) s += fi.vec[(size_t)a * nI + p] * tmp[(size_t)a * nF + q];
              rot[c + 3 * ((size_t)lI + (size_t)fi.nlevels *
                               ((size_t)lF + (size_t)ff.nlevels *
                                                 ((size_t)dI + (size_t)md * dF)))] =
                  (double)s;
            }
        }
    }
  });
}

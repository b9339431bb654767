// Readers of the TROVE checkpoint files the reference's host classes load,
// and a loader that puts their contents into an engine context.
//
//   j0_extfield.chk          dipole, Fortran sequential records
//                            (TroveDipole.cpp:9-119, Util.cpp:89-109)
//   j0eigen_quanta<J>.chk    contracted basis: cases, lambdas, rotational
//                            quanta, irreducible transformation (text)
//                            (TroveBasisSet.cpp:27-292), correlated with the
//                            J=0 basis (TroveBasisSet.cpp:294-681) and the
//                            ijterms prefix sums (:688-709)
//   j0eigen_descr<J>_<g>.chk states: record, level, energy (text)
//                            (TroveStates.cpp:15-131)
//   j0eigen_vectors<J>_<g>.chk eigenvector records, rec_len doubles each at
//                            record * rec_len (EigenVector.cpp:36,88)
//
// Host code only: nothing here touches the GPU except the final uploads of
// gain_trove_load, which go through the engine's own entry points.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "engine_internal.hpp"

namespace {

// ---- Fortran sequential unformatted records ----------------------------
// Each record is [int32 n][n bytes][int32 n].  gfortran writes a record
// longer than 2^31-9 bytes as subrecords; every subrecord but the last has
// a negative leading marker (its length negated), which is how Util.cpp:
// 89-109 recognises them (0x80000009 = -(2^31-9)).
std::vector<char> read_record(FILE* f, const char* what) {
  std::vector<char> out;
  for (;;) {
    int32_t head = 0, tail = 0;
    if (fread(&head, 4, 1, f) != 1)
      fail(GAIN_ERR_ARG, "%s: truncated record marker", what);
    const bool more = head < 0;
    const int64_t len = more ? -(int64_t)head : (int64_t)head;
    const size_t off = out.size();
    out.resize(off + (size_t)len);
    if (len > 0 && fread(out.data() + off, 1, (size_t)len, f) != (size_t)len)
      fail(GAIN_ERR_ARG, "%s: truncated record", what);
    if (fread(&tail, 4, 1, f) != 1)
      fail(GAIN_ERR_ARG, "%s: truncated record marker", what);
    if (!more) break;
  }
  return out;
}

struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) : f(fopen(path, mode)) {}
  ~File() {
    if (f) fclose(f);
  }
};

std::string trim(const std::string& s) {
  const size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  const size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

// Whitespace-separated integers / doubles of one line, in order.
struct Fields {
  const char* p;
  const char* what;
  explicit Fields(const std::string& line, const char* w)
      : p(line.c_str()), what(w) {}
  long i() {
    char* e = nullptr;
    const long v = strtol(p, &e, 10);
    if (e == p) fail(GAIN_ERR_ARG, "%s: expected an integer", what);
    p = e;
    return v;
  }
  double d() {
    char* e = nullptr;
    const double v = strtod(p, &e);
    if (e == p) fail(GAIN_ERR_ARG, "%s: expected a number", what);
    p = e;
    return v;
  }
};

bool next_line(std::ifstream& in, std::string& line) {
  return (bool)std::getline(in, line);
}

void need_line(std::ifstream& in, std::string& line, const char* what) {
  if (!next_line(in, line)) fail(GAIN_ERR_ARG, "%s: unexpected end", what);
}

}  // namespace

// ---- contracted basis of one J --------------------------------------------
struct gain_trove_basis {
  int J = 0, nrepres = 0;
  // as read
  int ncases = 0, nlambdas = 0, ncontr = 0, nclasses = 0;
  std::vector<int> nlam;          // lambdas per case
  std::vector<int> cspace;        // [case][0..nclasses]
  std::vector<int> icoeffs;       // [contr][0..nclasses] (case-major order)
  std::vector<int> icontr2icase;  // [2][ncontr]: case, lambda
  std::vector<int> case_first;    // first contracted index of each case
  int nrot = 0, nrotdeg = 0;
  std::vector<int> rot_k, rot_tau;  // [ilevel + ideg*nrot]
  int matsize = 0;
  std::vector<int> ntot, sdeg;
  std::vector<std::vector<int>> N;
  std::vector<std::vector<double>> repres;
  // derived (Correlate, ComputeijTerms)
  std::vector<int> vib, tau, kstart, kblock, ijterms;
};

namespace {

void parse_quanta(const char* path, int nrepres, const int* sdeg,
                  gain_trove_basis& b) {
  std::ifstream in(path);
  if (!in) fail(GAIN_ERR_ARG, "cannot open %s", path);
  std::string line;
  need_line(in, line, path);
  if (trim(line) != "Start Primitive basis set")
    fail(GAIN_ERR_ARG, "%s: bad header", path);
  need_line(in, line, path);
  {
    Fields f(line, path);
    b.ncases = (int)f.i();
    b.nlambdas = (int)f.i();
    b.ncontr = (int)f.i();
    b.nclasses = (int)f.i();
  }
  GAIN_CHECK_ARG(b.ncases > 0 && b.nlambdas > 0 && b.ncontr > 0 &&
                     b.nclasses >= 0,
                 "%s: bad sizes", path);
  const int nc1 = b.nclasses + 1;
  b.nlam.resize(b.ncases);
  b.cspace.resize((size_t)b.ncases * nc1);
  b.icoeffs.resize((size_t)b.ncontr * nc1);
  b.icontr2icase.resize((size_t)2 * b.ncontr);
  b.case_first.resize(b.ncases);
  int icontr = 0;
  for (int c = 0; c < b.ncases; ++c) {
    need_line(in, line, path);
    Fields f(line, path);
    b.nlam[c] = (int)f.i();
    GAIN_CHECK_ARG(b.nlam[c] >= 1 && b.nlam[c] <= b.nlambdas,
                   "%s: case %d has %d lambdas", path, c + 1, b.nlam[c]);
    for (int q = 0; q < nc1; ++q)
      b.cspace[(size_t)c * nc1 + q] = (int)f.i() - 1;
    b.case_first[c] = icontr;
    for (int l = 0; l < b.nlam[c]; ++l) {
      need_line(in, line, path);
      Fields g(line, path);
      const int iroot = (int)g.i() - 1;
      if (iroot != icontr)
        fail(GAIN_ERR_ARG, "%s: wrong indexing icontr = %d iroot = %d", path,
             icontr, iroot);
      GAIN_CHECK_ARG(icontr < b.ncontr, "%s: more functions than ncontr",
                     path);
      for (int q = 0; q < nc1; ++q)
        b.icoeffs[(size_t)icontr * nc1 + q] = (int)g.i() - 1;
      b.icontr2icase[icontr] = c;
      b.icontr2icase[icontr + b.ncontr] = l;
      ++icontr;
    }
  }
  if (icontr != b.ncontr) fail(GAIN_ERR_ARG, "%s: wrong indexing", path);
  need_line(in, line, path);
  {
    Fields f(line, path);
    b.nrot = (int)f.i();
    b.nrotdeg = (int)f.i();
  }
  GAIN_CHECK_ARG(b.nrot > 0 && b.nrotdeg > 0, "%s: bad rotational sizes",
                 path);
  b.rot_k.assign((size_t)b.nrot * b.nrotdeg, -1);
  b.rot_tau.assign((size_t)b.nrot * b.nrotdeg, -1);
  for (;;) {
    need_line(in, line, path);
    if (trim(line) == "End Primitive basis set") break;
    Fields f(line, path);
    const int ic = (int)f.i() - 1, il = (int)f.i() - 1;
    f.i();  // j
    const int k = (int)f.i(), t = (int)f.i();
    GAIN_CHECK_ARG(ic >= 0 && ic < b.nrot && il >= 0 && il < b.nrotdeg,
                   "%s: rotational index out of range", path);
    b.rot_k[ic + (size_t)il * b.nrot] = k;
    b.rot_tau[ic + (size_t)il * b.nrot] = t;
  }
  need_line(in, line, path);
  if (trim(line) != "Start irreducible transformation")
    fail(GAIN_ERR_ARG, "%s: wrong sym-footer", path);
  need_line(in, line, path);
  b.matsize = (int)Fields(line, path).i();
  GAIN_CHECK_ARG(b.matsize >= 1, "%s: bad mat_size", path);
  need_line(in, line, path);
  {
    Fields f(line, path);
    b.ntot.resize(nrepres);
    for (int g = 0; g < nrepres; ++g) b.ntot[g] = (int)f.i();
  }
  b.sdeg.assign(sdeg, sdeg + nrepres);
  b.N.resize(nrepres);
  b.repres.resize(nrepres);
  for (int g = 0; g < nrepres; ++g) {
    b.N[g].resize(b.ncases);
    for (int c = 0; c < b.ncases; ++c) {
      need_line(in, line, path);
      b.N[g][c] = (int)Fields(line, path).i();
    }
    const size_t nt = b.ntot[g], sd = sdeg[g];
    b.repres[g].assign(nt * sd * b.matsize, 0.0);
    for (size_t t = 0; t < nt; ++t)
      for (size_t d = 0; d < sd; ++d) {
        need_line(in, line, path);
        Fields f(line, path);
        for (int i = 0; i < b.matsize; ++i)
          b.repres[g][t + d * nt + (size_t)i * nt * sd] = f.d();
      }
  }
  need_line(in, line, path);
  if (trim(line) != "End irreducible transformation")
    fail(GAIN_ERR_ARG, "%s: wrong irrep-footer", path);
}

// J=0 correlation and K-blocks (TroveBasisSet.cpp:294-681): each contracted
// function (case, lambda) of J is the J=0 function whose vibrational part
// (contractive_space(1:nclasses) and icoeffs(1:nclasses)) is the same; its
// K and tau come from the rotational level of the case and the rotational
// degeneracy index of the lambda.
void correlate(gain_trove_basis& b, const gain_trove_basis& j0) {
  const int nc1 = b.nclasses + 1;
  GAIN_CHECK_ARG(j0.nclasses == b.nclasses,
                 "J=0 basis has %d classes, J=%d has %d", j0.nclasses, b.J,
                 b.nclasses);
  // key of a J=0 function -> its contracted index
  std::map<std::vector<int>, int> vib_of;
  for (int c = 0; c < j0.ncases; ++c)
    for (int l = 0; l < j0.nlam[c]; ++l) {
      const int r = j0.case_first[c] + l;
      std::vector<int> key(2 * b.nclasses);
      for (int q = 0; q < b.nclasses; ++q) {
        key[q] = j0.cspace[(size_t)c * nc1 + 1 + q];
        key[b.nclasses + q] = j0.icoeffs[(size_t)r * nc1 + 1 + q];
      }
      vib_of.emplace(key, r);  // first match, as the reference's search
    }
  b.vib.resize(b.ncontr);
  b.tau.resize(b.ncontr);
  std::vector<int> kv(b.ncontr);
  for (int r = 0; r < b.ncontr; ++r) {
    const int c = b.icontr2icase[r];
    std::vector<int> key(2 * b.nclasses);
    for (int q = 0; q < b.nclasses; ++q) {
      key[q] = b.cspace[(size_t)c * nc1 + 1 + q];
      key[b.nclasses + q] = b.icoeffs[(size_t)r * nc1 + 1 + q];
    }
    auto it = vib_of.find(key);
    if (it == vib_of.end())
      fail(GAIN_ERR_ARG,
           "J=%d: function %d has no J=0 counterpart (checkpoints?)", b.J, r);
    b.vib[r] = it->second;
    const int ilevel = b.cspace[(size_t)c * nc1];
    const int ideg = b.icoeffs[(size_t)r * nc1];
    GAIN_CHECK_ARG(ilevel >= 0 && ilevel < b.nrot && ideg >= 0 &&
                       ideg < b.nrotdeg,
                   "J=%d: rotational index out of range", b.J);
    const int k = b.rot_k[ilevel + (size_t)ideg * b.nrot];
    const int t = b.rot_tau[ilevel + (size_t)ideg * b.nrot];
    GAIN_CHECK_ARG(k >= 0 && k <= b.J && t >= 0,
                   "J=%d: missing rotational quanta", b.J);
    kv[r] = k;
    b.tau[r] = (2 * k + t) & 1;
  }
  // K-blocks: runs of equal K, in increasing K (TroveBasisSet.cpp:659-681)
  b.kstart.assign(b.J + 1, 0);
  b.kblock.assign(b.J + 1, 0);
  for (int r = 0; r < b.ncontr; ++r) {
    if (r > 0 && kv[r] < kv[r - 1])
      fail(GAIN_ERR_ARG, "J=%d: functions are not ordered by K", b.J);
    b.kblock[kv[r]]++;
  }
  for (int k = 1; k <= b.J; ++k) b.kstart[k] = b.kstart[k - 1] + b.kblock[k - 1];
  // ijterms: prefix sums of N per irrep (TroveBasisSet.cpp:688-709)
  b.ijterms.assign((size_t)b.ncases * b.nrepres, 0);
  for (int g = 0; g < b.nrepres; ++g) {
    int acc = 0;
    for (int c = 0; c < b.ncases; ++c) {
      b.ijterms[c + (size_t)g * b.ncases] = acc;
      acc += b.N[g][c];
    }
  }
}

// States of one (J, gamma) descr file passing the reference's state filter
// (States.cpp:41-53: lower or upper window, isym_do).
struct StatesOut {
  int nroots = 0, dim_basis = 0;
  std::vector<int> irec, ilevel;
  std::vector<double> energy;
};

StatesOut parse_descr(const char* path, int gamma, const gain_params& p) {
  std::ifstream in(path);
  if (!in) fail(GAIN_ERR_ARG, "cannot open %s", path);
  std::string line;
  for (;;) {
    if (!next_line(in, line)) fail(GAIN_ERR_ARG, "malformed descr file %s", path);
    if (trim(line) == "Start Quantum numbers and energies") break;
  }
  StatesOut s;
  need_line(in, line, path);
  need_line(in, line, path);
  need_line(in, line, path);
  {
    Fields f(line, path);
    s.nroots = (int)f.i();
    s.dim_basis = (int)f.i();
  }
  bool complete = false;
  while (next_line(in, line)) {
    const std::string t = trim(line);
    if (t == "End Quantum numbers and energies") {
      complete = true;
      break;
    }
    if (t.empty()) continue;
    Fields f(line, path);
    const int irec = (int)f.i() - 1;
    const int ig = (int)f.i() - 1;
    if (ig != gamma)
      fail(GAIN_ERR_ARG, "%s: state of irrep %d in the file of irrep %d", path,
           ig + 1, gamma + 1);
    const int ilevel = (int)f.i() - 1;
    f.i();  // ideg
    const double e = f.d();
    const double eu = e - p.zpe;
    const bool lower = eu <= p.erange_lower[1] && eu >= p.erange_lower[0] &&
                       p.isym_do[gamma];
    const bool upper = eu <= p.erange_upper[1] && eu >= p.erange_upper[0] &&
                       p.isym_do[gamma];
    if (!(lower || upper)) continue;
    s.irec.push_back(irec);
    s.ilevel.push_back(ilevel);
    s.energy.push_back(e);
  }
  if (!complete)
    fail(GAIN_ERR_ARG, "%s: description file not complete (no footer)", path);
  return s;
}

void read_records(const char* path, const std::vector<int>& irec,
                  int rec_len, double* out) {
  File f(path, "rb");
  if (!f.f) fail(GAIN_ERR_ARG, "cannot open %s", path);
  for (size_t i = 0; i < irec.size(); ++i) {
    const int64_t off = (int64_t)irec[i] * rec_len * (int64_t)sizeof(double);
    if (fseeko(f.f, (off_t)off, SEEK_SET) != 0 ||
        fread(out + i * (size_t)rec_len, sizeof(double), (size_t)rec_len,
              f.f) != (size_t)rec_len)
      fail(GAIN_ERR_ARG, "%s: record %d out of the file", path, irec[i] + 1);
  }
}

}  // namespace

// ---- C ABI ----------------------------------------------------------------
extern "C" int gain_trove_read_dipole(const char* path, int* nc,
                                      double* piece) {
  return guarded([&] {
    GAIN_CHECK_ARG(path && nc, "null argument");
    File f(path, "rb");
    if (!f.f) fail(GAIN_ERR_ARG, "checkpoint file %s not found", path);
    std::vector<char> r = read_record(f.f, path);
    if (r.size() < 20 || memcmp(r.data(), "Start external field", 20) != 0)
      fail(GAIN_ERR_ARG, "%s: bogus header", path);
    r = read_record(f.f, path);
    GAIN_CHECK_ARG(r.size() >= 4, "%s: bad ncontr record", path);
    int32_t n = 0;
    memcpy(&n, r.data(), 4);
    GAIN_CHECK_ARG(n > 0, "%s: ncontr = %d", path, n);
    *nc = n;
    if (!piece) return;
    const size_t n2 = (size_t)n * n;
    for (int a = 0; a < 3; ++a) {
      r = read_record(f.f, path);
      int32_t imu = 0;
      if (r.size() >= 4) memcpy(&imu, r.data(), 4);
      if (imu != a + 1)
        fail(GAIN_ERR_ARG, "%s: bogus imu %d /= %d", path, imu, a + 1);
      r = read_record(f.f, path);
      if (r.size() != n2 * sizeof(double))
        fail(GAIN_ERR_ARG, "%s: component %d has %zu bytes, expected %zu",
             path, a + 1, r.size(), n2 * sizeof(double));
      const double* m = reinterpret_cast<const double*>(r.data());
      // file: m[i + f*n] (column-major); piece: [f + i*n + a*n*n]
      // (TroveDipole.cpp:110, one piece covering every f)
      double* dst = piece + (size_t)a * n2;
      for (size_t i = 0; i < (size_t)n; ++i)
        for (size_t fF = 0; fF < (size_t)n; ++fF)
          dst[fF + i * n] = m[i + fF * n];
    }
    r = read_record(f.f, path);
    if (r.size() < 18 || memcmp(r.data(), "End external field", 18) != 0)
      fail(GAIN_ERR_ARG, "%s: bogus footer", path);
  });
}

extern "C" int gain_trove_read_basis(const char* path, const char* path_j0,
                                     int J, int nrepres, const int* sdeg,
                                     gain_trove_basis** out) {
  return guarded([&] {
    GAIN_CHECK_ARG(path && path_j0 && sdeg && out, "null argument");
    GAIN_CHECK_ARG(J >= 0 && nrepres > 0 && nrepres <= GAIN_MAX_IRREPS,
                   "bad J or irrep count");
    *out = nullptr;
    auto b = std::make_unique<gain_trove_basis>();
    b->J = J;
    b->nrepres = nrepres;
    parse_quanta(path, nrepres, sdeg, *b);
    gain_trove_basis j0;
    j0.J = 0;
    j0.nrepres = nrepres;
    parse_quanta(path_j0, nrepres, sdeg, j0);
    correlate(*b, j0);
    *out = b.release();
  });
}

extern "C" int gain_trove_basis_free(gain_trove_basis* b) {
  delete b;
  return GAIN_OK;
}

extern "C" int gain_trove_basis_sizes(const gain_trove_basis* b, int* nprim,
                                      int* maxsymcoeffs, int* matsize,
                                      int* ntot) {
  return guarded([&] {
    GAIN_CHECK_ARG(b, "null basis");
    if (nprim) *nprim = b->ncontr;
    if (maxsymcoeffs) *maxsymcoeffs = b->ncases;
    if (matsize) *matsize = b->matsize;
    if (ntot)
      for (int g = 0; g < b->nrepres; ++g) ntot[g] = b->ntot[g];
  });
}

extern "C" int gain_trove_basis_get(const gain_trove_basis* b,
                                    const char* name, int gamma, void* dst) {
  return guarded([&] {
    GAIN_CHECK_ARG(b && name && dst, "null argument");
    const std::string k(name);
    // (an empty array, e.g. the repres of an irrep with no functions, has
    // no data pointer: copy nothing)
    auto ints = [&](const std::vector<int>& v) {
      if (!v.empty()) memcpy(dst, v.data(), v.size() * sizeof(int));
    };
    if (k == "kstart") {
      ints(b->kstart);
    } else if (k == "kblock") {
      ints(b->kblock);
    } else if (k == "tau") {
      ints(b->tau);
    } else if (k == "vib") {
      ints(b->vib);
    } else if (k == "ijterms") {
      ints(b->ijterms);
    } else if (k == "icontr") {
      ints(b->icontr2icase);
    } else if (k == "N" || k == "repres") {
      GAIN_CHECK_ARG(gamma >= 0 && gamma < b->nrepres, "irrep out of range");
      if (k == "N")
        ints(b->N[gamma]);
      else if (!b->repres[gamma].empty())
        memcpy(dst, b->repres[gamma].data(),
               b->repres[gamma].size() * sizeof(double));
    } else {
      fail(GAIN_ERR_ARG, "unknown basis array %s", name);
    }
  });
}

extern "C" int gain_trove_upload_basis(gain_ctx* c, int jind,
                                       const gain_trove_basis* b) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(b, "null basis");
    GAIN_CHECK_ARG(b->nrepres == c->nrepres, "basis has %d irreps, context %d",
                   b->nrepres, c->nrepres);
    int st = gain_upload_basis(c, jind, b->J, b->ncontr, b->kstart.data(),
                               b->kblock.data(), b->tau.data(), b->vib.data());
    if (st != GAIN_OK) {
      const std::string m = gain_last_error();
      fail(st, "%s", m.c_str());
    }
    std::vector<const int*> N(b->nrepres);
    std::vector<const double*> R(b->nrepres);
    for (int g = 0; g < b->nrepres; ++g) {
      N[g] = b->N[g].data();
      R[g] = b->repres[g].data();
    }
    st = gain_upload_inflation(c, jind, b->ncases, b->ijterms.data(),
                               b->icontr2icase.data(), b->matsize,
                               b->ntot.data(), b->sdeg.data(), N.data(),
                               R.data());
    if (st != GAIN_OK) {
      const std::string m = gain_last_error();
      fail(st, "%s", m.c_str());
    }
  });
}

extern "C" int gain_trove_read_states(const char* path, int gamma,
                                      const gain_params* p, int* nstates,
                                      int* rec_len, int* irec, int* ilevel,
                                      double* energy) {
  return guarded([&] {
    GAIN_CHECK_ARG(path && p && nstates, "null argument");
    GAIN_CHECK_ARG(gamma >= 0 && gamma < p->nrepres, "irrep out of range");
    StatesOut s = parse_descr(path, gamma, *p);
    *nstates = (int)s.irec.size();
    if (rec_len) *rec_len = s.dim_basis;
    if (s.irec.empty()) return;  // no state passes: nothing to copy
    if (irec) memcpy(irec, s.irec.data(), s.irec.size() * sizeof(int));
    if (ilevel) memcpy(ilevel, s.ilevel.data(), s.ilevel.size() * sizeof(int));
    if (energy)
      memcpy(energy, s.energy.data(), s.energy.size() * sizeof(double));
  });
}

extern "C" int gain_trove_read_records(const char* path, int n,
                                       const int* irec, int rec_len,
                                       double* out) {
  return guarded([&] {
    GAIN_CHECK_ARG(path && (n == 0 || (irec && out)) && n >= 0 &&
                       rec_len > 0,
                   "bad arguments");
    read_records(path, std::vector<int>(irec, irec + n), rec_len, out);
  });
}

extern "C" int gain_set_block_levels(gain_ctx* c, int jind, int gamma,
                                     const int* levels) {
  return guarded([&] {
    need(c);
    auto it = c->blocks.find({jind, gamma});
    if (it == c->blocks.end())
      fail(GAIN_ERR_STATE, "block (%d, %d) not uploaded", jind, gamma);
    if (levels)
      it->second.label.assign(levels, levels + it->second.n);
    else
      it->second.label.clear();
  });
}

extern "C" int gain_trove_load(gain_ctx* c, const char* dir,
                               const int* jvals, int nj) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(dir && jvals && nj > 0, "bad arguments");
    if (!c->params_set) fail(GAIN_ERR_STATE, "gain_set_params not called");
    const gain_params& p = c->params;
    const std::string d(dir);
    auto path = [&](const std::string& name) { return d + "/" + name; };
    auto call = [&](int st) {
      if (st != GAIN_OK) {
        const std::string m = gain_last_error();
        fail(st, "%s", m.c_str());
      }
    };
    // dipole (TroveDipole.cpp:9-119)
    const std::string dip = path("j0_extfield.chk");
    int nc = 0;
    call(gain_trove_read_dipole(dip.c_str(), &nc, nullptr));
    std::vector<double> mu((size_t)3 * nc * nc);
    call(gain_trove_read_dipole(dip.c_str(), &nc, mu.data()));
    call(gain_upload_dipole(c, mu.data(), nc, 0, nc));
    mu.clear();
    mu.shrink_to_fit();
    const std::string q0 = path("j0eigen_quanta0.chk");
    for (int j = 0; j < nj; ++j) {
      const int J = jvals[j];
      char name[128];
      snprintf(name, sizeof(name), "j0eigen_quanta%d.chk", J);
      const std::string q = path(name);
      gain_trove_basis* b = nullptr;
      call(gain_trove_read_basis(q.c_str(), q0.c_str(), J, p.nrepres,
                                 p.sym_degen, &b));
      std::unique_ptr<gain_trove_basis> hold(b);
      call(gain_trove_upload_basis(c, j, b));
      for (int g = 0; g < p.nrepres; ++g) {
        if (!p.isym_do[g]) continue;
        snprintf(name, sizeof(name), "j0eigen_descr%d_%d.chk", J, g + 1);
        const std::string ds = path(name);
        StatesOut s = parse_descr(ds.c_str(), g, p);
        if (s.irec.empty()) continue;
        GAIN_CHECK_ARG(s.dim_basis == b->ntot[g],
                       "%s: record length %d, basis has %d symmetrised "
                       "functions",
                       ds.c_str(), s.dim_basis, b->ntot[g]);
        snprintf(name, sizeof(name), "j0eigen_vectors%d_%d.chk", J, g + 1);
        const std::string vs = path(name);
        std::vector<double> rec(s.irec.size() * (size_t)s.dim_basis);
        read_records(vs.c_str(), s.irec, s.dim_basis, rec.data());
        call(gain_upload_block(c, j, g, (int)s.irec.size(), s.dim_basis,
                               s.energy.data(), rec.data()));
        call(gain_set_block_levels(c, j, g, s.ilevel.data()));
      }
    }
  });
}

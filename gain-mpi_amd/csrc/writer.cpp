// Transition list writer: the loop of main.cpp:238-338 over every uploaded
// (J, gamma) block, with the line format of Output::OutputLinestrength
// (Output.cpp:52-112).  Host code; the line strengths come from the batched
// engine (gain_batched_lines).
//
// Order of the output: initial states by energy, and for each initial state
// its final states by energy -- the energy-sorted state list
// (TroveStates.cpp:162) walked as main.cpp walks it.  The reference prints
// the finals of one initial state from several OpenMP threads, so its order
// inside one initial state is not fixed; this one is.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "engine_internal.hpp"
#include "kernels.hpp"

namespace {

struct StateRef {
  double e;
  int jind, g, level;
};

// One emitted line of one initial state before formatting.
struct Line {
  int64_t final_rank;  // position of the final state in the sorted list
  int jindF, gF, levelF;
  double nu, A, I;
  double ls[9];
  int nls;
};

int format_line(char* buf, int len, double nu, int levelF, int jF, int gF,
                int levelI, int jI, int gI, double A, bool with_I, double I,
                const double* ls, int nls) {
  // "%12.6f %8d %4d %4d <- %8d %4d %4d %16.9E " (Output.cpp:99), the
  // intensity "%16.9E " (:103), the components "[ %s ] " with " %16.9E"
  // each (:64,106), then " || \n" (:110)
  int n = snprintf(buf, len, "%12.6f %8d %4d %4d <- %8d %4d %4d %16.9E ", nu,
                   levelF + 1, jF, gF + 1, levelI + 1, jI, gI + 1, A);
  auto more = [&](int k) {
    if (k < 0 || n + k >= len) fail(GAIN_ERR_ARG, "line buffer too small");
    n += k;
  };
  if (n < 0 || n >= len) fail(GAIN_ERR_ARG, "line buffer too small");
  if (with_I) more(snprintf(buf + n, len - n, "%16.9E ", I));
  if (nls >= 0) {
    more(snprintf(buf + n, len - n, "[ "));
    for (int q = 0; q < nls; ++q)
      more(snprintf(buf + n, len - n, " %16.9E", ls[q]));
    more(snprintf(buf + n, len - n, " ] "));
  }
  more(snprintf(buf + n, len - n, " || \n"));
  return n;
}

}  // namespace

extern "C" int gain_format_line(char* buf, int len, double nu, int levelF,
                                int jF, int gammaF, int levelI, int jI,
                                int gammaI, double A, int with_intensity,
                                double intensity, const double* ls, int nls) {
  return guarded([&] {
    GAIN_CHECK_ARG(buf && len > 0, "bad buffer");
    GAIN_CHECK_ARG(nls >= -1 && nls <= 9 && (nls <= 0 || ls),
                   "bad line strength components");
    format_line(buf, len, nu, levelF, jF, gammaF, levelI, jI, gammaI, A,
                with_intensity != 0, intensity, ls, nls);
  });
}

extern "C" int gain_write_transitions(gain_ctx* c,
                                      const gain_output_opts* opts,
                                      const char* path,
                                      int64_t* nlines_out) {
  FILE* out = nullptr;
  const int rc = guarded([&] {
    need(c);
    if (!c->params_set) fail(GAIN_ERR_STATE, "gain_set_params not called");
    if (!c->d_mu) fail(GAIN_ERR_STATE, "dipole not uploaded");
    gain_output_opts o{};
    if (opts) o = *opts;
    const int B = o.batch > 0 ? o.batch : 64;
    GAIN_CHECK_ARG(o.precision >= GAIN_PREC_FP64 &&
                       o.precision <= GAIN_PREC_I8CRT14,
                   "unknown precision %d", o.precision);
    const gain_params& p = c->params;
    const int md = c->maxdeg;

    // energy-sorted state list of every uploaded block
    std::vector<StateRef> states;
    for (const auto& kv : c->blocks)
      for (int s = 0; s < kv.second.n; ++s)
        states.push_back({kv.second.energy[s], kv.first.first,
                          kv.first.second, s});
    std::stable_sort(states.begin(), states.end(),
                     [](const StateRef& a, const StateRef& b) {
                       return a.e < b.e;
                     });
    std::map<std::pair<int, int>, std::vector<int64_t>> rank_of;
    for (const auto& kv : c->blocks)
      rank_of[kv.first].assign(kv.second.n, -1);
    for (size_t r = 0; r < states.size(); ++r)
      rank_of[{states[r].jind, states[r].g}][states[r].level] = (int64_t)r;

    // final blocks reachable from each initial block
    auto finals_of = [&](int jindI, int gI) {
      std::vector<std::pair<int, int>> f;
      const int jI = c->bases[jindI].J;
      for (const auto& kv : c->blocks) {
        const int jindF = kv.first.first, gF = kv.first.second;
        if (block_pair_ok(c, jI, gI, c->bases[jindF].J, gF))
          f.push_back(kv.first);
      }
      return f;
    };

    // initial states in list order (FilterLowerState, States.cpp:46-53)
    std::vector<size_t> initials;
    for (size_t r = 0; r < states.size(); ++r) {
      const StateRef& s = states[r];
      if (filter_lower(c, s.e, s.g) && !finals_of(s.jind, s.g).empty())
        initials.push_back(r);
    }

    if (path) {
      out = fopen(path, "wb");
      if (!out) fail(GAIN_ERR_ARG, "cannot open %s", path);
    } else {
      out = stdout;
    }
    int64_t nlines = 0;
    std::vector<double> hls, hA, hI;
    std::vector<uint8_t> hfl;
    std::vector<char> text(4096);
    for (size_t w0 = 0; w0 < initials.size(); w0 += B) {
      const size_t w1 = std::min(initials.size(), w0 + (size_t)B);
      std::vector<std::vector<Line>> lines(w1 - w0);
      // the window's initial states, grouped by block
      std::map<std::pair<int, int>, std::vector<std::pair<int, size_t>>> grp;
      for (size_t w = w0; w < w1; ++w) {
        const StateRef& s = states[initials[w]];
        grp[{s.jind, s.g}].push_back({s.level, w - w0});
      }
      for (const auto& gkv : grp) {
        const int jindI = gkv.first.first, gI = gkv.first.second;
        std::vector<int> levels;
        for (const auto& lw : gkv.second) levels.push_back(lw.first);
        const int nI = (int)levels.size();
        const Block& blkI = c->blocks.at(gkv.first);
        const int ndegI = p.sym_degen[gI];
        for (const auto& fb : finals_of(jindI, gI)) {
          const int jindF = fb.first, gF = fb.second;
          const Block& blkF = c->blocks.at(fb);
          if (!blkI.vec || !blkF.vec)
            fail(GAIN_ERR_STATE,
                 "block (J index %d, irrep %d) has no eigenvectors",
                 blkI.vec ? jindF : jindI, blkI.vec ? gF : gI);
          const int nF = blkF.n;
          const size_t np = (size_t)nI * nF;
          hA.resize(np);
          hI.resize(np);
          hfl.resize(np);
          if (o.full_line) hls.resize(np * md * md);
          gain_lines res{};
          res.ls = o.full_line ? hls.data() : nullptr;
          res.A = hA.data();
          res.I = hI.data();
          res.flags = hfl.data();
          res.on_device = 0;
          const int st = gain_batched_lines(c, jindI, gI, levels.data(), nI,
                                            jindF, gF, o.precision, &res);
          if (st != GAIN_OK) {
            const std::string msg = gain_last_error();
            fail(st, "%s", msg.c_str());
          }
          const int ndegF = p.sym_degen[gF];
          const auto& rk = rank_of.at(fb);
          for (int i = 0; i < nI; ++i) {
            auto& dst = lines[gkv.second[i].second];
            const double eI = blkI.energy[levels[i]];
            for (int F = 0; F < nF; ++F) {
              const size_t q = (size_t)i * nF + F;
              if (!(hfl[q] & 2)) continue;
              Line L{};
              L.final_rank = rk[F];
              L.jindF = jindF;
              L.gF = gF;
              L.levelF = F;
              L.nu = blkF.energy[F] - eI;
              L.A = hA[q];
              L.I = hI[q];
              if (o.full_line)
                for (int dF = 0; dF < ndegF; ++dF)
                  for (int dI = 0; dI < ndegI; ++dI)
                    if (degfilter(c, gI, gF, dI, dF))
                      L.ls[L.nls++] = hls[q * md * md + dI + dF * md];
              dst.push_back(L);
            }
          }
        }
      }
      for (size_t w = w0; w < w1; ++w) {
        const StateRef& s = states[initials[w]];
        auto& v = lines[w - w0];
        std::sort(v.begin(), v.end(), [](const Line& a, const Line& b) {
          return a.final_rank < b.final_rank;
        });
        const Block& bI = c->blocks.at({s.jind, s.g});
        const int lvI = bI.label.empty() ? s.level : bI.label[s.level];
        for (const Line& L : v) {
          const Block& bF = c->blocks.at({L.jindF, L.gF});
          const int lvF = bF.label.empty() ? L.levelF : bF.label[L.levelF];
          format_line(text.data(), (int)text.size(), L.nu, lvF,
                      c->bases[L.jindF].J, L.gF, lvI, c->bases[s.jind].J,
                      s.g, L.A, o.compute_intensity != 0, L.I, L.ls,
                      o.full_line ? L.nls : -1);
          if (fputs(text.data(), out) < 0)
            fail(GAIN_ERR_ARG, "write failed");
          ++nlines;
        }
      }
    }
    if (nlines_out) *nlines_out = nlines;
  });
  if (out && out != stdout) fclose(out);
  if (out == stdout) fflush(stdout);
  return rc;
}

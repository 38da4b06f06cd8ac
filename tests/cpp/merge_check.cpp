// CPU check of the sweep's op-merging rules (quaru_amd/csrc/qs_sweep_merge.h).
//
// Random runs of gates of every kind a sweep holds are applied to a random
// state one by one, and again as the merged op list that the sweep encodes
// (multiplexed 2x2 ops in place of the gates they absorbed).  Both must agree
// to rounding.  Prints "ok <trials> <merged>" and exits 0, or prints the
// failing trial and exits 1.
#include <complex>
#include <cstdio>
#include <random>
#include <vector>

#include "qs_sweep_merge.h"

using C = std::complex<double>;
using qs::Gate;
using qs::cd;

static C c_of(const cd &v) { return C(v.re, v.im); }
static cd cd_of(const C &v) { return cd{v.real(), v.imag()}; }

static void pair_apply(std::vector<C> &psi, uint64_t i0, uint64_t i1, const cd *m) {
  const C x = psi[i0], y = psi[i1];
  psi[i0] = c_of(m[0]) * x + c_of(m[1]) * y;
  psi[i1] = c_of(m[2]) * x + c_of(m[3]) * y;
}

static void apply_gate(std::vector<C> &psi, int n, const Gate &g) {
  const uint64_t N = 1ull << n;
  if (g.kind == Gate::kDense && g.na == 1) {
    const uint64_t bt = 1ull << g.act[0];
    for (uint64_t i = 0; i < N; ++i)
      if (!(i & bt) && (i & g.cmask) == g.cval) pair_apply(psi, i, i | bt, g.m.data());
  } else if (g.kind == Gate::kDense && g.na == 2) {
    const uint64_t b0 = 1ull << g.act[0], b1 = 1ull << g.act[1];
    for (uint64_t i = 0; i < N; ++i) {
      if ((i & b0) || (i & b1) || (i & g.cmask) != g.cval) continue;
      const uint64_t idx[4] = {i, i | b0, i | b1, i | b0 | b1};
      C x[4], y[4];
      for (int k = 0; k < 4; ++k) x[k] = psi[idx[k]];
      for (int r = 0; r < 4; ++r) {
        y[r] = 0;
        for (int k = 0; k < 4; ++k) y[r] += c_of(g.m[4 * r + k]) * x[k];
      }
      for (int k = 0; k < 4; ++k) psi[idx[k]] = y[k];
    }
  } else if (g.kind == Gate::kDiag) {
    for (uint64_t i = 0; i < N; ++i) {
      if ((i & g.cmask) != g.cval) continue;
      int e = 0;
      for (int j = 0; j < g.na; ++j) e |= (int)((i >> g.act[j]) & 1) << j;
      psi[i] *= c_of(g.m[e]);
    }
  } else if (g.kind == Gate::kSparse) {
    for (uint64_t i = 0; i < N; ++i)
      for (size_t e = 0; e < g.spat.size(); ++e)
        if ((i & g.smask) == g.spat[e]) psi[i] *= c_of(g.m[e]);
  }
}

static void apply_mux(std::vector<C> &psi, int n, const qs::merge::SweepOp &o) {
  const uint64_t N = 1ull << n, bt = 1ull << o.t;
  for (uint64_t i = 0; i < N; ++i) {
    if (i & bt) continue;
    const bool on = (i & o.pm) == o.pv;
    pair_apply(psi, i, i | bt, on ? o.m1 : o.m0);
  }
}

static void unitary2(std::mt19937_64 &rng, cd *m) {
  // Haar-ish 2x2: Gram-Schmidt of a complex Gaussian matrix
  std::normal_distribution<double> nd;
  C a(nd(rng), nd(rng)), b(nd(rng), nd(rng)), c(nd(rng), nd(rng)), d(nd(rng), nd(rng));
  double na = std::sqrt(std::norm(a) + std::norm(c));
  a /= na;
  c /= na;
  C p = std::conj(a) * b + std::conj(c) * d;
  b -= p * a;
  d -= p * c;
  double nb = std::sqrt(std::norm(b) + std::norm(d));
  b /= nb;
  d /= nb;
  m[0] = cd_of(a);
  m[1] = cd_of(b);
  m[2] = cd_of(c);
  m[3] = cd_of(d);
}

// Applies `gates` one by one and as the merged op list; returns the largest
// amplitude difference, or -1 when the merged list is not encodable.
static double run_case(const std::vector<Gate> &gates, int n, const int *slot, uint64_t regq,
                       const std::vector<C> &psi, long *merged_count, int trial) {
  std::vector<C> ref = psi, got = psi;
  for (const Gate &g : gates) apply_gate(ref, n, g);
  std::vector<qs::merge::SweepOp> items(gates.size());
  for (size_t k = 0; k < gates.size(); ++k) items[k].g = &gates[k];
  qs::merge::merge_sweep_ops(items, slot, regq);
  for (const auto &o : items) {
    if (!o.alive) {
      ++*merged_count;
      continue;
    }
    if (o.g) {
      apply_gate(got, n, *o.g);
    } else {
      if (o.t < 0 || slot[o.t] < 0) {
        std::printf("trial %d: merged op on qubit %d, which is not in the tile\n", trial, o.t);
        return -1;
      }
      if (o.pm != 0 && !qs::merge::mux_encodable(o.t, o.pm, slot, regq)) {
        std::printf("trial %d: merged op has no kernel form (t=%d pm=%llx)\n", trial, o.t,
                    (unsigned long long)o.pm);
        return -1;
      }
      if (o.pm & (1ull << o.t)) {
        std::printf("trial %d: pattern holds the target\n", trial);
        return -1;
      }
      apply_mux(got, n, o);
    }
  }
  double d = 0;
  for (size_t i = 0; i < psi.size(); ++i) d = std::max(d, std::abs(ref[i] - got[i]));
  return d;
}

static std::vector<C> random_state(std::mt19937_64 &rng, int n) {
  std::normal_distribution<double> nd;
  std::vector<C> psi(1ull << n);
  double nrm = 0;
  for (C &v : psi) {
    v = C(nd(rng), nd(rng));
    nrm += std::norm(v);
  }
  for (C &v : psi) v /= std::sqrt(nrm);
  return psi;
}

static Gate dense1(int t, const cd *m, uint64_t cmask = 0, uint64_t cval = 0) {
  Gate g;
  g.kind = Gate::kDense;
  g.na = 1;
  g.act[0] = t;
  g.m.assign(m, m + 4);
  g.cmask = cmask;
  g.cval = cval;
  return g;
}

static Gate diag1(int t, cd a, cd b) {
  Gate g;
  g.kind = Gate::kDiag;
  g.na = 1;
  g.act[0] = t;
  g.m = {a, b};
  return g;
}

// Fixed sequences that once merged wrongly.  U(3); CX(4->3); RZ(4); H(4);
// CX(4->3): RZ and H merge into one plain op on 4 that is no longer diagonal,
// so the second CX must not fold back into the first one across it.  The
// same with the plain op on 4 built from two diagonal 2x2s and then a dense
// one (the merged-plain branch), and with a CZ-type phase pattern.
static int fixed_cases(std::mt19937_64 &rng) {
  const int n = 9;
  const cd X[4] = {cd{0, 0}, cd{1, 0}, cd{1, 0}, cd{0, 0}};
  const double r = 0.7071067811865476;
  const cd H[4] = {cd{r, 0}, cd{r, 0}, cd{r, 0}, cd{-r, 0}};
  cd U[4];
  unitary2(rng, U);
  const cd e0 = cd_of(std::polar(1.0, -0.4)), e1 = cd_of(std::polar(1.0, 0.4));
  std::vector<std::vector<Gate>> cases;
  cases.push_back({dense1(3, U), dense1(3, X, 1ull << 4, 1ull << 4), diag1(4, e0, e1),
                   dense1(4, H), dense1(3, X, 1ull << 4, 1ull << 4)});
  cases.push_back({dense1(3, U), dense1(3, X, 1ull << 4, 1ull << 4), diag1(4, e0, e1),
                   diag1(4, e1, e0), dense1(4, H), dense1(3, X, 1ull << 4, 1ull << 4)});
  cases.push_back({dense1(3, U), dense1(3, X, 1ull << 4, 0), diag1(4, e0, e1), dense1(4, U),
                   dense1(3, X, 1ull << 4, 0)});
  // layouts: both on register slots, the target on a lane bit, both on lanes
  const int layouts[3][2] = {{0, 1}, {qs::merge::kFirstLaneSlot, 0},
                             {qs::merge::kFirstLaneSlot, qs::merge::kFirstLaneSlot + 1}};
  int failures = 0;
  for (size_t c = 0; c < cases.size(); ++c)
    for (const auto &lay : layouts) {
      int slot[64];
      for (int q = 0; q < 64; ++q) slot[q] = -1;
      slot[3] = lay[0];
      slot[4] = lay[1];
      uint64_t regq = 0;
      for (int q : {3, 4})
        if (slot[q] < qs::merge::kFirstLaneSlot) regq |= 1ull << q;
      long merged = 0;
      const double d = run_case(cases[c], n, slot, regq, random_state(rng, n), &merged, -1);
      if (d < 0 || d > 1e-12) {
        std::printf("fixed case %zu (slots %d,%d): max diff %.3e\n", c, lay[0], lay[1], d);
        ++failures;
      }
    }
  return failures;
}

int main() {
  const int n = 9;  // qubits 0..7 in the tile, qubit 8 outside it
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> ud(0.0, 6.283185307179586);
  if (fixed_cases(rng)) return 1;
  long merged_total = 0;
  const int trials = 4000;
  for (int trial = 0; trial < trials; ++trial) {
    // slot table: 3 register slots, 5 lane bits, qubit 8 outside
    int perm[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    std::shuffle(perm, perm + 8, rng);
    int slot[64];
    for (int q = 0; q < 64; ++q) slot[q] = -1;
    uint64_t regq = 0;
    for (int k = 0; k < 8; ++k) {
      slot[perm[k]] = k < 3 ? k : qs::merge::kFirstLaneSlot + (k - 3);
      if (k < 3) regq |= 1ull << perm[k];
    }
    const int ngates = 4 + (int)(rng() % 40);
    // Half the trials draw every qubit from 4 of the 9 and every pattern
    // value from one fixed choice, so that gates meet on the same qubits
    // with the same patterns (the cases where merges chain).
    const bool dense_trial = trial & 1;
    int few[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
    std::shuffle(few, few + 9, rng);
    const uint64_t fixed_val = rng();
    std::vector<Gate> gates(ngates);
    for (Gate &g : gates) {
      int q[4];
      {
        int all[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
        const int pool = dense_trial ? 4 : 9;
        if (dense_trial)
          for (int j = 0; j < 9; ++j) all[j] = few[j];
        std::shuffle(all, all + pool, rng);
        for (int k = 0; k < 4; ++k) q[k] = all[k];
      }
      // dense targets must be tile qubits
      auto tileq = [&]() {
        if (dense_trial)
          for (int k = 0; k < 4; ++k)
            if (slot[q[k]] >= 0 && (rng() & 1)) return q[k];
        return perm[rng() % 8];
      };
      const int kind = (int)(rng() % 10);
      auto ctl = [&](Gate &gg, int t, int nc) {
        gg.cmask = gg.cval = 0;
        for (int c = 0, k = 0; c < nc && k < 4; ++k) {
          if (q[k] == t) continue;
          gg.cmask |= 1ull << q[k];
          if (dense_trial ? ((fixed_val >> q[k]) & 1) : (rng() & 1)) gg.cval |= 1ull << q[k];
          ++c;
        }
      };
      if (kind <= 2 || kind == 3) {  // dense 1q, plain or controlled (X or U)
        g.kind = Gate::kDense;
        g.na = 1;
        g.act[0] = tileq();
        g.m.resize(4);
        if (kind == 3 && (rng() & 1)) {
          g.m = {cd{0, 0}, cd{1, 0}, cd{1, 0}, cd{0, 0}};
        } else {
          unitary2(rng, g.m.data());
        }
        if (kind >= 2) ctl(g, g.act[0], 1 + (int)(rng() % 2));
      } else if (kind == 4) {  // diag 1q, plain or controlled
        g.kind = Gate::kDiag;
        g.na = 1;
        g.act[0] = q[0];
        g.m = {cd_of(std::polar(1.0, ud(rng))), cd_of(std::polar(1.0, ud(rng)))};
        if (rng() & 1) ctl(g, q[0], 1);
      } else if (kind == 5 || kind == 6) {  // phase on one pattern (CZ, CCZ)
        g.kind = Gate::kSparse;
        const int w = 2 + (int)(rng() % 2);
        g.smask = 0;
        uint64_t pat = 0;
        for (int k = 0; k < w; ++k) {
          g.smask |= 1ull << q[k];
          if (kind == 5 || (dense_trial ? ((fixed_val >> q[k]) & 1) : (rng() & 1)))
            pat |= 1ull << q[k];
        }
        g.spat = {pat};
        g.m = {(rng() & 1) ? cd{-1, 0} : cd_of(std::polar(1.0, ud(rng)))};
      } else if (kind == 7) {  // phase on two patterns: not mergeable
        g.kind = Gate::kSparse;
        g.smask = (1ull << q[0]) | (1ull << q[1]);
        g.spat = {1ull << q[0], g.smask};
        g.m = {cd{0, 1}, cd{-1, 0}};
      } else if (kind == 8) {  // dense 2q on tile qubits
        g.kind = Gate::kDense;
        g.na = 2;
        g.act[0] = tileq();
        do g.act[1] = tileq();
        while (g.act[1] == g.act[0]);
        g.m.assign(16, cd{0, 0});
        // a random permutation with phases: unitary and non-diagonal
        int p[4] = {0, 1, 2, 3};
        std::shuffle(p, p + 4, rng);
        for (int r = 0; r < 4; ++r) g.m[4 * r + p[r]] = cd_of(std::polar(1.0, ud(rng)));
      } else {  // diag 2q
        g.kind = Gate::kDiag;
        g.na = 2;
        g.act[0] = q[0];
        g.act[1] = q[1];
        for (int k = 0; k < 4; ++k) g.m.push_back(cd_of(std::polar(1.0, ud(rng))));
      }
    }
    // Dense trials follow some plain diagonal 2x2s with a dense 2x2 on the
    // same qubit: the two merge into one plain op that is no longer diagonal.
    if (dense_trial)
      for (size_t k = 0; k + 1 < gates.size(); ++k) {
        const Gate &g = gates[k];
        if (g.kind == Gate::kDiag && g.na == 1 && g.cmask == 0 && slot[g.act[0]] >= 0 &&
            (rng() & 1)) {
          cd m[4];
          unitary2(rng, m);
          gates[k + 1] = dense1(g.act[0], m);
        }
      }
    // Every fourth trial instead draws from a small menu of gates on three
    // tile qubits a, b, c with fixed patterns: chains of plain ops on a
    // pattern qubit between two ops of the same pattern come up often.
    if (trial % 4 == 3) {
      const int a = perm[0], b = perm[1], c = perm[2];
      const cd X[4] = {cd{0, 0}, cd{1, 0}, cd{1, 0}, cd{0, 0}};
      const uint64_t vb = (fixed_val & 1) << b, vc = ((fixed_val >> 1) & 1) << c;
      for (Gate &g : gates) {
        cd m[4];
        unitary2(rng, m);
        const int pick = (int)(rng() % 7);
        const int qa = (rng() & 1) ? a : b;
        switch (pick) {
          case 0: g = dense1(qa, m); break;
          case 1: {
            const int abc[3] = {a, b, c};
            g = diag1(abc[rng() % 3], cd_of(std::polar(1.0, ud(rng))),
                      cd_of(std::polar(1.0, ud(rng))));
            break;
          }
          case 2: g = dense1(a, X, 1ull << b, vb); break;
          case 3: g = dense1(b, m, 1ull << c, vc); break;
          case 4:
            g = Gate();
            g.kind = Gate::kSparse;
            g.smask = (1ull << a) | (1ull << b);
            g.spat = {g.smask};
            g.m = {cd{-1, 0}};
            break;
          case 5: g = dense1(c, m); break;
          default: g = dense1(a, X, (1ull << b) | (1ull << c), vb | vc); break;
        }
      }
    }
    const std::vector<C> psi = random_state(rng, n);
    const double d = run_case(gates, n, slot, regq, psi, &merged_total, trial);
    if (d < 0) return 1;
    if (d > 1e-12) {
      std::printf("trial %d: max diff %.3e over %d gates\n", trial, d, ngates);
      return 1;
    }
  }
  std::printf("ok %d %ld\n", trials, merged_total);
  return 0;
}

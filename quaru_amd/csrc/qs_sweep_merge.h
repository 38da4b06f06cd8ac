// qs_sweep_merge.h -- host side of the fused sweep: merging one-qubit ops of
// a sweep into multiplexed 2x2 ops.  Host code only (no kernels); included by
// qs_sweep.hip, and by the CPU check of the merge rules in tests/cpp.
#pragma once

#include <cstdint>
#include <deque>
#include <vector>

#include "qs_internal.h"

namespace qs {
namespace merge {

// Register slots are 0..kLaneSlot-1 in the slot table; lane bits kLaneSlot + k,
// wave bits kWaveSlot + k.
constexpr int kFirstLaneSlot = 8;
constexpr int kFirstWaveSlot = 16;

// Qubits a gate touches non-diagonally (x) and diagonally / as control (z).
inline void touch_masks(const Gate &g, uint64_t &x, uint64_t &z) {
  x = 0;
  z = 0;
  switch (g.kind) {
    case Gate::kDense:
    case Gate::kPerm:
      for (int j = 0; j < g.na; ++j) x |= 1ull << g.act[j];
      z = g.cmask;
      break;
    case Gate::kDiag:
      for (int j = 0; j < g.na; ++j) z |= 1ull << g.act[j];
      z |= g.cmask;
      break;
    case Gate::kSparse:
      z = g.smask;
      break;
    default:
      break;
  }
}

// Inside one sweep, a CNOT / Toffoli target, a controlled 2x2, or a phase on
// one control pattern (CZ, CCZ) next to a plain 2x2 on the same qubit t
// becomes one op: "apply M1 to t where the pattern holds, M0 elsewhere".
// Two ops that are adjacent on t (no op in between touches t) are merged at
// the position of the one that may move: a plain 2x2 acts on t alone, so it
// moves freely past ops that do not touch t; an op with a pattern moves only
// if no op in between changes a pattern qubit.  The merged op costs what one
// 2x2 costs, so the op it absorbs is free.

// A sweep op before encoding: a gate, or a multiplexed 2x2 (g == nullptr).
struct SweepOp {
  const Gate *g = nullptr;
  int t = -1;
  uint64_t pm = 0, pv = 0;  // pattern; pm == 0: a plain 2x2 (m1)
  cd m0[4], m1[4];
  bool alive = true;
};

inline cd cmul(const cd &a, const cd &b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline cd cadd(const cd &a, const cd &b) { return {a.re + b.re, a.im + b.im}; }

// c = a * b for row-major 2x2 matrices.
inline void mul2(const cd *a, const cd *b, cd *c) {
  cd r[4];
  r[0] = cadd(cmul(a[0], b[0]), cmul(a[1], b[2]));
  r[1] = cadd(cmul(a[0], b[1]), cmul(a[1], b[3]));
  r[2] = cadd(cmul(a[2], b[0]), cmul(a[3], b[2]));
  r[3] = cadd(cmul(a[2], b[1]), cmul(a[3], b[3]));
  for (int k = 0; k < 4; ++k) c[k] = r[k];
}

// True when a row-major 2x2 has no off-diagonal entries (exact zeros).
inline bool diagonal2(const cd *m) {
  return m[1].re == 0 && m[1].im == 0 && m[2].re == 0 && m[2].im == 0;
}

// A 2x2 on one qubit with no pattern.
inline bool plain_2x2(const Gate &g, cd *u) {
  if (g.na != 1 || g.cmask != 0 || g.bmask != 0) return false;
  if (g.kind == Gate::kDense) {
    for (int k = 0; k < 4; ++k) u[k] = g.m[k];
    return true;
  }
  if (g.kind == Gate::kDiag && g.m.size() == 2) {
    u[0] = g.m[0];
    u[1] = u[2] = cd{0, 0};
    u[3] = g.m[1];
    return true;
  }
  return false;
}

// The gate as "2x2 `mt` on qubit t where (idx & pm) == pv, identity
// elsewhere", with t not in pm; `choice` picks t among a phase pattern's
// qubits.  Returns false when the gate has no such form.
inline bool pattern_2x2(const Gate &g, int choice, int &t, uint64_t &pm, uint64_t &pv, cd *mt) {
  if (g.bmask != 0) return false;
  if ((g.kind == Gate::kDense || g.kind == Gate::kDiag) && g.na == 1 && g.cmask != 0) {
    if (choice != 0) return false;
    t = g.act[0];
    pm = g.cmask;
    pv = g.cval;
    if (g.kind == Gate::kDense) {
      for (int k = 0; k < 4; ++k) mt[k] = g.m[k];
    } else {
      if (g.m.size() != 2) return false;
      mt[0] = g.m[0];
      mt[1] = mt[2] = cd{0, 0};
      mt[3] = g.m[1];
    }
    return true;
  }
  if (g.kind == Gate::kSparse && g.spat.size() == 1) {
    uint64_t rest = g.smask;
    for (int k = 0; k < choice && rest; ++k) rest &= rest - 1;
    if (!rest) return false;
    t = __builtin_ctzll(rest);
    const uint64_t bt = 1ull << t;
    pm = g.smask & ~bt;
    pv = g.spat[0] & ~bt;
    const bool on1 = (g.spat[0] & bt) != 0;
    mt[0] = on1 ? cd{1, 0} : g.m[0];
    mt[1] = mt[2] = cd{0, 0};
    mt[3] = on1 ? g.m[0] : cd{1, 0};
    return true;
  }
  return false;
}

// The kernel form of a multiplexed op on t with pattern pm: 1 = the pattern
// on no register slot (t anywhere); 2 = on register slots only, t on a
// register slot; 3 = on one register slot alone, t on a lane bit; 0 = none.
inline int mux_form(int t, uint64_t pm, const int *slot, uint64_t reg_qubits) {
  if (t < 0 || slot[t] < 0) return 0;
  // bits off the tile are decided once per tile (the op's umask)
  for (uint64_t b = pm; b; b &= b - 1)
    if (slot[__builtin_ctzll(b)] < 0) pm &= ~(b & (~b + 1));
  if ((pm & reg_qubits) == 0) return 1;
  if (pm & ~reg_qubits) return 0;
  if (slot[t] < kFirstLaneSlot) return 2;
  if (slot[t] >= kFirstWaveSlot) return 0;  // wave bits: the first form only
  return __builtin_popcountll(pm) == 1 ? 3 : 0;
}

inline bool mux_encodable(int t, uint64_t pm, const int *slot, uint64_t reg_qubits) {
  return mux_form(t, pm, slot, reg_qubits) != 0;
}

// Qubits a gate touches in any way (targets, controls, phase patterns).
inline uint64_t touched(const Gate &g) {
  uint64_t m = g.cmask | g.smask;
  for (int j = 0; j < g.na; ++j) m |= 1ull << g.act[j];
  return m;
}

// Batched 2x2s (one matrix per state of a batch, the adjoint's rotations) on
// the same qubit t with nothing in between that touches t: the later one is
// folded into the earlier, state by state (M2_b M1_b), so the sweep runs one
// batched op where it ran two.  The merged gates live in `store` (stable
// addresses for the items that point at them).
inline bool is_batched_2x2(const Gate &g) {
  return g.kind == Gate::kDense && g.na == 1 && g.bmask != 0 && g.cmask == 0;
}

inline void merge_batched(std::vector<SweepOp> &items, std::deque<Gate> &store) {
  auto batched = [](const Gate *g) { return g && is_batched_2x2(*g); };
  for (size_t i = 0; i < items.size(); ++i) {
    if (!items[i].alive || !batched(items[i].g)) continue;
    const int t = items[i].g->act[0];
    const uint64_t bt = 1ull << t;
    for (size_t k = i + 1; k < items.size(); ++k) {
      if (!items[k].alive) continue;
      const Gate *h = items[k].g;
      if (h == nullptr) break;
      if (!(touched(*h) & bt)) continue;
      if (batched(h) && h->act[0] == t && h->bshift == items[i].g->bshift &&
          h->bmask == items[i].g->bmask) {
        const Gate &a = *items[i].g;
        Gate mg = a;
        const size_t nb = a.m.size() / 4;
        for (size_t b = 0; b < nb; ++b) mul2(&h->m[4 * b], &a.m[4 * b], &mg.m[4 * b]);
        mg.logical = a.logical + h->logical;
        store.push_back(std::move(mg));
        items[i].g = &store.back();
        items[k].alive = false;
        continue;  // the merged op may absorb the next one on t as well
      }
      break;  // another gate on t
    }
  }
}

inline void merge_sweep_ops(std::vector<SweepOp> &items, const int *slot, uint64_t reg_qubits) {
  int last[64], lastx[64];  // last op touching q at all / non-diagonally
  for (int q = 0; q < 64; ++q) last[q] = lastx[q] = -1;
  auto is_plain = [&](const SweepOp &o, cd *u) {
    if (o.g == nullptr) {
      if (o.pm != 0) return false;
      for (int k = 0; k < 4; ++k) u[k] = o.m1[k];
      return true;
    }
    return plain_2x2(*o.g, u);
  };
  for (int i = 0; i < (int)items.size(); ++i) {
    SweepOp &G = items[i];
    if (!G.alive) continue;  // folded into an earlier batched op
    cd u[4], uj[4], mt[4];
    bool merged = false;
    if (is_plain(G, u)) {
      const int t = G.g->act[0];
      const int j = last[t];
      if (j >= 0 && items[j].alive) {
        SweepOp &J = items[j];
        int tj;
        uint64_t pm, pv;
        if (J.g == nullptr) {
          // multiplexed (or merged plain) op at j; t may also be one of its
          // pattern qubits, which is not a merge
          if (J.t == t) {  // u moves back to j
            mul2(u, J.m0, J.m0);
            mul2(u, J.m1, J.m1);
            // the op at j may have been diagonal on t; with u it may not be
            if (!diagonal2(u)) lastx[t] = j;
            merged = true;
          }
        } else if (slot[t] >= 0 && plain_2x2(*J.g, uj)) {
          // (two diagonal 2x2s need not be on a tile qubit; a merged op does)
          J.g = nullptr;
          J.t = t;
          J.pm = J.pv = 0;
          mul2(u, uj, J.m1);
          for (int k = 0; k < 4; ++k) J.m0[k] = J.m1[k];
          if (!diagonal2(J.m1)) lastx[t] = j;
          merged = true;
        } else {
          for (int choice = 0; choice < 8; ++choice) {
            if (!pattern_2x2(*J.g, choice, tj, pm, pv, mt)) break;
            if (tj != t || !mux_encodable(t, pm, slot, reg_qubits)) continue;
            J.g = nullptr;
            J.t = t;
            J.pm = pm;
            J.pv = pv;
            for (int k = 0; k < 4; ++k) J.m0[k] = u[k];
            mul2(u, mt, J.m1);
            // a phase pattern acted diagonally on t; the merged op does not
            lastx[t] = j;
            merged = true;
            break;
          }
        }
      }
      if (merged) {
        G.alive = false;
        continue;  // the merged op sits at j, which is still the last op on t
      }
    } else if (G.g != nullptr) {
      int t;
      uint64_t pm, pv;
      for (int choice = 0; choice < 8 && !merged; ++choice) {
        if (!pattern_2x2(*G.g, choice, t, pm, pv, mt)) break;
        const int j = last[t];
        if (j < 0 || !items[j].alive) continue;
        SweepOp &J = items[j];
        if (is_plain(J, uj)) {
          // the plain op moves forward to i
          if (!mux_encodable(t, pm, slot, reg_qubits)) continue;
          G.g = nullptr;
          G.t = t;
          G.pm = pm;
          G.pv = pv;
          for (int k = 0; k < 4; ++k) G.m0[k] = uj[k];
          mul2(mt, uj, G.m1);
          J.alive = false;
          merged = true;
          last[t] = lastx[t] = i;
          for (uint64_t b = pm; b; b &= b - 1) last[__builtin_ctzll(b)] = i;
        } else if (J.g == nullptr && J.t == t && J.pm == pm && J.pv == pv) {
          // same pattern: G moves back to j if no pattern qubit changed since
          bool still = true;
          for (uint64_t b = pm; b; b &= b - 1) still = still && lastx[__builtin_ctzll(b)] < j;
          if (!still) continue;
          mul2(mt, J.m1, J.m1);
          G.alive = false;
          merged = true;
        }
      }
      if (merged) continue;
    }
    if (G.g == nullptr) continue;  // not reached: merged ops continue above
    uint64_t x, z;
    touch_masks(*G.g, x, z);
    for (uint64_t b = x; b; b &= b - 1) last[__builtin_ctzll(b)] = lastx[__builtin_ctzll(b)] = i;
    for (uint64_t b = z; b; b &= b - 1) last[__builtin_ctzll(b)] = i;
  }
}

}  // namespace merge
}  // namespace qs

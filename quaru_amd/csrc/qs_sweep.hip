// qs_sweep.hip -- fused sweep kernel launch and its planner (see qs_sweep.h
// and qs_sweep_kernel.h).
//
// One sweep reads and writes the state once (2 * 2^n * s bytes, s = bytes
// per amplitude) and applies every gate packed into it.  This is how the
// engine beats one pass per gate: the reference's try_apply_fast
// (src/register.rs:155-189) walks the whole state once per gate.
//
// Planner.  A sweep's tile holds qubits 0..b-1 (the fixed low lane bits)
// plus A + 6 - b more qubits (the other lane bits and A register slots).
// Gates are taken in order; a later gate may move ahead of earlier, not yet scheduled gates
// only if it commutes with each of them -- on every shared qubit both act
// diagonally (a control or a diagonal factor), which makes both operators
// block-diagonal in that qubit's basis.  A dense gate enters the sweep if
// its active qubits are, or can still become, tile qubits; diagonal gates
// and controls need no tile qubit at all.  A sweep that would move fewer
// amplitudes than its gates do one by one is replaced by per-gate kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <type_traits>
#include <vector>

#include "qs_sweep.h"
#include "qs_sweep_kernel.h"
#include "qs_sweep_jit.h"
#include "qs_sweep_launch.h"
#include "qs_sweep_merge.h"

namespace qs {
namespace {

using namespace qs::sweep;
using namespace qs::merge;
static_assert(kFirstLaneSlot == (int)kLaneSlot && kFirstWaveSlot == (int)kWaveSlot,
              "slot numbering shared with the merge rules");

constexpr int kMaxOps = 512;     // op records per sweep
constexpr int kWindow = 384;     // look-ahead over unscheduled gates
constexpr size_t kMaxBlobBytes = 64 * 1024;

inline int popc(uint64_t x) { return __builtin_popcountll(x); }

int env_int(const char *name, int dflt) {
  const char *v = std::getenv(name);
  if (!v || !*v) return dflt;
  return std::atoi(v);
}

// Qubits a gate touches non-diagonally (x) and diagonally / as control (z).
void gate_masks(const Gate &g, uint64_t &x, uint64_t &z) {
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

bool is_x(const Gate &g) {
  return g.kind == Gate::kDense && g.na == 1 && g.bmask == 0 && g.m[0].re == 0 && g.m[0].im == 0 &&
         g.m[1].re == 1 && g.m[1].im == 0 && g.m[2].re == 1 && g.m[2].im == 0 &&
         g.m[3].re == 0 && g.m[3].im == 0;
}

// Gates the sweep kernel can apply (`lowb` = qubits fixed on the low lane bits).
bool sweepable(const Gate &g, int lowb) {
  switch (g.kind) {
    case Gate::kDense:
      if (g.na == 1) return true;
      // both qubits of a 4x4 go to register slots, which low qubits never are
      return g.na == 2 && g.act[0] >= lowb && g.act[1] >= lowb;
    case Gate::kDiag: return g.na <= 2;
    case Gate::kSparse: return true;
    default: return false;
  }
}

int op_count(const Gate &g) { return g.kind == Gate::kSparse ? (int)g.spat.size() : 1; }

int pool_count(const Gate &g) {
  switch (g.kind) {
    case Gate::kSparse: return 2 * (int)g.spat.size();
    case Gate::kDiag: return 2 * (int)g.m.size();
    case Gate::kDense: return is_x(g) ? 0 : 2 * (int)g.m.size();
    default: return 0;
  }
}

// Amplitudes a per-gate kernel touches.
double single_amps(const Gate &g, int n) {
  switch (g.kind) {
    case Gate::kDense:
    case Gate::kPerm:
    case Gate::kDiag: return std::ldexp(1.0, n - popc(g.cmask));
    case Gate::kSparse: return (double)g.spat.size() * std::ldexp(1.0, n - popc(g.smask));
    default: return 0.0;
  }
}

struct Plan {
  uint64_t other = 0;     // non-low tile qubits taken by gates
  uint64_t must_reg = 0;  // qubits of 4x4 gates: register slots
  int cnt[64] = {};       // dense gates per qubit
  std::vector<int> gates;
};

// Tile shape limits of one sweep: `lowb` fixed low qubits, A register
// slots, and at most `hcap` tile qubits at or above the page-stride qubit.
struct TileShape {
  int A, lowb, hcap;
  uint64_t high;  // qubits whose stride spans a GPU page or more
  int W;          // wave bits
  int WD;         // wave bits that may carry dense gates (the rest take fillers)
};

bool try_place(const Gate &g, const TileShape &ts, Plan &pl) {
  if (g.kind != Gate::kDense) return true;
  const uint64_t low = (1ull << ts.lowb) - 1ull;
  uint64_t need = 0;
  for (int j = 0; j < g.na; ++j) need |= 1ull << g.act[j];
  const uint64_t other = pl.other | (need & ~low);
  if (popc(other) > ts.A + (kLaneBits - ts.lowb) + ts.WD) return false;
  if (popc(other & ts.high) > ts.hcap) return false;
  const int A = ts.A;
  uint64_t mr = pl.must_reg;
  if (g.na == 2) {
    mr |= need;
    if (popc(mr) > A) return false;
  }
  pl.other = other;
  pl.must_reg = mr;
  for (int j = 0; j < g.na; ++j) ++pl.cnt[g.act[j]];
  return true;
}

// try_place for a tile whose non-low qubits are fixed to the set S.
bool fits_set(const Gate &g, uint64_t S, const TileShape &ts, Plan &pl) {
  if (g.kind != Gate::kDense) return true;
  const uint64_t low = (1ull << ts.lowb) - 1ull;
  uint64_t need = 0;
  for (int j = 0; j < g.na; ++j) need |= 1ull << g.act[j];
  if (need & ~low & ~S) return false;
  uint64_t mr = pl.must_reg;
  if (g.na == 2) {
    mr |= need;
    if (popc(mr) > ts.A) return false;
  }
  pl.other |= need & ~low;
  pl.must_reg = mr;
  for (int j = 0; j < g.na; ++j) ++pl.cnt[g.act[j]];
  return true;
}

// Moves the pattern bits of a controlled op that lie off the tile (`tmask`
// = the tile qubits) to umask: they are the same for the whole tile, so the
// kernel decides them once per tile and skips the op where they fail.
inline void split_tile_controls(OpRec &op, uint64_t tmask) {
  const uint64_t off = op.cmask & ~tmask;
  if (off == 0) return;
  op.umask = off;
  op.cmask &= tmask;
  op.code_fail = CODE_NONE;
  op.uoff = 0;
}

inline bool is_minus_one(const cd &v) { return v.re == -1.0 && v.im == 0.0; }

template <typename R>
void encode_gate(const Gate &g, const int *slot, uint64_t tmask, std::vector<OpRec> &ops,
                 std::vector<R> &pool) {
  auto put = [&](const cd &v) {
    pool.push_back((R)v.re);
    pool.push_back((R)v.im);
  };
  if (g.kind == Gate::kSparse) {
    for (size_t e = 0; e < g.spat.size(); ++e) {
      OpRec op{};
      op.type = OP_SPARSE;
      op.cmask = g.smask;
      op.cval = g.spat[e];
      op.moff = (uint32_t)pool.size();
      op.code = is_minus_one(g.m[e]) ? (uint32_t)CODE_NEG : op_code(op.type, 0, 0);
      put(g.m[e]);
      split_tile_controls(op, tmask);
      ops.push_back(op);
    }
    return;
  }
  OpRec op{};
  op.cmask = g.cmask;
  op.cval = g.cval;
  op.moff = (uint32_t)pool.size();
  if (g.kind == Gate::kDiag) {
    // na == 0 is a phase on the control pattern: a sparse op
    op.type = g.na == 0 ? OP_SPARSE : OP_DIAG;
    op.na = (uint32_t)g.na;
    for (int j = 0; j < g.na; ++j) op.act[j] = (uint32_t)g.act[j];
    for (const cd &v : g.m) put(v);
    op.code = (g.na == 0 && is_minus_one(g.m[0])) ? (uint32_t)CODE_NEG : op_code(op.type, 0, 0);
    split_tile_controls(op, tmask);
    ops.push_back(op);
    return;
  }
  if (g.na == 1) {
    op.type = is_x(g) ? OP_X1 : (g.cmask == 0 ? OP_U1 : OP_C1);
    op.s0 = (uint32_t)slot[g.act[0]];
    op.bshift = (uint32_t)g.bshift;
    op.bmask = g.bmask;
  } else {
    op.type = OP_DENSE2;
    op.s0 = (uint32_t)slot[g.act[0]];
    op.s1 = (uint32_t)slot[g.act[1]];
  }
  if (op.type != OP_X1)
    for (const cd &v : g.m) put(v);
  split_tile_controls(op, tmask);
  // controls all off the tile: where they hold, the plain 2x2
  if (op.type == OP_C1 && op.cmask == 0) op.type = OP_U1;
  op.code = op_code(op.type, op.s0, op.s1);
  ops.push_back(op);
}

template <typename R>
void encode_mux(const SweepOp &o, const int *slot, uint64_t reg_qubits, uint64_t tmask,
                std::vector<OpRec> &ops, std::vector<R> &pool) {
  OpRec op{};
  op.s0 = (uint32_t)slot[o.t];
  op.moff = (uint32_t)pool.size();
  auto put = [&](const cd &v) {
    pool.push_back((R)v.re);
    pool.push_back((R)v.im);
  };
  const uint64_t pin = o.pm & tmask;  // the pattern's tile bits
  if (o.pm == 0) {
    op.type = OP_U1;
    for (int k = 0; k < 4; ++k) put(o.m1[k]);
    op.code = op_code(op.type, op.s0, 0);
  } else if (pin == 0) {
    // the pattern lies off the tile: one plain 2x2 per tile, m1 where the
    // pattern holds (pool + 8), m0 elsewhere
    op.type = OP_U1;
    for (int k = 0; k < 4; ++k) put(o.m0[k]);
    for (int k = 0; k < 4; ++k) put(o.m1[k]);
    op.code = op_code(op.type, op.s0, 0);
    op.cval = o.pv;
    op.umask = o.pm;
    op.code_fail = op.code;
    op.uoff = 8;
  } else {
    op.type = OP_MUX;
    op.cmask = pin;
    op.cval = o.pv;
    for (int k = 0; k < 4; ++k) put(o.m0[k]);
    for (int k = 0; k < 4; ++k) put(o.m1[k]);
    const int form = mux_form(o.t, pin, slot, reg_qubits);
    if (form == 3) op.s1 = (uint32_t)slot[__builtin_ctzll(pin)];
    op.code = op_code(op.type, op.s0, op.s1,
                      form == 1 ? MUX_LANES : (form == 2 ? MUX_REGS : MUX_ONE_REG));
    if (o.pm != pin) {
      // where the off-tile bits fail, m0 on the whole tile
      op.umask = o.pm & ~tmask;
      op.code_fail = op_code(OP_U1, op.s0, 0);
      op.uoff = 0;
    }
  }
  ops.push_back(op);
}

template <typename R>
hipError_t launch_sweep_a(int A, int W, SweepRunner &run, const void *d_blob, uint64_t ntiles,
                          const std::vector<OpRec> &ops, const uint64_t *gb) {
  const bool xcd = sweep_xcd_chunks();
  // a kernel specialised to this op-code sequence, when one is compiled
  // (32-bit indices only; in the automatic mode only from 22 qubits, where a
  // pass over the state takes long enough to repay a compile)
  const int jm = run.n <= 32 ? jit_mode() : 0;
  if (jm == 2 || (jm == 1 && run.n >= 22)) {
    const bool c64 = std::is_same<R, float>::value;
    const std::vector<int> rp = jit_register_patterns(ops, gb, A);
    hipFunction_t fn = jit_kernel(jit_signature(ops, c64, A, W, rp), ops, c64, A, W, rp);
    if (fn) {
      uint64_t grid;
      uint32_t xcd_chunk;
      sweep_grid(W, ntiles, xcd, &grid, &xcd_chunk);
      void *psi = run.psi;
      const void *blob = d_blob;
      void *args[] = {&psi, &blob, &xcd_chunk};
      jit_count_launch();
      return hipModuleLaunchKernel(fn, (unsigned)grid, 1, 1, kThreads, 1, 1, 0, run.stream, args,
                                   nullptr);
    }
  }
  switch (A) {
    case 3: return launch_sweep_ra<R, 3>(W, run.n, run.stream, run.psi, d_blob, ntiles, xcd);
    case 4: return launch_sweep_ra<R, 4>(W, run.n, run.stream, run.psi, d_blob, ntiles, xcd);
    case 5: return launch_sweep_ra<R, 5>(W, run.n, run.stream, run.psi, d_blob, ntiles, xcd);
    default: return hipErrorInvalidValue;
  }
}

template <typename R>
hipError_t execute_t(SweepRunner &run, const std::vector<Gate> &gates, uint64_t *sweeps) {
  const int n = run.n;
  const int A = sweep_reg_bits(run.dtype_c64);
  // the workgroup's wave bits add tile qubits when the register is large
  // enough for a tile per workgroup
  const int W = (sweep_wave_bits(run.dtype_c64) && n >= kLaneBits + A + kWaveBits) ? kWaveBits : 0;
  const int T = kLaneBits + A + W;
  const size_t G = gates.size();
  // Batched gates keep their batch bits off the tile (the kernel reads the
  // batch index from the tile base): tile qubits come from [0, nt).
  int nt = n;
  for (const Gate &g : gates)
    if (g.bmask) nt = std::min(nt, g.bshift);
  if (nt < T) {
    for (const Gate &g : gates) {
      hipError_t e = run.single(g);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  const int lowb = sweep_low_bits(run.dtype_c64);
  const int nfree = kLaneBits - lowb;  // lane bits the planner assigns per sweep
  const uint64_t low = (1ull << lowb) - 1ull;
  const uint64_t allq = (n >= 64) ? ~0ull : ((1ull << n) - 1ull);
  // A wave reads 2^(high tile qubits) separate pages; on register slots they
  // cost one page per load instruction, on lane bits they multiply the pages
  // of every load instruction.  High qubits go to register slots first and at
  // most `sweep_high_lanes` of them to lane bits.
  const int hq = sweep_page_qubit(run.dtype_c64);
  // Each dense gate on a wave bit costs an exchange through LDS with four
  // barriers; how many wave bits take dense gates: sweep_dense_wave_bits.
  const int WD = std::min(W, sweep_dense_wave_bits());
  const TileShape ts{A, lowb, A + sweep_high_lanes() + WD,
                     hq >= 64 ? 0ull : (allq & ~((1ull << hq) - 1ull)), W, WD};
  std::vector<uint64_t> gx(G), gz(G);
  for (size_t i = 0; i < G; ++i) gate_masks(gates[i], gx[i], gz[i]);
  std::vector<char> done(G, 0);
  size_t first = 0;
  std::vector<OpRec> ops;
  std::vector<R> pool;
  std::vector<char> blob;
  static const int max_gates = env_int("QS_SWEEP_MAX_GATES", 1 << 30);
  static const int search_iters = std::max(0, env_int("QS_SWEEP_SEARCH", 8));
  static const int search_restarts = std::max(0, env_int("QS_SWEEP_RESTARTS", 2));
  while (true) {
    while (first < G && done[first]) ++first;
    if (first == G) break;
    if (!sweepable(gates[first], lowb)) {
      hipError_t e = run.single(gates[first]);
      if (e != hipSuccess) return e;
      done[first] = 1;
      continue;
    }
    // ---- collect the gates of one sweep.  fixed = false: the tile's
    // non-low qubits are taken first come, first served; fixed = true: they
    // are the set S.
    auto collect = [&](bool fixed, uint64_t S, Plan &pl) {
      pl = Plan();
      int nops = 0, npool = 0;
      uint64_t bx = 0, bz = 0;
      int scanned = 0;
      for (size_t i = first; i < G && scanned < kWindow; ++i) {
        if (done[i]) continue;
        ++scanned;
        const Gate &g = gates[i];
        const bool movable = !(gx[i] & (bx | bz)) && !(gz[i] & bx);
        const bool room =
            (int)pl.gates.size() < max_gates && nops + op_count(g) <= kMaxOps &&
            blob_bytes<R>(nops + op_count(g), npool + pool_count(g)) <= kMaxBlobBytes;
        bool ok = movable && room && sweepable(g, lowb);
        if (ok) ok = fixed ? fits_set(g, S, ts, pl) : try_place(g, ts, pl);
        if (ok) {
          pl.gates.push_back((int)i);
          nops += op_count(g);
          npool += pool_count(g);
        } else {
          bx |= gx[i];
          bz |= gz[i];
          if (!sweepable(g, lowb)) bx |= gz[i];
          if ((bx & allq) == allq) break;
        }
      }
    };
    Plan pl;
    collect(false, 0, pl);
    // ---- local search over the tile's qubit set: add a qubit while there
    // is room, else swap one in for one out; keep a change when it schedules
    // more gates.  First come, first served fills the tile with the qubits of
    // the earliest gates, which is often not the set that reaches furthest
    // into the circuit.  The search runs from that set and from a few
    // pseudo-random full sets (seeded by the sweep number, so plans are
    // reproducible) and keeps the best result.
    if (search_iters > 0) {
      const int cap = A + nfree + WD;
      Plan cur, cand;
      auto climb = [&](uint64_t S, Plan &best) {
        for (int it = 0; it < search_iters; ++it) {
          bool improved = false;
          for (int qi = lowb; qi < nt; ++qi) {
            if ((S >> qi) & 1ull) continue;
            if (popc(S) < cap) {
              const uint64_t S2 = S | (1ull << qi);
              if (popc(S2 & ts.high) > ts.hcap) continue;
              collect(true, S2, cand);
              if (cand.gates.size() > best.gates.size()) {
                std::swap(best, cand);
                S = S2;
                improved = true;
              }
              continue;
            }
            for (int qo = lowb; qo < nt; ++qo) {
              if (!((S >> qo) & 1ull)) continue;
              const uint64_t S2 = (S & ~(1ull << qo)) | (1ull << qi);
              if (popc(S2 & ts.high) > ts.hcap) continue;
              collect(true, S2, cand);
              if (cand.gates.size() > best.gates.size()) {
                std::swap(best, cand);
                S = S2;
                improved = true;
                break;  // S changed: continue with the next qubit to bring in
              }
            }
          }
          if (!improved) break;
        }
      };
      climb(pl.other, pl);
      const int pool_q = nt - lowb;
      for (int k = 0; k < search_restarts && pool_q > cap; ++k) {
        uint64_t rs = (uint64_t)*sweeps * 0x9E3779B97F4A7C15ull + (uint64_t)(k + 1) * 0xD1B54A32D192ED03ull;
        uint64_t S = 0;
        for (int tries = 0; popc(S) < cap && tries < 64 * cap; ++tries) {
          rs += 0x9E3779B97F4A7C15ull;  // SplitMix64 step
          uint64_t z = rs;
          z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
          z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
          z ^= z >> 31;
          const int q = lowb + (int)(z % (uint64_t)pool_q);
          const uint64_t S2 = S | (1ull << q);
          if (popc(S2 & ts.high) <= ts.hcap) S = S2;
        }
        collect(true, S, cur);
        climb(S, cur);
        if (cur.gates.size() > pl.gates.size()) std::swap(pl, cur);
      }
    }
    // ---- hold back a plain 2x2 on qubit t whose next gate on t (a CNOT or
    // Toffoli target, a controlled 2x2, a phase pattern, another 2x2) is
    // left for a later sweep: nothing else of this sweep touches t after it,
    // and the later sweep, which has t on its tile anyway, merges the two
    // into one op.  Only for patterns of at most 2 qubits: a wide pattern
    // (a multi-controlled phase over the whole register) has no kernel form
    // to merge into, and holding back every gate in front of it would starve
    // the sweeps.
    if (sweep_merge_ops() && pl.gates.size() > 1) {
      std::vector<char> in_sweep_last(64, 0);
      std::vector<int> keep;
      keep.reserve(pl.gates.size());
      uint64_t later_touch = 0;  // qubits touched by gates after the current one in this sweep
      std::vector<char> drop(pl.gates.size(), 0);
      for (int k = (int)pl.gates.size() - 1; k >= 0; --k) {
        const int gi = pl.gates[k];
        const Gate &g = gates[gi];
        cd u[4];
        const uint64_t touch = gx[gi] | gz[gi];
        const bool bat = is_batched_2x2(g);
        if ((bat || plain_2x2(g, u)) && !(later_touch & (1ull << g.act[0]))) {
          const uint64_t bt = 1ull << g.act[0];
          // the next gate on t outside this sweep
          for (size_t i = (size_t)gi + 1; i < G; ++i) {
            if (done[i] || !((gx[i] | gz[i]) & bt)) continue;
            if (std::find(pl.gates.begin(), pl.gates.end(), (int)i) != pl.gates.end()) break;
            int tt;
            uint64_t pm, pv;
            cd mt[4];
            bool mergeable;
            if (bat) {
              // a batched 2x2 merges only with another one on t (merge_batched)
              const Gate &h = gates[i];
              mergeable = is_batched_2x2(h) && h.act[0] == g.act[0] && h.bshift == g.bshift &&
                          h.bmask == g.bmask;
            } else {
              mergeable = plain_2x2(gates[i], mt);
              for (int choice = 0; choice < 8 && !mergeable; ++choice) {
                if (!pattern_2x2(gates[i], choice, tt, pm, pv, mt)) break;
                mergeable = tt == g.act[0] && popc(pm) <= 2;
              }
            }
            if (mergeable) drop[k] = 1;
            break;
          }
        }
        later_touch |= touch;
      }
      for (size_t k = 0; k < pl.gates.size(); ++k)
        if (!drop[k]) keep.push_back(pl.gates[k]);
      if (!keep.empty() && keep.size() < pl.gates.size()) pl.gates.swap(keep);
    }
    for (int gi : pl.gates) done[gi] = 1;
    // ---- sweep or single kernels?
    double single_cost = 0.0;
    for (int gi : pl.gates) single_cost += single_amps(gates[gi], n);
    if (single_cost <= std::ldexp(1.0, n)) {
      for (int gi : pl.gates) {
        hipError_t e = run.single(gates[gi]);
        if (e != hipSuccess) return e;
      }
      continue;
    }
    // ---- roles: register slots for the most-used qubits (4x4 qubits
    // first), the rest of the gate qubits on the free lane bits, the least
    // used on the wave bits (their ops exchange through LDS), then fillers
    std::vector<int> regs, lanes, waves;
    for (int q = 0; q < n; ++q)
      if ((pl.must_reg >> q) & 1ull) regs.push_back(q);
    std::vector<int> rest;
    for (int q = 0; q < n; ++q)
      if (((pl.other & ~pl.must_reg) >> q) & 1ull) rest.push_back(q);
    // high qubits first (register slots), then by use count
    std::stable_sort(rest.begin(), rest.end(), [&](int a, int b) {
      const bool ha = (ts.high >> a) & 1ull, hb = (ts.high >> b) & 1ull;
      if (ha != hb) return ha;
      return pl.cnt[a] > pl.cnt[b];
    });
    std::vector<int> rest2;
    for (int q : rest) {
      if ((int)regs.size() < A)
        regs.push_back(q);
      else
        rest2.push_back(q);
    }
    std::stable_sort(rest2.begin(), rest2.end(),
                     [&](int a, int b) { return pl.cnt[a] > pl.cnt[b]; });
    for (int q : rest2) {
      if ((int)lanes.size() < nfree)
        lanes.push_back(q);
      else
        waves.push_back(q);
    }
    uint64_t P = low | pl.other;
    for (int q = lowb; q < nt &&
                       ((int)regs.size() < A || (int)lanes.size() < nfree || (int)waves.size() < W);
         ++q) {
      if ((P >> q) & 1ull) continue;
      P |= 1ull << q;
      if ((int)lanes.size() < nfree)
        lanes.push_back(q);
      else if ((int)regs.size() < A)
        regs.push_back(q);
      else
        waves.push_back(q);
    }
    SweepHeader h{};
    {
      int b = 0;
      for (int q = 0; q < 64 && b < T; ++q)
        if ((P >> q) & 1ull) h.tpos[b++] = (uint32_t)q;
    }
    int slot[64];
    for (int q = 0; q < 64; ++q) slot[q] = -1;
    for (int i = 0; i < lowb; ++i) {
      h.lg[i] = 1ull << i;
      slot[i] = (int)kLaneSlot + i;
    }
    for (int i = 0; i < nfree; ++i) {
      h.lg[lowb + i] = 1ull << lanes[i];
      slot[lanes[i]] = (int)kLaneSlot + lowb + i;
    }
    for (int j = 0; j < A; ++j) {
      h.gb[j] = 1ull << regs[j];
      slot[regs[j]] = j;
    }
    h.nwave = (uint32_t)W;
    for (int k = 0; k < W; ++k) {
      h.wb[k] = 1ull << waves[k];
      slot[waves[k]] = (int)kWaveSlot + k;
    }
    ops.clear();
    pool.clear();
    std::vector<SweepOp> items(pl.gates.size());
    for (size_t k = 0; k < pl.gates.size(); ++k) items[k].g = &gates[pl.gates[k]];
    uint64_t reg_qubits = 0;
    for (int j = 0; j < A; ++j) reg_qubits |= 1ull << regs[j];
    std::deque<Gate> merged_store;
    if (sweep_merge_ops()) {
      merge_batched(items, merged_store);
      merge_sweep_ops(items, slot, reg_qubits);
    }
    for (const SweepOp &o : items) {
      if (!o.alive) {
        if (run.merged) ++*run.merged;
        continue;
      }
      if (o.g)
        encode_gate<R>(*o.g, slot, P, ops, pool);
      else
        encode_mux<R>(o, slot, reg_qubits, P, ops, pool);
    }
    h.nops = (uint32_t)ops.size();
    if (run.sweep_ops) *run.sweep_ops += ops.size();
    // QS_SWEEP_TRACE=1: one line per planned sweep on stderr, the tile's
    // roles (free lane bits, register slots, wave bits) and its op count
    static const bool trace = env_int("QS_SWEEP_TRACE", 0) != 0;
    if (trace) {
      std::fprintf(stderr, "qs_sweep n %d lanes", n);
      for (int i = 0; i < nfree; ++i) std::fprintf(stderr, " %d", lanes[i]);
      std::fprintf(stderr, " regs");
      for (int j = 0; j < A; ++j) std::fprintf(stderr, " %d", regs[j]);
      std::fprintf(stderr, " waves");
      for (int k = 0; k < W; ++k) std::fprintf(stderr, " %d", waves[k]);
      std::fprintf(stderr, " ops %u\n", h.nops);
    }
    h.npool = (uint32_t)pool.size();
    h.ntiles = 1ull << (n - T);
    const size_t hbytes = (sizeof(SweepHeader) + 15) & ~(size_t)15;
    const size_t bbytes = blob_bytes<R>((int)h.nops, (int)h.npool);
    blob.assign(bbytes, 0);
    std::memcpy(blob.data(), &h, sizeof(SweepHeader));
    if (!ops.empty()) std::memcpy(blob.data() + hbytes, ops.data(), ops.size() * sizeof(OpRec));
    if (!pool.empty())
      std::memcpy(blob.data() + hbytes + ops.size() * sizeof(OpRec), pool.data(),
                  pool.size() * sizeof(R));
    if (run.host_sweep) {
      hipError_t e = run.host_sweep(blob.data(), blob.size());
      if (e != hipSuccess) return e;
      ++*sweeps;
      continue;
    }
    void *d_blob = nullptr;
    hipError_t e = run.upload(blob.data(), blob.size(), &d_blob);
    if (e != hipSuccess) return e;
    if (run.mark) run.mark(true);
    e = launch_sweep_a<R>(A, W, run, d_blob, h.ntiles, ops, h.gb);
    if (run.mark) run.mark(false);
    if (e != hipSuccess) return e;
    ++*sweeps;
  }
  return hipSuccess;
}

}  // namespace

struct SweepEngine {
  int dummy = 0;
};

SweepEngine *sweep_engine_new() { return new SweepEngine(); }
void sweep_engine_free(SweepEngine *e) { delete e; }

// First qubit whose index stride (2^q amplitudes) spans 2 MiB, the GPU's
// large-page size: 17 for complex128, 18 for complex64.
int sweep_page_qubit(bool c64) {
  return env_int(c64 ? "QS_SWEEP_PAGE_Q_C64" : "QS_SWEEP_PAGE_Q_C128", c64 ? 18 : 17);
}

// High qubits allowed on lane bits per sweep.  Measured on the 30-qubit
// random circuit, capping them costs more in extra sweeps than it saves per
// sweep, so the default is no cap.
int sweep_high_lanes() { return std::max(0, env_int("QS_SWEEP_HIGH_LANES", kLaneBits)); }

// Measured on the 30-qubit circuit at three waves per SIMD: complex128 with
// 3 fixed low qubits (128-byte runs, one more free lane bit) needs 81 sweeps
// of 6.73 ms instead of 92 of 6.39 ms (+7% gates/s); complex64 keeps 4
// (128-byte runs; 3 gives 70 sweeps of 4.52 ms instead of 79 of 3.46 ms).
int sweep_low_bits(bool c64) {
  const int b = env_int(c64 ? "QS_SWEEP_LOW_C64" : "QS_SWEEP_LOW_C128", c64 ? 4 : 3);
  return std::min(kLaneBits, std::max(0, b));
}

// Workgroup tiles with wave bits (QS_SWEEP_WAVES_C128 / _C64, default on).
bool sweep_wave_bits(bool c64) {
  static const bool on128 = env_int("QS_SWEEP_WAVES_C128", 1) != 0;
  // complex64 is bound by arithmetic more than complex128: measured on the
  // 30-qubit circuit, 68 workgroup-tile sweeps of 3.94 ms (4880 gates/s)
  // against 79 wave-tile sweeps of 3.45 ms (4780-4790 gates/s)
  static const bool on64 = env_int("QS_SWEEP_WAVES_C64", 1) != 0;
  return c64 ? on64 : on128;
}

// Wave bits open to dense gates (QS_SWEEP_DENSE_WAVES, 0..2, default 2).
// With the specialised kernels both pay on the 30-qubit circuit: 62 sweeps
// of 6.94 ms (3065 gates/s) against 70 of 6.58 ms (2847) with one open and
// 81 of 6.60 ms (2467) with none; complex64 5717 / 5528 / 5073 gates/s.
int sweep_dense_wave_bits() {
  static const int d = std::min(kWaveBits, std::max(0, env_int("QS_SWEEP_DENSE_WAVES", 2)));
  return d;
}

// Merging of one-qubit ops inside a sweep (QS_SWEEP_MERGE, default on).
bool sweep_merge_ops() {
  static const bool on = env_int("QS_SWEEP_MERGE", 1) != 0;
  return on;
}

// Contiguous tile ranges per XCD (see xcd_block in qs_sweep_kernel.h).
bool sweep_xcd_chunks() {
  static const bool on = env_int("QS_SWEEP_XCD", 1) != 0;
  return on;
}

// Measured on the 30-qubit circuit: complex128 at A = 4 (3 waves per SIMD)
// and complex64 at A = 5 (11 tile qubits: 79 sweeps instead of 92, 3 waves).
int sweep_reg_bits(bool c64) {
  const int a = env_int(c64 ? "QS_SWEEP_A_C64" : "QS_SWEEP_A_C128", c64 ? 5 : 4);
  return std::min(5, std::max(3, a));
}

hipError_t sweep_execute(SweepEngine *eng, SweepRunner &run, const std::vector<Gate> &gates,
                         uint64_t *sweeps, uint64_t *kernels) {
  (void)eng;
  (void)kernels;
  if (run.dtype_c64) return execute_t<float>(run, gates, sweeps);
  return execute_t<double>(run, gates, sweeps);
}

}  // namespace qs

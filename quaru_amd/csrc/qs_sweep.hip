// qs_sweep.hip -- fused sweep kernel and its planner (see qs_sweep.h).
//
// One sweep reads and writes the state once (2 * 2^n * s bytes, s = bytes
// per amplitude) and applies every gate packed into it.  This is how the
// engine beats one-pass-per-gate: the reference's try_apply_fast
// (src/register.rs:155-189) walks the whole state once per gate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "qs_sweep.h"
#include "qs_sweep_device.h"

namespace qs {
namespace {

using namespace qs::sweep;
constexpr int kLowBits = 4;      // qubits 0..3 always belong to the tile
constexpr int kMaxOps = 512;     // op records per sweep
constexpr int kWindow = 384;     // look-ahead over unscheduled gates

// ---------------------------------------------------------------------------
// planner

inline int popc(uint64_t x) { return __builtin_popcountll(x); }

// Qubits a gate touches non-diagonally (x) and diagonally / as control (z).
void gate_masks(const Gate &g, uint64_t &x, uint64_t &z) {
  x = 0;
  z = 0;
  switch (g.kind) {
    case Gate::kDense:
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
  return g.kind == Gate::kDense && g.na == 1 && g.m[0].re == 0 && g.m[0].im == 0 &&
         g.m[1].re == 1 && g.m[1].im == 0 && g.m[2].re == 1 && g.m[2].im == 0 &&
         g.m[3].re == 0 && g.m[3].im == 0;
}

// Gates the sweep kernel can apply.
bool sweepable(const Gate &g) {
  switch (g.kind) {
    case Gate::kDense: return g.na == 1 || g.na == 2;
    case Gate::kDiag: return g.na <= 4;
    case Gate::kSparse: return true;
    default: return false;
  }
}

int op_count(const Gate &g) {
  if (g.kind == Gate::kSparse) return (int)g.spat.size();
  return 1;
}

// Amplitudes a per-gate kernel touches.
double single_amps(const Gate &g, int n) {
  switch (g.kind) {
    case Gate::kDense:
    case Gate::kDiag: return std::ldexp(1.0, n - popc(g.cmask));
    case Gate::kSparse: return (double)g.spat.size() * std::ldexp(1.0, n - popc(g.smask));
    default: return 0.0;
  }
}

struct PlanPhase {
  uint64_t regs = 0;
  std::vector<int> order;  // register qubits in slot order
  std::vector<int> gates;
};

bool try_place(const Gate &g, uint64_t &P, int T, int A, PlanPhase &ph) {
  if (g.kind != Gate::kDense) return true;
  uint64_t need = 0;
  for (int j = 0; j < g.na; ++j) need |= 1ull << g.act[j];
  const uint64_t new_regs = need & ~ph.regs;
  const uint64_t new_tile = need & ~P;
  if ((int)ph.order.size() + popc(new_regs) > A) return false;
  if (popc(P) + popc(new_tile) > T) return false;
  for (int j = 0; j < g.na; ++j) {
    const int q = g.act[j];
    if (!((ph.regs >> q) & 1ull)) {
      ph.regs |= 1ull << q;
      ph.order.push_back(q);
    }
  }
  P |= new_tile;
  return true;
}

template <typename R>
void encode_gate(const Gate &g, const PlanPhase &ph, std::vector<SweepOp<R>> &ops) {
  auto slot = [&](int q) {
    for (size_t j = 0; j < ph.order.size(); ++j)
      if (ph.order[j] == q) return (int)j;
    return -1;
  };
  if (g.kind == Gate::kSparse) {
    for (size_t e = 0; e < g.spat.size(); ++e) {
      SweepOp<R> op{};
      op.type = OP_SPARSE;
      op.cmask = g.smask;
      op.cval = g.spat[e];
      op.m[0] = (R)g.m[e].re;
      op.m[1] = (R)g.m[e].im;
      ops.push_back(op);
    }
    return;
  }
  SweepOp<R> op{};
  op.cmask = g.cmask;
  op.cval = g.cval;
  if (g.kind == Gate::kDiag) {
    op.type = OP_DIAG;
    op.na = (uint8_t)g.na;
    for (int j = 0; j < g.na; ++j) op.act[j] = (uint8_t)g.act[j];
    for (size_t i = 0; i < g.m.size(); ++i) {
      op.m[2 * i] = (R)g.m[i].re;
      op.m[2 * i + 1] = (R)g.m[i].im;
    }
    if (g.na == 0) {
      // a controlled phase: same as a sparse op on the control pattern
      op.type = OP_SPARSE;
    }
    ops.push_back(op);
    return;
  }
  // dense
  if (g.na == 1) {
    op.type = is_x(g) ? OP_X1 : OP_DENSE1;
    op.s0 = (uint8_t)slot(g.act[0]);
  } else {
    op.type = OP_DENSE2;
    op.s0 = (uint8_t)slot(g.act[0]);
    op.s1 = (uint8_t)slot(g.act[1]);
  }
  const size_t d = (size_t)1 << g.na;
  for (size_t i = 0; i < d * d; ++i) {
    op.m[2 * i] = (R)g.m[i].re;
    op.m[2 * i + 1] = (R)g.m[i].im;
  }
  ops.push_back(op);
}

int env_int(const char *name, int dflt) {
  const char *v = std::getenv(name);
  if (!v || !*v) return dflt;
  return std::atoi(v);
}

template <typename R, int T>
hipError_t launch_sweep(SweepRunner &run, const SweepHeader *d_h, const SweepOp<R> *d_ops) {
  using V = typename SV2<R>::T;
  const size_t lds = sizeof(V) << T;
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void *)k_sweep<R, T>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  const uint64_t ntiles = 1ull << (run.n - T);
  hipLaunchKernelGGL((k_sweep<R, T>), dim3((unsigned)ntiles), dim3(kThreads), lds, run.stream,
                     (V *)run.psi, d_h, d_ops);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_sweep_t(int T, SweepRunner &run, const SweepHeader *d_h,
                          const SweepOp<R> *d_ops) {
  switch (T) {
    case 10: return launch_sweep<R, 10>(run, d_h, d_ops);
    case 11: return launch_sweep<R, 11>(run, d_h, d_ops);
    case 12: return launch_sweep<R, 12>(run, d_h, d_ops);
    default: return hipErrorInvalidValue;
  }
}

template <typename R>
hipError_t execute_t(SweepRunner &run, const std::vector<Gate> &gates, uint64_t *sweeps,
                     uint64_t *kernels) {
  const int n = run.n;
  const int T = sweep_tile_bits(run.dtype_c64);
  const int A = T - 8;
  const size_t G = gates.size();
  if (n < T + 1) {
    for (const Gate &g : gates) {
      hipError_t e = run.single(g);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  const int L = std::min(kLowBits, T - A);
  const uint64_t low = (1ull << L) - 1ull;
  const uint64_t allq = (n >= 64) ? ~0ull : ((1ull << n) - 1ull);
  std::vector<uint64_t> gx(G), gz(G);
  for (size_t i = 0; i < G; ++i) gate_masks(gates[i], gx[i], gz[i]);
  std::vector<char> done(G, 0);
  size_t first = 0;
  std::vector<SweepOp<R>> ops;
  std::vector<char> blob;
  while (true) {
    while (first < G && done[first]) ++first;
    if (first == G) break;
    if (!sweepable(gates[first])) {
      hipError_t e = run.single(gates[first]);
      if (e != hipSuccess) return e;
      done[first] = 1;
      continue;
    }
    // ---- build one sweep
    uint64_t P = low;
    std::vector<PlanPhase> phases;
    int nops = 0;
    for (int pi = 0; pi < kMaxPhases - 2; ++pi) {
      PlanPhase ph;
      uint64_t bx = 0, bz = 0;
      int scanned = 0;
      for (size_t i = first; i < G && scanned < kWindow; ++i) {
        if (done[i]) continue;
        ++scanned;
        const Gate &g = gates[i];
        const bool movable = !(gx[i] & (bx | bz)) && !(gz[i] & bx);
        bool placed = false;
        if (movable && sweepable(g) && nops + op_count(g) <= kMaxOps)
          placed = try_place(g, P, T, A, ph);
        if (placed) {
          ph.gates.push_back((int)i);
          done[i] = 1;
          nops += op_count(g);
        } else {
          bx |= gx[i];
          bz |= gz[i];
          if (!sweepable(g)) bx |= gz[i];
          if ((bx & allq) == allq) break;
        }
      }
      if (ph.gates.empty()) break;
      phases.push_back(std::move(ph));
    }
    // ---- sweep or single kernels?
    double single_cost = 0.0;
    for (const PlanPhase &ph : phases)
      for (int gi : ph.gates) single_cost += single_amps(gates[gi], n);
    const double sweep_cost = std::ldexp(1.0, n);
    if (single_cost <= sweep_cost) {
      for (const PlanPhase &ph : phases)
        for (int gi : ph.gates) {
          hipError_t e = run.single(gates[gi]);
          if (e != hipSuccess) return e;
        }
      continue;
    }
    // ---- complete the tile to T qubits with the lowest free qubits
    for (int q = 0; q < n && popc(P) < T; ++q) P |= 1ull << q;
    SweepHeader h{};
    h.tile = make_fixed(P, 0);
    int lbit[64];
    for (int b = 0; b < T; ++b) {
      h.tpos[b] = h.tile.pos[b];
      lbit[h.tile.pos[b]] = b;
    }
    // Register slots: the phase's own qubits first, then filler tile qubits
    // taken from the highest local bits, so that the thread bits cover the
    // low local bits (coalesced HBM access, conflict-free LDS).
    auto complete = [&](PlanPhase &ph) {
      for (int b = T - 1; b >= 0 && (int)ph.order.size() < A; --b) {
        const int q = h.tpos[b];
        if (!((ph.regs >> q) & 1ull)) {
          ph.regs |= 1ull << q;
          ph.order.push_back(q);
        }
      }
    };
    const uint64_t lowlocal = low;  // local bits 0..L-1 are qubits 0..L-1
    auto touches_low = [&](const PlanPhase &ph) { return (ph.regs & lowlocal) != 0; };
    for (PlanPhase &ph : phases) complete(ph);
    if (touches_low(phases.front())) {
      PlanPhase io;
      complete(io);
      phases.insert(phases.begin(), io);
    }
    if (touches_low(phases.back())) {
      PlanPhase io;
      complete(io);
      phases.push_back(io);
    }
    ops.clear();
    h.nphase = (int)phases.size();
    for (int p = 0; p < h.nphase; ++p) {
      PlanPhase &ph = phases[p];
      for (int j = 0; j < A; ++j) h.ph[p].reg[j] = (uint8_t)lbit[ph.order[j]];
      h.ph[p].op_begin = (uint16_t)ops.size();
      for (int gi : ph.gates) encode_gate<R>(gates[gi], ph, ops);
      h.ph[p].op_end = (uint16_t)ops.size();
    }
    // one upload: the header, padded to 256 bytes, then the op records
    const size_t hbytes = (sizeof(SweepHeader) + 255) & ~(size_t)255;
    blob.assign(hbytes + ops.size() * sizeof(SweepOp<R>), 0);
    std::memcpy(blob.data(), &h, sizeof(SweepHeader));
    if (!ops.empty()) std::memcpy(blob.data() + hbytes, ops.data(), ops.size() * sizeof(SweepOp<R>));
    void *d_blob = nullptr;
    hipError_t e = run.upload(blob.data(), blob.size(), &d_blob);
    if (e != hipSuccess) return e;
    if (run.mark) run.mark(true);
    e = launch_sweep_t<R>(T, run, (const SweepHeader *)d_blob,
                          (const SweepOp<R> *)((const char *)d_blob + hbytes));
    if (run.mark) run.mark(false);
    if (e != hipSuccess) return e;
    ++*sweeps;
  }
  (void)kernels;
  return hipSuccess;
}

}  // namespace

struct SweepEngine {
  int dummy = 0;
};

SweepEngine *sweep_engine_new() { return new SweepEngine(); }
void sweep_engine_free(SweepEngine *e) { delete e; }

int sweep_tile_bits(bool c64) {
  const int t = env_int(c64 ? "QS_SWEEP_T_C64" : "QS_SWEEP_T_C128", c64 ? 12 : 11);
  return std::min(12, std::max(10, t));
}

hipError_t sweep_execute(SweepEngine *eng, SweepRunner &run, const std::vector<Gate> &gates,
                         uint64_t *sweeps, uint64_t *kernels) {
  (void)eng;
  if (run.dtype_c64) return execute_t<float>(run, gates, sweeps, kernels);
  return execute_t<double>(run, gates, sweeps, kernels);
}

}  // namespace qs

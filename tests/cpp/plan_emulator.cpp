// CPU emulation of the fused sweep, for checks of the sweep planner.
//
// emu_run() classifies gate matrices with the engine's own classifier
// (qs::classify_matrix), runs the engine's planner (qs::sweep_execute) with a
// host hook in place of the kernel launch, and executes every sweep
// descriptor the planner produced on a host copy of the state, decoding it
// the way k_sweep does: dispatch on the op code alone, slot -> qubit through
// the header's register and lane bit tables, and each op's pattern evaluated
// only on the bits that op's kernel form looks at.  Gates the planner sends
// to per-gate kernels are applied by a plain host loop.  The descriptor is
// also checked for consistency (tile qubit tables, pool bounds).
//
// Built by tests/test_plan_emulation.py against quaru_amd/libquaru_hip.so.
#include <hip/hip_runtime.h>

#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "qs_internal.h"
#include "qs_sweep.h"
#include "qs_sweep_kernel.h"

using namespace qs;
using namespace qs::sweep;
using C = std::complex<double>;

namespace {

std::string g_err;

void apply_gate_host(std::vector<C> &psi, int n, const Gate &g) {
  const uint64_t N = 1ull << n;
  if (g.kind == Gate::kDense && g.bmask) {  // batched 2x2: matrix of state b
    const uint64_t bt = 1ull << g.act[0];
    for (uint64_t i = 0; i < N; ++i) {
      if (i & bt) continue;
      const qs::cd *m = g.m.data() + 4 * ((i >> g.bshift) & g.bmask);
      const C x = psi[i], y = psi[i | bt];
      psi[i] = C(m[0].re, m[0].im) * x + C(m[1].re, m[1].im) * y;
      psi[i | bt] = C(m[2].re, m[2].im) * x + C(m[3].re, m[3].im) * y;
    }
    return;
  }
  if (g.kind == Gate::kDense) {
    const int K = g.na, D = 1 << K;
    uint64_t am = 0;
    for (int j = 0; j < K; ++j) am |= 1ull << g.act[j];
    std::vector<uint64_t> idx(D);
    std::vector<C> x(D);
    for (uint64_t b = 0; b < N; ++b) {
      if ((b & am) || (b & g.cmask) != g.cval) continue;
      for (int i = 0; i < D; ++i) {
        uint64_t p = b;
        for (int j = 0; j < K; ++j)
          if ((i >> j) & 1) p |= 1ull << g.act[j];
        idx[i] = p;
        x[i] = psi[p];
      }
      for (int i = 0; i < D; ++i) {
        C s = 0;
        for (int j = 0; j < D; ++j) s += C(g.m[(size_t)i * D + j].re, g.m[(size_t)i * D + j].im) * x[j];
        psi[idx[i]] = s;
      }
    }
  } else if (g.kind == Gate::kDiag) {
    for (uint64_t i = 0; i < N; ++i) {
      if ((i & g.cmask) != g.cval) continue;
      size_t e = 0;
      for (int j = 0; j < g.na; ++j) e |= (size_t)((i >> g.act[j]) & 1) << j;
      psi[i] *= C(g.m[e].re, g.m[e].im);
    }
  } else if (g.kind == Gate::kSparse) {
    for (uint64_t i = 0; i < N; ++i)
      for (size_t e = 0; e < g.spat.size(); ++e)
        if ((i & g.smask) == g.spat[e]) psi[i] *= C(g.m[e].re, g.m[e].im);
  } else if (g.kind == Gate::kPerm) {
    const std::vector<C> src = psi;
    uint64_t am = 0;
    for (int j = 0; j < g.na; ++j) am |= 1ull << g.act[j];
    for (uint64_t i = 0; i < N; ++i) {
      if ((i & g.cmask) != g.cval) continue;
      uint32_t r = 0;
      for (int j = 0; j < g.na; ++j) r |= (uint32_t)((i >> g.act[j]) & 1) << j;
      uint64_t s = i & ~am;
      for (int j = 0; j < g.na; ++j)
        if ((g.perm[r] >> j) & 1) s |= 1ull << g.act[j];
      psi[i] = C(g.m[r].re, g.m[r].im) * src[s];
    }
  }
}

bool fail(const std::string &m) {
  if (g_err.empty()) g_err = m;
  return false;
}

uint64_t *g_hist = nullptr;  // op count per code, when set

template <typename R>
bool emulate(std::vector<C> &psi, int n, int A, const void *blob, size_t bytes) {
  SweepHeader h;
  std::memcpy(&h, blob, sizeof h);
  const size_t hb = (sizeof(SweepHeader) + 15) & ~(size_t)15;
  if (blob_bytes<R>((int)h.nops, (int)h.npool) != bytes) return fail("blob size");
  std::vector<OpRec> ops(h.nops);
  std::memcpy(ops.data(), (const char *)blob + hb, h.nops * sizeof(OpRec));
  std::vector<R> pool(h.npool);
  std::memcpy(pool.data(), (const char *)blob + hb + h.nops * sizeof(OpRec), h.npool * sizeof(R));
  if (h.nwave != 0 && h.nwave != (uint32_t)kWaveBits) return fail("wave bit count");
  const int W = (int)h.nwave;
  const int T = kLaneBits + A + W;
  uint64_t tmask = 0, smask = 0, regmask = 0;
  for (int i = 0; i < T; ++i) {
    if (i > 0 && h.tpos[i] <= h.tpos[i - 1]) return fail("tile qubits not ascending");
    if ((int)h.tpos[i] >= n) return fail("tile qubit out of range");
    tmask |= 1ull << h.tpos[i];
  }
  for (int i = 0; i < kLaneBits; ++i) {
    if (__builtin_popcountll(h.lg[i]) != 1) return fail("lane bit table");
    smask |= h.lg[i];
  }
  for (int j = 0; j < A; ++j) {
    if (__builtin_popcountll(h.gb[j]) != 1) return fail("register bit table");
    smask |= h.gb[j];
    regmask |= h.gb[j];
  }
  for (int k = 0; k < W; ++k) {
    if (__builtin_popcountll(h.wb[k]) != 1) return fail("wave bit table");
    smask |= h.wb[k];
  }
  if (smask != tmask || __builtin_popcountll(tmask) != T) return fail("tile tables disagree");
  if (h.ntiles != (1ull << (n - T))) return fail("tile count");
  const bool plan_only = psi.empty();
  const uint64_t N = plan_only ? 0 : 1ull << n;
  auto val = [&](uint32_t off, int k) {
    return C((double)pool[off + 2 * k], (double)pool[off + 2 * k + 1]);
  };
  // an op with pattern bits off the tile runs as `code` (values at moff +
  // uoff) on the tiles where they hold and as `code_fail` (values at moff)
  // on the others
  uint64_t sel_mask = 0, sel_val = 0;
  bool sel_want = true;
  auto sel = [&](uint64_t i) { return (((i ^ sel_val) & sel_mask) == 0) == sel_want; };
  for (const OpRec &o0 : ops) {
    if (g_hist && o0.code < 128) ++g_hist[o0.code];
    if (o0.umask & tmask) return fail("uniform pattern bits on the tile");
    if (o0.umask == 0 && (o0.code_fail != 0 || o0.uoff != 0)) return fail("stray fail code");
    for (int pass = 0; pass < (o0.umask ? 2 : 1); ++pass) {
    OpRec o = o0;
    sel_mask = o0.umask;
    sel_val = o0.cval;
    sel_want = pass == 0;
    if (o0.umask) {
      if (pass == 0)
        o.moff += o0.uoff;
      else
        o.code = o0.code_fail;
    }
    const uint32_t code = o.code;
    if (code == CODE_NONE) continue;
    if (code == CODE_NEG) {
      if (!plan_only && (size_t)o.moff + 2 > pool.size()) return fail("pool bounds");
      for (uint64_t i = 0; i < N; ++i)
        if (sel(i) && (i & o.cmask) == (o.cval & o.cmask)) psi[i] = -psi[i];
      continue;
    }
    int kind = 0, j = 0;
    int lane_code = -1;  // lane bit a lane-target code names
    if (code >= CODE_U1 && code < CODE_U1 + kLaneCodes) kind = OP_U1, j = code - CODE_U1;
    else if (code >= CODE_C1 && code < CODE_C1 + kLaneCodes) kind = OP_C1, j = code - CODE_C1;
    else if (code >= CODE_X1 && code < CODE_X1 + kLaneCodes) kind = OP_X1, j = code - CODE_X1;
    else if (code == CODE_SPARSE) kind = OP_SPARSE;
    else if (code == CODE_DIAG) kind = OP_DIAG;
    else if (code >= CODE_DENSE2 && code < CODE_MUXL) kind = OP_DENSE2, j = code - CODE_DENSE2;
    else if (code >= CODE_MUXL && code < CODE_MUXL + kLaneCodes) kind = 100, j = code - CODE_MUXL;
    else if (code >= CODE_MUXR && code < CODE_MUXR + kMaxA) kind = 101, j = code - CODE_MUXR;
    else if (code >= CODE_MUXRL && code < CODE_MUXRL + kMaxA * kLaneBits)
      kind = 102, j = (code - CODE_MUXRL) / kLaneBits, lane_code = (code - CODE_MUXRL) % kLaneBits;
    else if (code == CODE_U1W) kind = OP_U1, j = -1;
    else if (code == CODE_C1W) kind = OP_C1, j = -1;
    else if (code == CODE_X1W) kind = OP_X1, j = -1;
    else if (code == CODE_MUXW) kind = 100, j = -1;
    else return fail("unknown op code " + std::to_string(code));
    if (kind != 102 && kind != OP_DENSE2 && j >= kMaxA) lane_code = j - kMaxA, j = kMaxA;
    if (lane_code >= 0 && lane_code != (int)o.s0 - (int)kLaneSlot)
      return fail("lane bit of the code and of the op slot disagree");
    const uint64_t cm = o.cmask;
    const uint64_t cv = o.cval & cm;
    auto slot_qubit = [&](int jj, uint32_t s0) -> int {
      if (jj < 0) {  // wave bit
        const int k = (int)s0 - (int)kWaveSlot;
        if (k < 0 || k >= W) return -3;
        return __builtin_ctzll(h.wb[k]);
      }
      if (jj < kMaxA) {
        if (jj >= A) return -1;  // the kernel's switch skips such codes
        return __builtin_ctzll(h.gb[jj]);
      }
      const int k = (int)s0 - (int)kLaneSlot;
      if (k < 0 || k >= kLaneBits) return -2;
      return __builtin_ctzll(h.lg[k]);
    };
    auto need_pool = [&](uint32_t cnt) {
      return (size_t)o.moff + cnt <= pool.size() ? true : fail("pool bounds");
    };
    if (kind == 102) {
      // lane target; the pattern is register slot j's bit alone
      if (j >= A) return fail("pattern slot has no qubit");
      const int t = slot_qubit(kMaxA, o.s0);
      if (t < 0) return fail("op slot has no qubit");
      if (!need_pool(16)) return false;
      const uint64_t bt = 1ull << t, bj = h.gb[j];
      C m0[4], m1[4];
      for (int k = 0; k < 4; ++k) {
        m0[k] = val(o.moff, k);
        m1[k] = val(o.moff, 4 + k);
      }
      for (uint64_t i = 0; i < N; ++i) {
        if ((i & bt) || !sel(i)) continue;
        const bool on = (i & bj) == (o.cval & bj);
        const C *m = on ? m1 : m0;
        const C x = psi[i], y = psi[i | bt];
        psi[i] = m[0] * x + m[1] * y;
        psi[i | bt] = m[2] * x + m[3] * y;
      }
      continue;
    }
    if (kind == OP_U1 || kind == OP_C1 || kind == OP_X1 || kind == 100 || kind == 101) {
      const int t = slot_qubit(j, o.s0);
      if (t < 0) return fail("op slot has no qubit");
      const uint64_t bt = 1ull << t;
      if (cm & bt) return fail("pattern holds the target");
      C m0[4], m1[4];
      if ((kind == OP_U1 || kind == OP_C1) && o.bmask) {
        // batched 2x2: the batch bits must stay off the tile
        if (!need_pool(8 * (o.bmask + 1))) return false;
        if (tmask & ((uint64_t)o.bmask << o.bshift)) return fail("batch bit on the tile");
        for (uint64_t i = 0; i < N; ++i) {
          if ((i & bt) || !sel(i) || (kind == OP_C1 && (i & cm) != cv)) continue;
          const uint32_t mo = o.moff + 8 * (uint32_t)((i >> o.bshift) & o.bmask);
          const C x = psi[i], y = psi[i | bt];
          psi[i] = val(mo, 0) * x + val(mo, 1) * y;
          psi[i | bt] = val(mo, 2) * x + val(mo, 3) * y;
        }
        continue;
      }
      if (kind == OP_U1 || kind == OP_C1) {
        if (!need_pool(8)) return false;
        for (int k = 0; k < 4; ++k) m0[k] = m1[k] = val(o.moff, k);
      } else if (kind == OP_X1) {
        m1[0] = m1[3] = 0;
        m1[1] = m1[2] = 1;
      } else {
        if (!need_pool(16)) return false;
        for (int k = 0; k < 4; ++k) {
          m0[k] = val(o.moff, k);
          m1[k] = val(o.moff, 4 + k);
        }
      }
      for (uint64_t i = 0; i < N; ++i) {
        if ((i & bt) || !sel(i)) continue;
        bool on;
        if (kind == OP_U1) on = true;
        else if (kind == 100) on = (i & cm & ~regmask) == (cv & ~regmask);  // lane_ok only
        else if (kind == 101) on = (i & cm & regmask) == (cv & regmask);    // register part only
        else on = (i & cm) == cv;
        const bool mux = kind == 100 || kind == 101;
        if (!on && !mux) continue;  // controlled forms: identity where the pattern fails
        const C *m = on ? m1 : m0;
        const C x = psi[i], y = psi[i | bt];
        psi[i] = m[0] * x + m[1] * y;
        psi[i | bt] = m[2] * x + m[3] * y;
      }
    } else if (kind == OP_DENSE2) {
      const int s0 = j / kMaxA, s1 = j % kMaxA;
      if (s0 >= A || s1 >= A || s0 == s1) return fail("dense2 slots");
      if (!need_pool(32)) return false;
      const uint64_t b0 = h.gb[s0], b1 = h.gb[s1];
      for (uint64_t i = 0; i < N; ++i) {
        if ((i & b0) || (i & b1) || !sel(i) || (i & cm) != cv) continue;
        const uint64_t id[4] = {i, i | b0, i | b1, i | b0 | b1};
        C x[4], y[4];
        for (int k = 0; k < 4; ++k) x[k] = psi[id[k]];
        for (int r = 0; r < 4; ++r) {
          y[r] = 0;
          for (int k = 0; k < 4; ++k) y[r] += val(o.moff, 4 * r + k) * x[k];
        }
        for (int k = 0; k < 4; ++k) psi[id[k]] = y[k];
      }
    } else if (kind == OP_SPARSE) {
      if (!need_pool(2)) return false;
      const C v = val(o.moff, 0);
      for (uint64_t i = 0; i < N; ++i)
        if (sel(i) && (i & cm) == cv) psi[i] *= v;
    } else if (kind == OP_DIAG) {
      if (o.na < 1 || o.na > 2) return fail("diag arity");
      if (!need_pool(2u << o.na)) return false;
      const int q0 = (int)o.act[0], q1 = o.na > 1 ? (int)o.act[1] : (int)o.act[0];
      for (uint64_t i = 0; i < N; ++i) {
        if (!sel(i) || (i & cm) != cv) continue;
        const int b0 = (i >> q0) & 1, b1 = o.na > 1 ? (int)((i >> q1) & 1) : 0;
        psi[i] *= val(o.moff, b0 + 2 * b1);
      }
    }
    }
  }
  return true;
}

}  // namespace

extern "C" {

// psi: 2^n complex128 values, {re, im} interleaved, updated in place.  Gate
// g has k[g] targets (concatenated in `targets`) and a 2^k x 2^k complex128
// row-major matrix (concatenated in `mats`).  stats[0..3] = sweeps, gates
// run by the host per-gate path, sweep op records, merged gates.  Returns 0,
// or 1 with a message in err.
// psi == NULL: plan only (descriptors are decoded and checked, no state).
// hist (128 entries, may be NULL) counts the sweep op records per op code.
// cmasks/cvals (may be NULL): controls of each gate beyond its matrix, as
// qs_apply_controlled passes them to the classifier.
// emu_run2 also takes batched 2x2s: gate g with bmasks[g] != 0 (and k = 1)
// carries bmasks[g] + 1 matrices (state b = (index >> bshift) & bmask uses
// matrix b), as the adjoint's forward pass builds them.
int emu_run2(int n, int c64, double *psi, int ngates, const double *mats, const uint64_t *targets,
             const int *ks, const uint64_t *cmasks, const uint64_t *cvals, const uint32_t *bmasks,
             int bshift, uint64_t *stats, uint64_t *hist, char *err, int errlen) {
  g_err.clear();
  g_hist = hist;
  std::vector<C> st(psi ? 1ull << n : 0);
  for (size_t i = 0; i < st.size(); ++i) st[i] = C(psi[2 * i], psi[2 * i + 1]);
  std::vector<Gate> gates;
  size_t moff = 0, toff = 0;
  for (int g = 0; g < ngates; ++g) {
    const int k = ks[g];
    std::vector<size_t> t(targets + toff, targets + toff + k);
    const uint32_t bm = bmasks ? bmasks[g] : 0u;
    if (bm && k == 1) {
      Gate G;
      G.kind = Gate::kDense;
      G.na = 1;
      G.act[0] = (int)t[0];
      G.bshift = bshift;
      G.bmask = bm;
      G.m.resize(4 * ((size_t)bm + 1));
      for (size_t e = 0; e < G.m.size(); ++e) G.m[e] = qs::cd{mats[moff + 2 * e], mats[moff + 2 * e + 1]};
      moff += 8 * ((size_t)bm + 1);
      toff += 1;
      gates.push_back(std::move(G));
      continue;
    }
    Gate G = qs::classify_matrix(mats + moff, t.data(), (size_t)k, cmasks ? cmasks[g] : 0,
                                 cvals ? cvals[g] : 0);
    if (G.kind != Gate::kIdentity) gates.push_back(std::move(G));
    moff += (size_t)2 << (2 * k);
    toff += (size_t)k;
  }
  const int A = qs::sweep_reg_bits(c64 != 0);
  uint64_t singles = 0, sweep_ops = 0, merged = 0, sweeps = 0, kernels = 0;
  SweepRunner run;
  run.n = n;
  run.dtype_c64 = c64 != 0;
  run.single = [&](const Gate &g) {
    if (!st.empty()) apply_gate_host(st, n, g);
    ++singles;
    return hipSuccess;
  };
  run.host_sweep = [&](const void *blob, size_t bytes) {
    const bool ok = c64 ? emulate<float>(st, n, A, blob, bytes) : emulate<double>(st, n, A, blob, bytes);
    return ok ? hipSuccess : hipErrorInvalidValue;
  };
  run.sweep_ops = &sweep_ops;
  run.merged = &merged;
  hipError_t e = qs::sweep_execute(nullptr, run, gates, &sweeps, &kernels);
  stats[0] = sweeps;
  stats[1] = singles;
  stats[2] = sweep_ops;
  stats[3] = merged;
  if (e != hipSuccess || !g_err.empty()) {
    std::snprintf(err, (size_t)errlen, "%s", g_err.empty() ? "planner error" : g_err.c_str());
    return 1;
  }
  for (size_t i = 0; psi && i < st.size(); ++i) {
    psi[2 * i] = st[i].real();
    psi[2 * i + 1] = st[i].imag();
  }
  return 0;
}

int emu_run(int n, int c64, double *psi, int ngates, const double *mats, const uint64_t *targets,
            const int *ks, const uint64_t *cmasks, const uint64_t *cvals, uint64_t *stats,
            uint64_t *hist, char *err, int errlen) {
  return emu_run2(n, c64, psi, ngates, mats, targets, ks, cmasks, cvals, nullptr, 0, stats, hist,
                  err, errlen);
}

}  // extern "C"

// qs_sweep_kernel.h -- device side of the fused sweep engine: the sweep
// descriptor layout and the sweep kernel.  Included by qs_sweep.hip only.
//
// A sweep is one pass over the state.  A tile is owned by one wave (64
// lanes, 2^(6+A) amplitudes) or by the 4 waves of a workgroup (2^(8+A)
// amplitudes, the wave index supplying two wave bits); every lane keeps 2^A
// amplitudes in registers.  The tile index bits are assigned as
//   lane bits 0..b-1 <- qubits 0..b-1 (2^b lanes read 2^b consecutive
//                       amplitudes: b = 3 for complex128 and b = 4 for
//                       complex64 by default, 128-byte runs),
//   lane bits b..5   <- more tile qubits chosen per sweep,
//   register slots  <- A more tile qubits,
//   wave bits        <- 0 or 2 more tile qubits,
// and the remaining index bits come from the tile id.  A gate on a register
// qubit pairs registers inside a lane; a gate on a lane qubit fetches the
// partner lane's amplitude with VALU cross-lane moves (DPP, permlane swaps;
// lane_xor) and each lane computes its own output; a gate on a wave bit
// exchanges with the partner wave through LDS (op_wave).  The tile itself
// never goes through LDS: a sweep is load, ops, store.
//
// Controls are split into the part on register slots, which is the same for
// the whole wave (scalar arithmetic), and the part on every other index bit,
// which is one predicate per lane and op.
//
// The descriptor is read through the constant address space with 32/64-bit
// fields only, so the compiler keeps every descriptor read on scalar loads.
// Index arithmetic runs in 32 bits when the state has at most 2^32
// amplitudes (I = uint32_t) and in 64 bits otherwise.  The state is streamed
// with nontemporal loads and stores: every amplitude is touched once per
// sweep, so caching it would only evict other data.
// Amplitudes live in per-lane register arrays re[]/im[] of plain scalars
// (complex structs selected as a whole would be spilled to scratch).
#pragma once

#ifdef __HIPCC_RTC__
// compiled at run time (hipRTC) for one sweep's op list: the runtime's own
// headers, fixed-width types from its namespace
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
#else
#include <hip/hip_runtime.h>

#include <type_traits>

#include "qs_internal.h"
#endif

#define QS_CONST __attribute__((address_space(4)))

namespace qs {
namespace sweep {

constexpr int kThreads = 256;  // 4 waves per workgroup
constexpr int kWave = 64;
constexpr int kLaneBits = 6;
// Wave bits: with kWaveBits = 2 the 4 waves of a workgroup share one tile of
// 2^(6+A+2) amplitudes, the wave index supplying two more tile qubits
// (gates on them exchange through LDS); with 0 each wave has a tile of its
// own.
constexpr int kWaveBits = 2;
constexpr int kMaxA = 5;

enum : uint32_t {
  OP_U1 = 1,      // uncontrolled 2x2 on one qubit (register slot or lane bit)
  OP_C1 = 2,      // controlled 2x2 on one qubit
  OP_X1 = 3,      // (controlled) exchange: Pauli X / CNOT / Toffoli targets
  OP_DENSE2 = 4,  // 4x4 on two register slots (matrix bit 0 <-> s0)
  OP_SPARSE = 5,  // multiply by pool value where (idx & cmask) == cval
  OP_DIAG = 6,    // multiply by pool[pattern of 1-2 act bits] where (idx & cmask) == cval
  OP_MUX = 7      // 2x2 on one qubit: pool[0..3] where (idx & cmask) != cval, pool[4..7] where ==
};

// Slot encoding in s0/s1: 0..A-1 register slot, kLaneSlot + k lane bit k,
// kWaveSlot + k wave bit k.
constexpr uint32_t kLaneSlot = 8;
constexpr uint32_t kWaveSlot = 16;

// The kernel's dispatch code: type, register slots and lane bit in one
// number, so every case runs with its slots and lane bit known at compile
// time.  U1/C1/X1/MUXL: base + slot j (j < kMaxA) or base + kMaxA + lane
// bit; DENSE2: base + s0 * kMaxA + s1.
constexpr uint32_t kLaneCodes = kMaxA + kLaneBits;
enum : uint32_t {
  CODE_NONE = 0,
  CODE_U1 = 1,                      // 1 .. 11
  CODE_C1 = CODE_U1 + kLaneCodes,   // 12 .. 22
  CODE_X1 = CODE_C1 + kLaneCodes,   // 23 .. 33
  CODE_SPARSE = CODE_X1 + kLaneCodes,
  CODE_DIAG = CODE_SPARSE + 1,
  CODE_DENSE2 = CODE_DIAG + 1,      // + s0 * kMaxA + s1
  // MUX whose pattern has no register-slot bit: the matrix is chosen once
  // per lane.  + slot or kMaxA + lane bit
  CODE_MUXL = CODE_DENSE2 + kMaxA * kMaxA,
  // MUX on a register slot whose pattern lies on register slots only: the
  // matrix is chosen per register pair, uniformly over the wave.  + slot
  CODE_MUXR = CODE_MUXL + kLaneCodes,
  // MUX on a lane qubit whose pattern is one register slot JC:
  // + JC * kLaneBits + lane bit
  CODE_MUXRL = CODE_MUXR + kMaxA,
  // ops on a wave bit (the slot's wave bit is s0 - kWaveSlot)
  CODE_U1W = CODE_MUXRL + kMaxA * kLaneBits,
  CODE_C1W = CODE_U1W + 1,
  CODE_X1W = CODE_C1W + 1,
  CODE_MUXW = CODE_X1W + 1,  // pattern on no register slot
  // multiply by -1 where (idx & cmask) == cval (a sparse op whose value is
  // -1: CZ, CCZ and their kin): sign flips instead of complex products
  CODE_NEG = CODE_MUXW + 1
};

// Kernel forms of a multiplexed 2x2 (OP_MUX), by where its pattern lies.
enum : uint32_t {
  MUX_NONE = 0,    // no kernel form
  MUX_LANES = 1,   // pattern on no register slot; target anywhere
  MUX_REGS = 2,    // pattern on register slots only; target a register slot
  MUX_ONE_REG = 3  // pattern one register slot (s1); target a lane qubit
};

// `mux_form`: one of MUX_LANES, MUX_REGS, MUX_ONE_REG for OP_MUX.
__host__ __device__ inline uint32_t op_code(uint32_t type, uint32_t s0, uint32_t s1,
                                            uint32_t mux_form = MUX_LANES) {
  const uint32_t j0 = s0 >= kLaneSlot ? (uint32_t)kMaxA + (s0 - kLaneSlot) : s0;
  if (s0 >= kWaveSlot) {
    switch (type) {
      case OP_U1: return CODE_U1W;
      case OP_C1: return CODE_C1W;
      case OP_X1: return CODE_X1W;
      case OP_MUX: return mux_form == MUX_LANES ? CODE_MUXW : CODE_NONE;
      default: return CODE_NONE;
    }
  }
  switch (type) {
    case OP_MUX:
      if (mux_form == MUX_ONE_REG) return CODE_MUXRL + s1 * kLaneBits + (s0 - kLaneSlot);
      return (mux_form == MUX_REGS ? CODE_MUXR : CODE_MUXL) + j0;
    case OP_U1: return CODE_U1 + j0;
    case OP_C1: return CODE_C1 + j0;
    case OP_X1: return CODE_X1 + j0;
    case OP_SPARSE: return CODE_SPARSE;
    case OP_DIAG: return CODE_DIAG;
    case OP_DENSE2: return CODE_DENSE2 + s0 * kMaxA + s1;  // both register slots
    default: return CODE_NONE;
  }
}

// 80-byte op record; complex values live in a pool of R scalars
// ({re, im} interleaved, row-major) at pool[moff].
// A batched 2x2 (bmask != 0) takes its matrix at pool[moff + 8 b] for the
// tile's batch index b = (tile base >> bshift) & bmask; the batch bits are
// never tile qubits, so b is uniform over the tile.
// Pattern bits off the tile (umask) are the same for every amplitude of a
// tile: where (tile base & umask) == (cval & umask) the op runs as `code`
// with its values at pool[moff + uoff], elsewhere as `code_fail` (CODE_NONE:
// skipped) at pool[moff].  cmask then holds the pattern's tile bits only.
struct OpRec {
  uint32_t type, s0, s1, na;
  uint32_t act[4];
  uint32_t moff, code;
  uint64_t cmask, cval;
  uint32_t bshift, bmask;
  uint64_t umask;
  uint32_t code_fail, uoff;
};

struct SweepHeader {
  uint32_t tpos[16];  // the tile qubits, ascending (tile-id bits fill the rest)
  uint32_t nops, npool;
  uint64_t ntiles;
  uint64_t gb[8];     // global bit of register slot j
  uint64_t lg[8];     // global bit of lane bit i (i < 6)
  uint64_t wb[2];     // global bit of wave bit k (k < nwave)
  uint32_t nwave, pad;
};

template <typename R>
__host__ __device__ inline size_t blob_bytes(int nops, int npool) {
  const size_t h = (sizeof(SweepHeader) + 15) & ~(size_t)15;
  const size_t o = (size_t)nops * sizeof(OpRec);
  const size_t p = ((size_t)npool * sizeof(R) + 15) & ~(size_t)15;
  return h + o + p;
}

// Native two-element vectors for the nontemporal builtins.
template <typename R>
struct SNT {
  typedef R T __attribute__((ext_vector_type(2)));
};

template <typename R>
struct SV2;
template <>
struct SV2<double> {
  using T = double2;
};
template <>
struct SV2<float> {
  using T = float2;
};

// Base index of a tile: insert a zero at each of the T tile positions.
template <int T>
__device__ __forceinline__ uint64_t tile_base(uint64_t p, const QS_CONST SweepHeader &h) {
#pragma unroll
  for (int i = 0; i < T; ++i) {
    const uint64_t lo = p & ((1ull << h.tpos[i]) - 1ull);
    p = ((p ^ lo) << 1) | lo;
  }
  return p;
}

template <int A, typename I>
struct Regs {
  I gb[A];  // global bit of register slot j (wave-uniform)
  I mask;   // OR of gb[]
};

template <int A, typename I>
__device__ __forceinline__ I goff(const Regs<A, I> &g, int r) {
  I o = 0;
#pragma unroll
  for (int j = 0; j < A; ++j)
    if ((r >> j) & 1) o |= g.gb[j];
  return o;
}

template <typename I>
struct Ctx {
  I gbase;  // tile base | this lane's lane-bit and wave-bit offsets (no register bits)
  I tbase;  // tile base (uniform)
  int lane;
  int wave;  // wave index in the workgroup
};

// on(r) = lane_ok && ((goff(r) & cmr) == cvr): the second factor is uniform
// (in a specialised kernel a constant, see make_pred).
template <typename I>
struct Pred {
  bool lane_ok;
  I cmr, cvr;
};

// RP >= 0: the op's pattern on the register slots known when the kernel
// was generated, in slot space (register r meets it if (r & rm) == rv, rm in
// bits 0..7 of RP, rv in bits 8..15), so that the per-register tests fold
// away; RP < 0: the pattern's global bits, compared with each register's
// offset at run time.
template <int A, int RP, typename I>
__device__ __forceinline__ Pred<I> make_pred(const QS_CONST OpRec &op, const Ctx<I> &c,
                                             const Regs<A, I> &g) {
  const I cm = (I)op.cmask, cv = (I)op.cval;
  const I cmo = cm & ~g.mask;
  Pred<I> p;
  if constexpr (RP < 0) {
    p.cmr = cm & g.mask;
    p.cvr = cv & g.mask;
  } else {
    p.cmr = 0;
    p.cvr = 0;
  }
  p.lane_ok = (c.gbase & cmo) == (cv & cmo);
  return p;
}

template <int A, int RP, typename I>
__device__ __forceinline__ bool reg_on(const Pred<I> &p, const Regs<A, I> &g, int r) {
  if constexpr (RP >= 0)
    return ((uint32_t)r & ((uint32_t)RP & 0xFFu)) == (((uint32_t)RP >> 8) & 0xFFu);
  else
    return (goff<A>(g, r) & p.cmr) == p.cvr;
}

template <int A, int RP, typename I>
__device__ __forceinline__ bool pred_on(const Pred<I> &p, const Regs<A, I> &g, int r) {
  return p.lane_ok && reg_on<A, RP>(p, g, r);
}

// Partner lane's value (lane ^ 2^K) by VALU cross-lane instructions: DPP
// for lane bits 0..3 (bit 2 takes two moves: row_half_mirror is lane ^ 7 in
// groups of 8, quad_perm [3,2,1,0] is lane ^ 3), the gfx950 half swaps for
// bits 4 and 5 (swapping a value with itself leaves the lower half's values
// in the first result and the upper half's in the second).  ds_bpermute,
// which these replace, goes through the CU's LDS crossbar: measured 24
// cycles per wave instruction per SIMD at full occupancy, against 4.5 for
// these.
template <int K>
__device__ __forceinline__ int lane_xor(int v, bool hi) {
  static_assert(K >= 0 && K < kLaneBits, "lane bit");
  if constexpr (K == 0) {
    return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
  } else if constexpr (K == 1) {
    return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
  } else if constexpr (K == 2) {
    const int m = __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    return __builtin_amdgcn_mov_dpp(m, 0x1B, 0xF, 0xF, false);         // quad_perm [3,2,1,0]
  } else if constexpr (K == 3) {
    return __builtin_amdgcn_mov_dpp(v, 0x128, 0xF, 0xF, false);  // row_ror:8
  } else if constexpr (K == 4) {
    const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return hi ? p[0] : p[1];
  } else {
    const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return hi ? p[0] : p[1];
  }
}

template <int K>
__device__ __forceinline__ double lane_xor(double v, bool hi) {
  const long long b = __double_as_longlong(v);
  const int lo = lane_xor<K>((int)(b & 0xffffffffll), hi);
  const int h = lane_xor<K>((int)(b >> 32), hi);
  return __longlong_as_double(((long long)h << 32) | (unsigned int)lo);
}

template <int K>
__device__ __forceinline__ float lane_xor(float v, bool hi) {
  return __int_as_float(lane_xor<K>(__float_as_int(v), hi));
}

// The two amplitudes of a lane's pair on lane bit K.  K < 4: the lane's own
// amplitude (u) and its partner's (w), one DPP move.  K = 4, 5
// (kBitOrder<K>): the amplitude with the bit clear (u) and the one with it
// set (w), which the half swap of a value with itself hands every lane
// directly; ops then weight them with row `hi` of their matrix, so no
// per-register select is needed.
template <int K>
constexpr bool kBitOrder = K >= 4;

template <int K>
__device__ __forceinline__ void lane_pair(int v, bool hi, int &u, int &w) {
  if constexpr (K < 4) {
    u = v;
    w = lane_xor<K>(v, hi);
  } else if constexpr (K == 4) {
    const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    u = p[0];
    w = p[1];
  } else {
    const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    u = p[0];
    w = p[1];
  }
}

template <int K>
__device__ __forceinline__ void lane_pair(double v, bool hi, double &u, double &w) {
  const long long b = __double_as_longlong(v);
  int ul, uh, wl, wh;
  lane_pair<K>((int)(b & 0xffffffffll), hi, ul, wl);
  lane_pair<K>((int)(b >> 32), hi, uh, wh);
  u = __longlong_as_double(((long long)uh << 32) | (unsigned int)ul);
  w = __longlong_as_double(((long long)wh << 32) | (unsigned int)wl);
}

template <int K>
__device__ __forceinline__ void lane_pair(float v, bool hi, float &u, float &w) {
  int ui, wi;
  lane_pair<K>(__float_as_int(v), hi, ui, wi);
  u = __int_as_float(ui);
  w = __int_as_float(wi);
}

// Lane bits 4 and 5 by row transposition: half_swap<K>(x, y) swaps the
// K-bit-set lanes of x with the K-bit-clear lanes of y (gfx950's
// v_permlane16_swap / v_permlane32_swap of two different registers).
// Afterwards every lane holds in (x, y) the amplitude pair (bit K clear,
// bit K set) of register x's values if its own bit K is clear, of register
// y's if it is set; applying it again to the pair's two outputs hands every
// lane its own new values of both registers.  Two swaps per 32 bits per two
// registers, with no per-lane coefficient select: each lane computes both
// outputs of one pair.
template <int K>
__device__ __forceinline__ void half_swap(int &x, int &y) {
  static_assert(K == 4 || K == 5, "half swaps move lane bits 4 and 5");
  if constexpr (K == 4) {
    const auto p = __builtin_amdgcn_permlane16_swap(x, y, false, false);
    x = p[0];
    y = p[1];
  } else {
    const auto p = __builtin_amdgcn_permlane32_swap(x, y, false, false);
    x = p[0];
    y = p[1];
  }
}

template <int K>
__device__ __forceinline__ void half_swap(double &x, double &y) {
  const long long bx = __double_as_longlong(x), by = __double_as_longlong(y);
  int xl = (int)(bx & 0xffffffffll), xh = (int)(bx >> 32);
  int yl = (int)(by & 0xffffffffll), yh = (int)(by >> 32);
  half_swap<K>(xl, yl);
  half_swap<K>(xh, yh);
  x = __longlong_as_double(((long long)xh << 32) | (unsigned int)xl);
  y = __longlong_as_double(((long long)yh << 32) | (unsigned int)yl);
}

template <int K>
__device__ __forceinline__ void half_swap(float &x, float &y) {
  int xi = __float_as_int(x), yi = __float_as_int(y);
  half_swap<K>(xi, yi);
  x = __int_as_float(xi);
  y = __int_as_float(yi);
}

template <typename R, int NA>
struct Amp {
  R re[NA];
  R im[NA];
};

// (o_re, o_im) += (m_re + i m_im) * (x_re + i x_im)
template <typename R>
__device__ __forceinline__ void cfma(R &ore, R &oim, R mr, R mi, R xr, R xi) {
  ore = fma(mr, xr, ore);
  ore = fma(-mi, xi, ore);
  oim = fma(mr, xi, oim);
  oim = fma(mi, xr, oim);
}

// A descriptor value read on a scalar load and pinned in scalar registers.
// The empty asm (no instruction) keeps a per-lane choice between two
// descriptor values a select of two scalars: without it the compiler turns
// `c ? m[6] : m[0]` into one vector load from a per-lane address (a memory
// round trip inside the op loop), or into divergent branches around the two
// scalar loads.  Callers load both values first, then select.
__device__ __forceinline__ double sld(const QS_CONST double *p) {
  double v = *p;
  asm volatile("" : "+s"(v));
  return v;
}

__device__ __forceinline__ float sld(const QS_CONST float *p) {
  float v = *p;
  asm volatile("" : "+s"(v));
  return v;
}

// 2x2 on register slot J; kCtl selects the controlled form.
template <int J, bool kCtl, int A, int RP, typename R, typename I>
__device__ __forceinline__ void op_1q_reg(Amp<R, 1 << A> &a, const QS_CONST OpRec &op,
                                          const QS_CONST R *m, const Ctx<I> &c, const Regs<A, I> &g) {
  const R m0r = m[0], m0i = m[1], m1r = m[2], m1i = m[3];
  const R m2r = m[4], m2i = m[5], m3r = m[6], m3i = m[7];
  Pred<I> p{};
  if constexpr (kCtl) p = make_pred<A, RP>(op, c, g);
  constexpr int bj = 1 << J;
#pragma unroll
  for (int r = 0; r < (1 << A); ++r) {
    if ((r >> J) & 1) continue;
    const R xr = a.re[r], xi = a.im[r], yr = a.re[r | bj], yi = a.im[r | bj];
    R ur = 0, ui = 0, wr = 0, wi = 0;
    cfma(ur, ui, m0r, m0i, xr, xi);
    cfma(ur, ui, m1r, m1i, yr, yi);
    cfma(wr, wi, m2r, m2i, xr, xi);
    cfma(wr, wi, m3r, m3i, yr, yi);
    if constexpr (kCtl) {
      const bool on = pred_on<A, RP>(p, g, r);
      a.re[r] = on ? ur : xr;
      a.im[r] = on ? ui : xi;
      a.re[r | bj] = on ? wr : yr;
      a.im[r | bj] = on ? wi : yi;
    } else {
      a.re[r] = ur;
      a.im[r] = ui;
      a.re[r | bj] = wr;
      a.im[r | bj] = wi;
    }
  }
}

// A 2x2 on lane bit K = 4 or 5 by half swaps, register r paired with
// r | 2^JP.  coef(r, t) gives the pair's matrix (row-major {re, im} in
// t[0..7]), the same for r and r | 2^JP; on(r, hi) whether the op acts on
// this lane's pair (register r if bit K is clear, r | 2^JP if set).
template <int K, int JP, int A, typename R, typename Coef, typename On>
__device__ __forceinline__ void lane_2x2_swapped(Amp<R, 1 << A> &a, bool hi, Coef &&coef,
                                                 On &&on) {
  constexpr int bp = 1 << JP;
#pragma unroll
  for (int r = 0; r < (1 << A); ++r) {
    if ((r >> JP) & 1) continue;
    R ur = a.re[r], ui = a.im[r], wr = a.re[r | bp], wi = a.im[r | bp];
    half_swap<K>(ur, wr);
    half_swap<K>(ui, wi);
    R t[8];
    coef(r, t);
    R xr = 0, xi = 0, yr = 0, yi = 0;
    cfma(xr, xi, t[0], t[1], ur, ui);
    cfma(xr, xi, t[2], t[3], wr, wi);
    cfma(yr, yi, t[4], t[5], ur, ui);
    cfma(yr, yi, t[6], t[7], wr, wi);
    const bool act = on(r, hi);
    xr = act ? xr : ur;
    xi = act ? xi : ui;
    yr = act ? yr : wr;
    yi = act ? yi : wi;
    half_swap<K>(xr, yr);
    half_swap<K>(xi, yi);
    a.re[r] = xr;
    a.im[r] = xi;
    a.re[r | bp] = yr;
    a.im[r | bp] = yi;
  }
}

// 2x2 on a lane qubit: a lane with the bit clear holds x and gets the
// partner's y: out = m00 x + m01 y; a lane with the bit set holds y and
// computes out = m10 x + m11 y from the partner's x.
template <bool kCtl, int K, int A, int RP, typename R, typename I>
__device__ __forceinline__ void op_1q_lane(Amp<R, 1 << A> &a, const QS_CONST OpRec &op,
                                           const QS_CONST R *m, const Ctx<I> &c, const Regs<A, I> &g) {
  const bool hi = (c.lane >> K) & 1;
  if constexpr (K >= 4) {
    Pred<I> p{};
    if constexpr (kCtl) p = make_pred<A, RP>(op, c, g);
    lane_2x2_swapped<K, 0, A>(
        a, hi,
        [&](int, R *t) {
#pragma unroll
          for (int q = 0; q < 8; ++q) t[q] = m[q];
        },
        [&](int r, bool h) {
          if constexpr (kCtl) return p.lane_ok && (h ? pred_on<A, RP>(p, g, r | 1) : pred_on<A, RP>(p, g, r));
          return true;
        });
    return;
  }
  R mv[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) mv[q] = sld(m + q);
  // coefficients of the pair values u, w: row hi of the matrix, in the
  // order lane_pair delivers them
  constexpr bool ord = kBitOrder<K>;
  const R sr = hi ? (ord ? mv[4] : mv[6]) : mv[0], si = hi ? (ord ? mv[5] : mv[7]) : mv[1];
  const R pr = hi ? (ord ? mv[6] : mv[4]) : mv[2], pi = hi ? (ord ? mv[7] : mv[5]) : mv[3];
  Pred<I> p{};
  if constexpr (kCtl) p = make_pred<A, RP>(op, c, g);
#pragma unroll
  for (int r = 0; r < (1 << A); ++r) {
    const R xr = a.re[r], xi = a.im[r];
    R vr, vi, yr, yi;
    lane_pair<K>(xr, hi, vr, yr);
    lane_pair<K>(xi, hi, vi, yi);
    R ur = 0, ui = 0;
    cfma(ur, ui, sr, si, vr, vi);
    cfma(ur, ui, pr, pi, yr, yi);
    if constexpr (kCtl) {
      const bool on = pred_on<A, RP>(p, g, r);
      a.re[r] = on ? ur : xr;
      a.im[r] = on ? ui : xi;
    } else {
      a.re[r] = ur;
      a.im[r] = ui;
    }
  }
}

// (u, w) = t * (x, y) for the 2x2 `t` = {t0 .. t3} in row-major order.
template <typename R>
__device__ __forceinline__ void mat2(R &ur, R &ui, R &wr, R &wi, const R *tr, const R *ti, R xr,
                                     R xi, R yr, R yi) {
  ur = 0;
  ui = 0;
  wr = 0;
  wi = 0;
  cfma(ur, ui, tr[0], ti[0], xr, xi);
  cfma(ur, ui, tr[1], ti[1], yr, yi);
  cfma(wr, wi, tr[2], ti[2], xr, xi);
  cfma(wr, wi, tr[3], ti[3], yr, yi);
}

// Multiplexed 2x2 on register slot J: matrix m[0..3] where the op's pattern
// fails, m[4..7] where it holds.  kRegPattern: the pattern lies on register
// slots only, so the choice is uniform per register pair (scalar selects);
// otherwise it lies on no register slot and each lane chooses once.
template <int J, bool kRegPattern, int A, int RP, typename R, typename I>
__device__ __forceinline__ void op_mux_reg(Amp<R, 1 << A> &a, const QS_CONST OpRec &op,
                                           const QS_CONST R *m, const Ctx<I> &c,
                                           const Regs<A, I> &g) {
  const Pred<I> p = make_pred<A, RP>(op, c, g);
  constexpr int bj = 1 << J;
  R lr[4], li[4];
  // both matrices on scalar loads once; register patterns choose per pair
  // with scalar selects (a load per pair would hold 2^(A-1) matrices in
  // scalar registers at once)
  R v0r[4], v0i[4], v1r[4], v1i[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v0r[k] = sld(m + 2 * k);
    v0i[k] = sld(m + 2 * k + 1);
    v1r[k] = sld(m + 8 + 2 * k);
    v1i[k] = sld(m + 9 + 2 * k);
  }
  if constexpr (!kRegPattern) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      lr[k] = p.lane_ok ? v1r[k] : v0r[k];
      li[k] = p.lane_ok ? v1i[k] : v0i[k];
    }
  }
#pragma unroll
  for (int r = 0; r < (1 << A); ++r) {
    if ((r >> J) & 1) continue;
    R tr[4], ti[4];
    if constexpr (kRegPattern) {
      const bool hold = reg_on<A, RP>(p, g, r);  // the same for the whole wave
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        tr[k] = hold ? v1r[k] : v0r[k];
        ti[k] = hold ? v1i[k] : v0i[k];
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        tr[k] = lr[k];
        ti[k] = li[k];
      }
    }
    R ur, ui, wr, wi;
    mat2(ur, ui, wr, wi, tr, ti, a.re[r], a.im[r], a.re[r | bj], a.im[r | bj]);
    a.re[r] = ur;
    a.im[r] = ui;
    a.re[r | bj] = wr;
    a.im[r | bj] = wi;
  }
}

// Multiplexed 2x2 on a lane qubit; the pattern has no register-slot bit, so
// each lane picks its matrix row once for the whole op.
template <int K, int A, int RP, typename R, typename I>
__device__ __forceinline__ void op_mux_lane(Amp<R, 1 << A> &a, const QS_CONST OpRec &op,
                                            const QS_CONST R *m, const Ctx<I> &c,
                                            const Regs<A, I> &g) {
  const Pred<I> p = make_pred<A, RP>(op, c, g);
  const bool hi = (c.lane >> K) & 1;
  // this lane's output row: the coefficient of its own amplitude (s) and of
  // the partner's (p); selects keep the matrix on scalar loads
  const bool on = p.lane_ok;
  if constexpr (K >= 4) {
    // both lanes of a pair see the same pattern (it has no bit on K)
    R t8[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const R v0 = sld(m + q), v1 = sld(m + 8 + q);
      t8[q] = on ? v1 : v0;
    }
    lane_2x2_swapped<K, 0, A>(
        a, hi,
        [&](int, R *t) {
#pragma unroll
          for (int q = 0; q < 8; ++q) t[q] = t8[q];
        },
        [](int, bool) { return true; });
    return;
  }
  R mv[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) mv[q] = sld(m + q);
  // coefficients of the pair values u, w (see lane_pair)
  constexpr bool ord = kBitOrder<K>;
  const int su = ord ? 4 : 6, sw = ord ? 6 : 4;
  const R sr = hi ? (on ? mv[8 + su] : mv[su]) : (on ? mv[8] : mv[0]);
  const R si = hi ? (on ? mv[9 + su] : mv[su + 1]) : (on ? mv[9] : mv[1]);
  const R pr = hi ? (on ? mv[8 + sw] : mv[sw]) : (on ? mv[10] : mv[2]);
  const R pi = hi ? (on ? mv[9 + sw] : mv[sw + 1]) : (on ? mv[11] : mv[3]);
#pragma unroll
  for (int r = 0; r < (1 << A); ++r) {
    R xr, xi, yr, yi;
    lane_pair<K>(a.re[r], hi, xr, yr);
    lane_pair<K>(a.im[r], hi, xi, yi);
    R ur = 0, ui = 0;
    cfma(ur, ui, sr, si, xr, xi);
    cfma(ur, ui, pr, pi, yr, yi);
    a.re[r] = ur;
    a.im[r] = ui;
  }
}

// Multiplexed 2x2 on a lane qubit whose pattern is register slot JC alone:
// the register pairs with slot bit JC set use one matrix, the others the
// other, so each lane picks its two row pairs once and every register's
// choice is fixed at compile time.
template <int JC, int K, int A, int RP, typename R, typename I>
__device__ __forceinline__ void op_mux_lane_reg(Amp<R, 1 << A> &a, const QS_CONST OpRec &op,
                                                const QS_CONST R *m, const Ctx<I> &c,
                                                const Regs<A, I> &g) {
  const bool hi = (c.lane >> K) & 1;
  // the pattern holds where slot JC's bit equals its value in cval
  const bool v1 = ((I)op.cval & g.gb[JC]) != 0;
  const QS_CONST R *mb1 = v1 ? m + 8 : m;  // matrix where the slot bit is 1
  const QS_CONST R *mb0 = v1 ? m : m + 8;  // where it is 0
  if constexpr (K >= 4) {
    // pair registers across a slot other than JC: both share JC's bit, so
    // the pair's matrix is fixed at compile time
    lane_2x2_swapped<K, JC == 0 ? 1 : 0, A>(
        a, hi,
        [&](int r, R *t) {
          const QS_CONST R *mb = ((r >> JC) & 1) ? mb1 : mb0;
#pragma unroll
          for (int q = 0; q < 8; ++q) t[q] = mb[q];
        },
        [](int, bool) { return true; });
    return;
  }
  R w0[8], w1[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    w0[q] = sld(mb0 + q);
    w1[q] = sld(mb1 + q);
  }
  // coefficients of the pair values u, w (see lane_pair)
  constexpr bool ord = kBitOrder<K>;
  constexpr int su = ord ? 4 : 6, sw = ord ? 6 : 4;
  R sr[2], si[2], pr[2], pi[2];
  sr[0] = hi ? w0[su] : w0[0];
  si[0] = hi ? w0[su + 1] : w0[1];
  pr[0] = hi ? w0[sw] : w0[2];
  pi[0] = hi ? w0[sw + 1] : w0[3];
  sr[1] = hi ? w1[su] : w1[0];
  si[1] = hi ? w1[su + 1] : w1[1];
  pr[1] = hi ? w1[sw] : w1[2];
  pi[1] = hi ? w1[sw + 1] : w1[3];
#pragma unroll
  for (int r = 0; r < (1 << A); ++r) {
    R xr, xi, yr, yi;
    lane_pair<K>(a.re[r], hi, xr, yr);
    lane_pair<K>(a.im[r], hi, xi, yi);
    const int b = (r >> JC) & 1;
    R ur = 0, ui = 0;
    cfma(ur, ui, sr[b], si[b], xr, xi);
    cfma(ur, ui, pr[b], pi[b], yr, yi);
    a.re[r] = ur;
    a.im[r] = ui;
  }
}

// LDS exchange buffer of the wave-bit ops: two regions of 4 waves x NR
// registers x {re, im} x 64 lanes, one region per exchange round, alternating.
template <typename R, int NR>
__device__ __forceinline__ R *wave_xbuf() {
  __shared__ R buf[2 * 4 * NR * 2 * kWave];
  return buf;
}

// Registers exchanged per round: 4 complex128 or 8 complex64 values per lane
// (32 KiB of LDS per workgroup), and at least two rounds per op.
template <typename R, int NA>
constexpr int wave_round_regs() {
  constexpr int cap = sizeof(R) == 8 ? 4 : 8;
  return cap < NA / 2 ? cap : NA / 2;
}

// An op on wave bit k (slot kWaveSlot + k): the partner amplitude sits in
// the same lane and register of wave w ^ (1 << k).  Every wave writes a round
// of its registers to LDS, the workgroup synchronises, and every wave reads
// its partner's values and computes its own outputs, like a lane op.  The
// rounds alternate between two LDS regions and their count is even, so one
// barrier per round suffices: a region is rewritten only after the barrier
// that every wave reaches after reading it.  All waves of a workgroup run the
// same op list, so they meet the same barriers.
// KIND: 0 uncontrolled 2x2, 1 controlled 2x2, 2 (controlled) exchange,
// 3 multiplexed 2x2 whose pattern has no register bit.
template <int KIND, int A, int RP, typename R, typename I>
__device__ __forceinline__ void op_wave(Amp<R, 1 << A> &a, const QS_CONST OpRec &op,
                                        const QS_CONST R *m, const Ctx<I> &c,
                                        const Regs<A, I> &g) {
  constexpr int NA = 1 << A;
  constexpr int NR = wave_round_regs<R, NA>();
  static_assert((NA / NR) % 2 == 0, "an even number of exchange rounds");
  R *xb = wave_xbuf<R, NR>();
  const int k = (int)op.s0 - (int)kWaveSlot;
  const int w = c.wave;
  const bool hi = (w >> k) & 1;  // the same for the whole wave
  const int pw = w ^ (1 << k);
  Pred<I> pd{};
  if constexpr (KIND != 0) pd = make_pred<A, RP>(op, c, g);
  R sr = 0, si = 0, pr = 0, pi = 0;  // coefficients of this wave's own and partner amplitude
  if constexpr (KIND == 0 || KIND == 1) {
    R mv[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) mv[q] = sld(m + q);
    sr = hi ? mv[6] : mv[0];
    si = hi ? mv[7] : mv[1];
    pr = hi ? mv[4] : mv[2];
    pi = hi ? mv[5] : mv[3];
  } else if constexpr (KIND == 3) {
    R mv[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) mv[q] = sld(m + q);
    const bool on = pd.lane_ok;
    sr = hi ? (on ? mv[14] : mv[6]) : (on ? mv[8] : mv[0]);
    si = hi ? (on ? mv[15] : mv[7]) : (on ? mv[9] : mv[1]);
    pr = hi ? (on ? mv[12] : mv[4]) : (on ? mv[10] : mv[2]);
    pi = hi ? (on ? mv[13] : mv[5]) : (on ? mv[11] : mv[3]);
  }
#pragma unroll
  for (int r0 = 0; r0 < NA; r0 += NR) {
    R *reg = xb + ((r0 / NR) & 1) * (4 * NR * 2 * kWave);
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      reg[((w * NR + j) * 2 + 0) * kWave + c.lane] = a.re[r0 + j];
      reg[((w * NR + j) * 2 + 1) * kWave + c.lane] = a.im[r0 + j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      const int r = r0 + j;
      const R yr = reg[((pw * NR + j) * 2 + 0) * kWave + c.lane];
      const R yi = reg[((pw * NR + j) * 2 + 1) * kWave + c.lane];
      const R xr = a.re[r], xi = a.im[r];
      if constexpr (KIND == 2) {
        const bool on = pred_on<A, RP>(pd, g, r);
        a.re[r] = on ? yr : xr;
        a.im[r] = on ? yi : xi;
      } else {
        R ur = 0, ui = 0;
        cfma(ur, ui, sr, si, xr, xi);
        cfma(ur, ui, pr, pi, yr, yi);
        if constexpr (KIND == 1) {
          const bool on = pred_on<A, RP>(pd, g, r);
          a.re[r] = on ? ur : xr;
          a.im[r] = on ? ui : xi;
        } else {
          a.re[r] = ur;
          a.im[r] = ui;
        }
      }
    }
  }
}

template <int J, int A, int RP, typename R, typename I>
__device__ __forceinline__ void op_x1_reg(Amp<R, 1 << A> &a, const QS_CONST OpRec &op,
                                          const Ctx<I> &c, const Regs<A, I> &g) {
  const Pred<I> p = make_pred<A, RP>(op, c, g);
  constexpr int bj = 1 << J;
#pragma unroll
  for (int r = 0; r < (1 << A); ++r) {
    if ((r >> J) & 1) continue;
    const R xr = a.re[r], xi = a.im[r], yr = a.re[r | bj], yi = a.im[r | bj];
    const bool on = pred_on<A, RP>(p, g, r);
    a.re[r] = on ? yr : xr;
    a.im[r] = on ? yi : xi;
    a.re[r | bj] = on ? xr : yr;
    a.im[r | bj] = on ? xi : yi;
  }
}

template <int K, int A, int RP, typename R, typename I>
__device__ __forceinline__ void op_x1_lane(Amp<R, 1 << A> &a, const QS_CONST OpRec &op,
                                           const Ctx<I> &c, const Regs<A, I> &g) {
  const bool hi = (c.lane >> K) & 1;
  const Pred<I> p = make_pred<A, RP>(op, c, g);
  if constexpr (K >= 4) {
    // (controlled) exchange by half swaps: each lane swaps its pair's two
    // amplitudes where the pattern holds (registers r, r | 1 paired)
#pragma unroll
    for (int r = 0; r < (1 << A); r += 2) {
      R ur = a.re[r], ui = a.im[r], wr = a.re[r | 1], wi = a.im[r | 1];
      half_swap<K>(ur, wr);
      half_swap<K>(ui, wi);
      const bool on = p.lane_ok && (hi ? pred_on<A, RP>(p, g, r | 1) : pred_on<A, RP>(p, g, r));
      R xr = on ? wr : ur, xi = on ? wi : ui, yr = on ? ur : wr, yi = on ? ui : wi;
      half_swap<K>(xr, yr);
      half_swap<K>(xi, yi);
      a.re[r] = xr;
      a.im[r] = xi;
      a.re[r | 1] = yr;
      a.im[r | 1] = yi;
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < (1 << A); ++r) {
    const R yr = lane_xor<K>(a.re[r], hi), yi = lane_xor<K>(a.im[r], hi);
    const bool on = pred_on<A, RP>(p, g, r);
    a.re[r] = on ? yr : a.re[r];
    a.im[r] = on ? yi : a.im[r];
  }
}

template <int J0, int J1, int A, int RP, typename R, typename I>
__device__ __forceinline__ void op_dense2(Amp<R, 1 << A> &a, const QS_CONST OpRec &op,
                                          const QS_CONST R *m, const Ctx<I> &c, const Regs<A, I> &g) {
  const Pred<I> p = make_pred<A, RP>(op, c, g);
  constexpr int b0 = 1 << J0, b1 = 1 << J1;
#pragma unroll
  for (int r = 0; r < (1 << A); ++r) {
    if ((r & b0) || (r & b1)) continue;
    const int i0 = r, i1 = r | b0, i2 = r | b1, i3 = r | b0 | b1;
    const R xr[4] = {a.re[i0], a.re[i1], a.re[i2], a.re[i3]};
    const R xi[4] = {a.im[i0], a.im[i1], a.im[i2], a.im[i3]};
    R yr[4], yi[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      R sr = 0, si = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) cfma(sr, si, m[8 * i + 2 * j], m[8 * i + 2 * j + 1], xr[j], xi[j]);
      yr[i] = sr;
      yi[i] = si;
    }
    const bool on = pred_on<A, RP>(p, g, r);
    a.re[i0] = on ? yr[0] : xr[0];
    a.im[i0] = on ? yi[0] : xi[0];
    a.re[i1] = on ? yr[1] : xr[1];
    a.im[i1] = on ? yi[1] : xi[1];
    a.re[i2] = on ? yr[2] : xr[2];
    a.im[i2] = on ? yi[2] : xi[2];
    a.re[i3] = on ? yr[3] : xr[3];
    a.im[i3] = on ? yi[3] : xi[3];
  }
}

template <int A, int RP, typename R, typename I>
__device__ __forceinline__ void op_sparse(Amp<R, 1 << A> &a, const QS_CONST OpRec &op,
                                          const QS_CONST R *m, const Ctx<I> &c, const Regs<A, I> &g) {
  const R mr = m[0], mi = m[1];
  const Pred<I> p = make_pred<A, RP>(op, c, g);
#pragma unroll
  for (int r = 0; r < (1 << A); ++r) {
    const bool on = pred_on<A, RP>(p, g, r);
    const R xr = a.re[r], xi = a.im[r];
    R sr = 0, si = 0;
    cfma(sr, si, mr, mi, xr, xi);
    a.re[r] = on ? sr : xr;
    a.im[r] = on ? si : xi;
  }
}

// Sign flip where the pattern holds: the high word of each value carries
// the sign, so one select per amplitude and two integer xors.
template <int A, int RP, typename R, typename I>
__device__ __forceinline__ void op_neg(Amp<R, 1 << A> &a, const QS_CONST OpRec &op,
                                       const Ctx<I> &c, const Regs<A, I> &g) {
  const Pred<I> p = make_pred<A, RP>(op, c, g);
#pragma unroll
  for (int r = 0; r < (1 << A); ++r) {
    const uint32_t flip = pred_on<A, RP>(p, g, r) ? 0x80000000u : 0u;
    if constexpr (sizeof(R) == 8) {
      a.re[r] = __hiloint2double(__double2hiint(a.re[r]) ^ (int)flip, __double2loint(a.re[r]));
      a.im[r] = __hiloint2double(__double2hiint(a.im[r]) ^ (int)flip, __double2loint(a.im[r]));
    } else {
      a.re[r] = __int_as_float(__float_as_int(a.re[r]) ^ (int)flip);
      a.im[r] = __int_as_float(__float_as_int(a.im[r]) ^ (int)flip);
    }
  }
}

template <int A, int RP, typename R, typename I>
__device__ __forceinline__ void op_diag(Amp<R, 1 << A> &a, const QS_CONST OpRec &op,
                                        const QS_CONST R *m, const Ctx<I> &c, const Regs<A, I> &g) {
  const Pred<I> p = make_pred<A, RP>(op, c, g);
  const int na = op.na;
  const int q0 = op.act[0], q1 = na > 1 ? op.act[1] : op.act[0];
  R dv[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) dv[k] = (k < 4 || na > 1) ? sld(m + k) : (R)0;
#pragma unroll
  for (int r = 0; r < (1 << A); ++r) {
    const I gi = c.gbase | goff<A>(g, r);
    const bool b0 = (gi >> q0) & 1u;
    const bool b1 = na > 1 && ((gi >> q1) & 1u);
    const R vr = b1 ? (b0 ? dv[6] : dv[4]) : (b0 ? dv[2] : dv[0]);
    const R vi = b1 ? (b0 ? dv[7] : dv[5]) : (b0 ? dv[3] : dv[1]);
    const bool on = pred_on<A, RP>(p, g, r);
    const R xr = a.re[r], xi = a.im[r];
    R sr = 0, si = 0;
    cfma(sr, si, vr, vi, xr, xi);
    a.re[r] = on ? sr : xr;
    a.im[r] = on ? si : xi;
    // without this the scheduler selects the factors of every register
    // ahead of the multiplies, holding 2^A extra complex values live
    if (r & 1) __builtin_amdgcn_sched_barrier(0);
  }
}

// One op, its form fixed at compile time: CODE is the planner's op code
// (type and register slots folded together, see op_code).  Codes naming a
// register slot the kernel does not have (j >= A) or a wave bit of a
// kernel without wave bits compile to nothing, like CODE_NONE.
template <uint32_t CODE, int A, int W, int RP = -1, typename R, typename I>
__device__ __forceinline__ void apply_code(Amp<R, 1 << A> &a, const QS_CONST OpRec &op,
                                           const QS_CONST R *m, const Ctx<I> &c,
                                           const Regs<A, I> &g) {
  constexpr uint32_t C = CODE;
  if constexpr (C >= CODE_U1 && C < CODE_U1 + kMaxA) {
    if constexpr (C - CODE_U1 < (uint32_t)A) op_1q_reg<C - CODE_U1, false, A, RP>(a, op, m, c, g);
  } else if constexpr (C >= CODE_U1 + kMaxA && C < CODE_C1) {
    op_1q_lane<false, C - CODE_U1 - kMaxA, A, RP>(a, op, m, c, g);
  } else if constexpr (C >= CODE_C1 && C < CODE_C1 + kMaxA) {
    if constexpr (C - CODE_C1 < (uint32_t)A) op_1q_reg<C - CODE_C1, true, A, RP>(a, op, m, c, g);
  } else if constexpr (C >= CODE_C1 + kMaxA && C < CODE_X1) {
    op_1q_lane<true, C - CODE_C1 - kMaxA, A, RP>(a, op, m, c, g);
  } else if constexpr (C >= CODE_X1 && C < CODE_X1 + kMaxA) {
    if constexpr (C - CODE_X1 < (uint32_t)A) op_x1_reg<C - CODE_X1, A, RP, R, I>(a, op, c, g);
  } else if constexpr (C >= CODE_X1 + kMaxA && C < CODE_SPARSE) {
    op_x1_lane<C - CODE_X1 - kMaxA, A, RP, R, I>(a, op, c, g);
  } else if constexpr (C == CODE_SPARSE) {
    op_sparse<A, RP>(a, op, m, c, g);
  } else if constexpr (C == CODE_DIAG) {
    op_diag<A, RP>(a, op, m, c, g);
  } else if constexpr (C >= CODE_DENSE2 && C < CODE_MUXL) {
    constexpr int J0 = (int)((C - CODE_DENSE2) / kMaxA), J1 = (int)((C - CODE_DENSE2) % kMaxA);
    if constexpr (J0 < A && J1 < A && J0 != J1) op_dense2<J0, J1, A, RP>(a, op, m, c, g);
  } else if constexpr (C >= CODE_MUXL && C < CODE_MUXL + kMaxA) {
    if constexpr (C - CODE_MUXL < (uint32_t)A) op_mux_reg<C - CODE_MUXL, false, A, RP>(a, op, m, c, g);
  } else if constexpr (C >= CODE_MUXL + kMaxA && C < CODE_MUXR) {
    op_mux_lane<C - CODE_MUXL - kMaxA, A, RP>(a, op, m, c, g);
  } else if constexpr (C >= CODE_MUXR && C < CODE_MUXRL) {
    if constexpr (C - CODE_MUXR < (uint32_t)A) op_mux_reg<C - CODE_MUXR, true, A, RP>(a, op, m, c, g);
  } else if constexpr (C >= CODE_MUXRL && C < CODE_U1W) {
    constexpr int JC = (int)((C - CODE_MUXRL) / kLaneBits), LK = (int)((C - CODE_MUXRL) % kLaneBits);
    if constexpr (JC < A) op_mux_lane_reg<JC, LK, A, RP>(a, op, m, c, g);
  } else if constexpr (C >= CODE_U1W && C <= CODE_MUXW) {
    if constexpr (W > 0) op_wave<(int)(C - CODE_U1W), A, RP>(a, op, m, c, g);
  } else if constexpr (C == CODE_NEG) {
    op_neg<A, RP, R, I>(a, op, c, g);
  }
}

// The op's matrix pool entry and its code for this tile: pattern bits off
// the tile (umask) choose between `code` (values at moff + uoff) and
// `code_fail`; a batched op's matrix follows the tile's batch index.  The
// tile base is the same for every lane of a wave: it is read from the first
// lane, so that the choices stay scalar.
template <typename R, typename I>
__device__ __forceinline__ uint32_t op_setup(const QS_CONST OpRec &op, const QS_CONST R *pool,
                                             const Ctx<I> &c, const QS_CONST R *&m) {
  m = pool + op.moff;
  uint32_t code = op.code;
  if (op.umask) {
    const uint32_t tok = __builtin_amdgcn_readfirstlane(
        (uint32_t)((((uint64_t)c.tbase ^ op.cval) & op.umask) == 0));
    if (tok)
      m += op.uoff;
    else
      code = op.code_fail;
  }
  if (op.bmask)
    m += 8 * __builtin_amdgcn_readfirstlane(
                 (uint32_t)(((uint64_t)c.tbase >> op.bshift) & op.bmask));
  return code;
}

// One op of the sweep on one wave's tile held in registers.
//
// One flat switch over the op's code (type and register slots folded
// together by the planner, see op_code) rather than nested switches on type
// and slot: each nesting level is a join where the compiler copied the whole
// amplitude array between register sets, three whole-array copies per op.
template <int A, int W, typename R, typename I>
__device__ __forceinline__ void run_op(Amp<R, 1 << A> &a, const QS_CONST OpRec &op,
                                       const QS_CONST R *pool, const Ctx<I> &c,
                                       const Regs<A, I> &g) {
  const QS_CONST R *m;
  const uint32_t code = op_setup<R, I>(op, pool, c, m);
#define QS_CASE(K) \
  case (K): apply_code<(K), A, W>(a, op, m, c, g); break;
#define QS_CASE8(K) QS_CASE(K) QS_CASE(K + 1) QS_CASE(K + 2) QS_CASE(K + 3) \
  QS_CASE(K + 4) QS_CASE(K + 5) QS_CASE(K + 6) QS_CASE(K + 7)
  switch (code) {
    QS_CASE8(1) QS_CASE8(9) QS_CASE8(17) QS_CASE8(25) QS_CASE8(33) QS_CASE8(41)
    QS_CASE8(49) QS_CASE8(57) QS_CASE8(65) QS_CASE8(73) QS_CASE8(81) QS_CASE8(89)
    QS_CASE8(97) QS_CASE8(105)
    default: break;
  }
#undef QS_CASE8
#undef QS_CASE
}

// Every op of the sweep on one wave's tile held in registers.
template <int A, int W, typename R, typename I>
__device__ __forceinline__ void run_ops(Amp<R, 1 << A> &a, const QS_CONST OpRec *ops,
                                        const QS_CONST R *pool, int nops, const Ctx<I> &c,
                                        const Regs<A, I> &g) {
  for (int o = 0; o < nops; ++o) run_op<A, W>(a, ops[o], pool, c, g);
}

// The descriptor's parts, as seen through the constant address space.
template <typename R>
struct Blob {
  const QS_CONST SweepHeader *h;
  const QS_CONST OpRec *ops;
  const QS_CONST R *pool;
};

template <typename R>
__device__ __forceinline__ Blob<R> open_blob(const char *blob_g) {
  const QS_CONST char *blob = (const QS_CONST char *)blob_g;
  Blob<R> b;
  b.h = (const QS_CONST SweepHeader *)blob;
  const size_t hb = (sizeof(SweepHeader) + 15) & ~(size_t)15;
  b.ops = (const QS_CONST OpRec *)(blob + hb);
  b.pool = (const QS_CONST R *)(blob + hb + (size_t)b.h->nops * sizeof(OpRec));
  return b;
}

template <int A, typename I>
__device__ __forceinline__ Regs<A, I> reg_bits(const QS_CONST SweepHeader &h) {
  Regs<A, I> g;
  g.mask = 0;
#pragma unroll
  for (int j = 0; j < A; ++j) {
    g.gb[j] = (I)h.gb[j];
    g.mask |= g.gb[j];
  }
  return g;
}

// This lane's offsets from its six lane bits.
template <typename I>
__device__ __forceinline__ I lane_offset(const QS_CONST SweepHeader &h, int lane) {
  I lo = 0;
#pragma unroll
  for (int i = 0; i < kLaneBits; ++i) lo |= ((lane >> i) & 1) ? (I)h.lg[i] : (I)0;
  return lo;
}

// Workgroups go to the 8 XCDs round robin (workgroup w to XCD w mod 8).
// With `xcd_chunk` = grid / 8 the workgroup order is permuted so that each
// XCD walks one contiguous eighth of the tiles: its L2 and address
// translation caches then see the pages of one region instead of every
// region at once.  xcd_chunk = 0 keeps the plain order.
__device__ __forceinline__ uint64_t xcd_block(uint32_t xcd_chunk) {
  const uint32_t w = blockIdx.x;
  if (xcd_chunk == 0) return w;
  return (uint64_t)(w & 7u) * xcd_chunk + (w >> 3);
}

// This wave's offsets from the workgroup's wave bits.
template <int W, typename I>
__device__ __forceinline__ I wave_offset(const QS_CONST SweepHeader &h, int wave) {
  I wo = 0;
#pragma unroll
  for (int k = 0; k < W; ++k) wo |= ((wave >> k) & 1) ? (I)h.wb[k] : (I)0;
  return wo;
}

// Waves per SIMD the sweep kernels are built for: complex128 at A = 4 and
// complex64 at A = 5 fit in 168 VGPRs, three waves per SIMD (measured: 6.40
// ms vs 6.67 ms per 30-qubit complex128 sweep at two waves); the few values
// that do not fit are spilled outside the op loop.  The 64-bit index
// variants (more than 32 qubits) need more registers and keep the
// compiler's own choice.
template <typename R, int A, typename I>
constexpr int sweep_waves_per_eu() {
  return ((A <= 4 || sizeof(R) == 4) && sizeof(I) == 4) ? 3 : 1;
}

// One tile per wave (W = 0) or per workgroup (W = kWaveBits): load, run
// `ops(a, blob, ctx, regs)` on the tile in registers, store.
template <typename R, int A, typename I, int W, typename F>
__device__ __forceinline__ void sweep_tile(typename SV2<R>::T *__restrict__ psi,
                                           const char *__restrict__ blob_g, uint32_t xcd_chunk,
                                           F &&ops) {
  constexpr int T = kLaneBits + A + W;
  constexpr int NA = 1 << A;
  const Blob<R> b = open_blob<R>(blob_g);
  const QS_CONST SweepHeader &h = *b.h;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const uint64_t tile =
      W == 0 ? xcd_block(xcd_chunk) * (kThreads / kWave) + wave : xcd_block(xcd_chunk);
  if (tile >= h.ntiles) return;  // whole waves (W = 0) or whole workgroups
  const Regs<A, I> g = reg_bits<A, I>(h);
  Ctx<I> c;
  c.tbase = (I)tile_base<T>(tile, h);
  c.gbase = c.tbase | lane_offset<I>(h, lane) | wave_offset<W, I>(h, wave);
  c.wave = wave;
  c.lane = lane;
  Amp<R, NA> a;
  {
    typename SNT<R>::T v[NA];
    const typename SNT<R>::T *src = reinterpret_cast<const typename SNT<R>::T *>(psi);
#pragma unroll
    for (int r = 0; r < NA; ++r) v[r] = __builtin_nontemporal_load(src + (c.gbase | goff<A>(g, r)));
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      a.re[r] = v[r].x;
      a.im[r] = v[r].y;
    }
  }
  ops(a, b, c, g);
  typename SNT<R>::T *dst = reinterpret_cast<typename SNT<R>::T *>(psi);
  // the store addresses are recomputed from an opaque copy of the tile base:
  // keeping the 2^A load addresses alive across the op loop costs 2^(A+1)
  // registers
  I sbase = c.gbase;
  asm volatile("" : "+v"(sbase));
#pragma unroll
  for (int r = 0; r < NA; ++r) {
    typename SNT<R>::T v;
    v.x = a.re[r];
    v.y = a.im[r];
    __builtin_nontemporal_store(v, dst + (sbase | goff<A>(g, r)));
  }
}

// The generic sweep kernel: the op list is data, one switch per op.
template <typename R, int A, typename I, int W>
struct SweepOps {
  __device__ __forceinline__ void operator()(Amp<R, 1 << A> &a, const Blob<R> &b,
                                             const Ctx<I> &c, const Regs<A, I> &g) const {
    run_ops<A, W>(a, b.ops, b.pool, (int)b.h->nops, c, g);
  }
};

template <typename R, int A, typename I, int W>
__global__ __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(sweep_waves_per_eu<R, A, I>()))) void k_sweep(
    typename SV2<R>::T *__restrict__ psi, const char *__restrict__ blob_g, uint32_t xcd_chunk) {
  sweep_tile<R, A, I, W>(psi, blob_g, xcd_chunk, SweepOps<R, A, I, W>());
}

}  // namespace sweep
}  // namespace qs

// qs_sweep_device.h -- device side of the fused sweep engine: the sweep
// descriptor layout and the sweep kernel.  Included by qs_sweep.hip only.
//
// Amplitudes live in per-thread register arrays re[]/im[] of plain scalars
// (complex structs in registers would be spilled to scratch by the compiler
// whenever they are selected as a whole).
#pragma once

#include <hip/hip_runtime.h>

#include "qs_internal.h"

namespace qs {
namespace sweep {

constexpr int kThreads = 256;
constexpr int kMaxPhases = 16;

enum : uint32_t {
  OP_DENSE1 = 1,  // 2x2 on one register slot
  OP_DENSE2 = 2,  // 4x4 on two register slots (matrix bit 0 <-> s0)
  OP_X1 = 3,      // exchange on one register slot (Pauli X / CNOT / Toffoli)
  OP_SPARSE = 4,  // multiply by m[0] where (idx & cmask) == cval
  OP_DIAG = 5     // multiply by m[pattern of act bits] where (idx & cmask) == cval
};

template <typename R>
struct SweepOp {
  uint32_t type;
  uint8_t s0, s1, na, pad0;
  uint8_t act[4];
  uint32_t pad1;
  uint64_t cmask, cval;
  R m[32];  // row-major complex matrix, {re, im} interleaved
};

struct PhaseDesc {
  uint8_t reg[8];  // local bit of register slot j
  uint16_t op_begin, op_end;
  uint32_t pad;
};

struct SweepHeader {
  Fixed tile;        // tile qubits, sorted; expanding blockIdx gives the tile base
  uint8_t tpos[16];  // local bit -> global qubit
  int nphase;
  PhaseDesc ph[kMaxPhases];
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

__device__ __forceinline__ uint64_t sw_expand(uint64_t p, const Fixed &f) {
  for (int i = 0; i < f.nf; ++i) {
    const uint64_t lo = p & ((1ull << f.pos[i]) - 1ull);
    p = ((p ^ lo) << 1) | lo;
  }
  return p | f.set;
}

// Register slot <-> local/global bit offsets for the current phase.
template <int A>
struct Map {
  uint32_t lthr;   // local index bits owned by this thread
  uint64_t gthr;   // the same bits as global index bits
  uint32_t lb[A];  // local bit of slot j
  uint64_t gb[A];  // global bit of slot j
};

template <int T, int A>
__device__ __forceinline__ void make_map(const SweepHeader &h, int p, int t, Map<A> &mp) {
  uint32_t regmask = 0;
#pragma unroll
  for (int j = 0; j < A; ++j) {
    const int b = h.ph[p].reg[j];
    regmask |= 1u << b;
    mp.lb[j] = 1u << b;
    mp.gb[j] = 1ull << h.tpos[b];
  }
  uint32_t l = 0;
  int i = 0;
#pragma unroll
  for (int b = 0; b < T; ++b) {
    if (!((regmask >> b) & 1u)) {
      if ((t >> i) & 1) l |= 1u << b;
      ++i;
    }
  }
  uint64_t g = 0;
#pragma unroll
  for (int b = 0; b < T; ++b)
    if ((l >> b) & 1u) g |= 1ull << h.tpos[b];
  mp.lthr = l;
  mp.gthr = g;
}

template <int A>
__device__ __forceinline__ uint64_t goff(const Map<A> &mp, int r) {
  uint64_t o = 0;
#pragma unroll
  for (int j = 0; j < A; ++j)
    if ((r >> j) & 1) o |= mp.gb[j];
  return o;
}

template <int A>
__device__ __forceinline__ uint32_t loff(const Map<A> &mp, int r) {
  uint32_t o = 0;
#pragma unroll
  for (int j = 0; j < A; ++j)
    if ((r >> j) & 1) o |= mp.lb[j];
  return o;
}

// LDS slot of local index l: the low 4 bits are XOR-ed with bits 4..7, so
// both lanes-along-the-low-bits and lanes-along-bits-4..7 layouts are free
// of bank conflicts.
__device__ __forceinline__ uint32_t swz(uint32_t l) { return l ^ ((l >> 4) & 15u); }

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

template <int J, int A, typename R>
__device__ __forceinline__ void op_dense1(Amp<R, 1 << A> &a, const SweepOp<R> &op, uint64_t gbase,
                                          const Map<A> &mp) {
  const R m0r = op.m[0], m0i = op.m[1], m1r = op.m[2], m1i = op.m[3];
  const R m2r = op.m[4], m2i = op.m[5], m3r = op.m[6], m3i = op.m[7];
  const uint64_t cm = op.cmask, cv = op.cval;
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
    const bool on = cm == 0 || (((gbase | goff<A>(mp, r)) & cm) == cv);
    a.re[r] = on ? ur : xr;
    a.im[r] = on ? ui : xi;
    a.re[r | bj] = on ? wr : yr;
    a.im[r | bj] = on ? wi : yi;
  }
}

template <int J, int A, typename R>
__device__ __forceinline__ void op_x1(Amp<R, 1 << A> &a, const SweepOp<R> &op, uint64_t gbase,
                                      const Map<A> &mp) {
  const uint64_t cm = op.cmask, cv = op.cval;
  constexpr int bj = 1 << J;
#pragma unroll
  for (int r = 0; r < (1 << A); ++r) {
    if ((r >> J) & 1) continue;
    const R xr = a.re[r], xi = a.im[r], yr = a.re[r | bj], yi = a.im[r | bj];
    const bool on = cm == 0 || (((gbase | goff<A>(mp, r)) & cm) == cv);
    a.re[r] = on ? yr : xr;
    a.im[r] = on ? yi : xi;
    a.re[r | bj] = on ? xr : yr;
    a.im[r | bj] = on ? xi : yi;
  }
}

template <int J0, int J1, int A, typename R>
__device__ __forceinline__ void op_dense2(Amp<R, 1 << A> &a, const SweepOp<R> &op, uint64_t gbase,
                                          const Map<A> &mp) {
  const uint64_t cm = op.cmask, cv = op.cval;
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
      for (int j = 0; j < 4; ++j)
        cfma(sr, si, op.m[8 * i + 2 * j], op.m[8 * i + 2 * j + 1], xr[j], xi[j]);
      yr[i] = sr;
      yi[i] = si;
    }
    const bool on = cm == 0 || (((gbase | goff<A>(mp, r)) & cm) == cv);
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

template <int A, typename R>
__device__ __forceinline__ void op_sparse(Amp<R, 1 << A> &a, const SweepOp<R> &op, uint64_t gbase,
                                          const Map<A> &mp) {
  const R mr = op.m[0], mi = op.m[1];
  const uint64_t cm = op.cmask, cv = op.cval;
#pragma unroll
  for (int r = 0; r < (1 << A); ++r) {
    const bool on = ((gbase | goff<A>(mp, r)) & cm) == cv;
    const R xr = a.re[r], xi = a.im[r];
    R sr = 0, si = 0;
    cfma(sr, si, mr, mi, xr, xi);
    a.re[r] = on ? sr : xr;
    a.im[r] = on ? si : xi;
  }
}

template <int A, typename R>
__device__ __forceinline__ void op_diag(Amp<R, 1 << A> &a, const SweepOp<R> &op, uint64_t gbase,
                                        const Map<A> &mp) {
  const uint64_t cm = op.cmask, cv = op.cval;
  const int na = op.na;
  if (na == 1) {
    const int q = op.act[0];
    const R d0r = op.m[0], d0i = op.m[1], d1r = op.m[2], d1i = op.m[3];
#pragma unroll
    for (int r = 0; r < (1 << A); ++r) {
      const uint64_t g = gbase | goff<A>(mp, r);
      const bool on = (g & cm) == cv;
      const bool hi = (g >> q) & 1ull;
      const R xr = a.re[r], xi = a.im[r];
      R sr = 0, si = 0;
      cfma(sr, si, hi ? d1r : d0r, hi ? d1i : d0i, xr, xi);
      a.re[r] = on ? sr : xr;
      a.im[r] = on ? si : xi;
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < (1 << A); ++r) {
    const uint64_t g = gbase | goff<A>(mp, r);
    int pat = 0;
    for (int j = 0; j < na; ++j) pat |= (int)((g >> op.act[j]) & 1ull) << j;
    const bool on = (g & cm) == cv;
    const R xr = a.re[r], xi = a.im[r];
    R sr = 0, si = 0;
    cfma(sr, si, op.m[2 * pat], op.m[2 * pat + 1], xr, xi);
    a.re[r] = on ? sr : xr;
    a.im[r] = on ? si : xi;
  }
}

template <int A, typename R>
__device__ __forceinline__ void dispatch_dense1(Amp<R, 1 << A> &a, const SweepOp<R> &op,
                                                uint64_t gbase, const Map<A> &mp) {
  switch (op.s0) {
    case 0: op_dense1<0, A>(a, op, gbase, mp); break;
    case 1: if constexpr (A > 1) op_dense1<1, A>(a, op, gbase, mp); break;
    case 2: if constexpr (A > 2) op_dense1<2, A>(a, op, gbase, mp); break;
    case 3: if constexpr (A > 3) op_dense1<3, A>(a, op, gbase, mp); break;
    case 4: if constexpr (A > 4) op_dense1<4, A>(a, op, gbase, mp); break;
    default: break;
  }
}

template <int A, typename R>
__device__ __forceinline__ void dispatch_x1(Amp<R, 1 << A> &a, const SweepOp<R> &op,
                                            uint64_t gbase, const Map<A> &mp) {
  switch (op.s0) {
    case 0: op_x1<0, A>(a, op, gbase, mp); break;
    case 1: if constexpr (A > 1) op_x1<1, A>(a, op, gbase, mp); break;
    case 2: if constexpr (A > 2) op_x1<2, A>(a, op, gbase, mp); break;
    case 3: if constexpr (A > 3) op_x1<3, A>(a, op, gbase, mp); break;
    case 4: if constexpr (A > 4) op_x1<4, A>(a, op, gbase, mp); break;
    default: break;
  }
}

template <int J0, int A, typename R>
__device__ __forceinline__ void dispatch_dense2_b(Amp<R, 1 << A> &a, const SweepOp<R> &op,
                                                  uint64_t gbase, const Map<A> &mp) {
  switch (op.s1) {
    case 0: if constexpr (J0 != 0) op_dense2<J0, 0, A>(a, op, gbase, mp); break;
    case 1: if constexpr (A > 1 && J0 != 1) op_dense2<J0, 1, A>(a, op, gbase, mp); break;
    case 2: if constexpr (A > 2 && J0 != 2) op_dense2<J0, 2, A>(a, op, gbase, mp); break;
    case 3: if constexpr (A > 3 && J0 != 3) op_dense2<J0, 3, A>(a, op, gbase, mp); break;
    case 4: if constexpr (A > 4 && J0 != 4) op_dense2<J0, 4, A>(a, op, gbase, mp); break;
    default: break;
  }
}

template <int A, typename R>
__device__ __forceinline__ void dispatch_dense2(Amp<R, 1 << A> &a, const SweepOp<R> &op,
                                                uint64_t gbase, const Map<A> &mp) {
  switch (op.s0) {
    case 0: dispatch_dense2_b<0, A>(a, op, gbase, mp); break;
    case 1: if constexpr (A > 1) dispatch_dense2_b<1, A>(a, op, gbase, mp); break;
    case 2: if constexpr (A > 2) dispatch_dense2_b<2, A>(a, op, gbase, mp); break;
    case 3: if constexpr (A > 3) dispatch_dense2_b<3, A>(a, op, gbase, mp); break;
    case 4: if constexpr (A > 4) dispatch_dense2_b<4, A>(a, op, gbase, mp); break;
    default: break;
  }
}

template <int A, typename R>
__device__ __forceinline__ void run_ops(Amp<R, 1 << A> &a, const SweepOp<R> *__restrict__ ops,
                                        int begin, int end, uint64_t gbase, const Map<A> &mp) {
  for (int o = begin; o < end; ++o) {
    const SweepOp<R> &op = ops[o];
    switch (op.type) {
      case OP_DENSE1: dispatch_dense1<A>(a, op, gbase, mp); break;
      case OP_X1: dispatch_x1<A>(a, op, gbase, mp); break;
      case OP_DENSE2: dispatch_dense2<A>(a, op, gbase, mp); break;
      case OP_SPARSE: op_sparse<A>(a, op, gbase, mp); break;
      case OP_DIAG: op_diag<A>(a, op, gbase, mp); break;
      default: break;
    }
  }
}

// One workgroup per tile.  Every tile's amplitudes are read from HBM once
// and written once; all phases in between run in registers and LDS.
template <typename R, int T>
__global__ __launch_bounds__(kThreads) void k_sweep(typename SV2<R>::T *__restrict__ psi,
                                                    const SweepHeader *__restrict__ hdr,
                                                    const SweepOp<R> *__restrict__ ops) {
  using V = typename SV2<R>::T;
  constexpr int A = T - 8;
  constexpr int NA = 1 << A;
  const SweepHeader &h = *hdr;
  extern __shared__ unsigned char lds_raw[];
  V *lds = reinterpret_cast<V *>(lds_raw);
  const int t = threadIdx.x;
  const uint64_t base = sw_expand((uint64_t)blockIdx.x, h.tile);
  Amp<R, NA> a;
  Map<A> mp;
  make_map<T, A>(h, 0, t, mp);
  {
    const uint64_t gb = base | mp.gthr;
    V v[NA];
#pragma unroll
    for (int r = 0; r < NA; ++r) v[r] = psi[gb | goff<A>(mp, r)];
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      a.re[r] = v[r].x;
      a.im[r] = v[r].y;
    }
  }
  run_ops<A>(a, ops, h.ph[0].op_begin, h.ph[0].op_end, base | mp.gthr, mp);
  const int nphase = h.nphase;
  for (int p = 1; p < nphase; ++p) {
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      V v;
      v.x = a.re[r];
      v.y = a.im[r];
      lds[swz(mp.lthr | loff<A>(mp, r))] = v;
    }
    __syncthreads();
    make_map<T, A>(h, p, t, mp);
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      const V v = lds[swz(mp.lthr | loff<A>(mp, r))];
      a.re[r] = v.x;
      a.im[r] = v.y;
    }
    __syncthreads();
    run_ops<A>(a, ops, h.ph[p].op_begin, h.ph[p].op_end, base | mp.gthr, mp);
  }
  {
    const uint64_t gb = base | mp.gthr;
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      V v;
      v.x = a.re[r];
      v.y = a.im[r];
      psi[gb | goff<A>(mp, r)] = v;
    }
  }
}

}  // namespace sweep
}  // namespace qs

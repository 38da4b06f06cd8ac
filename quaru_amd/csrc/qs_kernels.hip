// qs_kernels.hip -- per-gate kernels and FP64 reductions for gfx950.
//
// The per-gate kernels replace the hot loop of Register::try_apply_fast
// (src/register.rs:155-189): instead of filtering all 2^n indices and
// gathering each group through heap vectors, a thread expands a dense counter
// p over the free index bits into a group base index (inserting zeros at the
// sorted fixed bit positions: targets and controls), then reads, multiplies
// and writes the group in place.  Groups are disjoint, so in-place equals the
// reference's out-of-place copy.  Amplitudes whose control bits do not match
// are never touched, which is why a gate with c controls moves 2^(n-c)
// amplitudes instead of 2^n.
//
// Reductions (measure, norm, inner product, equality) accumulate in FP64 for
// both storage precisions, with a fixed grid so results are deterministic.
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "qs_internal.h"

namespace qs {

template <typename R>
struct V2;
template <>
struct V2<double> {
  using T = double2;
};
template <>
struct V2<float> {
  using T = float2;
};

Fixed make_fixed(uint64_t mask, uint64_t set) {
  Fixed f{};
  f.set = set;
  f.nf = 0;
  for (int b = 0; b < 64; ++b)
    if ((mask >> b) & 1ull) f.pos[f.nf++] = (uint8_t)b;
  return f;
}

__device__ __forceinline__ uint64_t expand(uint64_t p, const Fixed &f) {
  for (int i = 0; i < f.nf; ++i) {
    const uint64_t lo = p & ((1ull << f.pos[i]) - 1ull);
    p = ((p ^ lo) << 1) | lo;
  }
  return p | f.set;
}

template <typename T>
__device__ __forceinline__ T cmla(T acc, decltype(T::x) mr, decltype(T::x) mi, T a) {
  acc.x = fma(mr, a.x, acc.x);
  acc.x = fma(-mi, a.y, acc.x);
  acc.y = fma(mr, a.y, acc.y);
  acc.y = fma(mi, a.x, acc.y);
  return acc;
}

template <typename T>
__device__ __forceinline__ T czero() {
  T z;
  z.x = 0;
  z.y = 0;
  return z;
}

constexpr int kBlock = 256;

// Streaming access.  A state larger than the last-level cache (256 MiB
// Infinity Cache) is read and written once per gate, so the gate kernels
// load and store its amplitudes with the nontemporal hint.  Measured on a
// 16 GiB complex128 state, pair updates (i, i | 2^t) in place: 5.5-5.8 TB/s
// with the hint against 5.2-5.4 without for t >= 3, but 5.0 against 5.2 for
// t < 3, where a wave's two loads interleave inside the same 128-byte lines.
constexpr uint64_t kStreamBytes = 256ull << 20;

static bool streaming(int n, size_t amp_bytes) { return ((uint64_t)amp_bytes << n) > kStreamBytes; }

template <typename R>
struct NTV {
  typedef R T __attribute__((ext_vector_type(2)));
};

template <bool kNT, typename T>
__device__ __forceinline__ T ld(const T *p) {
  if constexpr (kNT) {
    using E = typename NTV<decltype(T::x)>::T;
    const E v = __builtin_nontemporal_load(reinterpret_cast<const E *>(p));
    T r;
    r.x = v.x;
    r.y = v.y;
    return r;
  } else {
    return *p;
  }
}

template <bool kNT, typename T>
__device__ __forceinline__ void st(T *p, T v) {
  if constexpr (kNT) {
    using E = typename NTV<decltype(T::x)>::T;
    E e;
    e.x = v.x;
    e.y = v.y;
    __builtin_nontemporal_store(e, reinterpret_cast<E *>(p));
  } else {
    *p = v;
  }
}

// Calls f(integral_constant<bool, nt>).
template <typename F>
static void nt_dispatch(bool nt, F &&f) {
  if (nt) f(std::integral_constant<bool, true>{});
  else f(std::integral_constant<bool, false>{});
}

// Read-only data every thread reads at the same address (gate matrices):
// through the constant address space the compiler keeps the reads scalar.
#define QS_KCONST __attribute__((address_space(4)))

static inline unsigned grid_for(uint64_t work, uint64_t per_block) {
  uint64_t g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// ---- state initialisation ---------------------------------------------------

template <typename R>
__global__ __launch_bounds__(kBlock) void k_init_basis(typename V2<R>::T *psi, uint64_t len,
                                                       uint64_t k) {
  using T = typename V2<R>::T;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * kBlock) {
    T v = czero<T>();
    if (i == k) v.x = 1;
    psi[i] = v;
  }
}

// Product state in the order of the kron fold of Register::new /
// try_from_qubits (src/register.rs:47-50, :94-96): amp = b_{n-1} * (... *
// (b_0 * 1)).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_init_product(typename V2<R>::T *psi, int n,
                                                         const double *__restrict__ q) {
  using T = typename V2<R>::T;
  const uint64_t len = 1ull << n;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * kBlock) {
    double ar = 1.0, ai = 0.0;
    for (int b = 0; b < n; ++b) {
      const int bit = (int)((i >> b) & 1ull);
      const double br = q[4 * b + 2 * bit], bi = q[4 * b + 2 * bit + 1];
      const double nr = br * ar - bi * ai;
      const double ni = br * ai + bi * ar;
      ar = nr;
      ai = ni;
    }
    T v;
    v.x = (R)ar;
    v.y = (R)ai;
    psi[i] = v;
  }
}

template <typename R>
hipError_t launch_init_basis(void *psi, int n, uint64_t k, hipStream_t s) {
  const uint64_t len = 1ull << n;
  unsigned g = grid_for(len, kBlock * 8);
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_init_basis<R>, dim3(g), dim3(kBlock), 0, s,
                     (typename V2<R>::T *)psi, len, k);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_init_product(void *psi, int n, const double *d_qubits, hipStream_t s) {
  const uint64_t len = 1ull << n;
  unsigned g = grid_for(len, kBlock * 8);
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_init_product<R>, dim3(g), dim3(kBlock), 0, s,
                     (typename V2<R>::T *)psi, n, d_qubits);
  return hipGetLastError();
}

// ---- dense gates --------------------------------------------------------------

template <typename R, int D>
struct MatArg {
  R re[D * D];
  R im[D * D];
};

// One active target (matrix bit 0 <-> qubit with stride `tbit`).  Each thread
// owns kItems units spaced one block apart, so a wave issues all its loads
// before the first multiply.  kX = 1: a unit is one pair (i, i | tbit).
// kX = 2: a unit is the two pairs at i and i | xbit, xbit a free index bit
// that fx also holds at 0, so a workgroup reads four streams (xbit 32 or 64
// KiB of amplitudes, as the host picks it; complex128 only).  Measured at 30 qubits,
// complex128, Haar gate, ms per launch: target 6-10 6.25-6.28 with kX = 1,
// 5.61-5.69 with kX = 2 and xbit = 2^12; target 13-29 6.00-6.10 against
// 5.37-5.53; target 11 5.27 against 5.22.  An xbit of 2^13 ran 6.0-6.9.
template <typename R, int kItems, bool kNT, int kX = 1>
__global__ __launch_bounds__(kBlock) void k_gate1(typename V2<R>::T *__restrict__ psi,
                                                  uint64_t npairs, Fixed fx, uint64_t tbit,
                                                  MatArg<R, 2> m, uint64_t xbit = 0) {
  using T = typename V2<R>::T;
  static_assert(kX == 1 || kX == 2, "kX: 1 or 2 pairs per unit");
  constexpr int NP = kItems * kX;
  const uint64_t base = (uint64_t)blockIdx.x * (kBlock * kItems) + threadIdx.x;
  T a0[NP], a1[NP];
  uint64_t i0[NP];
  bool ok[kItems];
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint64_t p = base + (uint64_t)it * kBlock;
    ok[it] = p < npairs;
    if (ok[it]) {
      const uint64_t i = expand(p, fx);
#pragma unroll
      for (int x = 0; x < kX; ++x) {
        const int j = it * kX + x;
        i0[j] = x ? (i | xbit) : i;
        a0[j] = ld<kNT>(psi + i0[j]);
        a1[j] = ld<kNT>(psi + (i0[j] | tbit));
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    if (ok[j / kX]) {
      T r0 = czero<T>(), r1 = czero<T>();
      r0 = cmla(r0, m.re[0], m.im[0], a0[j]);
      r0 = cmla(r0, m.re[1], m.im[1], a1[j]);
      r1 = cmla(r1, m.re[2], m.im[2], a0[j]);
      r1 = cmla(r1, m.re[3], m.im[3], a1[j]);
      st<kNT>(psi + i0[j], r0);
      st<kNT>(psi + (i0[j] | tbit), r1);
    }
  }
}

// Partner lane l ^ 2^K of a 32-bit value by VALU cross-lane moves: DPP
// inside rows of 16 lanes for K < 4 -- quad_perm [1,0,3,2] (K = 0),
// [2,3,0,1] (K = 1), row_shl:4 into banks 0 and 2 plus row_shr:4 into banks
// 1 and 3 (K = 2: two independent moves, each writing the lanes the other
// leaves alone), row_ror:8 (K = 3) -- and gfx950's half swaps for K = 4, 5
// (a value swapped with itself: the first result holds the bit-clear half's
// values, the second the bit-set half's).
template <int K>
__device__ __forceinline__ int lane_partner(int v) {
  const bool hi = (threadIdx.x >> K) & 1;
  if constexpr (K == 0) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);
  if constexpr (K == 1) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);
  if constexpr (K == 2) {
    const int lo = __builtin_amdgcn_update_dpp(0, v, 0x104, 0xF, 0x5, false);  // row_shl:4
    return __builtin_amdgcn_update_dpp(lo, v, 0x114, 0xF, 0xA, false);        // row_shr:4
  }
  if constexpr (K == 3) return __builtin_amdgcn_mov_dpp(v, 0x128, 0xF, 0xF, false);
  if constexpr (K == 4) {
    const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return hi ? p[0] : p[1];
  }
  const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
  return hi ? p[0] : p[1];
}

template <int K>
__device__ __forceinline__ double lane_partner(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = lane_partner<K>((int)(b & 0xffffffffll));
  const int hi = lane_partner<K>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int K>
__device__ __forceinline__ float lane_partner(float v) {
  return __int_as_float(lane_partner<K>(__float_as_int(v)));
}

// One active target on qubit K < 6, controls (if any) on index bits >= 6:
// the pairs lie inside runs of 64 consecutive amplitudes, so k_gate1's two
// loads per pair would each touch every line at part occupancy.  Here lane
// l of a wave owns amplitude l of a run of 64 consecutive amplitudes (every
// load and store instruction is one contiguous 1 KiB / 512 B access), gets
// its partner's amplitude by DPP, and computes its own output with row
// ((l >> K) & 1) of the matrix.  Chunk c of 64 amplitudes is expand(c, fx)
// << 6 with fx holding the controls relative to bit 6.
//
// Chunk order (kStreams == 0): wave w takes chunks w * kItems + it.  With
// kStreams = S = 2^SB (nchunks a multiple of 2^(5 + SB)), a workgroup reads
// S streams 32 chunks apart: unit u = block * 4 * kItems + it * 4 + wave
// gives chunks (u mod 32) | (u / 32) << (5 + SB) | s << 5 for s < S.
// Measured at 30 qubits, complex128, gate on qubit 0: S = 4, kItems = 4
// (16 loads in flight per lane, 64 KiB per workgroup in four 16 KiB runs)
// 5.28 ms; one stream of 4 chunks per wave 5.92 ms; S = 2 5.55-5.70 ms;
// S = 8 5.42-5.96 ms; streams 16, 64 or 128 chunks apart 5.38-5.93 ms.
template <typename R, int K, int kItems, bool kNT, int kStreams = 0>
__global__ __launch_bounds__(kBlock) void k_gate1_lane(typename V2<R>::T *__restrict__ psi,
                                                       uint64_t nchunks, Fixed fx,
                                                       MatArg<R, 2> m) {
  using T = typename V2<R>::T;
  constexpr int S = kStreams > 0 ? kStreams : 1;
  constexpr int SB = S == 1 ? 0 : (S == 2 ? 1 : (S == 4 ? 2 : 3));
  static_assert(S == (1 << SB), "kStreams: 0, 1, 2, 4 or 8");
  constexpr int NL = kItems * S;
  const int lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const bool hi = (lane >> K) & 1;
  // this lane's row: coefficient of its own amplitude (s) and the partner's (p)
  const R sr = hi ? m.re[3] : m.re[0], si = hi ? m.im[3] : m.im[0];
  const R pr = hi ? m.re[2] : m.re[1], pi = hi ? m.im[2] : m.im[1];
  T a[NL];
  uint64_t idx[NL];
  bool ok[NL];
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    if constexpr (kStreams == 0) {
      const uint64_t c = wave * kItems + it;
      ok[it] = c < nchunks;  // uniform over the wave: whole chunks only
      if (ok[it]) idx[it] = (expand(c, fx) << 6) | (uint64_t)lane;
    } else {
      // the workgroup's 4 * kItems * S chunks lie below nchunks as a whole or
      // not at all (the host launches nchunks / (4 * kItems * S) workgroups)
      const bool in = ((uint64_t)blockIdx.x + 1) * (4 * kItems * S) <= nchunks;
      const uint64_t u = (uint64_t)blockIdx.x * (4 * kItems) + it * 4 + (threadIdx.x >> 6);
      const uint64_t c0 = (u & 31ull) | ((u >> 5) << (5 + SB));
#pragma unroll
      for (int s = 0; s < S; ++s) {
        ok[it * S + s] = in;
        if (in) idx[it * S + s] = (expand(c0 | ((uint64_t)s << 5), fx) << 6) | (uint64_t)lane;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NL; ++j)
    if (ok[j]) a[j] = ld<kNT>(psi + idx[j]);
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    if (ok[j]) {
      T y;
      y.x = lane_partner<K>(a[j].x);
      y.y = lane_partner<K>(a[j].y);
      T r = czero<T>();
      r = cmla(r, sr, si, a[j]);
      r = cmla(r, pr, pi, y);
      st<kNT>(psi + idx[j], r);
    }
  }
}

// Two active targets: quadruples (b, b|t0, b|t1, b|t0|t1) and a 4x4 matrix.
template <typename R, int kItems, bool kNT>
__global__ __launch_bounds__(kBlock) void k_gate2(typename V2<R>::T *__restrict__ psi,
                                                  uint64_t nquads, Fixed fx, uint64_t t0,
                                                  uint64_t t1, MatArg<R, 4> m) {
  using T = typename V2<R>::T;
  const uint64_t base = (uint64_t)blockIdx.x * (kBlock * kItems) + threadIdx.x;
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint64_t p = base + (uint64_t)it * kBlock;
    if (p >= nquads) break;
    const uint64_t b = expand(p, fx);
    const uint64_t idx[4] = {b, b | t0, b | t1, b | t0 | t1};
    T a[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] = ld<kNT>(psi + idx[j]);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      T r = czero<T>();
#pragma unroll
      for (int j = 0; j < 4; ++j) r = cmla(r, m.re[4 * i + j], m.im[4 * i + j], a[j]);
      st<kNT>(psi + idx[i], r);
    }
  }
}

// Offset of row i of a group: the K bits of i deposited at the active
// positions (ax.pos[j] <- bit j of i).  Wave-uniform for uniform i.
template <int K>
__device__ __forceinline__ uint64_t row_offset(int i, const Fixed &ax) {
  uint64_t off = 0;
#pragma unroll
  for (int j = 0; j < K; ++j)
    if ((i >> j) & 1) off |= 1ull << ax.pos[j];
  return off;
}

// K = 3..5 active targets: one thread per group of D = 2^K amplitudes.  The
// thread loads its D amplitudes into registers, then computes and stores one
// output row at a time, out_i = sum_j M[i][j] a_j, so the D inputs are never
// overwritten before all rows are done (in place is safe: groups are
// disjoint).  The matrix (row-major {re, im}, the same for every thread) is
// read through the constant address space: its entries arrive on scalar
// loads and enter the FMAs as scalar operands, so the only per-amplitude
// memory traffic is the state itself.  Consecutive threads take consecutive
// groups, i.e. consecutive free index bits: with the low qubits free, each
// of the D loads of a wave is a run of 64 consecutive amplitudes.
// kReal: every matrix entry is real (Hadamard transforms, permutations,
// RY products), so a complex MAC is 2 FMAs instead of 4.
template <typename R, int K, bool kReal, bool kNT>
__global__ __launch_bounds__(kBlock) void k_gatek(typename V2<R>::T *__restrict__ psi,
                                                  uint64_t ngroups, Fixed fx, Fixed ax,
                                                  const R *__restrict__ m_g) {
  using T = typename V2<R>::T;
  constexpr int D = 1 << K;
  const QS_KCONST R *m = (const QS_KCONST R *)m_g;
  const uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (g >= ngroups) return;
  const uint64_t base = expand(g, fx);
  R ar[D], ai[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    const T v = ld<kNT>(psi + (base | row_offset<K>(j, ax)));
    ar[j] = v.x;
    ai[j] = v.y;
  }
  for (int i = 0; i < D; ++i) {
    const QS_KCONST R *row = m + 2 * D * i;
    // two partial sums per component: independent FMA chains
    R sr0 = 0, si0 = 0, sr1 = 0, si1 = 0;
#pragma unroll
    for (int j = 0; j < D; j += 2) {
      const R mr0 = row[2 * j], mr1 = row[2 * j + 2];
      sr0 = fma(mr0, ar[j], sr0);
      si0 = fma(mr0, ai[j], si0);
      sr1 = fma(mr1, ar[j + 1], sr1);
      si1 = fma(mr1, ai[j + 1], si1);
      if constexpr (!kReal) {
        const R mi0 = row[2 * j + 1], mi1 = row[2 * j + 3];
        sr0 = fma(-mi0, ai[j], sr0);
        si0 = fma(mi0, ar[j], si0);
        sr1 = fma(-mi1, ai[j + 1], sr1);
        si1 = fma(mi1, ar[j + 1], si1);
      }
    }
    T out;
    out.x = sr0 + sr1;
    out.y = si0 + si1;
    st<kNT>(psi + (base | row_offset<K>(i, ax)), out);
  }
}

// Complex64 5-qubit blocks on the matrix cores.  Per wave, 64 groups (32
// amplitudes each) as the real product OUT = W x IN, with W the 64 x 64 real
// form [[Mr, -Mi], [Mi, Mr]] of the 32 x 32 complex matrix and IN the 64 x 64
// matrix of the groups' {re, im} components, by v_mfma_f32_32x32x2f32: two
// column blocks of 32 groups x two output row blocks (re, im) x 32 steps of
// K = 2 = 128 instructions per 2048 amplitudes.  FP32 matrix issue is twice
// the non-packed FP32 vector rate (157.3 against 78.6 TFLOP/s on MI355X),
// which the complex VALU form of k_gatek<float, 5> needs to reach the memory
// time.
//
// Operand layouts of the instruction (one 32 x 32 block): lane l holds
// A[l % 32][l / 32] and B[l / 32][l % 32]; accumulator register v of lane l
// holds D[(v % 4) + 8 (v / 4) + 4 (l / 32)][l % 32].  Lane l loads row m of
// group l, so every load instruction reads 64 consecutive groups (the 6
// lowest index bits are free).  A half swap (v_permlane32_swap) of rows 2t
// and 2t + 1 then leaves rows 2t (lanes 0..31) and 2t + 1 (lanes 32..63) of
// groups 0..31 in the first register, of groups 32..63 in the second: the B
// operands of step t (real parts; step 16 + t: imaginary parts) of the two
// column blocks.  The same swap of the two column blocks' accumulators gives
// rows i and i + 4 of all 64 groups, stored 64 consecutive groups at a time.
// W's A operands (64 per lane, one per instruction) are read once per
// persistent wave from the row-major matrix.
typedef float qs_floatx16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ void half32_swap(float &x, float &y) {
  const auto p = __builtin_amdgcn_permlane32_swap(__float_as_int(x), __float_as_int(y), false,
                                                  false);
  x = __int_as_float(p[0]);
  y = __int_as_float(p[1]);
}

template <bool kNT>
__global__ __launch_bounds__(64) void k_gatek5_mfma(float2 *__restrict__ psi, uint64_t nchunks,
                                                    Fixed fx, Fixed ax,
                                                    const float *__restrict__ m) {
  const int lane = threadIdx.x;
  const int h = lane >> 5, j = lane & 31;
  float aop0[32], aop1[32];  // W rows 0..31 (re out) and 32..63 (im out)
#pragma unroll
  for (int t = 0; t < 32; ++t) {
    const int c = 2 * t + h;  // W column
    const int cc = c & 31;
    const float mr = m[2 * (32 * j + cc)], mi = m[2 * (32 * j + cc) + 1];
    aop0[t] = c < 32 ? mr : -mi;
    aop1[t] = c < 32 ? mi : mr;
  }
  for (uint64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const uint64_t base = expand(ch * 64 + (uint64_t)lane, fx);
    // b[0][t], b[1][t]: B operands of step t for column blocks 0 and 1
    float b0[32], b1[32];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const float2 x = ld<kNT>(psi + (base | row_offset<5>(2 * t, ax)));
      const float2 y = ld<kNT>(psi + (base | row_offset<5>(2 * t + 1, ax)));
      b0[t] = x.x;
      b1[t] = y.x;
      b0[16 + t] = x.y;
      b1[16 + t] = y.y;
      half32_swap(b0[t], b1[t]);
      half32_swap(b0[16 + t], b1[16 + t]);
    }
    qs_floatx16 d00 = {}, d01 = {}, d10 = {}, d11 = {};  // d[column block][re/im]
#pragma unroll
    for (int t = 0; t < 32; ++t) {
      d00 = __builtin_amdgcn_mfma_f32_32x32x2f32(aop0[t], b0[t], d00, 0, 0, 0);
      d01 = __builtin_amdgcn_mfma_f32_32x32x2f32(aop1[t], b0[t], d01, 0, 0, 0);
      d10 = __builtin_amdgcn_mfma_f32_32x32x2f32(aop0[t], b1[t], d10, 0, 0, 0);
      d11 = __builtin_amdgcn_mfma_f32_32x32x2f32(aop1[t], b1[t], d11, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      // rows i0 = (v % 4) + 8 (v / 4) and i0 + 4 of all 64 groups
      float r0 = d00[v], r1 = d10[v], i0 = d01[v], i1 = d11[v];
      half32_swap(r0, r1);
      half32_swap(i0, i1);
      const int row = (v & 3) + 8 * (v >> 2);
      float2 o;
      o.x = r0;
      o.y = i0;
      st<kNT>(psi + (base | row_offset<5>(row, ax)), o);
      o.x = r1;
      o.y = i1;
      st<kNT>(psi + (base | row_offset<5>(row + 4, ax)), o);
    }
  }
}

// The same with 32 groups per wave, for blocks whose lowest active bit is
// index bit 5 (so 64 consecutive groups would not be consecutive
// amplitudes): lane l = 32h + j serves group j and loads its rows 2t + h,
// which are directly the B operands of steps t (real parts) and 16 + t
// (imaginary parts); rows 2t and 2t + 1 then differ in bit 5, so each load
// instruction still reads 512 consecutive bytes.  Its outputs are rows
// (v % 4) + 8 (v / 4) + 4h of D0 (real) and D1 (imaginary).
template <bool kNT>
__global__ __launch_bounds__(64) void k_gatek5_mfma32(float2 *__restrict__ psi,
                                                      uint64_t nchunks, Fixed fx, Fixed ax,
                                                      const float *__restrict__ m) {
  const int lane = threadIdx.x;
  const int h = lane >> 5, j = lane & 31;
  float aop0[32], aop1[32];
#pragma unroll
  for (int t = 0; t < 32; ++t) {
    const int c = 2 * t + h;
    const int cc = c & 31;
    const float mr = m[2 * (32 * j + cc)], mi = m[2 * (32 * j + cc) + 1];
    aop0[t] = c < 32 ? mr : -mi;
    aop1[t] = c < 32 ? mi : mr;
  }
  for (uint64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const uint64_t base = expand(ch * 32 + (uint64_t)j, fx);
    float br[16], bi[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const float2 v = ld<kNT>(psi + (base | row_offset<5>(2 * t + h, ax)));
      br[t] = v.x;
      bi[t] = v.y;
    }
    qs_floatx16 d0 = {}, d1 = {};
#pragma unroll
    for (int t = 0; t < 32; ++t) {
      const float b = t < 16 ? br[t] : bi[t - 16];
      d0 = __builtin_amdgcn_mfma_f32_32x32x2f32(aop0[t], b, d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_32x32x2f32(aop1[t], b, d1, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      float2 o;
      o.x = d0[v];
      o.y = d1[v];
      st<kNT>(psi + (base | row_offset<5>((v & 3) + 8 * (v >> 2) + 4 * h, ax)), o);
    }
  }
}

// Whether complex64 5-qubit blocks use k_gatek5_mfma (QS_GATEK_MFMA, default
// on); the VALU form k_gatek<float, 5, false> otherwise.
static bool gatek_mfma() {
  static const bool on = [] {
    const char *v = std::getenv("QS_GATEK_MFMA");
    return !(v && v[0] == '0');
  }();
  return on;
}

// Bit layout of one staged block of k_gatek_staged: 64 groups x 2^K rows =
// 2^(6+K) slots.  Slot bit b sits at state index bit pa[b] (the active
// positions and the block's 6 lowest free positions, ascending), so the
// lanes of a staging load vary the lowest index bits the block has.  gslot[k]
// is the slot bit that carries bit k of the group number, jslot[k] the slot
// bit that carries matrix bit k.
struct StageMap {
  uint8_t pa[16];
  uint8_t gslot[6];
  uint8_t jslot[8];
};

// LDS position of slot s: the low 5 bits are XORed with the next 5, which
// spreads both the staging pattern (lanes = consecutive slots) and the
// compute pattern (lanes = groups) over the banks.
__device__ __forceinline__ int stage_phys(int s) { return s ^ ((s >> 5) & 31); }

// K = 3..5 active targets when low index bits are fixed (active or
// controls): the direct kernel's loads would then stride by 2^(lowest free
// bit) amplitudes per lane.  One wave stages 64 groups through LDS with
// loads over the block's lowest index bits, each lane computes its own group
// from registers (matrix on scalar loads, as in k_gatek), writes its outputs
// back to its own slots, and the wave stores the block the way it loaded it.
template <typename R, int K, bool kReal, bool kNT>
__global__ __launch_bounds__(64) void k_gatek_staged(typename V2<R>::T *__restrict__ psi,
                                                     uint64_t ngroups, Fixed fx, StageMap sm,
                                                     const R *__restrict__ m_g) {
  using T = typename V2<R>::T;
  constexpr int D = 1 << K;
  constexpr int kSlots = 64 * D;
  __shared__ T sl[kSlots];
  const QS_KCONST R *m = (const QS_KCONST R *)m_g;
  const int lane = threadIdx.x;
  const uint64_t g0 = (uint64_t)blockIdx.x * 64;
  if (g0 >= ngroups) return;  // whole blocks only: ngroups is a multiple of 64
  const uint64_t block_base = expand(g0, fx);
  // address bits of this lane's low 6 slot bits
  uint64_t lp = 0;
#pragma unroll
  for (int b = 0; b < 6; ++b)
    if ((lane >> b) & 1) lp |= 1ull << sm.pa[b];
#pragma unroll
  for (int r = 0; r < D; ++r) {
    uint64_t rp = 0;  // uniform: address bits of slot bits 6.. (value r)
#pragma unroll
    for (int b = 0; b < K; ++b)
      if ((r >> b) & 1) rp |= 1ull << sm.pa[6 + b];
    sl[stage_phys(r * 64 + lane)] = ld<kNT>(psi + (block_base | lp | rp));
  }
  __syncthreads();
  int tp = 0;  // this lane's group number on the slot bits
#pragma unroll
  for (int k = 0; k < 6; ++k)
    if ((lane >> k) & 1) tp |= 1 << sm.gslot[k];
  R ar[D], ai[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    int jp = 0;
#pragma unroll
    for (int k = 0; k < K; ++k)
      if ((j >> k) & 1) jp |= 1 << sm.jslot[k];
    const T v = sl[stage_phys(tp | jp)];
    ar[j] = v.x;
    ai[j] = v.y;
  }
  for (int i = 0; i < D; ++i) {
    const QS_KCONST R *row = m + 2 * D * i;
    R sr0 = 0, si0 = 0, sr1 = 0, si1 = 0;
#pragma unroll
    for (int j = 0; j < D; j += 2) {
      const R mr0 = row[2 * j], mr1 = row[2 * j + 2];
      sr0 = fma(mr0, ar[j], sr0);
      si0 = fma(mr0, ai[j], si0);
      sr1 = fma(mr1, ar[j + 1], sr1);
      si1 = fma(mr1, ai[j + 1], si1);
      if constexpr (!kReal) {
        const R mi0 = row[2 * j + 1], mi1 = row[2 * j + 3];
        sr0 = fma(-mi0, ai[j], sr0);
        si0 = fma(mi0, ar[j], si0);
        sr1 = fma(-mi1, ai[j + 1], sr1);
        si1 = fma(mi1, ar[j + 1], si1);
      }
    }
    int ip = 0;
#pragma unroll
    for (int k = 0; k < K; ++k)
      if ((i >> k) & 1) ip |= 1 << sm.jslot[k];
    T out;
    out.x = sr0 + sr1;
    out.y = si0 + si1;
    sl[stage_phys(tp | ip)] = out;  // own slot: already read into registers
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < D; ++r) {
    uint64_t rp = 0;
#pragma unroll
    for (int b = 0; b < K; ++b)
      if ((r >> b) & 1) rp |= 1ull << sm.pa[6 + b];
    st<kNT>(psi + (block_base | lp | rp), sl[stage_phys(r * 64 + lane)]);
  }
}

static uint64_t fx_mask(const Fixed &fx) {
  uint64_t m = 0;
  for (int i = 0; i < fx.nf; ++i) m |= 1ull << fx.pos[i];
  return m;
}

// The staged layout for a gate: active positions `ax` (matrix-bit order) and
// the lowest 6 free positions of `fx`.  Returns false when fewer than 6 free
// positions exist (tiny registers use the direct kernel).
static bool stage_map(const Fixed &fx, const Fixed &ax, int n, StageMap &sm) {
  uint64_t fixed = 0;
  for (int i = 0; i < fx.nf; ++i) fixed |= 1ull << fx.pos[i];
  uint64_t act = 0;
  for (int j = 0; j < ax.nf; ++j) act |= 1ull << ax.pos[j];
  uint64_t free6 = 0;
  int nfree = 0;
  for (int q = 0; q < n && nfree < 6; ++q)
    if (!((fixed >> q) & 1ull)) {
      free6 |= 1ull << q;
      ++nfree;
    }
  if (nfree < 6) return false;
  const uint64_t P = act | free6;
  int b = 0, gk = 0;
  uint8_t slot_of[64];
  for (int q = 0; q < 64; ++q)
    if ((P >> q) & 1ull) {
      slot_of[q] = (uint8_t)b;
      sm.pa[b++] = (uint8_t)q;
      if ((free6 >> q) & 1ull) sm.gslot[gk++] = (uint8_t)(b - 1);
    }
  // pa[] must list the slot bits low to high: slots 0..5 are the lowest six
  // positions of P and feed the lanes; the rest follow in order.
  for (int j = 0; j < ax.nf; ++j) sm.jslot[j] = slot_of[ax.pos[j]];
  return true;
}

// Any K: out-of-place from `src` (a copy of the state) into psi, one thread
// per (group, row).  Used for the dense k = n operations of the examples.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_gate_big(typename V2<R>::T *__restrict__ psi,
                                                     const typename V2<R>::T *__restrict__ src,
                                                     uint64_t ngroups, int K, Fixed fx, Fixed ax,
                                                     const typename V2<R>::T *__restrict__ m) {
  using T = typename V2<R>::T;
  const uint64_t D = 1ull << K;
  const uint64_t total = ngroups * D;
  for (uint64_t w = (uint64_t)blockIdx.x * kBlock + threadIdx.x; w < total;
       w += (uint64_t)gridDim.x * kBlock) {
    const uint64_t i = w % D, g = w / D;
    const uint64_t b = expand(g, fx);
    uint64_t oi = 0;
    for (int j = 0; j < K; ++j)
      if ((i >> j) & 1ull) oi |= 1ull << ax.pos[j];
    T acc = czero<T>();
    for (uint64_t j = 0; j < D; ++j) {
      uint64_t oj = 0;
      for (int q = 0; q < K; ++q)
        if ((j >> q) & 1ull) oj |= 1ull << ax.pos[q];
      const T mv = m[i * D + j];
      acc = cmla(acc, mv.x, mv.y, src[b | oj]);
    }
    psi[b | oi] = acc;
  }
}

template <typename R, int D>
static MatArg<R, D> to_matarg(const Gate &g) {
  MatArg<R, D> m;
  for (int e = 0; e < D * D; ++e) {
    m.re[e] = (R)g.m[e].re;
    m.im[e] = (R)g.m[e].im;
  }
  return m;
}

// Active positions in matrix-bit order (not sorted): pos[j] = act[j].
static Fixed active_positions(const Gate &g) {
  Fixed a{};
  a.nf = g.na;
  a.set = 0;
  for (int j = 0; j < g.na; ++j) a.pos[j] = (uint8_t)g.act[j];
  return a;
}

static uint64_t active_mask(const Gate &g) {
  uint64_t m = 0;
  for (int j = 0; j < g.na; ++j) m |= 1ull << g.act[j];
  return m;
}

// Calls f(integral_constant K, integral_constant real) for K = na in 3..5.
template <typename F>
static void gatek_dispatch(int na, bool real, F &&f) {
  using std::integral_constant;
  if (real) {
    if (na == 3) f(integral_constant<int, 3>{}, integral_constant<bool, true>{});
    else if (na == 4) f(integral_constant<int, 4>{}, integral_constant<bool, true>{});
    else f(integral_constant<int, 5>{}, integral_constant<bool, true>{});
  } else {
    if (na == 3) f(integral_constant<int, 3>{}, integral_constant<bool, false>{});
    else if (na == 4) f(integral_constant<int, 4>{}, integral_constant<bool, false>{});
    else f(integral_constant<int, 5>{}, integral_constant<bool, false>{});
  }
}

// d_mat layout: row-major D x D {re, im} in precision R.  `scratch` (2^n amplitudes) is needed
// only when K > 5.
template <typename R>
hipError_t launch_gate_dense(void *psi_v, int n, const Gate &g, void *d_mat, void *scratch,
                             hipStream_t s) {
  using T = typename V2<R>::T;
  T *psi = (T *)psi_v;
  const uint64_t amask = active_mask(g);
  const Fixed fx = make_fixed(g.cmask | amask, g.cval);
  const int nfree = n - fx.nf;
  const uint64_t nfreeN = 1ull << nfree;
  const bool big = streaming(n, sizeof(T));
  // pair partners closer than a 128-byte line: plain loads (see kStreamBytes)
  const auto near = [](int q) { return (sizeof(T) << q) < 128; };
  // QS_GATE1_LANE=0 keeps k_gate1 for every target qubit
  static const bool lane_form = [] {
    const char *v = std::getenv("QS_GATE1_LANE");
    return !(v && v[0] == '0');
  }();
  if (g.na == 1 && g.act[0] < 6 && n >= 6 && (g.cmask & 0x3Full) == 0 && lane_form) {
    // measured at 30 qubits, complex128, on qubit 0: four streams 32 chunks
    // apart per workgroup 5.28 ms, one run of 4 chunks per wave 5.92 ms,
    // k_gate1 6.28-6.54 ms.  States of fewer than 128 chunks take the one-run
    // order (the four-stream order needs chunk bits 5 and 6).
    constexpr int kItems = 4;
    constexpr int kStreams = 4;
    const Fixed fxc = make_fixed(g.cmask >> 6, g.cval >> 6);
    const uint64_t nchunks = 1ull << (n - 6 - fxc.nf);
    const bool streams = nchunks >= 128;
    const unsigned grid = streams ? (unsigned)(nchunks / (4 * kItems * kStreams))
                                  : grid_for(nchunks * 64, kBlock * kItems);
    const MatArg<R, 2> ma = to_matarg<R, 2>(g);
    nt_dispatch(big, [&](auto nt) {
      constexpr bool NT = decltype(nt)::value;
      auto go = [&](auto kk) {
        if (streams)
          hipLaunchKernelGGL((k_gate1_lane<R, decltype(kk)::value, kItems, NT, kStreams>),
                             dim3(grid), dim3(kBlock), 0, s, psi, nchunks, fxc, ma);
        else
          hipLaunchKernelGGL((k_gate1_lane<R, decltype(kk)::value, kItems, NT>), dim3(grid),
                             dim3(kBlock), 0, s, psi, nchunks, fxc, ma);
      };
      switch (g.act[0]) {
        case 0: go(std::integral_constant<int, 0>{}); break;
        case 1: go(std::integral_constant<int, 1>{}); break;
        case 2: go(std::integral_constant<int, 2>{}); break;
        case 3: go(std::integral_constant<int, 3>{}); break;
        case 4: go(std::integral_constant<int, 4>{}); break;
        default: go(std::integral_constant<int, 5>{}); break;
      }
    });
  } else if (g.na == 1) {
    constexpr int kItems = 4;
    // streaming complex128 states, target not on qubits 0..2: two pairs per
    // unit, the second xbit away; xbit is qubit 12 (64 KiB of amplitudes)
    // when free, else qubit 11.  Complex64 keeps one pair per unit: at 30
    // qubits the two forms measured within +-4% of each other, either way
    // depending on the target (wall time per gate, 2.57-3.01 ms).
    // QS_GATE1_STREAMS=0 keeps one pair per unit.
    static const bool streams_on = [] {
      const char *v = std::getenv("QS_GATE1_STREAMS");
      return !(v && v[0] == '0');
    }();
    const uint64_t used = g.cmask | amask;
    int xq = -1;
    if (sizeof(T) == 16 && big && !near(g.act[0]) && streams_on) {
      for (int q : {12, 11})
        if (xq < 0 && q < n && !((used >> q) & 1ull)) xq = q;
    }
    if (xq >= 0) {
      const Fixed fxx = make_fixed(used | (1ull << xq), g.cval);
      const uint64_t nunits = nfreeN >> 1;
      hipLaunchKernelGGL((k_gate1<R, kItems, true, 2>), dim3(grid_for(nunits, kBlock * kItems)),
                         dim3(kBlock), 0, s, psi, nunits, fxx, 1ull << g.act[0],
                         to_matarg<R, 2>(g), 1ull << xq);
    } else {
      nt_dispatch(big && !near(g.act[0]), [&](auto nt) {
        hipLaunchKernelGGL((k_gate1<R, kItems, decltype(nt)::value>),
                           dim3(grid_for(nfreeN, kBlock * kItems)), dim3(kBlock), 0, s, psi,
                           nfreeN, fx, 1ull << g.act[0], to_matarg<R, 2>(g), (uint64_t)0);
      });
    }
  } else if (g.na == 2) {
    constexpr int kItems = 2;
    nt_dispatch(big && !near(g.act[0]) && !near(g.act[1]), [&](auto nt) {
      hipLaunchKernelGGL((k_gate2<R, kItems, decltype(nt)::value>),
                         dim3(grid_for(nfreeN, kBlock * kItems)), dim3(kBlock), 0, s, psi, nfreeN,
                         fx, 1ull << g.act[0], 1ull << g.act[1], to_matarg<R, 4>(g));
    });
  } else if (g.na >= 3 && g.na <= 5) {
    const Fixed ax = active_positions(g);
    const R *m = (const R *)d_mat;  // row-major {re, im}
    bool real = true;
    for (const cd &v : g.m) real = real && v.im == 0.0;
    StageMap sm{};
    // fixed low bits (active or control on qubits 0..3) break the direct
    // kernel's 64-amplitude runs: stage through LDS instead
    if ((fx_mask(fx) & 0xFull) && stage_map(fx, ax, n, sm)) {
      const unsigned sgrid = (unsigned)(nfreeN / 64);
      auto staged = [&](auto kk, auto rr) {
        nt_dispatch(big, [&](auto nt) {
          hipLaunchKernelGGL((k_gatek_staged<R, decltype(kk)::value, decltype(rr)::value,
                                             decltype(nt)::value>),
                             dim3(sgrid), dim3(64), 0, s, psi, nfreeN, fx, sm, m);
        });
      };
      gatek_dispatch(g.na, real, staged);
      return hipGetLastError();
    }
    if constexpr (std::is_same<R, float>::value) {
      // the matrix cores when the lowest 6 (or 5) index bits are free
      const uint64_t low = fx_mask(fx);
      if (g.na == 5 && !real && gatek_mfma() && (low & 0x1Full) == 0 && nfreeN % 64 == 0) {
        const bool wide = (low & 0x20ull) == 0;
        const uint64_t nchunks = nfreeN / (wide ? 64 : 32);
        const unsigned mgrid = (unsigned)std::min<uint64_t>(nchunks, 8192);
        nt_dispatch(big, [&](auto nt) {
          if (wide)
            hipLaunchKernelGGL((k_gatek5_mfma<decltype(nt)::value>), dim3(mgrid), dim3(64), 0,
                               s, (float2 *)psi, nchunks, fx, ax, (const float *)m);
          else
            hipLaunchKernelGGL((k_gatek5_mfma32<decltype(nt)::value>), dim3(mgrid), dim3(64), 0,
                               s, (float2 *)psi, nchunks, fx, ax, (const float *)m);
        });
        return hipGetLastError();
      }
    }
    const unsigned grid = grid_for(nfreeN, kBlock);
    auto direct = [&](auto kk, auto rr) {
      nt_dispatch(big, [&](auto nt) {
        hipLaunchKernelGGL((k_gatek<R, decltype(kk)::value, decltype(rr)::value,
                                    decltype(nt)::value>),
                           dim3(grid), dim3(kBlock), 0, s, psi, nfreeN, fx, ax, m);
      });
    };
    gatek_dispatch(g.na, real, direct);
  } else {
    const Fixed ax = active_positions(g);
    hipError_t e = hipMemcpyAsync(scratch, psi, sizeof(T) << n, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    unsigned grid = grid_for(nfreeN << g.na, kBlock);
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(k_gate_big<R>, dim3(grid), dim3(kBlock), 0, s, psi, (const T *)scratch,
                       nfreeN, g.na, fx, ax, (const T *)d_mat);
  }
  return hipGetLastError();
}

// ---- permutation gates ------------------------------------------------------------

// Monomial gate (kPerm) out of place: every amplitude idx of the controlled
// subspace, with act pattern r, becomes ph[r] * src[idx with act pattern
// perm[r]].  `src` is a copy of the state; amplitudes outside the
// controlled subspace are left as they are.  Two passes over the state (the
// copy and the gather) whatever the arity, against 2^na complex products
// per amplitude for the dense form.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_perm(typename V2<R>::T *__restrict__ psi,
                                                 const typename V2<R>::T *__restrict__ src,
                                                 uint64_t count, Fixed cx, Fixed ax,
                                                 const uint32_t *__restrict__ perm,
                                                 const typename V2<R>::T *__restrict__ ph) {
  using T = typename V2<R>::T;
  uint64_t amask = 0;
  for (int j = 0; j < ax.nf; ++j) amask |= 1ull << ax.pos[j];
  for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < count;
       p += (uint64_t)gridDim.x * kBlock) {
    const uint64_t idx = expand(p, cx);
    uint32_t r = 0;
    for (int j = 0; j < ax.nf; ++j) r |= (uint32_t)((idx >> ax.pos[j]) & 1ull) << j;
    const uint32_t c = perm[r];
    uint64_t sidx = idx & ~amask;
    for (int j = 0; j < ax.nf; ++j)
      if ((c >> j) & 1u) sidx |= 1ull << ax.pos[j];
    const T f = ph[r];
    psi[idx] = cmla(czero<T>(), f.x, f.y, src[sidx]);
  }
}

size_t perm_table_phase_offset(int na) { return ((sizeof(uint32_t) << na) + 15) & ~(size_t)15; }

template <typename R>
hipError_t launch_gate_perm(void *psi, int n, const Gate &g, const void *d_tab, void *scratch,
                            hipStream_t s) {
  using T = typename V2<R>::T;
  const Fixed cx = make_fixed(g.cmask, g.cval);
  const Fixed ax = active_positions(g);
  hipError_t e = hipMemcpyAsync(scratch, psi, sizeof(T) << n, hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return e;
  const uint64_t count = 1ull << (n - cx.nf);
  unsigned grid = grid_for(count, kBlock * 4);
  if (grid > 262144) grid = 262144;
  const char *tab = (const char *)d_tab;
  hipLaunchKernelGGL(k_perm<R>, dim3(grid), dim3(kBlock), 0, s, (T *)psi, (const T *)scratch,
                     count, cx, ax, (const uint32_t *)tab,
                     (const T *)(tab + perm_table_phase_offset(g.na)));
  return hipGetLastError();
}

// ---- diagonal gates -------------------------------------------------------------

// Multiply every amplitude whose control bits match by d[pattern of act bits].
template <typename R, bool kNT>
__global__ __launch_bounds__(kBlock) void k_diag(typename V2<R>::T *__restrict__ psi,
                                                 uint64_t count, Fixed cx, Fixed ax,
                                                 const typename V2<R>::T *__restrict__ d) {
  using T = typename V2<R>::T;
  for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < count;
       p += (uint64_t)gridDim.x * kBlock) {
    const uint64_t idx = expand(p, cx);
    uint32_t pat = 0;
    for (int j = 0; j < ax.nf; ++j) pat |= (uint32_t)((idx >> ax.pos[j]) & 1ull) << j;
    const T dv = d[pat];
    T a = ld<kNT>(psi + idx);
    T r = czero<T>();
    r = cmla(r, dv.x, dv.y, a);
    st<kNT>(psi + idx, r);
  }
}

template <typename R>
hipError_t launch_gate_diag(void *psi, int n, const Gate &g, void *d_mat, hipStream_t s) {
  using T = typename V2<R>::T;
  const Fixed cx = make_fixed(g.cmask, g.cval);
  const Fixed ax = active_positions(g);
  const uint64_t count = 1ull << (n - cx.nf);
  unsigned grid = grid_for(count, kBlock * 4);
  if (grid > 262144) grid = 262144;
  nt_dispatch(streaming(n, sizeof(T)), [&](auto nt) {
    hipLaunchKernelGGL((k_diag<R, decltype(nt)::value>), dim3(grid), dim3(kBlock), 0, s, (T *)psi,
                       count, cx, ax, (const T *)d_mat);
  });
  return hipGetLastError();
}

// Up to kMaxSparse (pattern, value) exceptions of a diagonal that is 1
// elsewhere.
constexpr int kMaxSparse = 16;
template <typename R>
struct SparseArg {
  uint64_t pat[kMaxSparse];
  R re[kMaxSparse], im[kMaxSparse];
  int count;
};

template <typename R, bool kNT>
__global__ __launch_bounds__(kBlock) void k_sparse(typename V2<R>::T *__restrict__ psi,
                                                   uint64_t per, Fixed fx, SparseArg<R> sp) {
  using T = typename V2<R>::T;
  const uint64_t total = per * (uint64_t)sp.count;
  for (uint64_t w = (uint64_t)blockIdx.x * kBlock + threadIdx.x; w < total;
       w += (uint64_t)gridDim.x * kBlock) {
    const uint64_t e = w / per, p = w % per;
    const uint64_t idx = expand(p, fx) | sp.pat[e];
    T a = ld<kNT>(psi + idx);
    T r = czero<T>();
    r = cmla(r, sp.re[e], sp.im[e], a);
    st<kNT>(psi + idx, r);
  }
}

template <typename R>
hipError_t launch_gate_sparse(void *psi, int n, const Gate &g, hipStream_t s) {
  using T = typename V2<R>::T;
  SparseArg<R> sp{};
  sp.count = (int)g.spat.size();
  if (sp.count > kMaxSparse) return hipErrorInvalidValue;
  if (sp.count == 0) return hipSuccess;
  for (int e = 0; e < sp.count; ++e) {
    sp.pat[e] = g.spat[e];
    sp.re[e] = (R)g.m[e].re;
    sp.im[e] = (R)g.m[e].im;
  }
  const Fixed fx = make_fixed(g.smask, 0);
  const uint64_t per = 1ull << (n - fx.nf);
  unsigned grid = grid_for(per * sp.count, kBlock * 4);
  if (grid > 262144) grid = 262144;
  nt_dispatch(streaming(n, sizeof(T)), [&](auto nt) {
    hipLaunchKernelGGL((k_sparse<R, decltype(nt)::value>), dim3(grid), dim3(kBlock), 0, s,
                       (T *)psi, per, fx, sp);
  });
  return hipGetLastError();
}

// ---- reductions -------------------------------------------------------------------

template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double (*sh)[kBlock]) {
#pragma unroll
  for (int c = 0; c < N; ++c) sh[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int c = 0; c < N; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
    }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < N; ++c) v[c] = sh[c][0];
}

// Second stage: one block sums nblocks partials of N channels (fixed order).
template <int N>
__global__ __launch_bounds__(kBlock) void k_finish(const double *__restrict__ part, int nblocks,
                                                   double *__restrict__ out) {
  __shared__ double sh[N][kBlock];
  double v[N];
#pragma unroll
  for (int c = 0; c < N; ++c) v[c] = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += kBlock)
#pragma unroll
    for (int c = 0; c < N; ++c) v[c] += part[(size_t)b * N + c];
  block_sum<N>(v, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < N; ++c) out[c] = v[c];
}

static inline int reduce_grid(uint64_t len) {
  uint64_t g = (len + kBlock - 1) / kBlock;
  if (g > (uint64_t)kReduceBlocks) g = kReduceBlocks;
  if (g < 1) g = 1;
  return (int)g;
}

// Block order of the reductions over a state of at least kReduceBlocks *
// kRedTrip amplitudes: block b owns the len / gridDim.x consecutive
// amplitudes from b * len / gridDim.x and reads them in trips of kRedTrip
// (a contiguous 64 KiB of complex128, 16 nontemporal loads in flight per
// thread); each thread adds its amplitudes in index order.  Smaller states
// take the grid-stride order.  Measured at 2^30 complex128 amplitudes, read
// and split by one bit: 2.42-2.45 ms (0.88 of 8 TB/s) against 3.07 ms for the
// grid-stride order with four ordinary loads in flight (2.92 ms nontemporal).
constexpr int kRedLoads = 16;
constexpr uint64_t kRedTrip = (uint64_t)kRedLoads * kBlock;

static inline bool reduce_blocked(uint64_t len) {
  return len >= (uint64_t)kReduceBlocks * kRedTrip;
}

template <typename R, bool kBlocked>
__global__ __launch_bounds__(kBlock) void k_prob_split(const typename V2<R>::T *__restrict__ psi,
                                                       uint64_t len, int target,
                                                       double *__restrict__ part) {
  __shared__ double sh[2][kBlock];
  double v[2] = {0.0, 0.0};
  // each thread adds its amplitudes in index order (k_collapse_split relies
  // on the same order)
  if constexpr (kBlocked) {
    const uint64_t per = len / gridDim.x;
    const uint64_t b0 = (uint64_t)blockIdx.x * per + threadIdx.x;
    for (uint64_t w = b0; w < b0 + per; w += kRedTrip) {
      typename V2<R>::T a[kRedLoads];
#pragma unroll
      for (int k = 0; k < kRedLoads; ++k) a[k] = ld<true>(psi + w + (uint64_t)k * kBlock);
#pragma unroll
      for (int k = 0; k < kRedLoads; ++k) {
        const double re = (double)a[k].x, im = (double)a[k].y;
        const double p = re * re + im * im;
        if (((w + (uint64_t)k * kBlock) >> target) & 1ull)
          v[1] += p;
        else
          v[0] += p;
      }
    }
    block_sum<2>(v, sh);
    if (threadIdx.x == 0) {
      part[2 * blockIdx.x] = v[0];
      part[2 * blockIdx.x + 1] = v[1];
    }
    return;
  }
  // four loads in flight per trip
  const uint64_t S = (uint64_t)gridDim.x * kBlock;
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  for (; i + 3 * S < len; i += 4 * S) {
    typename V2<R>::T a[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) a[k] = psi[i + k * S];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double re = (double)a[k].x, im = (double)a[k].y;
      const double p = re * re + im * im;
      if (((i + k * S) >> target) & 1ull)
        v[1] += p;
      else
        v[0] += p;
    }
  }
  for (; i < len; i += S) {
    const auto a = psi[i];
    const double re = (double)a.x, im = (double)a.y;
    const double p = re * re + im * im;
    if ((i >> target) & 1ull)
      v[1] += p;
    else
      v[0] += p;
  }
  block_sum<2>(v, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = v[0];
    part[2 * blockIdx.x + 1] = v[1];
  }
}

template <typename R>
hipError_t launch_prob_split(const void *psi, int n, int target, double *d_part, double *d_out,
                             hipStream_t s) {
  const uint64_t len = 1ull << n;
  const int g = reduce_grid(len);
  if (reduce_blocked(len))
    hipLaunchKernelGGL((k_prob_split<R, true>), dim3(g), dim3(kBlock), 0, s,
                       (const typename V2<R>::T *)psi, len, target, d_part);
  else
    hipLaunchKernelGGL((k_prob_split<R, false>), dim3(g), dim3(kBlock), 0, s,
                       (const typename V2<R>::T *)psi, len, target, d_part);
  hipLaunchKernelGGL(k_finish<2>, dim3(1), dim3(kBlock), 0, s, (const double *)d_part, g, d_out);
  return hipGetLastError();
}

// The collapse of try_measure (src/register.rs:324-339): zero the amplitudes
// whose target bit differs from the outcome; divide the rest, component by
// component, by sqrt(total).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_collapse(typename V2<R>::T *__restrict__ psi,
                                                     uint64_t len, int target, int keep,
                                                     double sq) {
  using T = typename V2<R>::T;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * kBlock) {
    T a = psi[i];
    if ((int)((i >> target) & 1ull) != keep) {
      a = czero<T>();
    } else {
      a.x = (R)((double)a.x / sq);
      a.y = (R)((double)a.y / sq);
    }
    psi[i] = a;
  }
}

// The collapse of one measure fused with the probability split of the next
// (src/register.rs:297-339 twice): amplitudes whose bit `ctarget` differs
// from `keep` are zeroed, the rest divided by sqrt(total), and the squared
// moduli of the values stored (after rounding to R) are summed by bit
// `target`, over the same grid, in the same order and with the same block
// sums as k_prob_split on the collapsed state, so {prob_0, prob_1} are the
// same bits as a separate split would give.  One pass instead of two: the
// split's full read of the collapsed state is saved.
template <typename R, bool kBlocked>
__global__ __launch_bounds__(kBlock) void k_collapse_split(typename V2<R>::T *__restrict__ psi,
                                                           uint64_t len, int ctarget, int keep,
                                                           double sq, int target,
                                                           double *__restrict__ part) {
  using T = typename V2<R>::T;
  __shared__ double sh[2][kBlock];
  double v[2] = {0.0, 0.0};
  auto one = [&](uint64_t i, T a) {
    if ((int)((i >> ctarget) & 1ull) != keep) {
      a = czero<T>();
    } else {
      a.x = (R)((double)a.x / sq);
      a.y = (R)((double)a.y / sq);
    }
    st<kBlocked>(psi + i, a);
    const double re = (double)a.x, im = (double)a.y;
    const double p = re * re + im * im;
    if ((i >> target) & 1ull)
      v[1] += p;
    else
      v[0] += p;
  };
  // the trip structure and order of k_prob_split; only kept amplitudes are
  // read
  if constexpr (kBlocked) {
    const uint64_t per = len / gridDim.x;
    const uint64_t b0 = (uint64_t)blockIdx.x * per + threadIdx.x;
    for (uint64_t w = b0; w < b0 + per; w += kRedTrip) {
      T a[kRedLoads];
#pragma unroll
      for (int k = 0; k < kRedLoads; ++k) {
        const uint64_t i = w + (uint64_t)k * kBlock;
        a[k] = (int)((i >> ctarget) & 1ull) == keep ? ld<true>(psi + i) : czero<T>();
      }
#pragma unroll
      for (int k = 0; k < kRedLoads; ++k) one(w + (uint64_t)k * kBlock, a[k]);
    }
    block_sum<2>(v, sh);
    if (threadIdx.x == 0) {
      part[2 * blockIdx.x] = v[0];
      part[2 * blockIdx.x + 1] = v[1];
    }
    return;
  }
  const uint64_t S = (uint64_t)gridDim.x * kBlock;
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  for (; i + 3 * S < len; i += 4 * S) {
    T a[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      a[k] = (int)(((i + k * S) >> ctarget) & 1ull) == keep ? psi[i + k * S] : czero<T>();
#pragma unroll
    for (int k = 0; k < 4; ++k) one(i + k * S, a[k]);
  }
  for (; i < len; i += S) one(i, (int)((i >> ctarget) & 1ull) == keep ? psi[i] : czero<T>());
  block_sum<2>(v, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = v[0];
    part[2 * blockIdx.x + 1] = v[1];
  }
}

template <typename R>
hipError_t launch_collapse_split(void *psi, int n, int ctarget, int keep, double total,
                                 int target, double *d_part, double *d_out, hipStream_t s) {
  const uint64_t len = 1ull << n;
  const int g = reduce_grid(len);
  if (reduce_blocked(len))
    hipLaunchKernelGGL((k_collapse_split<R, true>), dim3(g), dim3(kBlock), 0, s,
                       (typename V2<R>::T *)psi, len, ctarget, keep, sqrt(total), target, d_part);
  else
    hipLaunchKernelGGL((k_collapse_split<R, false>), dim3(g), dim3(kBlock), 0, s,
                       (typename V2<R>::T *)psi, len, ctarget, keep, sqrt(total), target, d_part);
  hipLaunchKernelGGL(k_finish<2>, dim3(1), dim3(kBlock), 0, s, (const double *)d_part, g, d_out);
  return hipGetLastError();
}

// Deferred measurement, decided on the device: probs = {prob_0, prob_1} from
// k_prob_split, the rule of src/register.rs:317-321 (result = x > prob_0),
// the outcome into *slot and {keep, sqrt(total)} into coll for
// k_collapse_dev.
__global__ void k_measure_decide(const double *__restrict__ probs, double x,
                                 unsigned char *__restrict__ slot, double *__restrict__ coll) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const double p0 = probs[0], p1 = probs[1];
  const int res = x > p0;
  coll[0] = (double)res;
  coll[1] = sqrt(res ? p1 : p0);
  *slot = (unsigned char)res;
}

// k_collapse with its outcome and scale read from device memory.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_collapse_dev(typename V2<R>::T *__restrict__ psi,
                                                         uint64_t len, int target,
                                                         const double *__restrict__ coll) {
  using T = typename V2<R>::T;
  const int keep = (int)coll[0];
  const double sq = coll[1];
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * kBlock) {
    T a = psi[i];
    if ((int)((i >> target) & 1ull) != keep) {
      a = czero<T>();
    } else {
      a.x = (R)((double)a.x / sq);
      a.y = (R)((double)a.y / sq);
    }
    psi[i] = a;
  }
}

template <typename R>
hipError_t launch_measure_deferred(void *psi, int n, int target, double x, double *d_part,
                                   double *d_probs, double *d_coll, unsigned char *d_slot,
                                   hipStream_t s) {
  hipError_t e = launch_prob_split<R>(psi, n, target, d_part, d_probs, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_measure_decide, dim3(1), dim3(64), 0, s, (const double *)d_probs, x,
                     d_slot, d_coll);
  const uint64_t len = 1ull << n;
  unsigned grid = grid_for(len, kBlock * 4);
  if (grid > 262144) grid = 262144;
  hipLaunchKernelGGL(k_collapse_dev<R>, dim3(grid), dim3(kBlock), 0, s, (typename V2<R>::T *)psi,
                     len, target, (const double *)d_coll);
  return hipGetLastError();
}

// The collapse onto the amplitudes i with (i & mask) == pat (k_collapse:
// one target bit; the mask collapse of measure_sequence: every decided
// qubit), over chunks of 64 amplitudes in the stream order of k_gate1_lane
// (four streams 32 chunks apart per workgroup, 16 chunks per workgroup per
// stream); lane l takes amplitude l of each chunk, reads it only when it is
// kept.  nchunks is a power of two >= 128; the host launches
// nchunks / 64 workgroups.  Measured at 30 qubits, complex128, collapse on
// qubit 29 (half the state read, all of it written): 4.08 ms (6.3 TB/s)
// against 4.78 ms for the grid-stride k_collapse.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_collapse_streams(typename V2<R>::T *__restrict__ psi,
                                                             uint64_t nchunks, uint64_t mask,
                                                             uint64_t pat, double sq) {
  using T = typename V2<R>::T;
  constexpr int kItems = 4, S = 4;
  if (((uint64_t)blockIdx.x + 1) * (4 * kItems * S) > nchunks) return;
  const uint64_t lane = threadIdx.x & 63;
  uint64_t idx[kItems * S];
  T a[kItems * S];
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint64_t u = (uint64_t)blockIdx.x * (4 * kItems) + it * 4 + (threadIdx.x >> 6);
    const uint64_t c0 = (u & 31ull) | ((u >> 5) << 7);
#pragma unroll
    for (int st_ = 0; st_ < S; ++st_) idx[it * S + st_] = ((c0 | ((uint64_t)st_ << 5)) << 6) | lane;
  }
#pragma unroll
  for (int j = 0; j < kItems * S; ++j)
    a[j] = (idx[j] & mask) == pat ? ld<true>(psi + idx[j]) : czero<T>();
#pragma unroll
  for (int j = 0; j < kItems * S; ++j) {
    T r = a[j];
    r.x = (R)((double)r.x / sq);
    r.y = (R)((double)r.y / sq);
    st<true>(psi + idx[j], r);
  }
}

template <typename R>
hipError_t launch_collapse_pattern(void *psi, int n, uint64_t mask, uint64_t pat, double total,
                                   hipStream_t s) {
  const uint64_t nchunks = (1ull << n) >> 6;
  if (n < 13) return hipErrorInvalidValue;  // the stream order needs 128 chunks
  hipLaunchKernelGGL(k_collapse_streams<R>, dim3((unsigned)(nchunks / 64)), dim3(kBlock), 0, s,
                     (typename V2<R>::T *)psi, nchunks, mask, pat, sqrt(total));
  return hipGetLastError();
}

template <typename R>
hipError_t launch_collapse(void *psi, int n, int target, int keep, double total, hipStream_t s) {
  const uint64_t len = 1ull << n;
  if (streaming(n, sizeof(typename V2<R>::T))) {
    // states above the last-level cache: the stream order
    return launch_collapse_pattern<R>(psi, n, 1ull << target, (uint64_t)keep << target, total, s);
  }
  unsigned grid = grid_for(len, kBlock * 4);
  if (grid > 262144) grid = 262144;
  hipLaunchKernelGGL(k_collapse<R>, dim3(grid), dim3(kBlock), 0, s, (typename V2<R>::T *)psi, len,
                     target, keep, sqrt(total));
  return hipGetLastError();
}

template <typename R, bool kBlocked>
__global__ __launch_bounds__(kBlock) void k_norm(const typename V2<R>::T *__restrict__ psi,
                                                 uint64_t len, double *__restrict__ part) {
  __shared__ double sh[1][kBlock];
  double v[1] = {0.0};
  if constexpr (kBlocked) {
    const uint64_t per = len / gridDim.x;
    const uint64_t b0 = (uint64_t)blockIdx.x * per + threadIdx.x;
    for (uint64_t w = b0; w < b0 + per; w += kRedTrip) {
      typename V2<R>::T a[kRedLoads];
#pragma unroll
      for (int k = 0; k < kRedLoads; ++k) a[k] = ld<true>(psi + w + (uint64_t)k * kBlock);
#pragma unroll
      for (int k = 0; k < kRedLoads; ++k)
        v[0] += (double)a[k].x * (double)a[k].x + (double)a[k].y * (double)a[k].y;
    }
  } else {
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
         i += (uint64_t)gridDim.x * kBlock) {
      const auto a = psi[i];
      v[0] += (double)a.x * (double)a.x + (double)a.y * (double)a.y;
    }
  }
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

template <typename R>
hipError_t launch_norm(const void *psi, int n, double *d_part, double *d_out, hipStream_t s) {
  const uint64_t len = 1ull << n;
  const int g = reduce_grid(len);
  if (reduce_blocked(len))
    hipLaunchKernelGGL((k_norm<R, true>), dim3(g), dim3(kBlock), 0, s,
                       (const typename V2<R>::T *)psi, len, d_part);
  else
    hipLaunchKernelGGL((k_norm<R, false>), dim3(g), dim3(kBlock), 0, s,
                       (const typename V2<R>::T *)psi, len, d_part);
  hipLaunchKernelGGL(k_finish<1>, dim3(1), dim3(kBlock), 0, s, (const double *)d_part, g, d_out);
  return hipGetLastError();
}

template <typename Ra, typename Rb>
__global__ __launch_bounds__(kBlock) void k_inner(const typename V2<Ra>::T *__restrict__ a,
                                                  const typename V2<Rb>::T *__restrict__ b,
                                                  uint64_t len, double *__restrict__ part) {
  __shared__ double sh[2][kBlock];
  double v[2] = {0.0, 0.0};
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * kBlock) {
    const auto x = a[i];
    const auto y = b[i];
    const double xr = x.x, xi = x.y, yr = y.x, yi = y.y;
    v[0] += xr * yr + xi * yi;  // Re conj(x) y
    v[1] += xr * yi - xi * yr;  // Im conj(x) y
  }
  block_sum<2>(v, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = v[0];
    part[2 * blockIdx.x + 1] = v[1];
  }
}

template <typename Ra, typename Rb>
static void inner_launch(const void *a, const void *b, uint64_t len, int g, double *d_part,
                         hipStream_t s) {
  hipLaunchKernelGGL((k_inner<Ra, Rb>), dim3(g), dim3(kBlock), 0, s,
                     (const typename V2<Ra>::T *)a, (const typename V2<Rb>::T *)b, len, d_part);
}

hipError_t launch_inner(const void *a, bool a_c64, const void *b, bool b_c64, int n,
                        double *d_part, double *d_out, hipStream_t s) {
  const uint64_t len = 1ull << n;
  const int g = reduce_grid(len);
  if (!a_c64 && !b_c64) inner_launch<double, double>(a, b, len, g, d_part, s);
  if (!a_c64 && b_c64) inner_launch<double, float>(a, b, len, g, d_part, s);
  if (a_c64 && !b_c64) inner_launch<float, double>(a, b, len, g, d_part, s);
  if (a_c64 && b_c64) inner_launch<float, float>(a, b, len, g, d_part, s);
  hipLaunchKernelGGL(k_finish<2>, dim3(1), dim3(kBlock), 0, s, (const double *)d_part, g, d_out);
  return hipGetLastError();
}

// Count of amplitudes where !(|a_i - b_i| < tol) -- NaN counts as unequal,
// as `e.norm() < 1e-8` is false for NaN in PartialEq (src/register.rs:411).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_neq(const typename V2<R>::T *__restrict__ a,
                                                const typename V2<R>::T *__restrict__ b,
                                                uint64_t len, double tol,
                                                double *__restrict__ part) {
  __shared__ double sh[1][kBlock];
  double v[1] = {0.0};
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * kBlock) {
    const auto x = a[i];
    const auto y = b[i];
    const double dr = (double)x.x - (double)y.x, di = (double)x.y - (double)y.y;
    const double d = hypot(dr, di);
    if (!(d < tol)) v[0] += 1.0;
  }
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

template <typename R>
hipError_t launch_neq_count(const void *a, const void *b, int n, double tol, double *d_part,
                            double *d_out, hipStream_t s) {
  const uint64_t len = 1ull << n;
  const int g = reduce_grid(len);
  hipLaunchKernelGGL(k_neq<R>, dim3(g), dim3(kBlock), 0, s, (const typename V2<R>::T *)a,
                     (const typename V2<R>::T *)b, len, tol, d_part);
  hipLaunchKernelGGL(k_finish<1>, dim3(1), dim3(kBlock), 0, s, (const double *)d_part, g, d_out);
  return hipGetLastError();
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_probs(const typename V2<R>::T *__restrict__ psi,
                                                  uint64_t len, double *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * kBlock) {
    const auto a = psi[i];
    out[i] = (double)a.x * (double)a.x + (double)a.y * (double)a.y;
  }
}

template <typename R>
hipError_t launch_probabilities(const void *psi, int n, double *d_out, hipStream_t s) {
  const uint64_t len = 1ull << n;
  unsigned grid = grid_for(len, kBlock * 4);
  if (grid > 262144) grid = 262144;
  hipLaunchKernelGGL(k_probs<R>, dim3(grid), dim3(kBlock), 0, s, (const typename V2<R>::T *)psi,
                     len, d_out);
  return hipGetLastError();
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_from_c128(typename V2<R>::T *__restrict__ psi,
                                                      const double *__restrict__ src,
                                                      uint64_t count) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * kBlock) {
    typename V2<R>::T v;
    v.x = (R)src[2 * i];
    v.y = (R)src[2 * i + 1];
    psi[i] = v;
  }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_to_c128(double *__restrict__ dst,
                                                    const typename V2<R>::T *__restrict__ psi,
                                                    uint64_t count) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * kBlock) {
    const auto v = psi[i];
    dst[2 * i] = (double)v.x;
    dst[2 * i + 1] = (double)v.y;
  }
}

// Amplitudes at `count` given indices, widened to complex128.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_gather_amps(double *__restrict__ dst,
                                                        const typename V2<R>::T *__restrict__ psi,
                                                        const uint64_t *__restrict__ idx,
                                                        uint64_t count) {
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= count) return;
  const auto v = psi[idx[i]];
  dst[2 * i] = (double)v.x;
  dst[2 * i + 1] = (double)v.y;
}

template <typename R>
hipError_t launch_gather_amps(double *dst, const void *psi, const uint64_t *d_idx, uint64_t count,
                              hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_amps<R>, dim3(grid_for(count, kBlock)), dim3(kBlock), 0, s, dst,
                     (const typename V2<R>::T *)psi, d_idx, count);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_convert_from_c128(void *psi, const double *src, uint64_t count, hipStream_t s) {
  unsigned grid = grid_for(count, kBlock * 4);
  if (grid > 262144) grid = 262144;
  hipLaunchKernelGGL(k_from_c128<R>, dim3(grid), dim3(kBlock), 0, s, (typename V2<R>::T *)psi, src,
                     count);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_convert_to_c128(double *dst, const void *psi, uint64_t count, hipStream_t s) {
  unsigned grid = grid_for(count, kBlock * 4);
  if (grid > 262144) grid = 262144;
  hipLaunchKernelGGL(k_to_c128<R>, dim3(grid), dim3(kBlock), 0, s, dst,
                     (const typename V2<R>::T *)psi, count);
  return hipGetLastError();
}

// Histogram of |psi|^2 over the bit pattern of m qubits (bin bit j <->
// qubit pos_order[j]).  Block b owns a contiguous range of the free counter
// (the index bits outside the m qubits); thread t owns bins t, t + 256, ...
// and sums them over the block's range in a fixed order, so the result is
// deterministic without atomics.
struct PosList {
  uint8_t p[64];
};

// Histogram of |psi|^2 over the 2^m patterns of the qubits po.p[0..m):
// block b walks the free indices j of its chunk in order; for each j (the
// same for the whole block, so expand() runs on scalar registers) thread t
// loads the amplitudes of its BPT bins (h = t + 256 k) at once, so every
// thread keeps BPT loads in flight.  With fewer than 256 bins, the threads
// of bin h split the chunk S = 256 / 2^m ways (thread t: phase t / 2^m) and
// each phase is a row of its own.  part[(b * S + phase) * 2^m + h].
template <typename R, int BPT>
__global__ __launch_bounds__(kBlock) void k_set_hist(const typename V2<R>::T *__restrict__ psi,
                                                     Fixed fx, PosList po, int m, uint64_t nfree,
                                                     uint64_t per_block,
                                                     double *__restrict__ part) {
  const int H = 1 << m;
  const int S = H >= kBlock ? 1 : kBlock / H;
  const int phase = H >= kBlock ? 0 : (int)threadIdx.x / H;
  // bin h's offset is the deposit of its bits at po.p: the thread's part
  // (bits 0..7 of h, h < 256: one value per thread) ORed with bin k's part
  // (bits 8.. of t + 256 k: the same for every thread, kept on scalar
  // registers), so that a thread holds one 64-bit offset, not BPT
  const int lo_bits = m < 8 ? m : 8;
  const int tl = H >= kBlock ? (int)threadIdx.x : (int)threadIdx.x % H;
  uint64_t dlo = 0;
  for (int b = 0; b < lo_bits; ++b)
    if ((tl >> b) & 1) dlo |= 1ull << po.p[b];
  uint64_t dhi[BPT];
  double acc[BPT];
#pragma unroll
  for (int k = 0; k < BPT; ++k) {
    uint64_t d = 0;
    for (int b = lo_bits; b < m; ++b)
      if (((k * kBlock) >> b) & 1) d |= 1ull << po.p[b];
    dhi[k] = d;
    acc[k] = 0.0;
  }
  const uint64_t j0 = (uint64_t)blockIdx.x * per_block;
  uint64_t j1 = j0 + per_block;
  if (j1 > nfree) j1 = nfree;
  for (uint64_t jj = j0; jj < j1; jj += (uint64_t)S) {
    const uint64_t j = jj + (uint64_t)phase;
    if (j >= j1) break;
    const uint64_t e = expand(j, fx);
    typename V2<R>::T a[BPT];
#pragma unroll
    for (int k = 0; k < BPT; ++k) a[k] = psi[(e | dhi[k]) | dlo];
#pragma unroll
    for (int k = 0; k < BPT; ++k)
      acc[k] += (double)a[k].x * (double)a[k].x + (double)a[k].y * (double)a[k].y;
  }
#pragma unroll
  for (int k = 0; k < BPT; ++k) {
    const int h = H >= kBlock ? (int)threadIdx.x + k * kBlock : (int)threadIdx.x % H;
    part[((size_t)blockIdx.x * S + phase) * H + h] = acc[k];
  }
}

__global__ __launch_bounds__(kBlock) void k_hist_finish(const double *__restrict__ part,
                                                        int blocks, int H,
                                                        double *__restrict__ out) {
  for (int h = blockIdx.x * kBlock + threadIdx.x; h < H; h += gridDim.x * kBlock) {
    double v = 0.0;
    for (int b = 0; b < blocks; ++b) v += part[(size_t)b * H + h];
    out[h] = v;
  }
}

// d_part holds blocks * max(1, 256 / 2^m) rows of 2^m partial sums.
template <typename R>
hipError_t launch_set_histogram(const void *psi, int n, const Fixed &fx, const uint8_t *pos_order,
                                int m, double *d_part, int blocks, double *d_hist,
                                hipStream_t s) {
  const int H = 1 << m;
  if (m > 12) return hipErrorInvalidValue;
  const uint64_t nfree = 1ull << (n - m);
  const uint64_t per_block = (nfree + blocks - 1) / blocks;
  PosList po{};
  for (int j = 0; j < m; ++j) po.p[j] = pos_order[j];
  const int bpt = H >= kBlock ? H / kBlock : 1;
  const int rows = blocks * (H >= kBlock ? 1 : kBlock / H);
  using T = typename V2<R>::T;
  auto kern = bpt == 1 ? k_set_hist<R, 1>
              : bpt == 2 ? k_set_hist<R, 2>
              : bpt == 4 ? k_set_hist<R, 4>
              : bpt == 8 ? k_set_hist<R, 8>
                         : k_set_hist<R, 16>;
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), 0, s, (const T *)psi, fx, po, m, nfree,
                     per_block, d_part);
  hipLaunchKernelGGL(k_hist_finish, dim3((H + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     (const double *)d_part, rows, H, d_hist);
  return hipGetLastError();
}

// |psi|^2 of the amplitudes expand(j, fx) for j < count.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_gather_probs(const typename V2<R>::T *__restrict__ psi,
                                                         Fixed fx, uint64_t count,
                                                         double *__restrict__ out) {
  for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < count;
       j += (uint64_t)gridDim.x * kBlock) {
    const auto a = psi[expand(j, fx)];
    out[j] = (double)a.x * (double)a.x + (double)a.y * (double)a.y;
  }
}

template <typename R>
hipError_t launch_gather_probs(const void *psi, const Fixed &fx, uint64_t count, double *d_out,
                               hipStream_t s) {
  unsigned grid = grid_for(count, kBlock);
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(k_gather_probs<R>, dim3(grid), dim3(kBlock), 0, s,
                     (const typename V2<R>::T *)psi, fx, count, d_out);
  return hipGetLastError();
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_collapse_to(typename V2<R>::T *__restrict__ psi,
                                                        uint64_t len, uint64_t idx, double sq) {
  using T = typename V2<R>::T;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * kBlock) {
    if (i == idx) {
      T a = psi[i];
      a.x = (R)((double)a.x / sq);
      a.y = (R)((double)a.y / sq);
      psi[i] = a;
    } else {
      psi[i] = czero<T>();
    }
  }
}

template <typename R>
hipError_t launch_collapse_to(void *psi, int n, uint64_t idx, double total, hipStream_t s) {
  const uint64_t len = 1ull << n;
  unsigned grid = grid_for(len, kBlock * 4);
  if (grid > 262144) grid = 262144;
  hipLaunchKernelGGL(k_collapse_to<R>, dim3(grid), dim3(kBlock), 0, s, (typename V2<R>::T *)psi,
                     len, idx, sqrt(total));
  return hipGetLastError();
}

// One block per chunk; within a block a fixed tree order.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_chunk_sums(const typename V2<R>::T *__restrict__ psi,
                                                       uint64_t len, uint64_t chunk,
                                                       double *__restrict__ out) {
  __shared__ double sh[1][kBlock];
  double v[1] = {0.0};
  const uint64_t start = (uint64_t)blockIdx.x * chunk;
  uint64_t end = start + chunk;
  if (end > len) end = len;
  for (uint64_t i = start + threadIdx.x; i < end; i += kBlock) {
    const auto a = psi[i];
    v[0] += (double)a.x * (double)a.x + (double)a.y * (double)a.y;
  }
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}

template <typename R>
hipError_t launch_chunk_sums(const void *psi, int n, uint64_t chunk, double *d_out,
                             hipStream_t s) {
  const uint64_t len = 1ull << n;
  const uint64_t nchunks = (len + chunk - 1) / chunk;
  hipLaunchKernelGGL(k_chunk_sums<R>, dim3((unsigned)nchunks), dim3(kBlock), 0, s,
                     (const typename V2<R>::T *)psi, len, chunk, d_out);
  return hipGetLastError();
}

#define QS_INSTANTIATE(R)                                                                        \
  template hipError_t launch_init_basis<R>(void *, int, uint64_t, hipStream_t);                 \
  template hipError_t launch_init_product<R>(void *, int, const double *, hipStream_t);         \
  template hipError_t launch_gate_dense<R>(void *, int, const Gate &, void *, void *,            \
                                           hipStream_t);                                         \
  template hipError_t launch_gate_diag<R>(void *, int, const Gate &, void *, hipStream_t);      \
  template hipError_t launch_gate_sparse<R>(void *, int, const Gate &, hipStream_t);            \
  template hipError_t launch_gate_perm<R>(void *, int, const Gate &, const void *, void *,       \
                                          hipStream_t);                                          \
  template hipError_t launch_prob_split<R>(const void *, int, int, double *, double *,          \
                                           hipStream_t);                                         \
  template hipError_t launch_collapse<R>(void *, int, int, int, double, hipStream_t);           \
  template hipError_t launch_collapse_pattern<R>(void *, int, uint64_t, uint64_t, double,        \
                                                 hipStream_t);                                   \
  template hipError_t launch_collapse_split<R>(void *, int, int, int, double, int, double *,     \
                                               double *, hipStream_t);                           \
  template hipError_t launch_measure_deferred<R>(void *, int, int, double, double *, double *,  \
                                                 double *, unsigned char *, hipStream_t);       \
  template hipError_t launch_norm<R>(const void *, int, double *, double *, hipStream_t);       \
  template hipError_t launch_neq_count<R>(const void *, const void *, int, double, double *,    \
                                          double *, hipStream_t);                                \
  template hipError_t launch_probabilities<R>(const void *, int, double *, hipStream_t);        \
  template hipError_t launch_convert_from_c128<R>(void *, const double *, uint64_t,             \
                                                  hipStream_t);                                  \
  template hipError_t launch_convert_to_c128<R>(double *, const void *, uint64_t, hipStream_t); \
  template hipError_t launch_gather_amps<R>(double *, const void *, const uint64_t *, uint64_t,  \
                                            hipStream_t);                                       \
  template hipError_t launch_set_histogram<R>(const void *, int, const Fixed &, const uint8_t *, \
                                              int, double *, int, double *, hipStream_t);        \
  template hipError_t launch_gather_probs<R>(const void *, const Fixed &, uint64_t, double *,    \
                                             hipStream_t);                                       \
  template hipError_t launch_collapse_to<R>(void *, int, uint64_t, double, hipStream_t);        \
  template hipError_t launch_chunk_sums<R>(const void *, int, uint64_t, double *, hipStream_t);

QS_INSTANTIATE(double)
QS_INSTANTIATE(float)

}  // namespace qs

// qs_adjoint.hip -- adjoint-method gradients of <psi(theta)|H|psi(theta)>.
//
// The reference has no differentiation (quaru 0.2.0 only has the OpenQASM U
// gate as a parameterised operation, src/openqasm.rs:145-163, whose
// conventions the rotations here follow: RY(t) = u(t, 0, 0), RZ(t) =
// u(0, 0, t)).
//
// Batch layout: B parameter sets are padded to Bp = 2^nb states stored back
// to back, [Bp][2^n]; the batch index is then nb extra high index bits, so a
// fixed gate on qubit q < n is one launch of the ordinary gate kernels over
// n + nb "qubits".  Rotations carry one matrix per batch element.
//
// Forward:  psi_b = U_N ... U_1 |0>,  E_b = sum_x |psi_b(x)|^2 h(x) with
//           h(x) = sum_t c_t (-1)^popcount(x & mask_t).
// Backward (one reverse sweep, no per-parameter reruns):
//           lambda = h . psi; for i = N..1:
//             rotation i on parameter p:  g_p += Im <lambda| sigma |psi>
//               (= 2 Re <lambda| dU_i/dtheta U_i^dagger |psi>, with
//                dU/dtheta = -(i/2) sigma U),
//             then psi <- U_i^dagger psi, lambda <- U_i^dagger lambda
//           (the three steps are one fused pass over psi and lambda).
// Checkpoints: with checkpoint_every = L > 0 the forward keeps psi after
// every L-th layer, and the backward re-anchors psi to that copy when it
// reaches it, bounding the rounding drift of the uncompute (and, in
// complex64, keeping the reverse sweep within the forward's accuracy).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/quaru_hip.h"
#include "qs_internal.h"
#include "qs_sweep.h"

namespace {

template <typename R>
struct AV2;
template <>
struct AV2<double> {
  using T = double2;
};
template <>
struct AV2<float> {
  using T = float2;
};

constexpr int kBlock = 256;
constexpr int kItems = 4;

struct POp {
  int kind = 0;  // 0 fixed gate, 1 rotation
  int axis = 0;
  int target = 0;
  size_t param = 0;
  qs::Gate fwd, inv;
  int rot_slot = -1;
};

int fail_arg(const char *msg) { return qs::set_error(QS_ERR_ARGUMENT, 0, 0, msg); }

__device__ __forceinline__ uint64_t insert0(uint64_t p, int q) {
  const uint64_t lo = p & ((1ull << q) - 1ull);
  return ((p ^ lo) << 1) | lo;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_init_zero_states(typename AV2<R>::T *psi, int n,
                                                             uint64_t total) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * kBlock) {
    typename AV2<R>::T v;
    v.x = ((i & ((1ull << n) - 1ull)) == 0) ? (R)1 : (R)0;
    v.y = 0;
    psi[i] = v;
  }
}

__global__ __launch_bounds__(kBlock) void k_htable(double *h, int n,
                                                   const unsigned long long *__restrict__ masks,
                                                   const double *__restrict__ coeffs,
                                                   int nterms) {
  const uint64_t len = 1ull << n;
  for (uint64_t x = (uint64_t)blockIdx.x * kBlock + threadIdx.x; x < len;
       x += (uint64_t)gridDim.x * kBlock) {
    double v = 0.0;
    for (int t = 0; t < nterms; ++t)
      v += (__popcll((unsigned long long)x & masks[t]) & 1) ? -coeffs[t] : coeffs[t];
    h[x] = v;
  }
}

// Rotation with one 2x2 matrix per batch element (mats[b*4 + e], complex).
// Work item w is pair p = w mod 2^(n-1) of state b = w >> (n - 1).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_rot(typename AV2<R>::T *__restrict__ psi, int n,
                                                int q, const typename AV2<R>::T *__restrict__ mats,
                                                uint64_t total_pairs) {
  using T = typename AV2<R>::T;
  const uint64_t base = (uint64_t)blockIdx.x * (kBlock * kItems) + threadIdx.x;
  const uint64_t pmask = (1ull << (n - 1)) - 1ull;
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint64_t w = base + (uint64_t)it * kBlock;
    if (w >= total_pairs) break;
    const uint64_t b = w >> (n - 1), p = w & pmask;
    const uint64_t off = b << n;
    const uint64_t i0 = off | insert0(p, q), i1 = i0 | (1ull << q);
    const T m0 = mats[4 * b], m1 = mats[4 * b + 1], m2 = mats[4 * b + 2], m3 = mats[4 * b + 3];
    const T x = psi[i0], y = psi[i1];
    T u, v;
    u.x = m0.x * x.x - m0.y * x.y + m1.x * y.x - m1.y * y.y;
    u.y = m0.x * x.y + m0.y * x.x + m1.x * y.y + m1.y * y.x;
    v.x = m2.x * x.x - m2.y * x.y + m3.x * y.x - m3.y * y.y;
    v.y = m2.x * x.y + m2.y * x.x + m3.x * y.y + m3.y * y.x;
    psi[i0] = u;
    psi[i1] = v;
  }
}

// One reverse step of a rotation: accumulate Im <lam|sigma|psi> per batch
// element into part[b * nblk + blockIdx.x], then apply U^dagger to psi and
// lam.  grid = (nblk, Bp); block x covers pairs [x*per, (x+1)*per) of state b.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_rot_reverse(typename AV2<R>::T *__restrict__ psi,
                                                        typename AV2<R>::T *__restrict__ lam,
                                                        int n, int q, int axis,
                                                        const typename AV2<R>::T *__restrict__ mats,
                                                        uint64_t pairs_per_state, uint64_t per,
                                                        double *__restrict__ part) {
  using T = typename AV2<R>::T;
  __shared__ double sh[kBlock];
  const uint64_t b = blockIdx.y;
  const uint64_t off = b << n;
  const T m0 = mats[4 * b], m1 = mats[4 * b + 1], m2 = mats[4 * b + 2], m3 = mats[4 * b + 3];
  // U^dagger = [[conj m0, conj m2], [conj m1, conj m3]]
  double acc = 0.0;
  const uint64_t p0 = (uint64_t)blockIdx.x * per;
  uint64_t p1 = p0 + per;
  if (p1 > pairs_per_state) p1 = pairs_per_state;
  for (uint64_t p = p0 + threadIdx.x; p < p1; p += kBlock) {
    const uint64_t i0 = off | insert0(p, q), i1 = i0 | (1ull << q);
    const T x0 = psi[i0], x1 = psi[i1];
    const T l0 = lam[i0], l1 = lam[i1];
    // s = sigma . psi
    double s0r, s0i, s1r, s1i;
    if (axis == 0) {
      s0r = x1.x; s0i = x1.y; s1r = x0.x; s1i = x0.y;
    } else if (axis == 1) {
      s0r = x1.y; s0i = -(double)x1.x; s1r = -(double)x0.y; s1i = x0.x;
    } else {
      s0r = x0.x; s0i = x0.y; s1r = -(double)x1.x; s1i = -(double)x1.y;
    }
    // Im(conj(l0) s0 + conj(l1) s1)
    acc += (double)l0.x * s0i - (double)l0.y * s0r + (double)l1.x * s1i - (double)l1.y * s1r;
    T y0, y1, k0, k1;
    y0.x = m0.x * x0.x + m0.y * x0.y + m2.x * x1.x + m2.y * x1.y;
    y0.y = m0.x * x0.y - m0.y * x0.x + m2.x * x1.y - m2.y * x1.x;
    y1.x = m1.x * x0.x + m1.y * x0.y + m3.x * x1.x + m3.y * x1.y;
    y1.y = m1.x * x0.y - m1.y * x0.x + m3.x * x1.y - m3.y * x1.x;
    k0.x = m0.x * l0.x + m0.y * l0.y + m2.x * l1.x + m2.y * l1.y;
    k0.y = m0.x * l0.y - m0.y * l0.x + m2.x * l1.y - m2.y * l1.x;
    k1.x = m1.x * l0.x + m1.y * l0.y + m3.x * l1.x + m3.y * l1.y;
    k1.y = m1.x * l0.y - m1.y * l0.x + m3.x * l1.y - m3.y * l1.x;
    psi[i0] = y0;
    psi[i1] = y1;
    lam[i0] = k0;
    lam[i1] = k1;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[b * gridDim.x + blockIdx.x] = sh[0];
}

// grad[b * P + param] += sum_k part[b * nblk + k]   (grid = Bp blocks)
__global__ __launch_bounds__(kBlock) void k_grad_accum(const double *__restrict__ part, int nblk,
                                                       double *__restrict__ grad, size_t P,
                                                       size_t param) {
  __shared__ double sh[kBlock];
  const int b = blockIdx.x;
  double v = 0.0;
  for (int k = threadIdx.x; k < nblk; k += kBlock) v += part[(size_t)b * nblk + k];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) grad[(size_t)b * P + param] += sh[0];
}

// part[b * nblk + x] = sum over the block's chunk of |psi_b|^2 h
template <typename R>
__global__ __launch_bounds__(kBlock) void k_expect(const typename AV2<R>::T *__restrict__ psi,
                                                   int n, const double *__restrict__ h,
                                                   uint64_t per, double *__restrict__ part) {
  __shared__ double sh[kBlock];
  const uint64_t b = blockIdx.y;
  const uint64_t len = 1ull << n;
  const uint64_t x0 = (uint64_t)blockIdx.x * per;
  uint64_t x1 = x0 + per;
  if (x1 > len) x1 = len;
  double acc = 0.0;
  for (uint64_t x = x0 + threadIdx.x; x < x1; x += kBlock) {
    const auto a = psi[(b << n) | x];
    acc += ((double)a.x * (double)a.x + (double)a.y * (double)a.y) * h[x];
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[b * gridDim.x + blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(kBlock) void k_sum_rows(const double *__restrict__ part, int nblk,
                                                     double *__restrict__ out) {
  __shared__ double sh[kBlock];
  const int b = blockIdx.x;
  double v = 0.0;
  for (int k = threadIdx.x; k < nblk; k += kBlock) v += part[(size_t)b * nblk + k];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[b] = sh[0];
}

// lam = h . psi
template <typename R>
__global__ __launch_bounds__(kBlock) void k_apply_h(typename AV2<R>::T *__restrict__ lam,
                                                    const typename AV2<R>::T *__restrict__ psi,
                                                    int n, const double *__restrict__ h,
                                                    uint64_t total) {
  const uint64_t mask = (1ull << n) - 1ull;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * kBlock) {
    auto a = psi[i];
    const double hv = h[i & mask];
    a.x = (R)((double)a.x * hv);
    a.y = (R)((double)a.y * hv);
    lam[i] = a;
  }
}

__global__ __launch_bounds__(kBlock) void k_fill_zero(double *p, uint64_t count) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * kBlock)
    p[i] = 0.0;
}

// ---- fused reverse pass.  The reverse sweep runs in blocks of consecutive
// operations whose pair targets fit one tile of T qubits (the low qubits
// always among them, for contiguous loads).  A workgroup loads its tile of
// psi and of lambda into LDS, runs the block's reverse steps there (each
// rotation: the tile's share of Im <lambda|sigma|psi>, then U^dagger on
// both; each fixed gate: its inverse on both) and writes the tile back:
// one pass over psi and lambda per block instead of one per operation.
// Diagonal and phase gates act on any qubits (their bits off the tile are
// the same for the whole tile).
enum : uint32_t { REV_ROT = 1, REV_PHASE = 2, REV_DIAG = 3, REV_U1 = 4 };

struct RevRec {
  uint32_t kind;
  uint32_t k;      // tile bit of the target (REV_ROT, REV_U1)
  uint32_t axis;   // REV_ROT: rotation axis; REV_DIAG: number of act bits
  uint32_t rslot;  // REV_ROT: rotation slot of the matrix table
  uint32_t out;    // REV_ROT: row of the partial sums
  uint32_t moff;   // values at vals[moff..] as {re, im} pairs
  int32_t a0, a1;  // REV_DIAG: act bits: tile bit (>= 0) or -1 - global qubit
  uint64_t umask, uval;  // pattern bits off the tile (global index bits)
  uint64_t lmask, lval;  // pattern bits on the tile (tile bits)
};

constexpr int kRevMaxOps = 64;
constexpr int kRevMaxTile = 16;
constexpr int kRevMaxBits = 11;  // tile of at most 2^11 amplitudes

struct RevHeader {
  uint32_t T, nops, nrot, nvals;
  uint32_t tq[kRevMaxTile];  // tile qubits, ascending
};

// One reverse step of a rotation on the pair (x0, x1) of psi and (l0, l1) of
// lambda: acc += Im(conj(l0) s0 + conj(l1) s1) with s = sigma . psi, then
// U^dagger = [[conj m0, conj m2], [conj m1, conj m3]] on both pairs.
template <typename R>
__device__ __forceinline__ void rot_reverse_pair(typename AV2<R>::T &x0, typename AV2<R>::T &x1,
                                                 typename AV2<R>::T &l0, typename AV2<R>::T &l1,
                                                 uint32_t axis, const typename AV2<R>::T *m,
                                                 double &acc) {
  using T = typename AV2<R>::T;
  double s0r, s0i, s1r, s1i;
  if (axis == 0) {
    s0r = x1.x; s0i = x1.y; s1r = x0.x; s1i = x0.y;
  } else if (axis == 1) {
    s0r = x1.y; s0i = -(double)x1.x; s1r = -(double)x0.y; s1i = x0.x;
  } else {
    s0r = x0.x; s0i = x0.y; s1r = -(double)x1.x; s1i = -(double)x1.y;
  }
  acc += (double)l0.x * s0i - (double)l0.y * s0r + (double)l1.x * s1i - (double)l1.y * s1r;
  const T m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3];
  T y0, y1, k0, k1;
  y0.x = m0.x * x0.x + m0.y * x0.y + m2.x * x1.x + m2.y * x1.y;
  y0.y = m0.x * x0.y - m0.y * x0.x + m2.x * x1.y - m2.y * x1.x;
  y1.x = m1.x * x0.x + m1.y * x0.y + m3.x * x1.x + m3.y * x1.y;
  y1.y = m1.x * x0.y - m1.y * x0.x + m3.x * x1.y - m3.y * x1.x;
  k0.x = m0.x * l0.x + m0.y * l0.y + m2.x * l1.x + m2.y * l1.y;
  k0.y = m0.x * l0.y - m0.y * l0.x + m2.x * l1.y - m2.y * l1.x;
  k1.x = m1.x * l0.x + m1.y * l0.y + m3.x * l1.x + m3.y * l1.y;
  k1.y = m1.x * l0.y - m1.y * l0.x + m3.x * l1.y - m3.y * l1.x;
  x0 = y0;
  x1 = y1;
  l0 = k0;
  l1 = k1;
}

// Sum over the 64 lanes of a wave, every lane receiving the total: an xor
// butterfly on VALU cross-lane moves (DPP within rows of 16, gfx950's
// permlane swaps across them) instead of LDS-crossbar shuffles.
__device__ __forceinline__ int dpp_i(int v, int ctrl) {
  switch (ctrl) {
    case 0: return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    case 1: return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    case 2: return __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    default: return __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, false); // row_mirror
  }
}

__device__ __forceinline__ double wave_sum(double v) {
  // steps 1, 2: lane ^ 1, lane ^ 2; then every quad is uniform, so the half
  // mirror adds the other quad of each 8 and the row mirror the other 8
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i((int)(b & 0xffffffffll), c), hi = dpp_i((int)(b >> 32), c);
    v += __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
  // rows: every row of 16 is uniform; the swaps hand each lane the value of
  // the row 16 (then 32) lanes away
  const long long b16 = __double_as_longlong(v);
  const auto l16 = __builtin_amdgcn_permlane16_swap((int)(b16 & 0xffffffffll),
                                                    (int)(b16 & 0xffffffffll), false, false);
  const auto h16 = __builtin_amdgcn_permlane16_swap((int)(b16 >> 32), (int)(b16 >> 32), false,
                                                    false);
  const bool r16 = (threadIdx.x >> 4) & 1;
  const int lo16 = r16 ? l16[0] : l16[1], hi16 = r16 ? h16[0] : h16[1];
  v += __longlong_as_double(((long long)hi16 << 32) | (unsigned int)lo16);
  const long long b32 = __double_as_longlong(v);
  const auto l32 = __builtin_amdgcn_permlane32_swap((int)(b32 & 0xffffffffll),
                                                    (int)(b32 & 0xffffffffll), false, false);
  const auto h32 = __builtin_amdgcn_permlane32_swap((int)(b32 >> 32), (int)(b32 >> 32), false,
                                                    false);
  const bool r32 = (threadIdx.x >> 5) & 1;
  const int lo32 = r32 ? l32[0] : l32[1], hi32 = r32 ? h32[0] : h32[1];
  return v + __longlong_as_double(((long long)hi32 << 32) | (unsigned int)lo32);
}

// One workgroup per tile (grid = ntiles); dynamic LDS 2 x 2^T amplitudes;
// PER = max(1, 2^T / 256).  part[out * ntiles + tile] = the tile's share of
// rotation `out`'s gradient.
template <typename R, int PER>
__global__ __launch_bounds__(kBlock) void k_reverse_tile(typename AV2<R>::T *__restrict__ psi,
                                                         typename AV2<R>::T *__restrict__ lam,
                                                         int n, const char *__restrict__ blob,
                                                         const typename AV2<R>::T *__restrict__ rot,
                                                         uint64_t Bp, double *__restrict__ part,
                                                         uint64_t ntiles) {
  using T = typename AV2<R>::T;
  extern __shared__ unsigned char rev_smem[];
  __shared__ double red[4][kBlock / 64];
  const RevHeader &h = *(const RevHeader *)blob;
  const RevRec *recs = (const RevRec *)(blob + sizeof(RevHeader));
  const R *vals = (const R *)(blob + sizeof(RevHeader) + h.nops * sizeof(RevRec));
  const int Tt = (int)h.T;
  const uint32_t E = 1u << Tt;
  T *ps = (T *)rev_smem;
  T *ls = ps + E;
  const uint64_t tile = blockIdx.x;
  uint64_t base = tile;
  for (int i = 0; i < Tt; ++i) base = insert0(base, (int)h.tq[i]);
  const uint64_t b = base >> n;
  // Amplitude j = threadIdx.x + 256 i of the tile: the thread index gives
  // its low 8 tile bits, i the rest (the same for the whole workgroup).  All
  // loads of the tile are issued before the first is waited for.
  const int tb = Tt < 8 ? Tt : 8;
  uint64_t toff = 0;
  for (int i = 0; i < tb; ++i)
    if ((threadIdx.x >> i) & 1u) toff |= 1ull << h.tq[i];
  // PER amplitudes per thread (2^T / 256, at least 1: a tile of fewer than
  // 256 amplitudes leaves the other threads idle)
  const bool active = threadIdx.x < E;
  uint64_t ioff[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    uint64_t o = 0;
    for (int k = 8; k < Tt; ++k)
      if ((i >> (k - 8)) & 1) o |= 1ull << h.tq[k];
    ioff[i] = base | toff | o;
  }
  typedef R V __attribute__((ext_vector_type(2)));
  const V *psv = reinterpret_cast<const V *>(psi), *lav = reinterpret_cast<const V *>(lam);
  V *psl = reinterpret_cast<V *>(ps), *lsl = reinterpret_cast<V *>(ls);
  if (active) {
    V xv[PER], lv[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      xv[i] = psv[ioff[i]];
      lv[i] = lav[ioff[i]];
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      psl[threadIdx.x + i * kBlock] = xv[i];
      lsl[threadIdx.x + i * kBlock] = lv[i];
    }
  }
  __syncthreads();
  for (uint32_t o = 0; o < h.nops; ++o) {
    const RevRec &r = recs[o];
    if (r.kind == REV_ROT) {
      // two consecutive rotations on different tile bits run in one pass
      // over quads of amplitudes; rows use four buffers by row number, so a
      // buffer is rewritten only after a barrier that thread 0 passes after
      // reading it
      const bool two = o + 1 < h.nops && recs[o + 1].kind == REV_ROT && recs[o + 1].k != r.k;
      const RevRec &r2 = two ? recs[o + 1] : r;
      const T *mm = rot + ((uint64_t)r.rslot * Bp + b) * 4;
      const T *mm2 = rot + ((uint64_t)r2.rslot * Bp + b) * 4;
      const T m[4] = {mm[0], mm[1], mm[2], mm[3]};
      const T n2[4] = {mm2[0], mm2[1], mm2[2], mm2[3]};
      double acc = 0.0, acc2 = 0.0;
      if (two) {
        const uint32_t b1 = 1u << r.k, b2 = 1u << r2.k;
        const int klo = r.k < r2.k ? (int)r.k : (int)r2.k;
        const int khi = r.k < r2.k ? (int)r2.k : (int)r.k;
#pragma unroll
        for (int qi = 0; qi < (PER + 3) / 4; ++qi) {
          const uint32_t q = threadIdx.x + (uint32_t)qi * kBlock;
          if (q >= E / 4) break;
          const uint32_t j00 = (uint32_t)insert0(insert0(q, klo), khi);
          const uint32_t j01 = j00 | b1, j10 = j00 | b2, j11 = j00 | b1 | b2;
          T x00 = ps[j00], x01 = ps[j01], x10 = ps[j10], x11 = ps[j11];
          T l00 = ls[j00], l01 = ls[j01], l10 = ls[j10], l11 = ls[j11];
          rot_reverse_pair<R>(x00, x01, l00, l01, r.axis, m, acc);
          rot_reverse_pair<R>(x10, x11, l10, l11, r.axis, m, acc);
          rot_reverse_pair<R>(x00, x10, l00, l10, r2.axis, n2, acc2);
          rot_reverse_pair<R>(x01, x11, l01, l11, r2.axis, n2, acc2);
          ps[j00] = x00;
          ps[j01] = x01;
          ps[j10] = x10;
          ps[j11] = x11;
          ls[j00] = l00;
          ls[j01] = l01;
          ls[j10] = l10;
          ls[j11] = l11;
        }
      } else {
        const uint32_t bk = 1u << r.k;
#pragma unroll
        for (int pi = 0; pi < (PER + 1) / 2; ++pi) {
          const uint32_t p = threadIdx.x + (uint32_t)pi * kBlock;
          if (p >= E / 2) break;
          const uint32_t j0 = (uint32_t)insert0(p, (int)r.k), j1 = j0 | bk;
          T x0 = ps[j0], x1 = ps[j1], l0 = ls[j0], l1 = ls[j1];
          rot_reverse_pair<R>(x0, x1, l0, l1, r.axis, m, acc);
          ps[j0] = x0;
          ps[j1] = x1;
          ls[j0] = l0;
          ls[j1] = l1;
        }
      }
      acc = wave_sum(acc);
      double *rb = red[r.out & 3u];
      double *rb2 = red[r2.out & 3u];
      if ((threadIdx.x & 63u) == 0) rb[threadIdx.x >> 6] = acc;
      if (two) {
        acc2 = wave_sum(acc2);
        if ((threadIdx.x & 63u) == 0) rb2[threadIdx.x >> 6] = acc2;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        double t = 0.0, t2 = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) {
          t += rb[w];
          t2 += rb2[w];
        }
        part[(uint64_t)r.out * ntiles + tile] = t;
        if (two) part[(uint64_t)r2.out * ntiles + tile] = t2;
      }
      if (two) ++o;
    } else if (r.kind == REV_PHASE || r.kind == REV_DIAG) {
      // a run of consecutive phase and diagonal records: one LDS pass, each
      // amplitude multiplied by every factor whose pattern it meets
      uint32_t e = o + 1;
      while (e < h.nops && (recs[e].kind == REV_PHASE || recs[e].kind == REV_DIAG)) ++e;
#pragma unroll
      for (int ji = 0; ji < PER; ++ji) {
        const uint32_t j = threadIdx.x + (uint32_t)ji * kBlock;
        if (j >= E) break;
        T x = ps[j], l = ls[j];
        for (uint32_t t = o; t < e; ++t) {
          const RevRec &d = recs[t];
          if ((base & d.umask) != d.uval || (j & (uint32_t)d.lmask) != (uint32_t)d.lval) continue;
          uint32_t vo = d.moff;
          if (d.kind == REV_DIAG) {
            const uint32_t i0 =
                d.a0 >= 0 ? (j >> d.a0) & 1u : (uint32_t)((base >> (-1 - d.a0)) & 1ull);
            const uint32_t i1 =
                d.axis < 2 ? 0u
                           : (d.a1 >= 0 ? (j >> d.a1) & 1u : (uint32_t)((base >> (-1 - d.a1)) & 1ull));
            vo += 2u * (i0 + 2u * i1);
          }
          const R vr = vals[vo], vi = vals[vo + 1];
          T y, k;
          y.x = vr * x.x - vi * x.y;
          y.y = vr * x.y + vi * x.x;
          k.x = vr * l.x - vi * l.y;
          k.y = vr * l.y + vi * l.x;
          x = y;
          l = k;
        }
        ps[j] = x;
        ls[j] = l;
      }
      __syncthreads();
      o = e - 1;
    } else if (r.kind == REV_U1) {
      if ((base & r.umask) != r.uval) continue;  // the same for the whole workgroup
      const R *v = vals + r.moff;  // row-major 2x2 of the inverse gate
      const uint32_t bk = 1u << r.k;
#pragma unroll
      for (int pi = 0; pi < (PER + 1) / 2; ++pi) {
        const uint32_t p = threadIdx.x + (uint32_t)pi * kBlock;
        if (p >= E / 2) break;
        const uint32_t j0 = (uint32_t)insert0(p, (int)r.k), j1 = j0 | bk;
        if ((j0 & (uint32_t)r.lmask) != (uint32_t)r.lval) continue;
        T z[2][2] = {{ps[j0], ps[j1]}, {ls[j0], ls[j1]}};
        for (int s = 0; s < 2; ++s) {
          const T x0 = z[s][0], x1 = z[s][1];
          T y0, y1;
          y0.x = v[0] * x0.x - v[1] * x0.y + v[2] * x1.x - v[3] * x1.y;
          y0.y = v[0] * x0.y + v[1] * x0.x + v[2] * x1.y + v[3] * x1.x;
          y1.x = v[4] * x0.x - v[5] * x0.y + v[6] * x1.x - v[7] * x1.y;
          y1.y = v[4] * x0.y + v[5] * x0.x + v[6] * x1.y + v[7] * x1.x;
          z[s][0] = y0;
          z[s][1] = y1;
        }
        ps[j0] = z[0][0];
        ps[j1] = z[0][1];
        ls[j0] = z[1][0];
        ls[j1] = z[1][1];
      }
      __syncthreads();
    }
  }
  if (active) {
    V *pso = reinterpret_cast<V *>(psi), *lao = reinterpret_cast<V *>(lam);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      pso[ioff[i]] = psl[threadIdx.x + i * kBlock];
      lao[ioff[i]] = lsl[threadIdx.x + i * kBlock];
    }
  }
}

// grad[b * P + params[row]] += sum of the tiles of state b in row `row` of
// part (rows in order, so a parameter shared by two rotations of the block
// is summed without a race).  grid = Bp.
__global__ __launch_bounds__(kBlock) void k_rev_accum(const double *__restrict__ part,
                                                      uint32_t nrot, uint64_t ntiles,
                                                      uint64_t tiles_per_state,
                                                      const uint32_t *__restrict__ params,
                                                      double *__restrict__ grad, size_t P) {
  __shared__ double sh[kBlock];
  const uint64_t b = blockIdx.x;
  for (uint32_t row = 0; row < nrot; ++row) {
    const double *pr = part + (uint64_t)row * ntiles + b * tiles_per_state;
    double v = 0.0;
    for (uint64_t i = threadIdx.x; i < tiles_per_state; i += kBlock) v += pr[i];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) grad[b * P + params[row]] += sh[0];
    __syncthreads();
  }
}

unsigned grid_of(uint64_t work, uint64_t per, unsigned cap = 262144) {
  uint64_t g = (work + per - 1) / per;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

}  // namespace

struct qs_pcircuit {
  int n = 0;
  std::vector<POp> ops;
  std::vector<size_t> layer_ends;
  size_t nparams = 0;
  int nrot = 0;
};

struct qs_adjoint {
  qs_pcircuit circ;
  int n = 0, nb = 0;
  size_t B = 0, Bp = 1, P = 0;
  int dtype = QS_C128;
  size_t checkpoint_every = 0;
  hipStream_t stream = nullptr;
  void *psi = nullptr, *lam = nullptr;
  std::vector<void *> ckpt;
  std::vector<size_t> ckpt_at;  // checkpoint j holds psi after ops [0, ckpt_at[j])
  double *d_h = nullptr;
  void *d_rot = nullptr;        // [nrot][Bp][4] complex, precision of dtype
  void *d_gmat = nullptr;       // diagonal tables of fixed gates (forward, inverse)
  std::vector<size_t> gmat_off; // per op: offset of fwd table, inv table (bytes)
  double *d_part = nullptr;
  size_t part_count = 0;
  double *d_grad = nullptr;
  double *d_rpart = nullptr;  // partial sums of the tiled reverse blocks
  uint64_t rev_blocks = 0, rev_restores = 0;  // of the last backward
  // descriptor ring of the fused forward pass (device + pinned host)
  char *d_ring = nullptr, *h_ring = nullptr;
  size_t ring_cap = 0, ring_off = 0;
  qs::SweepEngine *engine = nullptr;
  uint64_t sweeps = 0, kernels = 0;
  size_t amp_bytes() const { return dtype == QS_C128 ? 16 : 8; }
  size_t state_bytes() const { return amp_bytes() * (Bp << n); }
};

namespace {

void free_ctx(qs_adjoint *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->psi) (void)hipFree(c->psi);
  if (c->lam) (void)hipFree(c->lam);
  for (void *p : c->ckpt) (void)hipFree(p);
  if (c->d_h) (void)hipFree(c->d_h);
  if (c->d_rot) (void)hipFree(c->d_rot);
  if (c->d_gmat) (void)hipFree(c->d_gmat);
  if (c->d_part) (void)hipFree(c->d_part);
  if (c->d_grad) (void)hipFree(c->d_grad);
  if (c->d_rpart) (void)hipFree(c->d_rpart);
  if (c->d_ring) (void)hipFree(c->d_ring);
  if (c->h_ring) (void)hipHostFree(c->h_ring);
  if (c->engine) qs::sweep_engine_free(c->engine);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// Copy host bytes into the descriptor ring; a slot is reused only after the
// stream has drained.
hipError_t ring_upload(qs_adjoint *c, const void *src, size_t bytes, void **dev) {
  const size_t need = (bytes + 255) & ~(size_t)255;
  if (need > c->ring_cap) return hipErrorInvalidValue;
  if (c->ring_off + need > c->ring_cap) {
    const hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return e;
    c->ring_off = 0;
  }
  char *h = c->h_ring + c->ring_off;
  char *d = c->d_ring + c->ring_off;
  std::memcpy(h, src, bytes);
  c->ring_off += need;
  *dev = d;
  return hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream);
}

// Whether the forward pass runs through the fused sweep engine
// (QS_ADJOINT_SWEEP, default on) or one kernel per operation.
bool adjoint_sweeps() {
  static const bool on = [] {
    const char *v = std::getenv("QS_ADJOINT_SWEEP");
    return !(v && v[0] == '0');
  }();
  return on;
}

// Apply a classified fixed gate over all Bp states (n + nb index bits).
template <typename R>
hipError_t apply_fixed(qs_adjoint *c, void *buf, const qs::Gate &g, const void *d_table) {
  const int ntot = c->n + c->nb;
  switch (g.kind) {
    case qs::Gate::kIdentity: return hipSuccess;
    case qs::Gate::kSparse: return qs::launch_gate_sparse<R>(buf, ntot, g, c->stream);
    case qs::Gate::kDiag:
      return qs::launch_gate_diag<R>(buf, ntot, g, const_cast<void *>(d_table), c->stream);
    case qs::Gate::kDense:
      return qs::launch_gate_dense<R>(buf, ntot, g, nullptr, nullptr, c->stream);
    case qs::Gate::kPerm:  // rejected when the gate is added
      return hipErrorInvalidValue;
  }
  return hipSuccess;
}

#define AJ(expr)                                          \
  do {                                                    \
    e = (expr);                                           \
    if (e != hipSuccess) return qs::set_hip_error(e, #expr); \
  } while (0)

// The forward pass through the fused sweep engine over n + nb qubits: the
// batch index is nb extra high qubits, each rotation a batched 2x2 whose
// matrix the kernel picks by the tile's batch bits (they are never tile
// qubits), each fixed gate an ordinary gate.  The operations between two
// checkpoints go to the planner as one list; the state is copied at each
// checkpoint.  Gates the planner leaves to per-gate kernels run here.
template <typename R>
int forward_sweeps(qs_adjoint *c, const double *thetas) {
  using T = typename AV2<R>::T;
  const int n = c->n, ntot = c->n + c->nb;
  const size_t Bp = c->Bp, B = c->B, P = c->P;
  hipError_t e = hipSuccess;
  if (!c->d_ring) {
    c->ring_cap = 8u << 20;
    AJ(hipMalloc((void **)&c->d_ring, c->ring_cap));
    AJ(hipHostMalloc((void **)&c->h_ring, c->ring_cap));
  }
  if (!c->engine) c->engine = qs::sweep_engine_new();
  qs::SweepRunner run;
  run.n = ntot;
  run.dtype_c64 = c->dtype == QS_C64;
  run.psi = c->psi;
  run.stream = c->stream;
  run.upload = [c](const void *src, size_t bytes, void **dev) {
    return ring_upload(c, src, bytes, dev);
  };
  run.single = [c, n, Bp](const qs::Gate &g) -> hipError_t {
    ++c->kernels;
    if (g.bmask) {
      std::vector<T> mats(4 * Bp);
      for (size_t k = 0; k < 4 * Bp; ++k) {
        mats[k].x = (R)g.m[k].re;
        mats[k].y = (R)g.m[k].im;
      }
      void *d = nullptr;
      hipError_t e2 = ring_upload(c, mats.data(), mats.size() * sizeof(T), &d);
      if (e2 != hipSuccess) return e2;
      const uint64_t tp = (1ull << (n - 1)) * Bp;
      hipLaunchKernelGGL(k_rot<R>, dim3(grid_of(tp, kBlock * kItems, 0xffffffffu)),
                         dim3(kBlock), 0, c->stream, (T *)c->psi, n, g.act[0], (const T *)d, tp);
      return hipGetLastError();
    }
    void *tab = nullptr;
    if (g.kind == qs::Gate::kDiag) {
      std::vector<R> v;
      for (const qs::cd &x : g.m) {
        v.push_back((R)x.re);
        v.push_back((R)x.im);
      }
      hipError_t e2 = ring_upload(c, v.data(), v.size() * sizeof(R), &tab);
      if (e2 != hipSuccess) return e2;
    }
    return apply_fixed<R>(c, c->psi, g, tab);
  };
  std::vector<qs::Gate> gates;
  size_t layer = 0;
  auto flush = [&]() -> int {
    if (gates.empty()) return QS_OK;
    hipError_t e2 = qs::sweep_execute(c->engine, run, gates, &c->sweeps, &c->kernels);
    gates.clear();
    if (e2 != hipSuccess) return qs::set_hip_error(e2, "adjoint forward sweeps");
    return QS_OK;
  };
  for (size_t i = 0; i < c->circ.ops.size(); ++i) {
    const POp &op = c->circ.ops[i];
    if (op.kind == 1) {
      qs::Gate g;
      g.kind = qs::Gate::kDense;
      g.na = 1;
      g.act[0] = op.target;
      g.m.resize(4 * Bp);
      for (size_t b = 0; b < Bp; ++b) {
        const double th = b < B ? thetas[b * P + op.param] : 0.0;
        const double cs = std::cos(th / 2.0), sn = std::sin(th / 2.0);
        qs::cd *m = g.m.data() + 4 * b;
        if (op.axis == QS_AXIS_X) {
          m[0] = {cs, 0};
          m[1] = {0, -sn};
          m[2] = {0, -sn};
          m[3] = {cs, 0};
        } else if (op.axis == QS_AXIS_Y) {
          m[0] = {cs, 0};
          m[1] = {-sn, 0};
          m[2] = {sn, 0};
          m[3] = {cs, 0};
        } else {
          m[0] = {cs, -sn};
          m[1] = {0, 0};
          m[2] = {0, 0};
          m[3] = {cs, sn};
        }
      }
      if (Bp > 1) {
        g.bshift = n;
        g.bmask = (uint32_t)(Bp - 1);
      }
      gates.push_back(std::move(g));
    } else if (op.fwd.kind != qs::Gate::kIdentity) {
      gates.push_back(op.fwd);
    }
    while (layer < c->circ.layer_ends.size() && c->circ.layer_ends[layer] == i + 1) {
      ++layer;
      if (c->checkpoint_every > 0 && layer % c->checkpoint_every == 0 &&
          layer < c->circ.layer_ends.size()) {
        const int rc = flush();
        if (rc != QS_OK) return rc;
        void *cp = nullptr;
        AJ(hipMalloc(&cp, c->state_bytes()));
        AJ(hipMemcpyAsync(cp, c->psi, c->state_bytes(), hipMemcpyDeviceToDevice, c->stream));
        c->ckpt.push_back(cp);
        c->ckpt_at.push_back(i + 1);
      }
    }
  }
  return flush();
}

template <typename R>
int forward_t(qs_adjoint *c, const double *thetas, const uint64_t *masks, const double *coeffs,
              size_t nterms, double *energies) {
  using T = typename AV2<R>::T;
  const int n = c->n;
  const size_t Bp = c->Bp, B = c->B, P = c->P;
  const uint64_t total = (uint64_t)Bp << n;
  hipError_t e = hipSuccess;
  AJ(hipMalloc(&c->psi, c->state_bytes()));
  // observable table
  {
    unsigned long long *d_m = nullptr;
    double *d_c = nullptr;
    AJ(hipMalloc((void **)&c->d_h, sizeof(double) << n));
    AJ(hipMalloc((void **)&d_m, sizeof(unsigned long long) * (nterms ? nterms : 1)));
    AJ(hipMalloc((void **)&d_c, sizeof(double) * (nterms ? nterms : 1)));
    if (nterms) {
      AJ(hipMemcpy(d_m, masks, sizeof(unsigned long long) * nterms, hipMemcpyHostToDevice));
      AJ(hipMemcpy(d_c, coeffs, sizeof(double) * nterms, hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(k_htable, dim3(grid_of(1ull << n, kBlock * 4)), dim3(kBlock), 0, c->stream,
                       c->d_h, n, d_m, d_c, (int)nterms);
    AJ(hipGetLastError());
    AJ(hipStreamSynchronize(c->stream));
    (void)hipFree(d_m);
    (void)hipFree(d_c);
  }
  // rotation matrices, one per (rotation, batch element)
  const int nrot = c->circ.nrot;
  std::vector<R> rot((size_t)std::max(nrot, 1) * Bp * 8, (R)0);
  for (const POp &op : c->circ.ops) {
    if (op.kind != 1) continue;
    for (size_t b = 0; b < Bp; ++b) {
      const double th = b < B ? thetas[b * P + op.param] : 0.0;
      const double cs = std::cos(th / 2.0), sn = std::sin(th / 2.0);
      double m[8];
      if (op.axis == QS_AXIS_X) {
        const double v[8] = {cs, 0, 0, -sn, 0, -sn, cs, 0};
        std::memcpy(m, v, sizeof(m));
      } else if (op.axis == QS_AXIS_Y) {
        const double v[8] = {cs, 0, -sn, 0, sn, 0, cs, 0};
        std::memcpy(m, v, sizeof(m));
      } else {
        const double v[8] = {cs, -sn, 0, 0, 0, 0, cs, sn};
        std::memcpy(m, v, sizeof(m));
      }
      R *dst = rot.data() + ((size_t)op.rot_slot * Bp + b) * 8;
      for (int k = 0; k < 8; ++k) dst[k] = (R)m[k];
    }
  }
  AJ(hipMalloc(&c->d_rot, rot.size() * sizeof(R)));
  AJ(hipMemcpy(c->d_rot, rot.data(), rot.size() * sizeof(R), hipMemcpyHostToDevice));
  // diagonal tables of the fixed gates
  {
    std::vector<R> tab;
    c->gmat_off.assign(c->circ.ops.size() * 2, 0);
    for (size_t i = 0; i < c->circ.ops.size(); ++i) {
      const POp &op = c->circ.ops[i];
      if (op.kind != 0) continue;
      const qs::Gate *gs[2] = {&op.fwd, &op.inv};
      for (int s = 0; s < 2; ++s) {
        c->gmat_off[2 * i + s] = tab.size() * sizeof(R);
        if (gs[s]->kind == qs::Gate::kDiag)
          for (const qs::cd &v : gs[s]->m) {
            tab.push_back((R)v.re);
            tab.push_back((R)v.im);
          }
      }
    }
    if (tab.empty()) tab.push_back((R)0);
    AJ(hipMalloc(&c->d_gmat, tab.size() * sizeof(R)));
    AJ(hipMemcpy(c->d_gmat, tab.data(), tab.size() * sizeof(R), hipMemcpyHostToDevice));
  }
  hipLaunchKernelGGL(k_init_zero_states<R>, dim3(grid_of(total, kBlock * 4)), dim3(kBlock), 0,
                     c->stream, (T *)c->psi, n, total);
  AJ(hipGetLastError());
  const uint64_t pps = 1ull << (n - 1);
  if (adjoint_sweeps()) {
    const int rc = forward_sweeps<R>(c, thetas);
    if (rc != QS_OK) return rc;
  }
  size_t layer = 0;
  for (size_t i = 0; i < c->circ.ops.size() && !adjoint_sweeps(); ++i) {
    const POp &op = c->circ.ops[i];
    if (op.kind == 1) {
      const T *mats = (const T *)c->d_rot + (size_t)op.rot_slot * Bp * 4;
      const uint64_t tp = pps * Bp;
      hipLaunchKernelGGL(k_rot<R>, dim3(grid_of(tp, kBlock * kItems, 0xffffffffu)), dim3(kBlock),
                         0, c->stream, (T *)c->psi, n, op.target, mats, tp);
      AJ(hipGetLastError());
    } else {
      AJ(apply_fixed<R>(c, c->psi, op.fwd, (const char *)c->d_gmat + c->gmat_off[2 * i]));
    }
    // checkpoint after every checkpoint_every-th layer (not after the last)
    while (layer < c->circ.layer_ends.size() && c->circ.layer_ends[layer] == i + 1) {
      ++layer;
      if (c->checkpoint_every > 0 && layer % c->checkpoint_every == 0 &&
          layer < c->circ.layer_ends.size()) {
        void *cp = nullptr;
        AJ(hipMalloc(&cp, c->state_bytes()));
        AJ(hipMemcpyAsync(cp, c->psi, c->state_bytes(), hipMemcpyDeviceToDevice, c->stream));
        c->ckpt.push_back(cp);
        c->ckpt_at.push_back(i + 1);
      }
    }
  }
  // energies
  const uint64_t len = 1ull << n;
  const uint64_t per = std::min<uint64_t>(len, 1ull << 14);
  const int nblk = (int)((len + per - 1) / per);
  c->part_count = std::max<size_t>((size_t)nblk, (size_t)((pps + 1023) / 1024)) * Bp;
  AJ(hipMalloc((void **)&c->d_part, c->part_count * sizeof(double)));
  double *d_e = nullptr;
  AJ(hipMalloc((void **)&d_e, Bp * sizeof(double)));
  hipLaunchKernelGGL(k_expect<R>, dim3(nblk, (unsigned)Bp), dim3(kBlock), 0, c->stream,
                     (const T *)c->psi, n, c->d_h, per, c->d_part);
  hipLaunchKernelGGL(k_sum_rows, dim3((unsigned)Bp), dim3(kBlock), 0, c->stream,
                     (const double *)c->d_part, nblk, d_e);
  AJ(hipGetLastError());
  std::vector<double> eh(Bp);
  AJ(hipMemcpyAsync(eh.data(), d_e, Bp * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  AJ(hipStreamSynchronize(c->stream));
  (void)hipFree(d_e);
  if (energies)
    for (size_t b = 0; b < B; ++b) energies[b] = eh[b];
  return QS_OK;
}

// Whether the reverse pass runs in tiled blocks (QS_ADJOINT_FUSED_REVERSE,
// default on) or one kernel per operation.
bool adjoint_fused_reverse() {
  static const bool on = [] {
    const char *v = std::getenv("QS_ADJOINT_FUSED_REVERSE");
    return !(v && v[0] == '0');
  }();
  return on;
}

// Tile qubits of the reverse blocks: QS_ADJOINT_TILE, default 10 for
// complex128 and 11 for complex64 (32 KiB of LDS for psi and lambda).
int reverse_tile_bits(int dtype) {
  const char *v = std::getenv("QS_ADJOINT_TILE");
  const int t = v ? std::atoi(v) : (dtype == QS_C128 ? 10 : 11);
  return std::max(1, std::min(t, kRevMaxBits));
}

// The reverse sweep in tiled blocks (k_reverse_tile): operations are taken
// in reverse order while their pair targets fit the tile; a checkpoint
// restore, a gate with no tile form (two or more dense target qubits) or a
// full record list ends a block.
template <typename R>
int reverse_blocks(qs_adjoint *c) {
  using T = typename AV2<R>::T;
  hipError_t e = hipSuccess;
  const int n = c->n;
  const size_t Bp = c->Bp, P = c->P;
  const int Tt = std::min(reverse_tile_bits(c->dtype), n);
  const int low = std::min(Tt, c->dtype == QS_C128 ? 4 : 5);  // 256-byte runs
  const uint64_t tiles_per_state = 1ull << (n - Tt);
  const uint64_t ntiles = tiles_per_state * Bp;
  // rows of partial sums per block: at most 32, within 64 MiB
  const uint64_t rmax = std::max<uint64_t>(1, std::min<uint64_t>(32, (64ull << 20) / (8 * ntiles)));
  if (!c->d_rpart) AJ(hipMalloc((void **)&c->d_rpart, rmax * ntiles * sizeof(double)));
  if (!c->d_ring) {
    c->ring_cap = 8u << 20;
    AJ(hipMalloc((void **)&c->d_ring, c->ring_cap));
    AJ(hipHostMalloc((void **)&c->h_ring, c->ring_cap));
  }
  const uint64_t lowset = (1ull << low) - 1ull;
  // the block under construction; qubit fields hold global qubits until
  // the tile is fixed
  struct Pend {
    RevRec r;
    uint64_t cmask, cval;  // global pattern
    int q0, q1;            // global target / act qubits
  };
  std::vector<Pend> pend;
  std::vector<R> vals;
  std::vector<uint32_t> params;
  uint64_t S = lowset;  // tile qubits taken
  auto put = [&](const qs::cd &v) {
    vals.push_back((R)v.re);
    vals.push_back((R)v.im);
  };
  auto flush = [&]() -> int {
    if (pend.empty()) return QS_OK;
    RevHeader hd{};
    hd.T = (uint32_t)Tt;
    uint64_t full = S;
    for (int q = 0; q < n && __builtin_popcountll(full) < Tt; ++q) full |= 1ull << q;
    int loc[64];
    int t = 0;
    for (int q = 0; q < 64; ++q) {
      loc[q] = -1;
      if ((full >> q) & 1ull) {
        hd.tq[t] = (uint32_t)q;
        loc[q] = t++;
      }
    }
    std::vector<RevRec> recs;
    for (const Pend &pd : pend) {
      RevRec r = pd.r;
      const uint64_t off = pd.cmask & ~full;
      r.umask = off;
      r.uval = pd.cval & off;
      r.lmask = r.lval = 0;
      for (uint64_t bits = pd.cmask & full; bits; bits &= bits - 1) {
        const int q = __builtin_ctzll(bits);
        r.lmask |= 1ull << loc[q];
        if ((pd.cval >> q) & 1ull) r.lval |= 1ull << loc[q];
      }
      if (r.kind == REV_ROT || r.kind == REV_U1) r.k = (uint32_t)loc[pd.q0];
      if (r.kind == REV_DIAG) {
        r.a0 = loc[pd.q0] >= 0 ? loc[pd.q0] : -1 - pd.q0;
        r.a1 = pd.q1 < 0 ? 0 : (loc[pd.q1] >= 0 ? loc[pd.q1] : -1 - pd.q1);
      }
      recs.push_back(r);
    }
    hd.nops = (uint32_t)recs.size();
    hd.nrot = (uint32_t)params.size();
    hd.nvals = (uint32_t)vals.size();
    std::vector<char> blob(sizeof(RevHeader) + recs.size() * sizeof(RevRec) +
                           vals.size() * sizeof(R) + params.size() * sizeof(uint32_t) + 16);
    char *w = blob.data();
    std::memcpy(w, &hd, sizeof hd);
    w += sizeof hd;
    std::memcpy(w, recs.data(), recs.size() * sizeof(RevRec));
    w += recs.size() * sizeof(RevRec);
    if (!vals.empty()) std::memcpy(w, vals.data(), vals.size() * sizeof(R));
    w += vals.size() * sizeof(R);
    const size_t poff = (size_t)(w - blob.data() + 3) & ~(size_t)3;
    if (!params.empty()) std::memcpy(blob.data() + poff, params.data(), params.size() * 4);
    void *d_blob = nullptr;
    AJ(ring_upload(c, blob.data(), blob.size(), &d_blob));
    const size_t lds = 2 * ((size_t)1 << Tt) * sizeof(T);
    auto kern = Tt <= 8 ? k_reverse_tile<R, 1>
                        : (Tt == 9 ? k_reverse_tile<R, 2>
                                   : (Tt == 10 ? k_reverse_tile<R, 4> : k_reverse_tile<R, 8>));
    hipLaunchKernelGGL(kern, dim3((unsigned)ntiles), dim3(kBlock), lds, c->stream, (T *)c->psi,
                       (T *)c->lam, n, (const char *)d_blob, (const T *)c->d_rot, (uint64_t)Bp,
                       c->d_rpart, ntiles);
    AJ(hipGetLastError());
    ++c->rev_blocks;
    if (!params.empty()) {
      hipLaunchKernelGGL(k_rev_accum, dim3((unsigned)Bp), dim3(kBlock), 0, c->stream,
                         (const double *)c->d_rpart, (uint32_t)params.size(), ntiles,
                         tiles_per_state, (const uint32_t *)((char *)d_blob + poff), c->d_grad,
                         P);
      AJ(hipGetLastError());
    }
    pend.clear();
    vals.clear();
    params.clear();
    S = lowset;
    return QS_OK;
  };
  // pair targets must be tile qubits
  auto fits = [&](int q) { return ((S >> q) & 1ull) || __builtin_popcountll(S) < Tt; };
  int ck = (int)c->ckpt.size() - 1;
  for (size_t ii = c->circ.ops.size(); ii-- > 0;) {
    if (ck >= 0 && c->ckpt_at[ck] == ii + 1) {
      int rc = flush();
      if (rc != QS_OK) return rc;
      AJ(hipMemcpyAsync(c->psi, c->ckpt[ck], c->state_bytes(), hipMemcpyDeviceToDevice,
                        c->stream));
      ++c->rev_restores;
      --ck;
    }
    const POp &op = c->circ.ops[ii];
    const qs::Gate &g = op.inv;
    int need = 1;  // records this operation adds
    bool tile_form = true;
    int pq = -1;  // its pair target
    if (op.kind == 1) {
      pq = op.target;
    } else if (g.kind == qs::Gate::kIdentity) {
      continue;
    } else if (g.kind == qs::Gate::kSparse) {
      need = (int)g.spat.size();
    } else if (g.kind == qs::Gate::kDiag) {
      tile_form = g.na >= 1 && g.na <= 2;
    } else if (g.kind == qs::Gate::kDense && g.na == 1) {
      pq = g.act[0];
    } else {
      tile_form = false;
    }
    if (!tile_form) {
      int rc = flush();
      if (rc != QS_OK) return rc;
      const void *tab = (const char *)c->d_gmat + c->gmat_off[2 * ii + 1];
      AJ(apply_fixed<R>(c, c->psi, g, tab));
      AJ(apply_fixed<R>(c, c->lam, g, tab));
      continue;
    }
    if ((int)pend.size() + need > kRevMaxOps || (pq >= 0 && !fits(pq)) ||
        (op.kind == 1 && params.size() >= rmax)) {
      int rc = flush();
      if (rc != QS_OK) return rc;
    }
    if (pq >= 0) S |= 1ull << pq;
    if (op.kind == 1) {
      Pend pd{};
      pd.r.kind = REV_ROT;
      pd.r.axis = (uint32_t)op.axis;
      pd.r.rslot = (uint32_t)op.rot_slot;
      pd.r.out = (uint32_t)params.size();
      pd.q0 = op.target;
      pd.q1 = -1;
      params.push_back((uint32_t)op.param);
      pend.push_back(pd);
    } else if (g.kind == qs::Gate::kSparse) {
      for (size_t k = 0; k < g.spat.size(); ++k) {
        Pend pd{};
        pd.r.kind = REV_PHASE;
        pd.r.moff = (uint32_t)vals.size();
        put(g.m[k]);
        pd.cmask = g.smask;
        pd.cval = g.spat[k];
        pd.q0 = pd.q1 = -1;
        pend.push_back(pd);
      }
    } else if (g.kind == qs::Gate::kDiag) {
      Pend pd{};
      pd.r.kind = REV_DIAG;
      pd.r.axis = (uint32_t)g.na;
      pd.r.moff = (uint32_t)vals.size();
      for (const qs::cd &v : g.m) put(v);
      pd.cmask = g.cmask;
      pd.cval = g.cval;
      pd.q0 = g.act[0];
      pd.q1 = g.na > 1 ? g.act[1] : -1;
      pend.push_back(pd);
    } else {
      Pend pd{};
      pd.r.kind = REV_U1;
      pd.r.moff = (uint32_t)vals.size();
      for (const qs::cd &v : g.m) put(v);
      pd.cmask = g.cmask;
      pd.cval = g.cval;
      pd.q0 = g.act[0];
      pd.q1 = -1;
      pend.push_back(pd);
    }
  }
  return flush();
}

template <typename R>
int backward_t(qs_adjoint *c, double *grads) {
  using T = typename AV2<R>::T;
  const int n = c->n;
  const size_t Bp = c->Bp, P = c->P;
  const uint64_t total = (uint64_t)Bp << n;
  hipError_t e = hipSuccess;
  if (!c->lam) AJ(hipMalloc(&c->lam, c->state_bytes()));
  if (!c->d_grad) AJ(hipMalloc((void **)&c->d_grad, Bp * std::max<size_t>(P, 1) * sizeof(double)));
  hipLaunchKernelGGL(k_fill_zero, dim3(grid_of(Bp * P, kBlock)), dim3(kBlock), 0, c->stream,
                     c->d_grad, (uint64_t)(Bp * P));
  hipLaunchKernelGGL(k_apply_h<R>, dim3(grid_of(total, kBlock * 4)), dim3(kBlock), 0, c->stream,
                     (T *)c->lam, (const T *)c->psi, n, c->d_h, total);
  AJ(hipGetLastError());
  const uint64_t pps = 1ull << (n - 1);
  const uint64_t per = std::min<uint64_t>(pps, 1024);
  const int nblk = (int)((pps + per - 1) / per);
  int ck = (int)c->ckpt.size() - 1;
  c->rev_blocks = c->rev_restores = 0;
  if (adjoint_fused_reverse()) {
    const int rc = reverse_blocks<R>(c);
    if (rc != QS_OK) return rc;
  }
  for (size_t ii = adjoint_fused_reverse() ? 0 : c->circ.ops.size(); ii-- > 0;) {
    if (ck >= 0 && c->ckpt_at[ck] == ii + 1) {
      AJ(hipMemcpyAsync(c->psi, c->ckpt[ck], c->state_bytes(), hipMemcpyDeviceToDevice,
                        c->stream));
      --ck;
    }
    const POp &op = c->circ.ops[ii];
    if (op.kind == 1) {
      const T *mats = (const T *)c->d_rot + (size_t)op.rot_slot * Bp * 4;
      hipLaunchKernelGGL(k_rot_reverse<R>, dim3(nblk, (unsigned)Bp), dim3(kBlock), 0, c->stream,
                         (T *)c->psi, (T *)c->lam, n, op.target, op.axis, mats, pps, per,
                         c->d_part);
      hipLaunchKernelGGL(k_grad_accum, dim3((unsigned)Bp), dim3(kBlock), 0, c->stream,
                         (const double *)c->d_part, nblk, c->d_grad, P, op.param);
      AJ(hipGetLastError());
    } else {
      const void *tab = (const char *)c->d_gmat + c->gmat_off[2 * ii + 1];
      AJ(apply_fixed<R>(c, c->psi, op.inv, tab));
      AJ(apply_fixed<R>(c, c->lam, op.inv, tab));
    }
  }
  std::vector<double> g(Bp * P);
  AJ(hipMemcpyAsync(g.data(), c->d_grad, g.size() * sizeof(double), hipMemcpyDeviceToHost,
                    c->stream));
  AJ(hipStreamSynchronize(c->stream));
  std::memcpy(grads, g.data(), c->B * P * sizeof(double));
  return QS_OK;
#undef AJ
}

}  // namespace

extern "C" {

int qs_pcircuit_new(size_t n, qs_pcircuit **out) {
  if (!out) return fail_arg("out is NULL");
  if (n < 1 || n > 30) return qs::set_error(QS_ERR_ARGUMENT, n, 0, "pcircuit needs 1..30 qubits");
  qs_pcircuit *c = new qs_pcircuit();
  c->n = (int)n;
  *out = c;
  return QS_OK;
}

void qs_pcircuit_free(qs_pcircuit *c) { delete c; }

int qs_pcircuit_gate(qs_pcircuit *c, const double *matrix, const size_t *targets, size_t k) {
  if (!c || !matrix || !targets) return fail_arg("NULL argument");
  if (k < 1 || k > 2) return qs::set_error(QS_INVALID_DIMENSIONS, k, k, "fixed gates: k = 1 or 2");
  std::vector<size_t> t(targets, targets + k);
  if (k == 2 && t[0] == t[1]) return qs::set_error(QS_INVALID_TARGET, t[0], 0, "duplicate target");
  for (size_t j = 0; j < k; ++j)
    if (t[j] >= (size_t)c->n) return qs::set_error(QS_INVALID_TARGET, t[j], 0, "target out of range");
  const size_t D = (size_t)1 << k;
  std::vector<double> inv(2 * D * D);
  for (size_t i = 0; i < D; ++i)
    for (size_t j = 0; j < D; ++j) {
      inv[2 * (i * D + j)] = matrix[2 * (j * D + i)];
      inv[2 * (i * D + j) + 1] = -matrix[2 * (j * D + i) + 1];
    }
  POp op;
  op.kind = 0;
  op.fwd = qs::classify_matrix(matrix, t.data(), k, 0, 0);
  op.inv = qs::classify_matrix(inv.data(), t.data(), k, 0, 0);
  if ((op.fwd.kind == qs::Gate::kDense && op.fwd.na > 2) ||
      (op.fwd.kind == qs::Gate::kDiag && op.fwd.na > 2) || op.fwd.kind == qs::Gate::kPerm)
    return qs::set_error(QS_INVALID_DIMENSIONS, k, k, "unsupported fixed gate");
  c->ops.push_back(std::move(op));
  return QS_OK;
}

int qs_pcircuit_rotation(qs_pcircuit *c, int axis, size_t target, size_t param) {
  if (!c) return fail_arg("NULL argument");
  if (axis < QS_AXIS_X || axis > QS_AXIS_Z) return fail_arg("axis must be 0 (X), 1 (Y) or 2 (Z)");
  if (target >= (size_t)c->n) return qs::set_error(QS_INVALID_TARGET, target, 0, "target out of range");
  POp op;
  op.kind = 1;
  op.axis = axis;
  op.target = (int)target;
  op.param = param;
  op.rot_slot = c->nrot++;
  c->nparams = std::max(c->nparams, param + 1);
  c->ops.push_back(std::move(op));
  return QS_OK;
}

int qs_pcircuit_layer(qs_pcircuit *c) {
  if (!c) return fail_arg("NULL argument");
  c->layer_ends.push_back(c->ops.size());
  return QS_OK;
}

size_t qs_pcircuit_nparams(const qs_pcircuit *c) { return c ? c->nparams : 0; }

int qs_adjoint_forward(const qs_pcircuit *c, const double *thetas, size_t batch,
                       const uint64_t *masks, const double *coeffs, size_t nterms, int dtype,
                       size_t checkpoint_every, qs_adjoint **ctx, double *energies) {
  if (ctx) *ctx = nullptr;
  if (!c || batch == 0 || (c->nparams > 0 && !thetas) || (nterms > 0 && (!masks || !coeffs)))
    return fail_arg("NULL argument or empty batch");
  if (dtype != QS_C128 && dtype != QS_C64) return fail_arg("bad dtype");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return qs::set_error(QS_ERR_NO_DEVICE, 0, 0, "no HIP device available");
  qs_adjoint *a = new qs_adjoint();
  a->circ = *c;
  a->n = c->n;
  a->B = batch;
  a->Bp = 1;
  a->nb = 0;
  while (a->Bp < batch) {
    a->Bp <<= 1;
    ++a->nb;
  }
  if (a->n + a->nb > qs::kMaxQubits) {
    delete a;
    return fail_arg("batch x state too large");
  }
  a->P = c->nparams;
  a->dtype = dtype;
  a->checkpoint_every = checkpoint_every;
  hipError_t e = hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete a;
    return qs::set_hip_error(e, "hipStreamCreate");
  }
  const int rc = dtype == QS_C128 ? forward_t<double>(a, thetas, masks, coeffs, nterms, energies)
                                  : forward_t<float>(a, thetas, masks, coeffs, nterms, energies);
  if (rc != QS_OK || !ctx) {
    free_ctx(a);
    return rc;
  }
  *ctx = a;
  return QS_OK;
}

int qs_adjoint_backward(qs_adjoint *ctx, double *grads) {
  if (!ctx || !grads) return fail_arg("NULL argument");
  if (ctx->P == 0) return QS_OK;
  return ctx->dtype == QS_C128 ? backward_t<double>(ctx, grads) : backward_t<float>(ctx, grads);
}

int qs_adjoint_stats(const qs_adjoint *ctx, uint64_t *sweeps, uint64_t *kernels) {
  if (!ctx || !sweeps || !kernels) return fail_arg("NULL argument");
  *sweeps = ctx->sweeps;
  *kernels = ctx->kernels;
  return QS_OK;
}

int qs_adjoint_reverse_stats(const qs_adjoint *ctx, uint64_t *blocks, uint64_t *restores) {
  if (!ctx || !blocks || !restores) return fail_arg("NULL argument");
  *blocks = ctx->rev_blocks;
  *restores = ctx->rev_restores;
  return QS_OK;
}

void qs_adjoint_free(qs_adjoint *ctx) { free_ctx(ctx); }

}  // extern "C"

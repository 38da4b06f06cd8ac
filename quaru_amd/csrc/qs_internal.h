// qs_internal.h -- types shared by the engine's translation units.
//
// The state vector is 2^n amplitudes, {re, im} interleaved, in basis-index
// order, qubit q = bit q (src/register.rs:35, :41-50).  complex128 registers
// store double2, complex64 registers float2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace qs {

constexpr int kMaxQubits = 40;
constexpr int kMaxFixed = 64;  // bit positions a kernel can hold fixed

struct cd {
  double re, im;
};

// A gate after classification on the host.
//  kDense : 2^na x 2^na matrix `m` on qubits act[0..na), where
//           (idx & cmask) == cval.
//  kDiag  : 2^na diagonal entries `m` on act[], same condition; na may be 0
//           (then `m[0]` multiplies every selected amplitude).
//  kSparse: a diagonal whose entries are 1 except the few listed in
//           (spat, m): amplitude idx is multiplied by m[e] when
//           (idx & smask) == spat[e]; smask covers act[] and the controls.
//  kPerm  : a monomial 2^na x 2^na matrix (one nonzero per row and column):
//           the amplitude at act pattern r becomes m[r] times the one at act
//           pattern perm[r], where (idx & cmask) == cval.
struct Gate {
  enum Kind : int { kIdentity = 0, kDense = 1, kDiag = 2, kSparse = 3, kPerm = 4 };
  Kind kind = kIdentity;
  uint64_t cmask = 0, cval = 0;
  int na = 0;
  int act[kMaxFixed] = {};
  std::vector<cd> m;
  uint64_t smask = 0;
  std::vector<uint64_t> spat;
  std::vector<uint32_t> perm;
  // Number of logical gates this entry stands for (apply_all counts n).
  int logical = 1;
  // Batched 2x2 (kDense, na = 1, no controls; the adjoint's rotations): the
  // state is 2^nb independent states stacked in the index bits from
  // `bshift` up, and state b uses m[4b .. 4b + 3].  bmask = 2^nb - 1; 0 for
  // an ordinary gate.
  int bshift = 0;
  uint32_t bmask = 0;
};

// Bit positions held fixed while the kernel enumerates the free index bits:
// the free counter p is expanded by inserting a zero at each pos[] (sorted
// ascending), then `set` is OR-ed in.
struct Fixed {
  uint64_t set;
  int nf;
  uint8_t pos[kMaxFixed];
};

Fixed make_fixed(uint64_t mask, uint64_t set);

// Classify a dense 2^k x 2^k complex128 matrix on `targets` (qs_register.hip);
// cmask/cval are controls already known.
Gate classify_matrix(const double *mat, const size_t *targets, size_t k, uint64_t cmask,
                     uint64_t cval);

// Record the thread's last error (read back by qs_last_error); returns code.
int set_error(int code, size_t a, size_t b, const std::string &msg);
int set_hip_error(hipError_t e, const char *where);

// ---- kernels (qs_kernels.hip) ---------------------------------------------
// All launches go to `s`; `psi` is the device state of 2^n amplitudes.
template <typename R>
hipError_t launch_init_basis(void *psi, int n, uint64_t k, hipStream_t s);
template <typename R>
hipError_t launch_init_product(void *psi, int n, const double *d_qubits, hipStream_t s);
template <typename R>
hipError_t launch_gate_dense(void *psi, int n, const Gate &g, void *d_mat, void *scratch,
                             hipStream_t s);
template <typename R>
hipError_t launch_gate_diag(void *psi, int n, const Gate &g, void *d_mat, hipStream_t s);
template <typename R>
hipError_t launch_gate_sparse(void *psi, int n, const Gate &g, hipStream_t s);
// kPerm: d_tab holds perm[2^na] (uint32) then, at byte offset
// perm_table_phase_offset(na), the 2^na factors in the register's precision;
// scratch receives a copy of the state.
size_t perm_table_phase_offset(int na);
template <typename R>
hipError_t launch_gate_perm(void *psi, int n, const Gate &g, const void *d_tab, void *scratch,
                            hipStream_t s);

// Reductions write FP64 results to device memory `d_out` (caller copies).
// d_part must hold kReduceBlocks * 2 doubles.
constexpr int kReduceBlocks = 1024;
template <typename R>
hipError_t launch_prob_split(const void *psi, int n, int target, double *d_part, double *d_out,
                             hipStream_t s);
template <typename R>
hipError_t launch_collapse(void *psi, int n, int target, int keep, double total, hipStream_t s);
// Keep the amplitudes i with (i & mask) == pat, divided by sqrt(total); zero
// the rest.  Stream order, for states of at least 2^13 amplitudes
// (hipErrorInvalidValue below).
template <typename R>
hipError_t launch_collapse_pattern(void *psi, int n, uint64_t mask, uint64_t pat, double total,
                                   hipStream_t s);
// launch_collapse(ctarget, keep, total) and launch_prob_split(target) in one
// pass, with the same {prob_0, prob_1} bits as the two in sequence.
template <typename R>
hipError_t launch_collapse_split(void *psi, int n, int ctarget, int keep, double total,
                                 int target, double *d_part, double *d_out, hipStream_t s);
// Measure and collapse with the outcome decided on the device (no host
// round trip): d_probs (2 doubles) and d_coll (2 doubles) are work space, the
// outcome (0 or 1) goes to *d_slot.
template <typename R>
hipError_t launch_measure_deferred(void *psi, int n, int target, double x, double *d_part,
                                   double *d_probs, double *d_coll, unsigned char *d_slot,
                                   hipStream_t s);
template <typename R>
hipError_t launch_norm(const void *psi, int n, double *d_part, double *d_out, hipStream_t s);
hipError_t launch_inner(const void *a, bool a_c64, const void *b, bool b_c64, int n,
                        double *d_part, double *d_out, hipStream_t s);
template <typename R>
hipError_t launch_neq_count(const void *a, const void *b, int n, double tol, double *d_part,
                            double *d_out, hipStream_t s);
template <typename R>
hipError_t launch_probabilities(const void *psi, int n, double *d_out, hipStream_t s);
template <typename R>
hipError_t launch_convert_from_c128(void *psi, const double *src, uint64_t count, hipStream_t s);
template <typename R>
hipError_t launch_convert_to_c128(double *dst, const void *psi, uint64_t count, hipStream_t s);
// psi[d_idx[i]] widened to complex128 into dst[2i], dst[2i+1]; every index < 2^n.
template <typename R>
hipError_t launch_gather_amps(double *dst, const void *psi, const uint64_t *d_idx, uint64_t count,
                              hipStream_t s);
// Histogram of |psi|^2 over the bit pattern of m qubits: bin bit j <->
// qubit pos_order[j]; fx holds those m positions fixed (set = 0).  d_part
// holds blocks x 2^m partials.
template <typename R>
hipError_t launch_set_histogram(const void *psi, int n, const Fixed &fx, const uint8_t *pos_order,
                                int m, double *d_part, int blocks, double *d_hist, hipStream_t s);
// |psi|^2 of the amplitudes expand(j, fx), j < count, into d_out.
template <typename R>
hipError_t launch_gather_probs(const void *psi, const Fixed &fx, uint64_t count, double *d_out,
                               hipStream_t s);
// Leave only amplitude idx, divided by sqrt(total); zero the rest.
template <typename R>
hipError_t launch_collapse_to(void *psi, int n, uint64_t idx, double total, hipStream_t s);
// Chunk sums of |psi|^2 for the CDF: chunk c covers [c*chunk, (c+1)*chunk).
template <typename R>
hipError_t launch_chunk_sums(const void *psi, int n, uint64_t chunk, double *d_out,
                             hipStream_t s);

// ---- fused sweep engine (qs_sweep.hip) --------------------------------------
struct SweepPlan;  // defined in qs_sweep.h
}  // namespace qs

// qs_sweep_launch.h -- launching the fused sweep kernel k_sweep.
//
// The kernel template (qs_sweep_kernel.h) is instantiated for two storage
// precisions, three register-bit counts A, two tile shapes W (a tile per
// wave, or a workgroup tile with two wave bits) and two index widths.  Each
// (precision, A) pair is compiled in a translation unit of its own
// (qs_sweep_launch.hip built once per pair), so the instantiations build in
// parallel; the planner (qs_sweep.hip) calls them through these functions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qs {
namespace sweep {

// Launch k_sweep<R, A, I, W> over `ntiles` wave tiles (W = 0) or workgroup
// tiles (W = 2) of an n-qubit state; I is 32-bit up to 32 qubits.  `d_blob`
// is the device copy of the sweep descriptor.  `xcd` gives each XCD one
// contiguous eighth of the tiles when the grid divides by 8.
template <typename R, int A>
hipError_t launch_sweep_ra(int W, int n, hipStream_t stream, void *psi, const void *d_blob,
                           uint64_t ntiles, bool xcd);

// Grid of a sweep over `ntiles` tiles: one workgroup per tile (W = 2) or
// per four wave tiles (W = 0); with `xcd`, each XCD walks one contiguous
// eighth of the tiles when the grid divides by 8 (xcd_chunk = grid / 8).
inline void sweep_grid(int W, uint64_t ntiles, bool xcd, uint64_t *grid, uint32_t *xcd_chunk) {
  const uint64_t per_block = W == 0 ? 4 : 1;
  *grid = (ntiles + per_block - 1) / per_block;
  *xcd_chunk = (xcd && *grid % 8 == 0) ? (uint32_t)(*grid / 8) : 0u;
}

#define QS_SWEEP_LAUNCH_DECL(R, A)                                                       \
  template <>                                                                            \
  hipError_t launch_sweep_ra<R, A>(int W, int n, hipStream_t stream, void *psi,          \
                                   const void *d_blob, uint64_t ntiles, bool xcd);
QS_SWEEP_LAUNCH_DECL(double, 3)
QS_SWEEP_LAUNCH_DECL(double, 4)
QS_SWEEP_LAUNCH_DECL(double, 5)
QS_SWEEP_LAUNCH_DECL(float, 3)
QS_SWEEP_LAUNCH_DECL(float, 4)
QS_SWEEP_LAUNCH_DECL(float, 5)
#undef QS_SWEEP_LAUNCH_DECL

}  // namespace sweep
}  // namespace qs

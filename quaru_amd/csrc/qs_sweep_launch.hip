// qs_sweep_launch.hip -- the k_sweep instantiations of one (precision,
// register bits) pair, chosen at compile time by QS_SWEEP_R (double or
// float) and QS_SWEEP_A (3, 4 or 5).  The build compiles this file once per
// pair (see qs_sweep_launch.h).
#include <hip/hip_runtime.h>

#include "qs_sweep_kernel.h"
#include "qs_sweep_launch.h"

#if !defined(QS_SWEEP_R) || !defined(QS_SWEEP_A)
#error "build with -DQS_SWEEP_R=double|float -DQS_SWEEP_A=3|4|5"
#endif

namespace qs {
namespace sweep {
namespace {

template <typename R, int A, int W>
hipError_t launch(int n, hipStream_t stream, void *psi, const void *d_blob, uint64_t ntiles,
                  bool xcd) {
  using V = typename SV2<R>::T;
  uint64_t grid;
  uint32_t xcd_chunk;
  sweep_grid(W, ntiles, xcd, &grid, &xcd_chunk);
  if (n <= 32)
    hipLaunchKernelGGL((k_sweep<R, A, uint32_t, W>), dim3((unsigned)grid), dim3(kThreads), 0,
                       stream, (V *)psi, (const char *)d_blob, xcd_chunk);
  else
    hipLaunchKernelGGL((k_sweep<R, A, uint64_t, W>), dim3((unsigned)grid), dim3(kThreads), 0,
                       stream, (V *)psi, (const char *)d_blob, xcd_chunk);
  return hipGetLastError();
}

}  // namespace

template <>
hipError_t launch_sweep_ra<QS_SWEEP_R, QS_SWEEP_A>(int W, int n, hipStream_t stream, void *psi,
                                                   const void *d_blob, uint64_t ntiles,
                                                   bool xcd) {
  if (W == kWaveBits)
    return launch<QS_SWEEP_R, QS_SWEEP_A, kWaveBits>(n, stream, psi, d_blob, ntiles, xcd);
  if (W == 0) return launch<QS_SWEEP_R, QS_SWEEP_A, 0>(n, stream, psi, d_blob, ntiles, xcd);
  return hipErrorInvalidValue;
}

}  // namespace sweep
}  // namespace qs

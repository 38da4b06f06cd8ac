// sweep_memory -- the memory time of the fused sweep's tile shapes.
//
// Runs the sweep kernel's own load/store path (sweep_tile in
// qs_sweep_kernel.h, complex128, 4 register slots, 2 wave bits) with no
// operations over a 30-qubit state (16 GiB), for tile role assignments:
//
//   sweep_memory PLANS   every complex128 line "qs_sweep n 30 lanes a b c regs
//                        r0 r1 r2 r3 waves w0 w1 ops K" of PLANS (the planner's
//                        trace, QS_SWEEP_TRACE=1), in order
//   sweep_memory         built-in shapes: a contiguous tile, one distant wave
//                        bit, windows of 8 tile qubits at each position
//
// Each shape: one untimed launch, then 4 launches timed with HIP events; the
// line gives ms per launch and the fraction of the 8 TB/s HBM peak for the
// algorithmic 2 x 2^30 x 16 B.  Lane bits 0..2 are qubits 0..2, as in every
// complex128 sweep.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "qs_sweep_kernel.h"
#include "qs_sweep_launch.h"

using namespace qs::sweep;

namespace {

constexpr int kN = 30;
constexpr int kA = 4;
constexpr int kW = 2;
constexpr int kFree = kLaneBits - 3;  // lane bits the planner assigns

struct NoOps {
  template <typename AA, typename BB, typename CC, typename GG>
  __device__ __forceinline__ void operator()(AA &, const BB &, const CC &, const GG &) const {}
};

__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3))) void k_sweep_memory(
    double2 *__restrict__ psi, const char *__restrict__ blob_g, uint32_t xcd_chunk) {
  sweep_tile<double, kA, uint32_t, kW>(psi, blob_g, xcd_chunk, NoOps());
}

#define CHECK(x)                                                                      \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) {                                                           \
      std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(1);                                                                   \
    }                                                                                 \
  } while (0)

struct Shape {
  std::string name;
  int q[kFree + kA + kW];  // free lane bits, register slots, wave bits
};

// Header of a no-op sweep for `s`; false if the roles are not 9 distinct
// qubits in 3..29.
bool header(const Shape &s, SweepHeader &h) {
  uint64_t P = 7;  // qubits 0..2: the fixed lane bits
  for (int q : s.q) {
    if (q < 3 || q >= kN || ((P >> q) & 1ull)) return false;
    P |= 1ull << q;
  }
  h = SweepHeader{};
  int b = 0;
  for (int q = 0; q < kN; ++q)
    if ((P >> q) & 1ull) h.tpos[b++] = (uint32_t)q;
  h.nops = 0;
  h.npool = 0;
  h.ntiles = 1ull << (kN - b);
  for (int i = 0; i < 3; ++i) h.lg[i] = 1ull << i;
  for (int i = 0; i < kFree; ++i) h.lg[3 + i] = 1ull << s.q[i];
  for (int j = 0; j < kA; ++j) h.gb[j] = 1ull << s.q[kFree + j];
  h.nwave = kW;
  for (int k = 0; k < kW; ++k) h.wb[k] = 1ull << s.q[kFree + kA + k];
  return b == kLaneBits + kA + kW;
}

std::vector<Shape> read_plans(const char *path) {
  std::vector<Shape> out;
  FILE *f = std::fopen(path, "r");
  if (!f) {
    std::fprintf(stderr, "cannot open %s\n", path);
    std::exit(2);
  }
  char line[1024];
  int k = 0;
  while (std::fgets(line, sizeof line, f)) {
    const char *p = std::strstr(line, "qs_sweep n ");
    if (!p) continue;
    Shape s;
    int n = 0, ops = 0;
    const int got = std::sscanf(p, "qs_sweep n %d lanes %d %d %d regs %d %d %d %d waves %d %d ops %d", &n,
                                &s.q[0], &s.q[1], &s.q[2], &s.q[3], &s.q[4], &s.q[5], &s.q[6], &s.q[7],
                                &s.q[8], &ops);
    if (got != 11 || n != kN) continue;  // complex128 plans of 30 qubits only
    s.name = "sweep" + std::to_string(k++) + "_ops" + std::to_string(ops);
    out.push_back(s);
  }
  std::fclose(f);
  return out;
}

std::vector<Shape> builtin() {
  std::vector<Shape> out;
  out.push_back({"contiguous", {3, 4, 5, 6, 7, 8, 9, 10, 11}});
  out.push_back({"contiguous_low_waves", {5, 6, 7, 8, 9, 10, 11, 3, 4}});
  for (int b = 12; b < kN; ++b) out.push_back({"wave_bit_" + std::to_string(b), {3, 4, 5, 6, 7, 8, 9, 10, b}});
  for (int s = 4; s + 8 <= kN; ++s)
    out.push_back({"window3_" + std::to_string(s), {3, s, s + 1, s + 2, s + 3, s + 4, s + 5, s + 6, s + 7}});
  for (int s = 5; s + 7 <= kN; ++s)
    out.push_back({"window34_" + std::to_string(s), {3, 4, s, s + 1, s + 2, s + 3, s + 4, s + 5, s + 6}});
  return out;
}

}  // namespace

int main(int argc, char **argv) {
  const std::vector<Shape> shapes = argc > 1 ? read_plans(argv[1]) : builtin();
  if (shapes.empty()) {
    std::fprintf(stderr, "no complex128 30-qubit plan lines\n");
    return 2;
  }
  const uint64_t len = 1ull << kN;
  double2 *psi = nullptr;
  char *d_blob = nullptr;
  CHECK(hipMalloc(&psi, len * sizeof(double2)));
  CHECK(hipMemset(psi, 0, len * sizeof(double2)));
  CHECK(hipMalloc(&d_blob, sizeof(SweepHeader)));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  const int reps = 4;
  double sum = 0.0;
  int count = 0;
  for (const Shape &s : shapes) {
    SweepHeader h;
    if (!header(s, h)) {
      std::fprintf(stderr, "%s: roles are not 9 distinct qubits in 3..%d\n", s.name.c_str(), kN - 1);
      return 2;
    }
    CHECK(hipMemcpy(d_blob, &h, sizeof h, hipMemcpyHostToDevice));
    uint64_t grid;
    uint32_t chunk;
    sweep_grid(kW, h.ntiles, true, &grid, &chunk);
    hipLaunchKernelGGL(k_sweep_memory, dim3((unsigned)grid), dim3(kThreads), 0, 0, psi, d_blob, chunk);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_sweep_memory, dim3((unsigned)grid), dim3(kThreads), 0, 0, psi, d_blob, chunk);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    ms /= reps;
    sum += ms;
    ++count;
    std::printf("%-24s lanes %2d %2d %2d regs %2d %2d %2d %2d waves %2d %2d  %.3f ms  frac %.3f\n",
                s.name.c_str(), s.q[0], s.q[1], s.q[2], s.q[3], s.q[4], s.q[5], s.q[6], s.q[7], s.q[8],
                ms, 2.0 * len * 16 / (ms * 1e-3) / 8e12);
    std::fflush(stdout);
  }
  std::printf("mean over %d shapes: %.3f ms, frac %.3f\n", count, sum / count,
              2.0 * len * 16 / (sum / count * 1e-3) / 8e12);
  CHECK(hipFree(d_blob));
  CHECK(hipFree(psi));
  return 0;
}

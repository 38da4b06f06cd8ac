// Process exit while specialised sweep kernels are compiling (CPU only: the
// compiles, in helper processes, need no GPU).  Queues compiles of several op-code sequences in
// the automatic mode, then returns from main half a second later, while
// they run; the process must exit normally.  With `wait` as argument it first waits for the
// compiles and prints how many compiled and failed.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/quaru_hip.h"
#include "qs_sweep_jit.h"
#include "qs_sweep_kernel.h"

using namespace qs::sweep;

int main(int argc, char **argv) {
  qs_sweep_jit_mode(1);
  for (int s = 0; s < 8; ++s) {
    std::vector<OpRec> ops;
    for (int i = 0; i < 6 + s; ++i) {
      OpRec o{};
      o.code = (uint32_t)(CODE_U1 + (i + s) % 4);  // 2x2 on register slot 0..3
      ops.push_back(o);
    }
    const std::string sig = jit_signature(ops, false, 4, kWaveBits);
    if (jit_kernel(sig, ops, false, 4, kWaveBits) != nullptr) {
      std::printf("a kernel was ready before its compile\n");
      return 1;
    }
  }
  if (argc > 1 && std::strcmp(argv[1], "wait") == 0) {
    qs_sweep_jit_wait();
    uint64_t compiled = 0, failed = 0, launches = 0;
    qs_sweep_jit_stats(&compiled, &failed, &launches);
    std::printf("compiled %lu failed %lu\n", (unsigned long)compiled, (unsigned long)failed);
  }
  // let the compiles start before the exit
  std::this_thread::sleep_for(std::chrono::milliseconds(500));
  std::printf("main returns\n");
  if (argc > 2 && std::strcmp(argv[1], "status") == 0) return std::atoi(argv[2]);
  return 0;
}

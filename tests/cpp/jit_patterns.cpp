// The register-slot patterns the specialised sweep kernels take as
// constants (jit_register_patterns) against the generic kernel's run-time
// test: for every op and register index r, (r & rm) == rv must hold exactly
// when (goff(r) & cm & mask) == (cv & mask), goff(r) being the OR of the
// slots' qubit bits selected by r.  Random slot qubits and patterns, CPU
// only.  Prints the number of cases checked; exit status 1 on a mismatch.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "qs_sweep_jit.h"
#include "qs_sweep_kernel.h"

using namespace qs::sweep;

int main() {
  std::mt19937_64 rng(12345);
  uint64_t checked = 0;
  for (int trial = 0; trial < 2000; ++trial) {
    const int A = 3 + (int)(rng() % 3);  // 3..5 register slots
    const int n = 16 + (int)(rng() % 17);  // 16..32 qubits
    uint64_t gb[8] = {};
    uint64_t used = 0, mask = 0;
    for (int j = 0; j < A; ++j) {
      int q;
      do q = (int)(rng() % (uint64_t)n); while ((used >> q) & 1ull);
      used |= 1ull << q;
      gb[j] = 1ull << q;
      mask |= gb[j];
    }
    std::vector<OpRec> ops(8);
    for (OpRec &o : ops) {
      o = OpRec{};
      o.cmask = rng() & ((1ull << n) - 1ull);
      // values mostly inside the mask, sometimes outside it
      o.cval = (rng() & o.cmask) | ((rng() % 8 == 0) ? (rng() & ((1ull << n) - 1ull)) : 0ull);
    }
    const std::vector<int> rp = jit_register_patterns(ops, gb, A);
    for (size_t i = 0; i < ops.size(); ++i) {
      const uint32_t rm = (uint32_t)rp[i] & 0xFFu, rv = ((uint32_t)rp[i] >> 8) & 0xFFu;
      const uint64_t cmr = ops[i].cmask & mask, cvr = ops[i].cval & mask;
      for (int r = 0; r < (1 << A); ++r) {
        uint64_t go = 0;
        for (int j = 0; j < A; ++j)
          if ((r >> j) & 1) go |= gb[j];
        const bool want = (go & cmr) == cvr;
        const bool got = ((uint32_t)r & rm) == rv;
        if (want != got) {
          std::printf("mismatch: trial %d op %zu r %d cmask %llx cval %llx\n", trial, i, r,
                      (unsigned long long)ops[i].cmask, (unsigned long long)ops[i].cval);
          return 1;
        }
        ++checked;
      }
    }
  }
  std::printf("checked %llu\n", (unsigned long long)checked);
  return 0;
}

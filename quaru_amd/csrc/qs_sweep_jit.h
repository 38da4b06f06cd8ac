// qs_sweep_jit.h -- sweep kernels specialised to one op list, compiled at
// run time with hipRTC.
//
// The generic sweep kernel (k_sweep) reads each op's code from the
// descriptor and dispatches through a switch: every op pays the dispatch and
// a whole-tile register copy at the loop's back edge.  A specialised kernel
// for a given sequence of op codes runs the same op functions (apply_code in
// qs_sweep_kernel.h) as straight-line code; matrices, slots' qubits and
// patterns still come from the descriptor, so one compiled kernel serves
// every sweep with the same code sequence (a repeated circuit, the layers of
// a circuit, the steps of an optimiser).
//
// Modes (QS_SWEEP_JIT, or qs_sweep_jit_mode): 0 off; 1 (default) for states
// of 22 qubits and more, a new code sequence launches the generic kernel and
// queues a compile (host threads start one helper process per kernel,
// bin/qs_sweep_compile beside the library), later sweeps with that sequence
// use the compiled kernel; 2 every sweep, compiled before its first launch.
// A failed compile leaves the sequence on the generic kernel.  At exit the
// helpers still running are stopped; nothing waits for them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace qs {
namespace sweep {

struct OpRec;

// Each op's pattern on the register slots, in the form the generated
// kernel takes it as a constant (make_pred in qs_sweep_kernel.h: slot mask
// in bits 0..7, slot values in bits 8..15), from the slots' qubits gb[0..A).
std::vector<int> jit_register_patterns(const std::vector<OpRec> &ops, const uint64_t *gb, int A);

// The key of a specialised kernel: precision, register bits, wave bits and
// each op's code (and its fail code, whether it is batched, and its
// register pattern when `rp` is given).
std::string jit_signature(const std::vector<OpRec> &ops, bool c64, int A, int W,
                          const std::vector<int> &rp = {});

// The compiled kernel for `sig` on the current device, or nullptr (not
// compiled yet, compiling, off or failed).  Queues the compile of a new
// signature; in mode 2 waits for it.  `rp` (empty: patterns read at run
// time) as for jit_signature.
hipFunction_t jit_kernel(const std::string &sig, const std::vector<OpRec> &ops, bool c64, int A,
                         int W, const std::vector<int> &rp = {});

int jit_mode();
int jit_set_mode(int mode);  // returns the previous mode
void jit_wait();             // until every queued compile has finished
void jit_stats(uint64_t *compiled, uint64_t *failed, uint64_t *launches);
void jit_count_launch();

}  // namespace sweep
}  // namespace qs

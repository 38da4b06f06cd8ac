// qs_sweep.h -- the fused sweep engine.
//
// A *sweep* is one pass over the state: each wave owns one tile of 2^T
// amplitudes (T = 6 + A; the index bits at the T tile qubits vary inside a
// tile, the others are fixed per wave), loads it once from HBM into
// registers, applies a run of gates to it, and stores it once.  Every lane
// keeps 2^A amplitudes in registers: a gate on a register qubit pairs
// registers inside a lane, a gate on a lane qubit exchanges with the partner
// lane.  Controls and diagonal gates need no data movement at all: every
// amplitude's global index is known, so they are predicates and multiplies.
//
// The planner packs the pending gate list into sweeps, moving a gate ahead of
// earlier ones only when it commutes with every gate it passes (on every
// shared qubit both act diagonally), so the result equals applying the list
// in order, and merges each sweep's one-qubit gates into multiplexed 2x2 ops
// (qs_sweep_merge.h).  Gates the sweep kernel does not handle (dense on 3+
// active qubits, large diagonal tables, permutations) run as single-gate
// kernels between sweeps.
#pragma once

#include <functional>
#include <vector>

#include "qs_internal.h"

namespace qs {

struct SweepEngine;
SweepEngine *sweep_engine_new();
void sweep_engine_free(SweepEngine *e);

struct SweepRunner {
  int n = 0;
  bool dtype_c64 = false;
  void *psi = nullptr;
  hipStream_t stream = nullptr;
  // Copy host bytes to a device buffer that stays valid until the stream
  // reaches the next launch.
  std::function<hipError_t(const void *, size_t, void **)> upload;
  // Launch one gate with the per-gate kernels.
  std::function<hipError_t(const Gate &)> single;
  // Called with true before and false after each sweep kernel launch
  // (timing hooks); may be empty.
  std::function<void(bool)> mark;
  // Op records run by sweep kernels, and gates merged into another op of
  // their sweep; incremented when non-null.
  uint64_t *sweep_ops = nullptr;
  uint64_t *merged = nullptr;
  // When set, each sweep's descriptor (the host bytes the kernel would read)
  // goes to this function instead of being uploaded and launched: the CPU
  // checks of the planner run the sweeps on a host copy of the state.
  std::function<hipError_t(const void *, size_t)> host_sweep;
};

// Apply `gates` in order (with commuting reorders) through sweeps and
// per-gate kernels.  Counters are incremented by the launches made.
hipError_t sweep_execute(SweepEngine *e, SweepRunner &run, const std::vector<Gate> &gates,
                         uint64_t *sweeps, uint64_t *kernels);

// Register bits per lane used for a dtype (a wave holds 2^(6 + bits)
// amplitudes of its tile).
int sweep_reg_bits(bool c64);

// Qubits 0..b-1 sit on lane bits 0..b-1 of every tile, so that each group of
// 2^b lanes reads 2^b consecutive amplitudes (defaults b = 3 for complex128,
// 4 for complex64: 128-byte runs); the other 6 - b lane bits take qubits the
// planner chooses per sweep.
int sweep_low_bits(bool c64);

// Tile qubits at or above sweep_page_qubit() each double the pages a wave
// touches; at most A + sweep_high_lanes() of them go into one sweep.
int sweep_page_qubit(bool c64);

// Whether sweeps use workgroup tiles whose 4 waves add two tile qubits
// (QS_SWEEP_WAVES_C128 / QS_SWEEP_WAVES_C64, both on by default).
bool sweep_wave_bits(bool c64);

// How many of a workgroup tile's wave bits take dense gates
// (QS_SWEEP_DENSE_WAVES, default 1); the others hold qubits that only
// controls and diagonal gates touch.
int sweep_dense_wave_bits();

// Whether one-qubit ops of a sweep are merged into multiplexed 2x2 ops
// (QS_SWEEP_MERGE, default on).
bool sweep_merge_ops();

// Whether sweep launches give each XCD one contiguous range of tiles
// (QS_SWEEP_XCD, default on).
bool sweep_xcd_chunks();
int sweep_high_lanes();

}  // namespace qs

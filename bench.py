#!/usr/bin/env python3
"""Benchmark: 30-qubit random-circuit gate application on one MI355X.

Workload: n = 30 qubits, 20 layers; per layer a Haar 2x2 on every qubit,
cnot(q, q+1) for q = 0..28, then 8 three-qubit gates on seeded distinct
triples (even -> Toffoli cnx([a,b],c), odd -> cnz([a,b],c)): 67 gates per
layer, 1340 per step.  One step applies the whole circuit through the GPU
register's drop-in ``apply`` (fused sweep engine, the default) and
synchronises.  The state (16 GiB complex128) stays in HBM.

Reported on one JSON line:
  value        gates/s (logical gate count / wall time), complex128
  roofline     the dominant kernel (the fused sweep), timed live with HIP
               events on the register's stream: algorithmic bytes per launch
               = 2 * 2^n * 16 (state read once, written once)
  c64          the same circuit stored in complex64 (FP64 reductions), gates/s
               and its infidelity against the complex128 result
  unfused      one kernel per gate for the first layer: gates/s and the
               per-gate kernel's roofline (2 * |changed set| * s per gate)
  adjoint      gradients/s of the 20-qubit RY/RZ + CZ-ring ansatz, batch 64,
               complex128 and complex64, against central finite differences
               on the same GPU
  cpu_baseline the CPU oracle (C restatement of try_apply_fast,
               src/register.rs:140-191) on a bounded sample of the same
               workload, single thread; a multi-core variant beside it
  grover32     (--grover32) the 32-qubit Grover workload: 5-qubit fused
               32 x 32 blocks vs one gate per qubit, complex128 and complex64

``--profile SECTION`` runs a 1 GiB copy (counter calibration) and then one
step of that section only, for rocprofv3 (c128 and c64: after one warmup
step, so the profiled step runs the specialised sweep kernels).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
CLASSES = ["sweep", "gate1", "gate2", "gatek", "big", "diag", "sparse", "perm"]

from quaru_amd import Register, circuits  # noqa: E402
from quaru_amd import operation as qops  # noqa: E402
from quaru_amd import _lib  # noqa: E402


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def recorded_traffic(name: str):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    (calibrated against copy kernels of known size) for this workload, as
    recorded in profiles/<name>; None if absent."""
    import re
    path = os.path.join(ROOT, "profiles", name)
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*(\d+)", text)
    return int(m.group(1)) if m else None


def timing(reg: Register, reset: bool = True):
    import ctypes
    ms = (ctypes.c_double * 8)()
    n = (ctypes.c_uint64 * 8)()
    rc = _lib.load().qs_timing_read(reg.handle, ms, n, 1 if reset else 0)
    assert rc == 0, rc
    return list(ms), list(n)


def timing_enable(reg: Register, on: bool) -> None:
    rc = _lib.load().qs_timing_enable(reg.handle, 1 if on else 0)
    assert rc == 0, rc


def device_sync() -> None:
    import torch
    if torch.cuda.is_available():
        torch.cuda.synchronize()


def run_steps(reg: Register, opl, steps: int) -> float:
    reg.sync()
    device_sync()
    t0 = time.perf_counter()
    for s in range(steps):
        for op in opl:
            reg.apply(op)
        reg.sync()
        log(f"  step {s + 1}/{steps} done")
    device_sync()
    return time.perf_counter() - t0


def jit_stats():
    """(compiled, failed, launches) of the specialised sweep kernels so far."""
    import ctypes
    lib = _lib.load()
    c, f, n = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    lib.qs_sweep_jit_stats(ctypes.byref(c), ctypes.byref(f), ctypes.byref(n))
    return c.value, f.value, n.value


def jit_wait():
    """Wait for the sweep kernels specialised during the warmup (compiled on
    host threads while the warmup ran the generic kernel) so that the timed
    steps run them; logs the totals."""
    t0 = time.perf_counter()
    _lib.load().qs_sweep_jit_wait()
    c, f, _ = jit_stats()
    log(f"specialised sweep kernels: {c} compiled, {f} failed "
        f"(waited {time.perf_counter() - t0:.1f} s)")


def timed_run(reg: Register, opl, steps: int, warmup: int):
    """Warm up, then time `steps` passes of `opl` with the kernel timers on.
    Returns (seconds, per-class kernel ms, per-class launches, stats)."""
    for i in range(warmup):
        for op in opl:
            reg.apply(op)
        reg.sync()
        log(f"warmup {i + 1}/{warmup} done")
    if warmup:
        jit_wait()
    reg.stats(reset=True)
    timing(reg)
    timing_enable(reg, True)
    jit0 = jit_stats()[2]
    dt = run_steps(reg, opl, steps)
    timing_enable(reg, False)
    ms, cnt = timing(reg)
    st = reg.stats(reset=True)
    st["specialised_launches"] = jit_stats()[2] - jit0
    return dt, ms, cnt, st


def cpu_baseline(w: circuits.Workload, threads: int):
    """Time the CPU oracle on a bounded sample of the workload: the first
    Haar gate on one thread, and one Haar, CNOT and Toffoli gate on
    `threads` threads (the OpenMP variant of the same algorithm)."""
    import ctypes

    import oracle
    n = w.n
    L = oracle.lib()
    state = np.zeros(1 << n, dtype=np.complex128)
    state[0] = 1.0
    haar = next(o for o in w.ops if o._label == "haar")
    cn = next(o for o in w.ops if o._label == "cnot")
    tof = next((o for o in w.ops if o._label == "cnx"), cn)

    def scalar(op, st=None, nq=None):
        st = state if st is None else st
        nq = n if nq is None else nq
        m = np.ascontiguousarray(op.matrix())
        t = np.ascontiguousarray(np.asarray(op.targets(), dtype=np.uint64))
        ea, eb = ctypes.c_size_t(0), ctypes.c_size_t(0)
        t0 = time.perf_counter()
        rc = L.orc_try_apply_fast(st.ctypes.data_as(ctypes.c_void_p), nq,
                                  m.ctypes.data_as(ctypes.c_void_p), m.shape[0], m.shape[1],
                                  t.ctypes.data_as(ctypes.c_void_p), t.size, ctypes.byref(ea),
                                  ctypes.byref(eb))
        assert rc == 0
        return time.perf_counter() - t0

    log(f"cpu baseline: scalar oracle, one Haar gate at {n} qubits")
    t_scalar = scalar(haar)
    scratch = np.empty_like(state)
    per = {}
    for name, op in (("haar", haar), ("cnot", cn), ("toffoli", tof)):
        t0 = time.perf_counter()
        oracle.apply_fast_omp(state, scratch, op.matrix(), op.targets(), threads)
        per[name] = time.perf_counter() - t0
        log(f"cpu baseline: omp {threads} threads {name} {per[name]:.2f} s")
    del scratch, state
    layer_gates = n + (n - 1) + (8 if tof is not cn else 0)
    layer_omp = n * per["haar"] + (n - 1) * per["cnot"] + (layer_gates - (2 * n - 1)) * per["toffoli"]
    # the 26-qubit random circuit: 20 layers = 520 Haar + 500 CNOT gates
    n26 = 26
    s26 = np.zeros(1 << n26, dtype=np.complex128)
    s26[0] = 1.0
    w26 = circuits.random_circuit(n26, 1, seed=2026, measure=False)
    h26 = next(o for o in w26.ops if o._label == "haar")
    c26 = next(o for o in w26.ops if o._label == "cnot")
    t26 = {}
    for name, op in (("haar", h26), ("cnot", c26)):
        t26[name] = scalar(op, s26, n26)
    del s26
    t26_circuit = 520 * t26["haar"] + 500 * t26["cnot"]
    log(f"cpu baseline: 26 qubits, one thread: haar {t26['haar']:.2f} s, "
        f"cnot {t26['cnot']:.2f} s")
    return {
        "value": 1.0 / t_scalar,
        "unit": "gates/s",
        "cores": 1,
        "nproc": os.cpu_count(),
        "kind": "port",
        "sample": (f"one Haar 2x2 gate of the {n}-qubit workload through the oracle's "
                   f"try_apply_fast restatement (filter scan, gather, 2x2 matvec, scatter, "
                   f"out-of-place copy), single thread: {t_scalar:.2f} s; CNOT/Toffoli "
                   f"cost the reference 1.1-1.3x a Haar gate, so gates/s is an upper bound"),
        "multicore": {"value": layer_gates / layer_omp, "unit": "gates/s", "cores": threads,
                      "nproc": os.cpu_count(),
                      "sample": "one Haar, one CNOT, one Toffoli gate (OpenMP variant), "
                                f"weighted by the per-layer gate mix, on every core this "
                                f"process may run on ({threads} of nproc {os.cpu_count()})",
                      "seconds": per},
        "random_26q": {"value": 1020.0 / t26_circuit, "unit": "gates/s", "cores": 1,
                        "sample": "one Haar and one CNOT gate of the 26-qubit random circuit "
                                  "(the oracle's try_apply_fast, one thread), extrapolated "
                                  "to its 520 Haar + 500 CNOT gates",
                        "seconds": t26, "circuit_seconds": t26_circuit},
    }


def adjoint_bench(n: int, layers: int, batch: int, fd_params: int, dtype: str):
    from quaru_amd.adjoint import Expectation, ParamCircuit
    ans = circuits.hardware_efficient_ansatz(n, layers)
    masks, coeffs = circuits.zz_chain_observable(n)
    th = circuits.seeded_thetas(batch, ans.nparams, seed=5)
    pc = ParamCircuit.from_ansatz(ans)
    ckpt = 2
    ex = Expectation(pc, masks, coeffs, dtype=dtype, checkpoint_every=ckpt)
    ex.forward(th)
    ex.backward()  # warmup
    jit_wait()
    device_sync()
    t0 = time.perf_counter()
    _, run = ex.record(th)
    t1 = time.perf_counter()
    fstats = run.stats()
    g = run.backward()
    device_sync()
    t2 = time.perf_counter()
    t_adj, t_fwd, t_bwd = t2 - t0, t1 - t0, t2 - t1
    # algorithmic HBM bytes of the reverse pass (S = the batch of states):
    # lambda = h psi (2S), each tiled block reads and writes psi and lambda
    # (4S), each checkpoint restore copies psi (2S)
    bp = 1 << max(0, (batch - 1).bit_length())
    S = float(bp << n) * (16 if dtype == "c128" else 8)
    rstats = run.reverse_stats
    bwd_bytes = 2 * S + rstats["blocks"] * 4 * S + rstats["restores"] * 2 * S
    fwd_bytes = fstats["sweeps"] * 2 * S
    # central finite differences (h = 1e-4, FP64) on the same GPU for a sample
    # of parameters: each needs two forward passes of the whole batch.
    h = 1e-4
    ex128 = Expectation(pc, masks, coeffs, dtype="c128")
    t0 = time.perf_counter()
    maxerr = 0.0
    for p in range(fd_params):
        tp, tm = th.copy(), th.copy()
        tp[:, p] += h
        tm[:, p] -= h
        fd = (ex128.energies(tp) - ex128.energies(tm)) / (2 * h)
        maxerr = max(maxerr, float(np.max(np.abs(fd - g[:, p]))))
    device_sync()
    t_fd_sample = time.perf_counter() - t0
    t_fd = t_fd_sample * ans.nparams / fd_params
    grads = batch * ans.nparams
    return {
        "workload": f"{n}-qubit RY/RZ + CZ-ring ansatz, {layers} layers, {ans.nparams} params, "
                    f"batch {batch}, checkpoint every {ckpt} layers",
        "dtype": dtype,
        "grads_per_s": grads / t_adj,
        "seconds": t_adj,
        "forward_seconds": t_fwd,
        "backward_seconds": t_bwd,
        "forward_sweeps": fstats["sweeps"],
        "forward_kernels": fstats["kernels"],
        "reverse_blocks": rstats["blocks"],
        "roofline": {
            "bound": "hbm", "kernel": "k_reverse_tile (tiled reverse blocks on psi and lambda), "
                                      "the reverse pass as a whole",
            "achieved": bwd_bytes / t_bwd / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": bwd_bytes / t_bwd / HBM_PEAK, "bytes": bwd_bytes,
            "note": "algorithmic bytes of the reverse pass / its wall time"},
        "forward_roofline": {
            "kernel": "k_sweep with batched rotations (fused forward pass)",
            "achieved": fwd_bytes / t_fwd / 1e9, "unit": "GB/s",
            "frac": fwd_bytes / t_fwd / HBM_PEAK, "bytes": fwd_bytes},
        "fd_grads_per_s": grads / t_fd,
        "fd_sample": f"{fd_params} parameters x 2 forward passes of the batch, extrapolated "
                     f"linearly to all {ans.nparams}",
        "speedup_vs_fd": t_fd / t_adj,
        "max_abs_diff_vs_fd": maxerr,
    }


def measure_bench(reg: Register, n: int):
    """`measure` (src/register.rs:291-342) on an n-qubit complex128 state,
    three in a row (targets 0, n/2, n-1) and a synchronisation: the first
    FP64 probability split reads the state (S = 2^n x 16 B); each later split
    runs fused with the previous measure's collapse, which reads the kept
    amplitudes (all of them for target 0, whose kept half is interleaved in
    every line; half for the others) and writes the state; the last collapse
    runs at the synchronisation.  S + 2S + 1.5S + 1.5S = 6S for three
    measures, 2S per measure.  Then
    a full readout (measure_sequence: one histogram pass over the first 12
    qubits, a gather of the consistent amplitudes, one collapse), twice.  A measure
    returns once its outcome is known and leaves the collapse queued, so the
    three measures are timed together up to a synchronisation, and so is the
    readout."""
    amp = 16
    reg.sync()
    device_sync()
    t0 = time.perf_counter()
    for target in (0, n // 2, n - 1):
        reg.measure(target, 0.5)
    reg.sync()
    t_m = (time.perf_counter() - t0) / 3
    byt = 2.0 * (1 << n) * amp
    # the first readout also loads its kernels and sizes the register's work
    # buffer; the second (on the collapsed state: the same passes) is the
    # steady-state time
    t_seq = []
    for _ in range(2):
        t0 = time.perf_counter()
        reg.measure_sequence(range(n), [0.5] * n)
        reg.sync()
        t_seq.append(time.perf_counter() - t0)
    return {"workload": f"{n}-qubit complex128 state",
            "measure": {"ms": t_m * 1e3, "achieved": byt / t_m / 1e9, "unit": "GB/s",
                        "frac": byt / t_m / HBM_PEAK, "bytes": byt,
                        "kernels": "k_prob_split, then k_collapse_split (a collapse fused "
                                   "with the next split), last k_collapse"},
            "measure_sequence_all_qubits_ms": t_seq[1] * 1e3,
            "measure_sequence_first_call_ms": t_seq[0] * 1e3}


def sweep_ops_bench(n: int, ks=(2, 16), reps: int = 3):
    """Sweeps of K Haar gates on qubits 3..10 of an n-qubit complex128 state
    (one sweep each; run with QS_SWEEP_MERGE=0 so that every gate stays one
    op): the per-sweep time against the op count, for counter runs."""
    rng = np.random.default_rng(1)
    r = Register.new([False] * n)
    res = {}
    for k in ks:
        gates = []
        for i in range(k):
            z = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
            u, _ = np.linalg.qr(z)
            gates.append(qops.Operation(u, [3 + i % 8]))
        dt, ms, cnt, st = timed_run(r, gates, reps, 1)
        avg = ms[0] / 1e3 / max(cnt[0], 1)
        res[f"k{k}"] = {"sweeps": st["sweeps"], "ops": st["sweep_ops"], "avg_ms": avg * 1e3,
                        "frac": 2.0 * (1 << n) * 16 / avg / HBM_PEAK if cnt[0] else None}
        log(f"sweep_ops k={k}: {res[f'k{k}']}")
    r.close()
    return res


def block5_bench(n: int, only=None):
    """Dense complex 32 x 32 blocks (random unitaries) on 5 qubits at n
    qubits, fusion off: complex64 on the FP32 matrix cores (k_gatek5_mfma),
    complex128 on the FP64 vector units (k_gatek<double, 5>).  Three blocks
    each on qubits 10..14 and 20..24; average kernel time and fraction of HBM
    peak (2 x 2^n x s bytes per launch)."""
    rng = np.random.default_rng(55)
    res = {"workload": f"complex 32 x 32 blocks, {n} qubits, qubits 10-14 and 20-24"}
    for dtype in ("c64", "c128"):
        if only and dtype != only:
            continue
        amp = 16 if dtype == "c128" else 8
        r = Register.new([False] * n, dtype).set_fusion(False)
        entry = {}
        for label, tg in (("q10_14", [10, 11, 12, 13, 14]), ("q20_24", [20, 21, 22, 23, 24])):
            z = rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))
            u, _ = np.linalg.qr(z)
            op = qops.Operation(u, tg)
            dt, ms, cnt, _ = timed_run(r, [op], 3, 1)
            avg = ms[3] / 1e3 / max(cnt[3], 1)
            byt = 2.0 * (1 << n) * amp
            entry[label] = {"avg_ms": avg * 1e3, "achieved": byt / avg / 1e9, "unit": "GB/s",
                            "frac": byt / avg / HBM_PEAK,
                            "tflops": (1 << n) * 32 * 8.0 / avg / 1e12}
        entry["kernel"] = ("k_gatek5_mfma (v_mfma_f32_32x32x2f32)" if dtype == "c64"
                           else "k_gatek<double, 5> (FP64 FMA)")
        res[dtype] = entry
        log(f"block5 {dtype}: " + ", ".join(f"{k} {v['avg_ms']:.2f} ms"
                                            for k, v in entry.items() if isinstance(v, dict)))
        r.close()
        del r
    return res


def grover_bench(n: int, iterations: int, seed: int, only=None):
    """The Grover workload at n qubits: H on all, then `iterations` x
    [phase flip of the seeded winner, diffusion].  'fused' applies every
    Hadamard layer as dense 32 x 32 blocks on 5 qubits, run by the k-qubit
    block kernel (fusion off); 'unfused' as n one-qubit gates (the sweep
    engine); 'fused_split' passes the same 32 x 32 blocks with fusion on, so
    the engine applies them as their one-qubit factors through the sweeps.
    Each run is checked against the closed form at the winner and at sampled
    indices."""
    res = {"workload": f"Grover, {n} qubits, {iterations} iterations, seed {seed}"}
    runs = [("c128_fused", "c128", True, False), ("c128_unfused", "c128", False, True),
            ("c128_fused_split", "c128", True, True), ("c64_fused", "c64", True, False),
            ("c64_unfused", "c64", False, True)]
    for name, dtype, fused, fusion in runs:
        if only and name != only:
            continue
        w, opl = circuits.grover_operations(n, iterations, seed=seed, fused=fused)
        amp = 16 if dtype == "c128" else 8
        r = Register.new([False] * n, dtype).set_fusion(fusion)
        dt, ms, cnt, st = timed_run(r, opl, 1, 0)
        aw, ao = circuits.grover_closed_form(n, iterations, w)
        rng = np.random.default_rng(1)
        idx = np.concatenate([[w], rng.integers(0, 1 << n, 32)]).astype(np.uint64)
        err = float(np.max(np.abs(r.amplitudes(idx) - np.where(idx == w, aw, ao))))
        entry = {"dtype": dtype, "ops": len(opl), "seconds": dt, "ops_per_s": len(opl) / dt,
                 "max_abs_err_vs_closed_form": err, "sweeps": st["sweeps"],
                 "kernel_ms": dict(zip(CLASSES, ms)), "launches": dict(zip(CLASSES, cnt))}
        if cnt[3]:
            avg = ms[3] / 1e3 / cnt[3]
            byt = 2.0 * (1 << n) * amp
            flops = (1 << n) * 32 * 8.0  # 32 complex MACs (8 real flops) per amplitude
            entry["gatek_roofline"] = {
                "kernel": "k_gatek<5> (dense 32 x 32 block on 5 qubits)",
                "avg_ms": avg * 1e3, "achieved": byt / avg / 1e9, "unit": "GB/s",
                "frac": byt / avg / HBM_PEAK, "tflops": flops / avg / 1e12,
                "traffic": (recorded_traffic("k_gatek5_grover32.md")
                            if dtype == "c128" and n == 32 else None)}
        res[name] = entry
        log(f"grover {name}: {len(opl) / dt:.1f} ops/s, err {err:.2e}")
        r.close()
        del r
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--layers", type=int, default=20)
    ap.add_argument("--triples", type=int, default=8)
    ap.add_argument("--seed", type=int, default=2025)
    ap.add_argument("--no-c64", action="store_true")
    ap.add_argument("--no-unfused", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-grover32", action="store_true",
                    help="skip the 32-qubit Grover workload (64 GiB complex128, 32 GiB "
                         "complex64): 5-qubit fused blocks vs one gate per qubit")
    ap.add_argument("--no-measure", action="store_true")
    ap.add_argument("--no-block5", action="store_true",
                    help="skip the 32-qubit complex 5-qubit block kernels")
    ap.add_argument("--cpu-threads", type=int, default=len(os.sched_getaffinity(0)),
                    help="threads of the multi-core CPU baseline (default: every core this "
                         "process may run on)")
    ap.add_argument("--profile", nargs="?", const="c128", default=None,
                    choices=["c128", "c64", "unfused", "measure", "adjoint", "grover_fused",
                             "grover_unfused", "grover_c64", "grover_c64_unfused",
                             "block5_c64", "block5_c128", "sweep_ops"],
                    help="short run for rocprofv3: a 1 GiB torch copy (counter "
                         "calibration), then one step of one section (c128, c64: "
                         "after one warmup step)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the state-vector path is built and measured on a "
              "single MI355X (multi-GPU qubit sharding is design only); use --gpus 1",
              file=sys.stderr)
        return 2
    sections = {"c128": True, "c64": not args.no_c64, "unfused": not args.no_unfused,
                "measure": not args.no_measure, "adjoint": not args.no_adjoint,
                "cpu": not args.no_cpu, "grover32": not args.no_grover32,
                "block5": not args.no_block5}
    if args.profile:
        # c128 / c64: one warmup step first, whose sweeps queue the compiles of
        # the specialised sweep kernels that the profiled step then runs
        args.steps, args.warmup = 1, (1 if args.profile in ("c128", "c64") else 0)
        sections = {k: False for k in sections}
        if args.profile.startswith("grover"):
            sections["grover32"] = True
        elif args.profile.startswith("block5"):
            sections["block5"] = True
        elif args.profile == "sweep_ops":
            pass
        else:
            sections[args.profile] = True
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("profile: calibration copy of 1 GiB done")
        del src, dst

    n = args.qubits
    w = circuits.random_circuit(n, args.layers, seed=args.seed, triples=args.triples,
                                measure=False)
    G = len(w.ops)
    log(f"workload: {n} qubits, {args.layers} layers, {G} gates, seed {args.seed}")
    out = {
        "metric": "30-qubit random-circuit gate application rate (c128, HBM-resident)",
        "value": None, "unit": "gates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c128", "data": "synthetic",
        "config": {"workload": f"random circuit, {n} qubits, {args.layers} layers of Haar 2x2 "
                               f"on every qubit + CNOT ladder + {args.triples} Toffoli/CCZ, "
                               f"seed {args.seed}",
                   "qubits": n, "layers": args.layers, "gates_per_step": G,
                   "state_bytes": (1 << n) * 16, "parallelism": "single GPU"},
    }

    # ---- complex128, fused (the drop-in default) ----
    reg = None
    if sections["c128"]:
        amp = 16
        reg = Register.new([False] * n)
        dt, ms, cnt, st = timed_run(reg, w.ops, args.steps, args.warmup)
        gates_per_s = G * args.steps / dt
        sweep_avg_s = ms[0] / 1e3 / max(cnt[0], 1)
        sweep_bytes = 2.0 * (1 << n) * amp
        achieved = sweep_bytes / sweep_avg_s if cnt[0] else 0.0
        alg_bytes = sum(2.0 * c * amp for c in w.changed) * args.steps
        norm = reg.norm_sqr()
        log(f"c128: {gates_per_s:.1f} gates/s, {cnt[0]} sweeps avg {sweep_avg_s * 1e3:.3f} ms "
            f"({achieved / 1e9:.0f} GB/s), norm {norm:.15f}")
        out["value"] = gates_per_s
        out["ms_per_step"] = dt / args.steps * 1e3
        out["roofline"] = {
            "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": achieved / HBM_PEAK,
            "traffic": (recorded_traffic("k_sweep_c128_30q.md") if n == 30 else None),
            "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/k_sweep_c128_30q.md",
            "kernel": "qs_sweep_jit / k_sweep (fused sweep, complex128; kernels specialised "
                      "to each sweep's op-code sequence)", "launches": cnt[0],
            "specialised_launches": st["specialised_launches"],
            "avg_ms": sweep_avg_s * 1e3, "bytes_per_launch": sweep_bytes}
        out["sweeps_per_step"] = st["sweeps"] / args.steps
        out["gates_per_sweep"] = G * args.steps / max(st["sweeps"], 1)
        out["ops_per_sweep"] = st["sweep_ops"] / max(st["sweeps"], 1)
        out["merged_per_step"] = st["merged"] / args.steps
        out["kernel_time_fraction"] = sum(ms) / 1e3 / dt
        out["effective_gb_per_s"] = alg_bytes / dt / 1e9
        out["effective_note"] = ("sum over gates of 2*|changed set|*16 B per wall second; fusion "
                                 "applies many gates per HBM pass, so this exceeds the HBM peak")
        out["norm_sqr"] = norm

    # ---- complex64 ----
    if sections["c64"]:
        r64 = Register.new([False] * n, "c64")
        dt64, ms64, cnt64, st64 = timed_run(r64, w.ops, args.steps, args.warmup)
        avg64 = ms64[0] / 1e3 / max(cnt64[0], 1)
        ach64 = 2.0 * (1 << n) * 8 / avg64 if cnt64[0] else 0.0
        g64 = G * args.steps / dt64
        entry = {"gates_per_s": g64,
                 "roofline": {"kernel": "qs_sweep_jit / k_sweep (fused sweep, complex64; "
                                        "specialised kernels)",
                              "specialised_launches": st64["specialised_launches"],
                              "achieved": ach64 / 1e9, "unit": "GB/s",
                              "frac": ach64 / HBM_PEAK, "avg_ms": avg64 * 1e3,
                              "launches": cnt64[0],
                              "traffic": (recorded_traffic("k_sweep_c64_30q.md")
                                          if n == 30 else None)},
                 "sweeps_per_step": st64["sweeps"] / args.steps,
                 "ops_per_sweep": st64["sweep_ops"] / max(st64["sweeps"], 1)}
        if reg is not None:
            entry["ratio_vs_c128"] = g64 / out["value"]
            entry["infidelity_vs_c128"] = r64.infidelity(reg)
        out["c64"] = entry
        log(f"c64: {g64:.1f} gates/s, infidelity {entry.get('infidelity_vs_c128')}")
        r64.close()
        del r64

    # ---- unfused: one kernel per gate, first layer ----
    if sections["unfused"]:
        amp = 16
        layer = w.ops[: n + (n - 1) + args.triples]
        ru = reg if reg is not None else Register.new([False] * n)
        ru.set_fusion(False)
        dtu, msu, cntu, _ = timed_run(ru, layer, 1, 0)
        g1 = msu[1] / 1e3 / max(cntu[1], 1)
        # dense 1q kernels in this layer are the Haar gates: 2 * 2^n * 16 B each;
        # CNOT / Toffoli (controlled) also run as 1q kernels on the controlled half.
        haar_bytes = 2.0 * (1 << n) * amp
        ctl = [op for op in layer if op._label in ("cnot", "cnx")]
        ctl_bytes = sum(2.0 * circuits._changed(op, n) * amp for op in ctl)
        g1_bytes = n * haar_bytes + ctl_bytes
        ach_u = g1_bytes / (msu[1] / 1e3) if cntu[1] else 0.0
        out["unfused"] = {"gates_per_s": len(layer) / dtu,
                          "roofline": {"kernel": "k_gate1 (one dense qubit, any controls)",
                                       "achieved": ach_u / 1e9, "unit": "GB/s",
                                       "frac": ach_u / HBM_PEAK, "avg_ms": g1 * 1e3,
                                       "launches": cntu[1]},
                          "kernel_ms": dict(zip(CLASSES, msu))}
        log(f"unfused: {len(layer) / dtu:.1f} gates/s, gate1 {ach_u / 1e9:.0f} GB/s")
        ru.set_fusion(True)
        if ru is not reg:
            ru.close()
    # ---- measurement: probability split + collapse, full readout ----
    if sections["measure"]:
        rm = reg
        if rm is None:
            rm = Register.new([False] * n)
            rm.apply_all(qops.hadamard(0))
        out["measure"] = measure_bench(rm, n)
        log(f"measure: {out['measure']['measure']['ms']:.2f} ms per measure, "
            f"{out['measure']['measure']['achieved']:.0f} GB/s")
        if rm is not reg:
            rm.close()
    if reg is not None:
        reg.close()
        del reg

    if sections["adjoint"]:
        out["adjoint"] = adjoint_bench(20, 10, 64, 4, "c128")
        out["adjoint_c64"] = adjoint_bench(20, 10, 64, 2, "c64")
        log(f"adjoint: {out['adjoint']['grads_per_s']:.0f} grads/s, "
            f"{out['adjoint']['speedup_vs_fd']:.0f}x vs FD")

    if sections["grover32"]:
        only = None
        if args.profile and args.profile.startswith("grover"):
            only = {"grover_fused": "c128_fused", "grover_unfused": "c128_unfused",
                    "grover_c64": "c64_fused", "grover_c64_unfused": "c64_unfused"}[args.profile]
        out["grover32"] = grover_bench(32, 12, seed=4, only=only)

    if args.profile == "sweep_ops":
        out["sweep_ops"] = sweep_ops_bench(n)

    if sections["block5"]:
        only = args.profile.split("_")[1] if args.profile and args.profile.startswith("block5") \
            else None
        out["block5"] = block5_bench(32, only=only)

    if sections["cpu"]:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_threads)

    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

/*
 * quaru_hip.h -- C ABI of the MI355X (gfx950) state-vector engine.
 *
 * The engine is a drop-in for quaru's gate-application path: the Rust
 * `Register` (src/register.rs:22-407) taking `&Operation`
 * (src/operation.rs:49-70).  Every entry point below names the reference
 * function it replaces.  The state vector lives in HBM on one GPU for the
 * register's whole life; per call only the small gate matrix and the targets
 * cross the bus.
 *
 * Conventions
 *  - Amplitudes are complex numbers stored {re, im} interleaved.  Host
 *    buffers are always complex128 (double pairs); a complex64 register
 *    converts on upload and download.
 *  - Qubit q is bit q of the basis index (src/register.rs:41-50).
 *  - A gate matrix is row-major, rows x cols complex128, and matrix-index bit
 *    j addresses qubit targets[j] (src/register.rs:165-176).
 *  - Every function returns a qs_status.  The OperationError variants of the
 *    reference (src/register.rs:10-20) keep their meaning and their order of
 *    precedence; their payloads, and a message for the other codes, are read
 *    with qs_last_error() on the same thread.
 *  - Calls on one register are not thread-safe (the reference takes
 *    `&mut self`).  Different registers may be used from different threads.
 *  - Gate calls are asynchronous: a HIP failure inside a kernel may be
 *    reported by a later call on the same register.  Every call that returns
 *    data to the host (measure, get_state, norm, ...) synchronises.
 */
#ifndef QUARU_HIP_H
#define QUARU_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qs_register qs_register;

/* Storage precision of a register. */
enum qs_dtype {
    QS_C128 = 0, /* complex128, the reference's Complex<f64> (src/math.rs:84-86) */
    QS_C64 = 1   /* complex64 storage; norms, probabilities and CDFs in FP64 */
};

enum qs_status {
    QS_OK = 0,
    QS_INVALID_TARGET = 1,     /* OperationError::InvalidTarget(a) */
    QS_INVALID_DIMENSIONS = 2, /* OperationError::InvalidDimensions(a, b) */
    QS_INVALID_ARITY = 3,      /* OperationError::InvalidArity(a) */
    QS_INVALID_QUBIT = 4,      /* OperationError::InvalidQubit(a) */
    QS_ERR_ARGUMENT = 16,      /* null pointer, bad dtype, size mismatch */
    QS_ERR_OUT_OF_MEMORY = 17, /* the device allocation failed */
    QS_ERR_HIP = 18,           /* any other HIP runtime failure */
    QS_ERR_NO_DEVICE = 19      /* no usable gfx950 device */
};

/* Payloads (a, b) and a message of the last failing call on this thread. */
int qs_last_error(size_t *a, size_t *b, const char **message);

/* Library version string. */
const char *qs_version(void);

/* ---- construction (src/register.rs:41-102) ------------------------------ */

/* Register::new (src/register.rs:41-56): bits[i] != 0 sets qubit i. */
int qs_register_new(const unsigned char *bits, size_t n, int dtype, qs_register **out);

/* Register::from_int (src/register.rs:58-63): the basis state |k> on n qubits. */
int qs_register_from_int(size_t n, size_t k, int dtype, qs_register **out);

/* Register::try_from_qubits (src/register.rs:81-102): `count` qubits, qubit i
 * given by lens[i] complex128 values at the running offset of `qubits`.  A
 * qubit with lens[i] != 2 or |(-1 + |a|^2) + |b|^2| >= 1e-6 gives
 * QS_INVALID_QUBIT(i) (is_qubit/is_one, :104-126). */
int qs_register_from_qubits(const double *qubits, const size_t *lens, size_t count, int dtype,
                            qs_register **out);

/* #[derive(Clone)] (src/register.rs:23): a deep copy on the same device. */
int qs_register_clone(qs_register *src, qs_register **out);

/* Drop.  Accepts NULL. */
void qs_register_free(qs_register *r);

/* Register::size (src/register.rs:404-406). */
size_t qs_register_size(const qs_register *r);

/* The register's qs_dtype. */
int qs_register_dtype(const qs_register *r);

/* ---- gates (src/register.rs:134-281) ------------------------------------ */

/* Register::try_apply_fast (src/register.rs:140-191); also serves
 * Register::try_apply (:193-250), which computes the same map.  `matrix` is
 * rows x cols complex128, `targets` has k entries.  Checks, in order:
 * rows != 2^k or cols != 2^k -> QS_INVALID_DIMENSIONS(rows, cols) (the shape
 * Operation::new guarantees, src/operation.rs:60-69); the smallest duplicate
 * target -> QS_INVALID_TARGET(dup) (:146-148, :418-429); the first target >= n
 * in targets order -> QS_INVALID_TARGET(t) (:149-153). */
int qs_apply(qs_register *r, const double *matrix, size_t rows, size_t cols,
             const size_t *targets, size_t k);

/* A structured Operation: `matrix` (2^k x 2^k) on `targets`, applied where
 * qubit controls[i] equals cvals[i] for every i (cvals == NULL: all 1).  The
 * result equals qs_apply of the dense matrix that to_controlled
 * (src/operation.rs:406-422) or cnx/cnz (:491-536) would build, without
 * materialising it; it is how cnx/cnz on 32 qubits are expressed.  Duplicate
 * or out-of-range qubits across targets + controls -> QS_INVALID_TARGET as in
 * qs_apply. */
int qs_apply_controlled(qs_register *r, const size_t *controls, const unsigned char *cvals,
                        size_t c, const double *matrix, const size_t *targets, size_t k);

/* A diagonal Operation given by its 2^k diagonal entries (complex128). */
int qs_apply_diagonal(qs_register *r, const double *diag, const size_t *targets, size_t k);

/* Register::try_apply_all (src/register.rs:267-281): the 2x2 `matrix` on
 * every qubit; arity != 1 -> QS_INVALID_ARITY(arity) (:269-271). */
int qs_apply_all(qs_register *r, const double *matrix, size_t rows, size_t cols, size_t arity);

/* ---- measurement (src/register.rs:291-342) ------------------------------ */

/* Register::try_measure (src/register.rs:297-342) with the uniform x in
 * [0, 1) supplied by the caller (the reference draws it from
 * rand::thread_rng, :317-318).  prob_0 and prob_1 are FP64 sums; *result =
 * x > prob_0; the other half is zeroed and the kept half divided by
 * sqrt(prob).  target >= n -> QS_INVALID_TARGET(target). */
int qs_measure(qs_register *r, size_t target, double uniform, int *result);

/* Device slots for the outcomes of deferred measurements. */
#define QS_MEASURE_SLOTS 64

/* qs_measure with the outcome decided on the device and kept in device slot
 * `slot` (< QS_MEASURE_SLOTS): no host round trip, nothing synchronises.
 * Same rule and arithmetic as qs_measure (x > prob_0 on the same FP64
 * sums), so the same uniform gives the same outcome.  Replaces the measure
 * loop of Register::measure_circuit (src/register.rs:368-375) when
 * Register::apply_circuit (src/register.rs:345-366) runs on the device: the
 * circuit's conditional operations never act on measured qubits
 * (src/operation.rs:302-321), so after the collapse a condition "target
 * measured v" holds exactly where qubit `target` is v, and each conditional
 * operation is applied as an operation controlled on those qubits.
 * target >= n -> QS_INVALID_TARGET(target); slot out of range ->
 * QS_ERR_ARGUMENT. */
int qs_measure_deferred(qs_register *r, size_t target, double uniform, size_t slot);

/* Synchronise and copy `count` deferred outcomes (0 or 1) from slot `first`
 * on.  No reference counterpart: the read-back of qs_measure_deferred. */
int qs_read_slots(qs_register *r, size_t first, size_t count, unsigned char *out);

/* `count` successive qs_measure calls on order[0..count) with uniforms
 * [0..count) -- the full-register readout of examples/shors.rs:219-222 and
 * examples/grovers.rs:156 -- done in one probability pass.  *index gets bit
 * order[j] = outcome j.  The state is left collapsed as the sequence of
 * measure calls leaves it. */
int qs_measure_sequence(qs_register *r, const size_t *order, size_t count,
                        const double *uniforms, uint64_t *index);

/* Inverse-CDF sampling without collapse: for each uniform u, the smallest
 * basis index x with sum_{i<=x} |psi_i|^2 > u * sum_i |psi_i|^2, with every
 * prefix sum in FP64. */
int qs_sample_cdf(qs_register *r, const double *uniforms, size_t count, uint64_t *indices);

/* ---- state access and reductions ---------------------------------------- */

/* The public `state` field (src/register.rs:35), read and written by the
 * examples' tests: 2^n complex128 values. */
int qs_get_state(qs_register *r, double *out);
int qs_set_state(qs_register *r, const double *in);

/* `state[indices[i]]` for i < count as complex128 into out[2i], out[2i+1]:
 * the element reads of the `state` field (examples/grovers.rs and the
 * examples' tests index it) without copying 2^n amplitudes to the host.
 * An index >= 2^n -> QS_ERR_ARGUMENT. */
int qs_get_amplitudes(qs_register *r, const uint64_t *indices, size_t count, double *out);

/* qs_get_amplitudes with no allocation of any kind (no reference
 * counterpart): for callers that cannot allocate, such as tight serving
 * loops that read a few amplitudes per step.  Indices and amplitudes are
 * staged through a fixed 1.5 MiB device array per GPU, held until the
 * process ends.  At most 65536 indices per call (more -> QS_ERR_ARGUMENT);
 * the same results as qs_get_amplitudes.  Calls from one process on one GPU
 * are serialised by a lock, since they share that GPU's array (which keeps
 * the last call's contents).  Gates still queued by earlier qs_apply calls
 * run first, as in every readout; call qs_sync beforehand to keep them (and
 * the state-sized scratch a dense gate on more than 5 qubits allocates on
 * first use) out of this call.  Synchronises. */
int qs_get_amplitudes_noalloc(qs_register *r, const uint64_t *indices, size_t count,
                              double *out);

/* sum_i |psi_i|^2 in FP64. */
int qs_norm_sqr(qs_register *r, double *out);

/* |psi_i|^2 for all i, in FP64 (print_probabilities, src/register.rs:379-384). */
int qs_probabilities(qs_register *r, double *out);

/* <a|b> = sum_i conj(a_i) b_i in FP64; a and b must have the same size. */
int qs_inner(qs_register *a, qs_register *b, double *re, double *im);

/* impl PartialEq (src/register.rs:409-413): *equal = every |a_i - b_i| < tol
 * (the reference uses tol = 1e-8). */
int qs_equal(qs_register *a, qs_register *b, double tol, int *equal);

/* Wait until every queued gate of the register has finished. */
int qs_sync(qs_register *r);

/* Gate fusion: 1 (default) lets the register defer gates and apply runs of
 * them in shared passes over the state (results equal to applying them one
 * by one, within rounding); 0 launches one kernel per gate. */
int qs_set_fusion(qs_register *r, int enabled);

/* Kernel launch statistics since the last reset: launches of the fused
 * sweep kernel, of the per-gate kernels, and the count of logical gates. */
int qs_stats(qs_register *r, uint64_t *sweeps, uint64_t *gate_kernels, uint64_t *gates,
             int reset);

/* Sweep op statistics since the last reset: op records executed by the
 * fused sweep kernel, and gates merged into a neighbouring one-qubit op of
 * their sweep (a CNOT / Toffoli target, controlled 2x2 or CZ-type phase next
 * to a 2x2 on the same qubit becomes one multiplexed 2x2).  No reference
 * counterpart: engine diagnostics. */
int qs_stats_sweep_ops(qs_register *r, uint64_t *ops, uint64_t *merged, int reset);

/* Kernel classes for qs_timing_read. */
enum qs_kernel_class {
    QS_K_SWEEP = 0,  /* fused sweep (a run of gates in one pass over the state) */
    QS_K_GATE1 = 1,  /* one dense active qubit (with any controls) */
    QS_K_GATE2 = 2,  /* two dense active qubits */
    QS_K_GATEK = 3,  /* 3..5 dense active qubits, LDS-staged block */
    QS_K_BIG = 4,    /* more than 5 dense active qubits */
    QS_K_DIAG = 5,   /* diagonal table */
    QS_K_SPARSE = 6, /* phases on a few basis patterns */
    QS_K_PERM = 7,   /* permutation with factors on more than 5 qubits (gather) */
    QS_K_COUNT = 8
};

/* Bracket every gate kernel launch of the register with HIP events on the
 * register's own stream (enable = 1), or stop (0). */
int qs_timing_enable(qs_register *r, int enable);

/* Summed kernel time in ms and launch counts per qs_kernel_class since the
 * last reset (arrays of QS_K_COUNT entries).  Synchronises. */
int qs_timing_read(qs_register *r, double *ms, uint64_t *launches, int reset);

/* ---- adjoint differentiation (no reference counterpart) ----------------- */

/* A parameterised circuit for <psi(theta)|H|psi(theta)> with H a sum of
 * Pauli-Z strings.  Rotations use the OpenQASM U conventions of
 * src/openqasm.rs:145-163: RY(t) = u(t, 0, 0), RZ(t) = u(0, 0, t). */
typedef struct qs_pcircuit qs_pcircuit;
typedef struct qs_adjoint qs_adjoint;

enum qs_axis { QS_AXIS_X = 0, QS_AXIS_Y = 1, QS_AXIS_Z = 2 };

int qs_pcircuit_new(size_t n, qs_pcircuit **out);
void qs_pcircuit_free(qs_pcircuit *c);
/* A fixed gate, as qs_apply (dense matrix, k <= 2). */
int qs_pcircuit_gate(qs_pcircuit *c, const double *matrix, const size_t *targets, size_t k);
/* A rotation exp(-i theta_p sigma_axis / 2) on `target`; theta_p is column
 * `param` of the parameter matrix. */
int qs_pcircuit_rotation(qs_pcircuit *c, int axis, size_t target, size_t param);
/* A layer boundary (checkpoints are taken every `checkpoint_every` layers). */
int qs_pcircuit_layer(qs_pcircuit *c);
/* Number of parameters (max param index + 1). */
size_t qs_pcircuit_nparams(const qs_pcircuit *c);

/* Forward pass for `batch` parameter sets: thetas is batch x nparams (row
 * major).  energies[b] = <psi_b|H|psi_b>, H = sum_t coeffs[t] Z^{masks[t]}.
 * The final states (and, when checkpoint_every > 0, one state every
 * checkpoint_every layers) stay on the device in *ctx for the backward. */
int qs_adjoint_forward(const qs_pcircuit *c, const double *thetas, size_t batch,
                       const uint64_t *masks, const double *coeffs, size_t nterms, int dtype,
                       size_t checkpoint_every, qs_adjoint **ctx, double *energies);

/* One reverse sweep: grads (batch x nparams) = dE_b / dtheta_{b,p}. */
int qs_adjoint_backward(qs_adjoint *ctx, double *grads);

/* Sweeps and per-gate kernels the fused forward pass of *ctx launched (no
 * reference counterpart; diagnostics). */
int qs_adjoint_stats(const qs_adjoint *ctx, uint64_t *sweeps, uint64_t *kernels);

/* Sweep kernels specialised to one op-code sequence, compiled at run time
 * with hipRTC (no reference counterpart; engine tuning).  mode 0: off; 1
 * (default, or QS_SWEEP_JIT): for states of 22 qubits and more, a new
 * sequence runs the generic kernel while its specialised kernel compiles in
 * a helper process (bin/qs_sweep_compile beside the library, or
 * QS_SWEEP_COMPILER); 2: every sequence, compiled before its first launch.
 * mode < 0 only queries.  Returns the previous mode. */
int qs_sweep_jit_mode(int mode);

/* Blocks until every queued compile has finished. */
int qs_sweep_jit_wait(void);

/* Kernels compiled, compiles failed (those sequences stay on the generic
 * kernel), and launches of compiled kernels, since the process started. */
int qs_sweep_jit_stats(uint64_t *compiled, uint64_t *failed, uint64_t *launches);

/* Tiled passes over psi and lambda, and checkpoint restores, of the last
 * backward of *ctx (no reference counterpart; diagnostics). */
int qs_adjoint_reverse_stats(const qs_adjoint *ctx, uint64_t *blocks, uint64_t *restores);

void qs_adjoint_free(qs_adjoint *ctx);

#ifdef __cplusplus
}
#endif
#endif /* QUARU_HIP_H */

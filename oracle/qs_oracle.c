/*
 * qs_oracle.c -- CPU FP64 restatement of quaru's state-vector hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file: it is the checker that tests/, the smoke check and the
 * cpu_baseline leg of bench.py compare the HIP engine against.
 *
 * Every function restates one function of the reference (quaru 0.2.0, Rust)
 * and cites the file:line it follows.  Complex numbers are {re, im} pairs of
 * doubles, interleaved, in basis-index order (src/register.rs:35,
 * src/math.rs:86).  Qubit q is bit q of the basis index (src/register.rs:41-50).
 * Compile with -ffp-contract=off so that a*b+c is never fused: Rust does not
 * contract either, so the per-operation rounding matches num-complex.
 *
 * Pinning: every exported function and the test in tests/test_oracle.py that
 * a dropped term or a swapped index fails.
 *   orc_operation_valid: test_validation_precedence (2^k x 2^k shapes only).
 *   orc_new / orc_from_int: test_new_is_basis_state, test_from_int (basis
 *       index closed forms).
 *   orc_try_from_qubits: test_from_qubits_kron_order (the kron ordering of
 *       tests/register_operation_tests.rs:25-40), test_from_qubits_invalid
 *       (the 1e-6 tolerance and the InvalidQubit(i) payload).
 *   orc_try_apply_fast: test_apply_fast_matches_dense_and_tensordot (against
 *       orc_try_apply, a different algorithm, and a numpy tensordot),
 *       test_transposed_matrix_is_caught, test_golden_random_12q and
 *       test_golden_toffoli_mini (numpy goldens), test_grover_closed_form,
 *       test_qft_add_identity, test_u_gate_modmul_exhaustive
 *       (examples/shors.rs:434-503), test_toffoli_cnx.
 *   orc_try_apply: test_apply_fast_matches_dense_and_tensordot,
 *       test_validation_precedence.
 *   orc_try_apply_all: test_apply_all_equals_each_qubit,
 *       test_apply_all_random_unitary_and_arity (tests/...:211-219).
 *   orc_try_measure: test_measure_basis_states_round_trip (tests/...:53-96),
 *       test_measure_rule_and_collapse (x > prob_0, strict, and the
 *       normalisation), test_bell_pairs_measure_equal,
 *       test_teleportation_outcome_independent, test_superdense_coding.
 *   orc_measure_sequence: test_golden_random_12q (sampled index 3733 and the
 *       collapsed state of the numpy golden).
 *   orc_apply_fast_omp (qs_oracle_omp.c): test_omp_variant_matches_scalar
 *       (bit-identical to orc_try_apply_fast).
 *   orc_expect_zstrings (qs_oracle_omp.c): test_expect_zstrings (numpy sum of
 *       |psi|^2 times the Z-string eigenvalues).
 *
 * Parity unpinned (nothing available here can fix them):
 *   - the uniform stream of measure: the reference draws x from the
 *     OS-seeded rand::thread_rng (src/register.rs:317-318), so no run of it
 *     is reproducible; the oracle and the engine take x as an argument and
 *     share only the rule res = x > prob_0.
 *   - bitwise rounding of ndarray's unrolled_dot for rows of length >= 8
 *     (see unrolled_dot below): restated from its published source, not run.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "qs_oracle.h"

typedef struct { double re, im; } cplx;

/* num-complex 0.4 Mul: (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re). */
static inline cplx cmul(cplx a, cplx b) {
    cplx r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}
static inline cplx cadd(cplx a, cplx b) {
    cplx r = {a.re + b.re, a.im + b.im};
    return r;
}
static inline cplx cld(const double *p, size_t i) {
    cplx r = {p[2 * i], p[2 * i + 1]};
    return r;
}
static inline void cst(double *p, size_t i, cplx v) {
    p[2 * i] = v.re;
    p[2 * i + 1] = v.im;
}

/*
 * ndarray 0.15.6 `numeric_util::unrolled_dot` (third party, not vendored; the
 * order below restates its published source): eight partial sums p0..p7 over
 * whole blocks of eight, then sum = ((((0 + (p0+p4)) + (p1+p5)) + (p2+p6)) +
 * (p3+p7)), then the tail added in order.  Reached from `Array2::dot(&Array1)`
 * (src/register.rs:183) through general_mat_vec_mul -> row.dot(x).  For length
 * below 8 this is the plain sequential sum.  Bitwise agreement with the Rust
 * build is unpinned (no toolchain here); agreement to 1e-15 relative is what
 * every test relies on.
 */
static cplx unrolled_dot(const double *row, const double *x, size_t len) {
    cplx z = {0.0, 0.0};
    cplx p[8];
    for (int i = 0; i < 8; ++i) p[i] = z;
    size_t i = 0;
    for (; i + 8 <= len; i += 8)
        for (int j = 0; j < 8; ++j) p[j] = cadd(p[j], cmul(cld(row, i + j), cld(x, i + j)));
    cplx sum = z;
    sum = cadd(sum, cadd(p[0], p[4]));
    sum = cadd(sum, cadd(p[1], p[5]));
    sum = cadd(sum, cadd(p[2], p[6]));
    sum = cadd(sum, cadd(p[3], p[7]));
    for (; i < len; ++i) sum = cadd(sum, cmul(cld(row, i), cld(x, i)));
    return sum;
}

/* src/register.rs:418-429: sort a copy, return the first value equal to its
 * predecessor. */
static int cmp_size(const void *a, const void *b) {
    size_t x = *(const size_t *)a, y = *(const size_t *)b;
    return (x > y) - (x < y);
}
static int get_duplicate(const size_t *v, size_t k, size_t *dup) {
    if (k < 2) return 0;
    size_t *c = (size_t *)malloc(k * sizeof(size_t));
    memcpy(c, v, k * sizeof(size_t));
    qsort(c, k, sizeof(size_t), cmp_size);
    int found = 0;
    for (size_t i = 1; i < k; ++i)
        if (c[i] == c[i - 1]) {
            *dup = c[i];
            found = 1;
            break;
        }
    free(c);
    return found;
}

/* src/operation.rs:60-69: Operation::new accepts only a 2^k x 2^k matrix. */
int orc_operation_valid(size_t rows, size_t cols, size_t k) {
    if (k >= 8 * sizeof(size_t)) return 0;
    size_t d = (size_t)1 << k;
    return rows == d && cols == d;
}

/*
 * Validation shared by try_apply_fast and try_apply (src/register.rs:141-153
 * and :195-207), in the reference's order:
 *   1. (rows, cols) == (k, k) -> InvalidDimensions(rows, cols).  This compares
 *      against the target count, not 2^k, so it never fires for a matrix that
 *      Operation::new accepted; restated as written.
 *   2. the smallest duplicated target -> InvalidTarget(dup).
 *   3. the first target (in targets order) >= n -> InvalidTarget(t).
 */
static int validate(size_t n, size_t rows, size_t cols, const size_t *t, size_t k, size_t *ea,
                    size_t *eb) {
    if (rows == k && cols == k) {
        *ea = rows;
        *eb = cols;
        return ORC_INVALID_DIMENSIONS;
    }
    size_t dup;
    if (get_duplicate(t, k, &dup)) {
        *ea = dup;
        return ORC_INVALID_TARGET;
    }
    for (size_t j = 0; j < k; ++j)
        if (t[j] >= n) {
            *ea = t[j];
            return ORC_INVALID_TARGET;
        }
    return ORC_OK;
}

/* Register::new (src/register.rs:41-56): fold kron(qubit(b), acc) over the
 * input bits, starting from [[1]]; to_qbit_vector is src/math.rs:13-18.  Bit i
 * of the input becomes bit i of the basis index. */
void orc_new(const unsigned char *bits, size_t n, double *state) {
    size_t len = 1;
    state[0] = 1.0;
    state[1] = 0.0;
    double *tmp = (double *)malloc(sizeof(double) * 2 * ((size_t)1 << n));
    for (size_t q = 0; q < n; ++q) {
        cplx b[2];
        b[0].re = bits[q] ? 0.0 : 1.0;
        b[0].im = 0.0;
        b[1].re = bits[q] ? 1.0 : 0.0;
        b[1].im = 0.0;
        /* kron(b (2x1), a (len x 1))[r*len + i] = b[r] * a[i] */
        for (size_t r = 0; r < 2; ++r)
            for (size_t i = 0; i < len; ++i) cst(tmp, r * len + i, cmul(b[r], cld(state, i)));
        len *= 2;
        memcpy(state, tmp, sizeof(double) * 2 * len);
    }
    free(tmp);
}

/* Register::from_int (src/register.rs:58-63) -> math::int_to_state
 * (src/math.rs:44-49): the basis vector e_k of length 2^n. */
void orc_from_int(size_t n, size_t k, double *state) {
    size_t len = (size_t)1 << n;
    memset(state, 0, sizeof(double) * 2 * len);
    state[2 * k] = 1.0;
}

/* Register::try_from_qubits (src/register.rs:81-102) with is_qubit/is_one
 * (:104-126): every qubit must have 2 entries with |(-1 + |a|^2) + |b|^2| <
 * 1e-6, else InvalidQubit(first bad index); then the same kron fold as new.
 * `lens[i]` is the element count of qubit i (the reference accepts any shape
 * and rejects len != 2). */
int orc_try_from_qubits(const double *qubits, const size_t *lens, size_t n, double *state,
                        size_t *ea) {
    size_t off = 0;
    for (size_t i = 0; i < n; ++i) {
        int ok = 0;
        if (lens[i] == 2) {
            double total = -1.0;
            for (size_t j = 0; j < 2; ++j) {
                cplx v = cld(qubits, off + j);
                total += v.re * v.re + v.im * v.im; /* num-complex norm_sqr */
            }
            ok = fabs(total) < 0.000001;
        }
        if (!ok) {
            *ea = i;
            return ORC_INVALID_QUBIT;
        }
        off += lens[i];
    }
    size_t len = 1;
    state[0] = 1.0;
    state[1] = 0.0;
    double *tmp = (double *)malloc(sizeof(double) * 2 * ((size_t)1 << n));
    for (size_t q = 0; q < n; ++q) {
        for (size_t r = 0; r < 2; ++r)
            for (size_t i = 0; i < len; ++i)
                cst(tmp, r * len + i, cmul(cld(qubits, 2 * q + r), cld(state, i)));
        len *= 2;
        memcpy(state, tmp, sizeof(double) * 2 * len);
    }
    free(tmp);
    return ORC_OK;
}

/*
 * Register::try_apply_fast (src/register.rs:140-191), the hot path.
 * For every base index b in ascending order whose target bits are all zero
 * (:156-157): p_i = b | sum_{j: bit j of i} 2^targets[j] for i < 2^k
 * (:163-176), gather a_i = psi[p_i] (:178-181), r = M . a with the
 * unrolled_dot order per row (:183), scatter psi'[p_i] = r_i into a copy of
 * the old state (:155, :185-189).
 */
int orc_try_apply_fast(double *state, size_t n, const double *m, size_t rows, size_t cols,
                       const size_t *targets, size_t k, size_t *ea, size_t *eb) {
    int rc = validate(n, rows, cols, targets, k, ea, eb);
    if (rc) return rc;
    size_t len = (size_t)1 << n, dim = (size_t)1 << k;
    double *next = (double *)malloc(sizeof(double) * 2 * len);
    memcpy(next, state, sizeof(double) * 2 * len);
    size_t *perm = (size_t *)malloc(sizeof(size_t) * dim);
    double *arr = (double *)malloc(sizeof(double) * 2 * dim);
    size_t tmask = 0;
    for (size_t j = 0; j < k; ++j) tmask |= (size_t)1 << targets[j];
    for (size_t index = 0; index < len; ++index) {
        if (index & tmask) continue;
        for (size_t i = 0; i < dim; ++i) {
            size_t p = index;
            for (size_t j = 0; j < k; ++j)
                if ((i >> j) & 1) p |= (size_t)1 << targets[j];
            perm[i] = p;
        }
        for (size_t i = 0; i < dim; ++i) cst(arr, i, cld(state, perm[i]));
        for (size_t i = 0; i < dim; ++i) cst(next, perm[i], unrolled_dot(m + 2 * i * dim, arr, dim));
    }
    memcpy(state, next, sizeof(double) * 2 * len);
    free(arr);
    free(perm);
    free(next);
    return ORC_OK;
}

/*
 * Register::try_apply (src/register.rs:193-250), the slow dense path, kept as
 * an independent brute-force check (n <= 10).  perm = targets followed by the
 * remaining qubits ascending (:211-216); permuted[i] = psi[j] with bit k of i
 * moved to bit perm[k] of j (:220-232); multiply by kron(eye(2^(n-k)), M)
 * (:235-238); permute back (:241-247).
 */
int orc_try_apply(double *state, size_t n, const double *m, size_t rows, size_t cols,
                  const size_t *targets, size_t k, size_t *ea, size_t *eb) {
    int rc = validate(n, rows, cols, targets, k, ea, eb);
    if (rc) return rc;
    size_t len = (size_t)1 << n, dim = (size_t)1 << k;
    size_t *perm = (size_t *)malloc(sizeof(size_t) * (n + 1));
    size_t plen = 0;
    for (size_t j = 0; j < k; ++j) perm[plen++] = targets[j];
    for (size_t i = 0; i < n; ++i) {
        int in = 0;
        for (size_t j = 0; j < k; ++j) in |= targets[j] == i;
        if (!in) perm[plen++] = i;
    }
    double *ps = (double *)malloc(sizeof(double) * 2 * len);
    for (size_t i = 0; i < len; ++i) {
        size_t j = 0;
        for (size_t b = 0; b < n; ++b) j |= ((i >> b) & 1) << perm[b];
        cst(ps, i, cld(state, j));
    }
    /* kron(eye(2^(n-k)), M): entry (r, c) = eye[r/dim][c/dim] * M[r%dim][c%dim];
     * built densely so that the matvec is the literal one. */
    double *big = (double *)malloc(sizeof(double) * 2 * len * len);
    for (size_t r = 0; r < len; ++r)
        for (size_t c = 0; c < len; ++c) {
            cplx e = {(r / dim == c / dim) ? 1.0 : 0.0, 0.0};
            cst(big, r * len + c, cmul(e, cld(m, (r % dim) * dim + (c % dim))));
        }
    double *out = (double *)malloc(sizeof(double) * 2 * len);
    for (size_t r = 0; r < len; ++r) cst(out, r, unrolled_dot(big + 2 * r * len, ps, len));
    for (size_t i = 0; i < len; ++i) {
        size_t j = 0;
        for (size_t b = 0; b < n; ++b) j |= ((i >> perm[b]) & 1) << b;
        cst(state, i, cld(out, j));
    }
    free(out);
    free(big);
    free(ps);
    free(perm);
    return ORC_OK;
}

/*
 * Register::try_apply_all (src/register.rs:267-281): InvalidArity(k) unless
 * k == 1; otherwise fold kron(acc, U) n times from [[1]] into a dense
 * 2^n x 2^n matrix and multiply.  Brute force: n <= 11.
 */
int orc_try_apply_all(double *state, size_t n, const double *u, size_t arity, size_t *ea) {
    if (arity != 1) {
        *ea = arity;
        return ORC_INVALID_ARITY;
    }
    size_t len = (size_t)1 << n;
    double *acc = (double *)malloc(sizeof(double) * 2 * len * len);
    double *tmp = (double *)malloc(sizeof(double) * 2 * len * len);
    acc[0] = 1.0;
    acc[1] = 0.0;
    size_t d = 1;
    for (size_t s = 0; s < n; ++s) {
        size_t nd = 2 * d;
        /* kron(A (d x d), U (2 x 2))[i*2+k][j*2+l] = A[i][j] * U[k][l] */
        for (size_t i = 0; i < d; ++i)
            for (size_t j = 0; j < d; ++j)
                for (size_t kk = 0; kk < 2; ++kk)
                    for (size_t l = 0; l < 2; ++l)
                        cst(tmp, (i * 2 + kk) * nd + (j * 2 + l),
                            cmul(cld(acc, i * d + j), cld(u, kk * 2 + l)));
        d = nd;
        double *sw = acc;
        acc = tmp;
        tmp = sw;
    }
    double *out = (double *)malloc(sizeof(double) * 2 * len);
    for (size_t r = 0; r < len; ++r) cst(out, r, unrolled_dot(acc + 2 * r * len, state, len));
    memcpy(state, out, sizeof(double) * 2 * len);
    free(out);
    free(tmp);
    free(acc);
    return ORC_OK;
}

/*
 * Register::try_measure (src/register.rs:297-342) with the uniform x supplied
 * by the caller instead of rand::thread_rng (:317-318).  prob_1 / prob_0 are
 * sequential FP64 sums of norm_sqr in index order (:303-315); res = x > prob_0
 * (strict, :321); the other half is zeroed and the kept amplitudes are divided
 * component-wise by sqrt(total_prob) (:323-339).
 */
int orc_try_measure(double *state, size_t n, size_t target, double x, int *res, size_t *ea) {
    if (target >= n) {
        *ea = target;
        return ORC_INVALID_TARGET;
    }
    size_t len = (size_t)1 << n;
    double p1 = 0.0, p0 = 0.0;
    for (size_t i = 0; i < len; ++i) {
        cplx s = cld(state, i);
        double p = s.re * s.re + s.im * s.im;
        if ((i >> target) & 1)
            p1 += p;
        else
            p0 += p;
    }
    int r = x > p0;
    double total = r ? p1 : p0;
    double sq = sqrt(total);
    for (size_t i = 0; i < len; ++i) {
        if ((int)((i >> target) & 1) != r) {
            cst(state, i, (cplx){0.0, 0.0});
        } else {
            cplx s = cld(state, i);
            s.re /= sq;
            s.im /= sq;
            cst(state, i, s);
        }
    }
    *res = r;
    return ORC_OK;
}

/* A full-register readout the way the examples do it (examples/shors.rs:
 * 219-222): measure(order[0]), measure(order[1]), ... with one uniform each.
 * Returns the sampled basis index (bit q = outcome of qubit q). */
size_t orc_measure_sequence(double *state, size_t n, const size_t *order, size_t count,
                            const double *uniforms) {
    size_t idx = 0;
    for (size_t j = 0; j < count; ++j) {
        int r = 0;
        size_t ea;
        orc_try_measure(state, n, order[j], uniforms[j], &r, &ea);
        if (r) idx |= (size_t)1 << order[j];
    }
    return idx;
}

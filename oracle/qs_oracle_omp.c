/*
 * qs_oracle_omp.c -- multi-core variant of the try_apply_fast restatement and
 * the diagonal Z-string expectation (TEST INFRASTRUCTURE ONLY: the checker and
 * the CPU baseline, never linked into the product).
 *
 * orc_apply_fast_omp keeps the reference's algorithm (src/register.rs:155-189):
 * filter the base indices whose target bits are zero, gather 2^k amplitudes,
 * multiply by M with the same unrolled_dot order as qs_oracle.c, scatter into
 * an out-of-place copy, then replace the state.  Only the loop over base
 * indices is split across threads (the reference's comment at :158 asks for
 * exactly that).  The caller passes the copy buffer so that timing loops do
 * not measure malloc of a multi-GiB array.
 * Pinned against orc_try_apply_fast (bit-identical: same per-group
 * arithmetic, disjoint groups).
 */
#include <stdint.h>
#include <string.h>

#include "qs_oracle.h"

typedef struct { double re, im; } cplx;

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

static cplx dot_row(const cplx *row, const cplx *x, size_t len) {
    cplx z = {0.0, 0.0}, p[8];
    for (int i = 0; i < 8; ++i) p[i] = z;
    size_t i = 0;
    for (; i + 8 <= len; i += 8)
        for (int j = 0; j < 8; ++j) p[j] = cadd(p[j], cmul(row[i + j], x[i + j]));
    cplx s = z;
    s = cadd(s, cadd(p[0], p[4]));
    s = cadd(s, cadd(p[1], p[5]));
    s = cadd(s, cadd(p[2], p[6]));
    s = cadd(s, cadd(p[3], p[7]));
    for (; i < len; ++i) s = cadd(s, cmul(row[i], x[i]));
    return s;
}

int orc_apply_fast_omp(double *state_d, double *scratch_d, size_t n, const double *m_d, size_t k,
                       const size_t *targets, int threads) {
    cplx *state = (cplx *)state_d, *next = (cplx *)scratch_d;
    const cplx *m = (const cplx *)m_d;
    size_t len = (size_t)1 << n, dim = (size_t)1 << k;
    if (k > 10) return -1;
    size_t tmask = 0;
    for (size_t j = 0; j < k; ++j) tmask |= (size_t)1 << targets[j];
    memcpy(next, state, sizeof(cplx) * len);
#pragma omp parallel for schedule(static) num_threads(threads)
    for (size_t index = 0; index < len; ++index) {
        if (index & tmask) continue;
        size_t perm[1024];
        cplx arr[1024];
        for (size_t i = 0; i < dim; ++i) {
            size_t p = index;
            for (size_t j = 0; j < k; ++j)
                if ((i >> j) & 1) p |= (size_t)1 << targets[j];
            perm[i] = p;
            arr[i] = state[p];
        }
        for (size_t i = 0; i < dim; ++i) next[perm[i]] = dot_row(m + i * dim, arr, dim);
    }
    memcpy(state, next, sizeof(cplx) * len);
    return 0;
}

/* <psi| H |psi> for H = sum_t c_t Z^{mask_t}: a Z string is diagonal with
 * eigenvalue (-1)^popcount(x & mask) on basis state x. */
double orc_expect_zstrings(const double *state, size_t n, const unsigned long long *masks,
                           const double *coeffs, size_t nterms) {
    size_t len = (size_t)1 << n;
    double e = 0.0;
    for (size_t x = 0; x < len; ++x) {
        double p = state[2 * x] * state[2 * x] + state[2 * x + 1] * state[2 * x + 1];
        double h = 0.0;
        for (size_t t = 0; t < nterms; ++t)
            h += (__builtin_popcountll((unsigned long long)x & masks[t]) & 1) ? -coeffs[t] : coeffs[t];
        e += p * h;
    }
    return e;
}

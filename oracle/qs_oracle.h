/* qs_oracle.h -- CPU FP64 restatement of quaru's hot path (test
 * infrastructure only; see qs_oracle.c for the reference citations). */
#ifndef QS_ORACLE_H
#define QS_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* OperationError variants, src/register.rs:10-20. */
enum {
    ORC_OK = 0,
    ORC_INVALID_TARGET = 1,
    ORC_INVALID_DIMENSIONS = 2,
    ORC_INVALID_ARITY = 3,
    ORC_INVALID_QUBIT = 4,
};

int orc_operation_valid(size_t rows, size_t cols, size_t k);
void orc_new(const unsigned char *bits, size_t n, double *state);
void orc_from_int(size_t n, size_t k, double *state);
int orc_try_from_qubits(const double *qubits, const size_t *lens, size_t n, double *state,
                        size_t *ea);
int orc_try_apply_fast(double *state, size_t n, const double *m, size_t rows, size_t cols,
                       const size_t *targets, size_t k, size_t *ea, size_t *eb);
int orc_try_apply(double *state, size_t n, const double *m, size_t rows, size_t cols,
                  const size_t *targets, size_t k, size_t *ea, size_t *eb);
int orc_try_apply_all(double *state, size_t n, const double *u, size_t arity, size_t *ea);
int orc_try_measure(double *state, size_t n, size_t target, double x, int *res, size_t *ea);
size_t orc_measure_sequence(double *state, size_t n, const size_t *order, size_t count,
                            const double *uniforms);

/* qs_oracle_omp.c: the same try_apply_fast algorithm split over OpenMP
 * threads (the multi-core CPU baseline), and the diagonal-observable
 * expectation used by the parameter-shift reference. */
int orc_apply_fast_omp(double *state, double *scratch, size_t n, const double *m, size_t k,
                       const size_t *targets, int threads);
double orc_expect_zstrings(const double *state, size_t n, const unsigned long long *masks,
                           const double *coeffs, size_t nterms);

#ifdef __cplusplus
}
#endif
#endif

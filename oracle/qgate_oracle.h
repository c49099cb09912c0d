/*
 * ORACLE — test infrastructure only. A CPU restatement of the reference qgate CPU runtime
 * (qgate/simulator/src/CPU*.cpp, BitPermTable.*, Parallel.*, GateMatrix.cpp), used by the
 * tests, the smoke check and the CPU-baseline leg of bench.py as the checker. It is never
 * part of the product path.
 *
 * The same C API is implemented twice:
 *   liboracle.so        this restatement (qgate_oracle.c),
 *   _ref/libqgate_ref.so the reference's own C++ CPU core, compiled from its sources by
 *                        oracle/Makefile (ref_driver.cpp binds it to this API).
 * prec: 0 = float64 (complex128 state), 1 = float32 (complex64 state).
 * Functions returning int return 0 on success, -1 on error (message: qgo_last_error()).
 */
#ifndef QGATE_ORACLE_H
#define QGATE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *qgo_last_error(void);
const char *qgo_kind(void);                 /* "port" or "reference" */
int qgo_num_workers(void);                  /* worker threads used above the 2^16 threshold */

void *qgo_states_new(int prec);
void qgo_states_delete(void *qs);
int qgo_states_n_lanes(const void *qs);
void *qgo_states_data(void *qs);            /* raw local amplitudes (complex, prec) */

int qgo_initialize_states(void *qs, int n_lanes);
int qgo_reset_states(void *qs);
int qgo_calc_probability(const void *qs, int lane, double *prob);
int qgo_join(void *qs, void *const *list, int n, int n_new_lanes);
int qgo_decohere(int value, double prob, void *qs, int lane);
int qgo_decohere_and_separate(int value, double prob, void *qs0, void *qs1, const void *qs, int lane);
int qgo_apply_reset(void *qs, int lane);
int qgo_apply_gate(const double *mat8, void *qs, int lane);
int qgo_apply_controlled_gate(const double *mat8, void *qs, const int *ctrl, int n_ctrl, int target);

/* FP64 gate matrix by type name, conjugate-transposed if adjoint (GateMatrix.cpp:13-158). */
int qgo_gate_matrix(const char *name, const double *args, int n_args, int adjoint, double *mat8);

int qgo_get_states(int prec, void *out, int mathop, const int *lane_tables, void *const *qs,
                   int n_qs, uint64_t empty_mask, int64_t n_states, int64_t start, int64_t step);
int qgo_prepare_prob_array(int prec, void *prob, const int *lane_tables, void *const *qs, int n_qs,
                           int n_lanes, int n_hidden);
void *qgo_pool_new(int prec, const int *lane_tables, void *const *qs, int n_qs, int n_lanes,
                   int n_hidden, const int *empty, int n_empty);
int qgo_pool_sample(void *pool, int64_t *obs, int n, const double *u);
void qgo_pool_delete(void *pool);
/* normalised cumulative probability array of a pool (n_states entries, pool precision) */
const void *qgo_pool_cdf(void *pool, int64_t *n_states);

#ifdef __cplusplus
}
#endif

#endif

/* ORACLE — test infrastructure only. C API of the CPU restatement of the
 * reference's dense state-vector path (see matrix.hpp, blocked.hpp,
 * qvd_oracle.cpp). Loaded by tests/ and bench.py's cpu_baseline leg. */
#ifndef QVD_ORACLE_H
#define QVD_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_op {
    const uint32_t *targets;
    uint32_t num_targets;
    const uint32_t *controls;
    const uint8_t *control_values; /* NULL = all true */
    uint32_t num_controls;
    const double *matrix; /* 2^k x 2^k complex128, row-major, (re, im) pairs */
} orc_op;

typedef struct orc_stats {
    uint64_t gates, kernels, passes, swaps;
    double gate_seconds;
} orc_stats;

const char *orc_last_error(void);
void orc_set_threads(int threads);
int orc_max_threads(void);
uint64_t orc_deposit(uint64_t value, uint64_t mask);
int orc_gate_matrix(const char *name, const double *params, uint32_t num_params, double *out, uint32_t capacity,
                    uint32_t *num_qubits);
int orc_gate_controls(const char *name, uint32_t *num_controls);
int orc_matrix_embed(const double *m, uint32_t k, const uint32_t *gate_qubits, const uint32_t *qubits, uint32_t n,
                     double *out);
int orc_matrix_controlled(const double *m, uint32_t k, uint32_t controls, uint64_t state, double *out);
int orc_reference_apply(double *state, uint32_t num_qubits, const double *matrix, const uint32_t *qubits, uint32_t k);
int orc_fuse(const orc_op *ops, uint64_t n, uint32_t max_qubits, uint32_t *out_count, uint32_t *out_targets,
             uint32_t *out_num_targets, uint32_t *out_num_controls, double *out_matrices, uint64_t matrix_capacity);
void *orc_sv_create(uint32_t num_qubits, int precision);
void orc_sv_destroy(void *h);
int orc_sv_set(void *h, const double *amps, uint64_t len);
int orc_sv_get(void *h, double *amps, uint64_t len);
int orc_sv_apply(void *h, const double *matrix, const uint32_t *targets, uint32_t k, const uint32_t *controls,
                 const uint8_t *values, uint32_t c);
int orc_sv_apply_ops(void *h, const orc_op *ops, uint64_t n, uint32_t max_fused_qubits, int32_t region_bits,
                     orc_stats *stats);
int orc_sv_norm_sqr(void *h, double *out);
int orc_sv_scale(void *h, double factor);
int orc_sv_probability_one(void *h, uint32_t qubit, double *out);
int orc_sv_probabilities(void *h, double *out, uint64_t len);
int orc_sv_measure(void *h, uint32_t qubit, double uniform, uint8_t *outcome);
int orc_sv_reset(void *h, uint32_t qubit, double uniform);
int orc_sv_sample(void *h, const double *uniforms, uint64_t shots, uint64_t *out);
int orc_sv_swap_physical(void *h, const uint32_t *pairs, uint32_t num_pairs);
int orc_sv_layout(void *h, uint32_t *out);
int orc_count_outcomes(const uint64_t *outcomes, uint64_t shots, const uint32_t *measure_qubits,
                       const uint32_t *measure_clbits, uint32_t num_measures, uint64_t *out_keys,
                       uint64_t *out_counts, uint64_t *out_distinct);
#ifdef __cplusplus
}
#endif
#endif

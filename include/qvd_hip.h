/*
 * qvd_hip.h — C ABI of the MI355X state-vector engine (libqvd_hip.so).
 *
 * This is the seam a Rust `qvd` crate binds through `extern "C"` in place of
 * its CPU state-vector backend. Every function returns an int status (0 = ok,
 * negative = error) and never aborts; the message of the last error on the
 * calling thread is qvd_last_error(). Calls are synchronous: when a function
 * returns, its GPU work is complete.
 *
 * Conventions (kept from the reference):
 *   - qubit q is bit q of an amplitude index (src/state.rs:12-16);
 *   - a matrix on targets [q0..q(k-1)] is row-major 2^k x 2^k complex128
 *     ((re, im) pairs), index bit j = qubit qj (src/matrix.rs:1-18);
 *   - controls are (qubit, value) pairs (src/fusion.rs:20-21);
 *   - amplitudes cross the ABI as complex128 in logical index order
 *     (src/state.rs:146-159; python/src/lib.rs:703-716).
 *
 * Ownership: the library owns device memory; the caller owns every host
 * array and keeps it alive for the duration of the call. A qvd_sv handle is
 * not safe for concurrent use (the reference's `&mut self`,
 * src/state.rs:181); distinct handles may be used from distinct threads.
 */
#ifndef QVD_HIP_H
#define QVD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define QVD_OK 0
#define QVD_EINVAL (-1)        /* invalid argument (the reference panics here) */
#define QVD_ENOMEM (-2)        /* device or host allocation failed */
#define QVD_EHIP (-3)          /* a HIP runtime call failed */
#define QVD_EUNSUPPORTED (-4)  /* gate wider than the kernels handle */

/* Amplitude precision: the reference's `Real` parameter (src/simd.rs:15-23). */
#define QVD_C64 0  /* complex64 ("single") */
#define QVD_C128 1 /* complex128 ("double") */

/* Planner choice for qvd_sv_apply_ops. */
#define QVD_PLANNER_HEURISTIC 0 /* restated fusion.rs + blocking.rs stage planning */
/* MLP cost model over the candidate plans that run the same arithmetic as
 * the heuristic plan (equal plan images, see qvd_plan_equivalent_variants):
 * the amplitudes are bit-identical to the heuristic's whichever it picks.
 * Falls back to the heuristic when no model is installed. */
#define QVD_PLANNER_LEARNED 1
/* MLP cost model over every plan variant; variants that order ops
 * differently round differently (complex64 results agree to about 1e-6
 * relative L2, not bit for bit). Heuristic fallback without a model. */
#define QVD_PLANNER_LEARNED_ANY 2

typedef struct qvd_sv qvd_sv; /* opaque: HBM amplitudes + a HIP stream */

/* One unitary op: `matrix` on `targets`, applied where every control holds.
 * Mirrors fusion::Op (src/fusion.rs:16-56). control_values may be NULL
 * (all controls true). */
typedef struct qvd_op {
    const uint32_t *targets;
    uint32_t num_targets;
    const uint32_t *controls;
    const uint8_t *control_values;
    uint32_t num_controls;
    const double *matrix;
} qvd_op;

/* Options of qvd_sv_apply_ops: simulator::Options (src/simulator.rs:37-66)
 * restricted to the state vector, plus the tile and planner knobs. */
typedef struct qvd_options {
    uint32_t max_fused_qubits; /* largest fused gate; 1 disables fusion (default 4) */
    int32_t tile_bits;         /* qubits per LDS-resident tile; <0 = no blocking, 0 = default */
    uint32_t planner;          /* QVD_PLANNER_* */
    uint32_t flags;            /* QVD_FLAG_* */
    uint32_t variant;          /* heuristic plan variant, 0..QVD_PLAN_VARIANTS-1 (0 = the heuristic);
                                  the learned planners choose it themselves */
    uint32_t fixed_lanes;      /* register-tile geometry of qvd_sv_apply_ops: the number F of lane bits
                                  pinned to qubits 0..F-1 (as qvd_circuit.fixed_lanes); 0 = the default, or
                                  the autotuner's choice with QVD_FLAG_AUTOTUNE. New in this engine. */
} qvd_options;

/* Plan variants (blocking.rs's lookahead with a reduced capacity, a greedy
 * choice of the next tile, register slots in order of need or for the
 * most-used qubits first, tiles chosen by the lookahead alone or among
 * regions of the circuit's coupling graph) and the per-stage features the
 * learned planner scores them by. */
#define QVD_PLAN_VARIANTS 32
#define QVD_PLAN_FEATURES 24

/* qvd_options.flags: time every kernel launch with HIP events on the
 * handle's stream and report the sums in qvd_stats. */
#define QVD_FLAG_TIMING 1u
/* qvd_options.flags: apply 5- and 6-qubit dense ops in complex64 with VALU
 * FMAs instead of the FP32 matrix cores (for comparisons). */
#define QVD_FLAG_DENSE_VALU 2u
/* qvd_options.flags: pick the register-tile geometry (qvd_circuit.fixed_lanes
 * or qvd_options.fixed_lanes left 0) by timing. Circuits of the same size and
 * gate layout form a class, and the fastest geometry is kept for this call
 * and for later ones of the process with that class. qvd_batch_create runs
 * each class as a batch of its own (up to enough circuits to fill the device
 * a few times) with every candidate geometry. qvd_sv_apply_ops runs the
 * class's plan once with each candidate on a scratch state of the same size
 * (allocated for the probes and freed; if it cannot be allocated the default
 * stays), the first time a class comes; the caller's state is untouched by
 * the probes. The geometry changes which code path a gate takes, so the
 * amplitudes can differ in the last bits between geometries. New in this
 * engine. */
#define QVD_FLAG_AUTOTUNE 4u
/* qvd_options.flags: visit every register tile. By default a stage skips the
 * tiles known to be zero: a qubit that no non-diagonal op has targeted since
 * the state was |0...0> (qvd_sv_create, qvd_sv_set_zero, a batch's start)
 * is still |0>, so every amplitude with its bit set is zero, and a stage
 * whose tile leaves that qubit out neither reads nor writes those tiles (its
 * ops are diagonal or controlled on that qubit, and keep them zero). The
 * amplitudes are the same either way, except that a zero may come out as
 * -0.0 when computed. */
#define QVD_FLAG_ALL_TILES 8u

/* What a run did: simulator::Stats (src/simulator.rs:70-93). */
typedef struct qvd_stats {
    uint64_t gates;   /* ops in */
    uint64_t kernels; /* ops after fusion */
    uint64_t passes;  /* full HBM passes: one per stage and per full-state op */
    uint64_t stages;  /* tile stages launched */
    double gate_seconds;
    double measure_seconds;
    double plan_seconds;          /* host planning inside gate_seconds */
    double stage_kernel_seconds;  /* sum of tile-stage kernel durations (QVD_FLAG_TIMING) */
    double other_kernel_seconds;  /* sum of full-state op kernel durations (QVD_FLAG_TIMING) */
    uint64_t plan_variant;        /* the plan variant that ran */
    uint64_t stage_bytes;         /* algorithmic HBM bytes of the register-tile stages: each reads and writes
                                     the tiles it visits (a stage from |0...0> only writes; amplitudes of
                                     qubits whose zeros were never written are not read) */
    uint64_t other_bytes;         /* algorithmic HBM bytes of the other passes (full-state ops, writing a
                                     pending |0...0> or unwritten zeros, a batch's initialisation and
                                     expectation values) */
    uint64_t skipped_tiles;       /* register tiles not visited because they are known to be zero */
    uint64_t fixed_lanes;         /* register-tile geometry the stages ran with (qvd_options.fixed_lanes) */
} qvd_stats;

const char *qvd_last_error(void);

/* Library / device information: number of visible devices and the name of
 * the architecture of device 0 (e.g. "gfx950"). */
int qvd_device_info(int *num_devices, char *arch, uint32_t arch_capacity);

/* state_vector_fits (src/simulator.rs:254-259): *fits = 1 when a state of
 * num_qubits amplitudes in `precision` takes at most 80 % of the current
 * device's HBM (the reference asks the same of host memory). *state_bytes and
 * *device_bytes report both sides of the test. */
int qvd_state_vector_fits(uint32_t num_qubits, int precision, int *fits, uint64_t *state_bytes,
                          uint64_t *device_bytes);

/* StateVector::zero (src/state.rs:39-49): |0...0> in HBM. The zeros are
 * written lazily: the next qvd_sv_apply_ops whose plan starts with a
 * register-tile stage makes that stage's tiles from |0...0> in registers (no
 * init pass, no loads). The tiles its stages skip as known to be zero may stay
 * unwritten through any number of such applies; any other call writes them
 * first. */
int qvd_sv_create(uint32_t num_qubits, int precision, qvd_sv **out);
int qvd_sv_destroy(qvd_sv *sv);
int qvd_sv_num_qubits(const qvd_sv *sv, uint32_t *out);
int qvd_sv_precision(const qvd_sv *sv, int *out);
/* Reset to |0...0> without reallocating, written lazily as for
 * qvd_sv_create. */
int qvd_sv_set_zero(qvd_sv *sv);

/* A copy of src's amplitudes in dst (same size and precision), device to
 * device: the branch copies of the batched shot loop (the reference re-runs
 * every shot from |0...0>, src/simulator.rs:383-425). */
int qvd_sv_copy(qvd_sv *dst, const qvd_sv *src);

/* StateVector::from_amplitudes / set (src/state.rs:52-63, 137-143):
 * `amplitudes` holds 2^n complex128 values in logical order. */
int qvd_sv_set_amplitudes(qvd_sv *sv, const double *amplitudes, uint64_t len);
/* StateVector::to_vec (src/state.rs:146-159), widened to complex128. */
int qvd_sv_get_amplitudes(const qvd_sv *sv, double *amplitudes, uint64_t len);

/* StateVector::apply (src/state.rs:181-198): one op in one HBM pass
 * (controls restrict the pass to the controlled subspace). */
int qvd_sv_apply(qvd_sv *sv, const double *matrix, const uint32_t *targets, uint32_t num_targets,
                 const uint32_t *controls, const uint8_t *control_values, uint32_t num_controls);

/* simulator::apply_ops (src/simulator.rs:122-147) with blocking::apply_blocked
 * (src/blocking.rs:230-302): fuse, plan LDS-resident stages, launch. */
int qvd_sv_apply_ops(qvd_sv *sv, const qvd_op *ops, uint64_t num_ops, const qvd_options *options,
                     qvd_stats *stats);

/* Plan `ops` for a state of `num_qubits` qubits and `precision` exactly as
 * qvd_sv_apply_ops would, without a device: fills gates, kernels, passes,
 * stages and plan_seconds of `stats`. */
int qvd_plan_ops(uint32_t num_qubits, int precision, const qvd_op *ops, uint64_t num_ops, const qvd_options *options,
                 qvd_stats *stats);

/* StateVector::norm_sqr (src/state.rs:172-177), accumulated in FP64. */
int qvd_sv_norm_sqr(const qvd_sv *sv, double *out);
/* StateVector::scale (src/state.rs:201-207). */
int qvd_sv_scale(qvd_sv *sv, double factor);
/* probability_one (src/measure.rs:43-65), FP64. */
int qvd_sv_probability_one(const qvd_sv *sv, uint32_t qubit, double *out);
/* probabilities (src/measure.rs:25-40): 2^n FP64 values in logical order,
 * computed on the device. */
int qvd_sv_probabilities(const qvd_sv *sv, double *out, uint64_t len);
/* <P_t> for Pauli strings P_t, FP64 on the device: the estimator's
 * statevector path (python/python/qevedo/simulator/qiskit.py:584-589, where
 * Qiskit's Statevector.expectation_value evaluates a SparsePauliOp term by
 * term). String t has X on the qubits of x_masks[t] only, Z on those of
 * z_masks[t] only and Y on those of both (bit q = qubit q). Strings with the
 * same X part share one pass over the state (up to 16 per pass). */
int qvd_sv_expect_pauli(const qvd_sv *sv, const uint64_t *x_masks, const uint64_t *z_masks, uint64_t num_terms,
                        double *out);
/* <Z_q> for every qubit (n values) and <Z_q Z_{q+1}> (n-1 values), FP64.
 * New in this engine; <Z_q> = 1 - 2 probability_one(q). */
int qvd_sv_expect_z(const qvd_sv *sv, double *z, double *zz);
/* measure (src/measure.rs:68-80) with the uniform draw supplied by the
 * caller: outcome = uniform < P(1); the state collapses and renormalises. */
int qvd_sv_measure(qvd_sv *sv, uint32_t qubit, double uniform, uint8_t *outcome);
/* reset (src/measure.rs:83-87): measure, then X on outcome 1. */
int qvd_sv_reset(qvd_sv *sv, uint32_t qubit, double uniform);
/* sample_unordered (src/measure.rs:104-174) with the caller's uniforms in
 * [0, 1): FP64 chunk sums in index order (chunks of 32768 amplitudes for
 * complex64, 16384 for complex128), sequential prefix, targets u * total in
 * ascending order, strict-< walk. out[i] is the logical basis index of the
 * i-th smallest target; bit-exact with the reference's algorithm on the same
 * amplitudes and uniforms. */
int qvd_sv_sample(const qvd_sv *sv, const double *uniforms, uint64_t shots, uint64_t *out);
/* qvd_sv_sample without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The sampler's buffers (chunk sums, prefix, bounds, uniforms,
 * targets, outcomes, sort temporaries, the walk's marks) come from one fixed
 * 64 MiB device array of the library, which every GPU it runs on holds until
 * the process ends; the walk's marks are spaced as finely as the array
 * allows (every 64 amplitudes, as in qvd_sv_sample, up to every chunk).
 * Requests of up to 1.7 * 10^6 shots on states of up to 33 qubits fit; one
 * whose buffers do not fit returns QVD_EINVAL and must use qvd_sv_sample.
 * It runs on the handle's stream only. The array is
 * shared by every handle on a GPU and keeps the last call's contents; calls
 * from one process are serialised, so calls that would overlap on one GPU
 * wait for each other. Same outcomes as qvd_sv_sample. */
int qvd_sv_sample_noalloc(const qvd_sv *sv, const double *uniforms, uint64_t shots, uint64_t *out);

/* Distance between two states of the same size (any precisions), in FP64
 * on the device: max_abs = max |a_i - b_i|, diff_sq = sum |a_i - b_i|^2,
 * ref_sq = sum |b_i|^2 (relative L2 error = sqrt(diff_sq / ref_sq)).
 * New in this engine: it checks a complex64 run against a complex128 run
 * without copying either state to the host. */
int qvd_sv_compare(const qvd_sv *a, const qvd_sv *b, double *max_abs, double *diff_sq, double *ref_sq);

/* Sampling statistics of the last qvd_sv_sample on this handle (seconds). */
int qvd_sv_last_sample_seconds(const qvd_sv *sv, double *out);

/* Gate::matrix (src/circuit.rs:133-202) by OpenQASM name: the full matrix
 * on the gate's qubits (controls first, on the low index bits). `matrix`
 * receives 4^k (re, im) pairs; capacity counts complex entries. */
int qvd_gate_matrix(const char *name, const double *params, uint32_t num_params, double *matrix, uint32_t capacity,
                    uint32_t *num_qubits);
/* gate_op (src/simulator.rs:106-119): the number of leading control qubits
 * and the matrix of the gate applied to the remaining qubits. */
int qvd_gate_lower(const char *name, const double *params, uint32_t num_params, uint32_t *num_controls,
                   uint32_t *num_targets, double *matrix, uint32_t capacity);

/* Features of the register-tile stages of the plan qvd_sv_apply_ops would
 * run with these options, QVD_PLAN_FEATURES floats per stage (row-major),
 * in launch order; `capacity` counts rows. New in this engine: the inputs of
 * the learned planner's cost model (the reference plans by a fixed rule,
 * src/blocking.rs:150-302). */
int qvd_plan_features(uint32_t num_qubits, int precision, const qvd_op *ops, uint64_t num_ops,
                      const qvd_options *options, float *features, uint64_t capacity, uint64_t *rows);
/* Which plan variants run exactly the arithmetic of variant options->variant
 * for these ops: bit v of *mask is set when variant v's plan image (every
 * register-tile program, stage descriptor, coefficient and full-state op,
 * byte for byte) equals that of options->variant, so it gives bit-identical
 * amplitudes. These are the candidates of QVD_PLANNER_LEARNED. New in this
 * engine (the reference has one planner, src/blocking.rs:230-302). */
int qvd_plan_equivalent_variants(uint32_t num_qubits, int precision, const qvd_op *ops, uint64_t num_ops,
                                 const qvd_options *options, uint32_t *mask);
/* Kernel time of each register-tile stage of the last qvd_sv_apply_ops run
 * with QVD_FLAG_TIMING, in launch order (seconds, HIP events). */
int qvd_sv_stage_seconds(const qvd_sv *sv, double *seconds, uint64_t capacity, uint64_t *count);

/* ---- The learned planner's cost model ------------------------------------
 * A 3-layer perceptron (inputs -> hidden -> hidden -> 1, ReLU, FP32) whose
 * forward, backward and Adam steps are HIP kernels of this library.
 * Parameters: W1[hidden x inputs], b1[hidden], W2[hidden x hidden],
 * b2[hidden], W3[hidden], b3 (row-major). New in this engine. */
typedef struct qvd_mlp qvd_mlp;
int qvd_mlp_create(uint32_t inputs, uint32_t hidden, qvd_mlp **out);
int qvd_mlp_destroy(qvd_mlp *mlp);
int qvd_mlp_num_params(const qvd_mlp *mlp, uint64_t *count);
int qvd_mlp_set_params(qvd_mlp *mlp, const float *params, uint64_t count);
int qvd_mlp_get_params(const qvd_mlp *mlp, float *params, uint64_t count);
/* y[r] for `rows` rows of x (row-major, inputs floats each). */
int qvd_mlp_forward(qvd_mlp *mlp, const float *x, uint64_t rows, float *y);
/* Gradient of the mean squared error (y - t)^2 over the rows with respect to
 * every parameter, and the loss. */
int qvd_mlp_gradient(qvd_mlp *mlp, const float *x, const float *t, uint64_t rows, float *grad, double *loss);
/* Adam (lr, betas 0.9 / 0.999) on shuffled minibatches; loss = final MSE. */
int qvd_mlp_train(qvd_mlp *mlp, const float *x, const float *t, uint64_t rows, uint32_t epochs, uint32_t batch,
                  float lr, uint64_t seed, double *loss);
/* Installs a copy of `mlp` (QVD_PLAN_FEATURES inputs) as the cost model of
 * QVD_PLANNER_LEARNED: the model sees (feature - mean) / scale and predicts
 * (log(seconds per amplitude) - y_mean) / y_scale. NULL removes the model,
 * and the learned planner then falls back to the heuristic. */
int qvd_planner_set_model(const qvd_mlp *mlp, const float *mean, const float *scale, float y_mean, float y_scale);

/* ---- Batches of circuits ------------------------------------------------
 * Many circuits, each on its own state, run together: the engine plans every
 * circuit into register-tile stages and runs all of them in one launch of the
 * stage kernel, whose workgroups take runs of tiles from a queue (stage r of
 * every circuit before stage r + 1 of any) and start a circuit's stage once
 * its previous stage has finished; the first stage starts from |0...0>
 * without reading the state. The launch and the expectation values are
 * captured once in a HIP graph at creation and replayed by every
 * qvd_batch_run. States smaller than one register tile
 * (14 qubits in complex64, 13 in complex128) are widened with ancilla qubits
 * in |0>, which no op touches. New in this engine: the reference runs such
 * circuits one at a time (run_on, src/simulator.rs:355-426); the estimator
 * path it serves is python/qevedo/qiskit.py:584-589. */
typedef struct qvd_batch qvd_batch;

typedef struct qvd_circuit {
    uint32_t num_qubits;
    const qvd_op *ops;
    uint64_t num_ops;
    /* Register-tile geometry of this circuit's stages: the number F of lane
     * bits pinned to qubits 0..F-1 (each access then reads 2^F consecutive
     * amplitudes; the other 6 - F lane bits hold any tile qubit). 0 = the
     * default (4 in complex64, 3 in complex128) or, with QVD_FLAG_AUTOTUNE,
     * the autotuner's choice; otherwise the default - 1 .. the default + 1. */
    uint32_t fixed_lanes;
} qvd_circuit;

/* Plans, uploads and captures `count` circuits. Every op must plan into
 * register-tile stages (ops of at most 4 targets, after fusion); a circuit
 * that needs a full-state pass is rejected with QVD_EUNSUPPORTED, and runs
 * through qvd_sv_apply_ops instead. */
int qvd_batch_create(const qvd_circuit *circuits, uint32_t count, int precision, const qvd_options *options,
                     qvd_batch **out);
int qvd_batch_destroy(qvd_batch *batch);
/* Runs every circuit from |0...0> and writes <Z_q> (num_qubits values per
 * circuit) to z and <Z_q Z_{q+1}> (num_qubits - 1 per circuit) to zz, both
 * concatenated in circuit order, FP64. stats: gates, kernels, passes (stage
 * levels: the most stages of any circuit), stages (register-tile stage
 * descriptors), gate_seconds (the graph replay, host clock). Fails with
 * QVD_EHIP if a stage waited more than 10 s for its previous one. */
int qvd_batch_run(qvd_batch *batch, double *z, double *zz, qvd_stats *stats);
/* Sizes of the z and zz arrays qvd_batch_run fills. */
int qvd_batch_outputs(const qvd_batch *batch, uint64_t *num_z, uint64_t *num_zz);
/* Copies circuit `index`'s state after the last qvd_batch_run (2^num_qubits
 * amplitudes in the batch's precision, interleaved re/im) to host memory
 * `out` of `capacity` amplitudes. No reference counterpart: for callers that
 * need the states themselves, and for checking a batch against single
 * states amplitude by amplitude. */
int qvd_batch_amplitudes(const qvd_batch *batch, uint32_t index, void *out, uint64_t capacity);
/* The register-tile geometry (fixed lane bits) each of the first `count`
 * circuits was built with, and its number of stages; either output may be
 * NULL. No reference counterpart. */
int qvd_batch_geometry(const qvd_batch *batch, uint32_t *fixed_lanes, uint32_t *stages, uint32_t count);

#ifdef __cplusplus
}
#endif
#endif /* QVD_HIP_H */

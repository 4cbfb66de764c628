// Host-callable launchers of the HIP kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "device.h"

namespace qvd {

struct DiagTargets {
    uint32_t q[MAX_GATE_QUBITS];
};

// Device buffers of one sampling call.
struct SampleScratch {
    double *sums = nullptr;              // per chunk
    double *prefix = nullptr;            // chunks + 1
    uint64_t *bounds = nullptr;          // chunks + 1
    double *targets = nullptr;           // shots: the sorted uniforms, then the targets, ascending
    void *sort_temp = nullptr;
    size_t sort_temp_bytes = 0;
    // The walk's running sum every 2^mark_bits amplitudes (6 <= mark_bits,
    // 2^mark_bits <= the chunk length): ceil(len / 2^mark_bits) values.
    uint32_t mark_bits = 6;
    double *marks = nullptr;
    uint32_t *walked = nullptr;  // per chunk: marks written
};

// The fixed device workspace of the allocation-free sampling entry point
// (qvd_sv_sample_noalloc): NOALLOC_SCRATCH_BYTES bytes of a file-scope
// __device__ array, which every GPU the library runs on holds until the
// process ends. *ptr receives its address on the current device, looked up
// once per device.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
hipError_t noalloc_scratch(void **ptr);

// Pauli strings per pass of the expectation kernel (one X part, up to this
// many Z parts).
constexpr uint32_t PAULI_TERMS = 16;

template <typename T2> void launch_init_zero(T2 *s, uint64_t len, hipStream_t st);
// Zero over every amplitude whose index has a bit of `stale` set.
template <typename T2> void launch_zero_stale(T2 *s, uint64_t len, uint64_t stale, hipStream_t st);
template <typename T2> void launch_probabilities(const T2 *s, uint64_t len, double *out, hipStream_t st);
template <typename T2>
void launch_expect_pauli(const T2 *s, uint64_t len, uint64_t x, const uint64_t *zs, uint32_t nterms, double *partial,
                         unsigned blocks, double *out, hipStream_t st);
template <typename T2> void launch_widen(const T2 *s, double2 *out, uint64_t count, hipStream_t st);
template <typename T2> void launch_narrow(const double2 *in, T2 *s, uint64_t count, hipStream_t st);
template <typename T2> void launch_scale(T2 *s, uint64_t len, double factor, hipStream_t st);
template <typename T2> void launch_dense(T2 *s, const GateDesc &d, const T2 *mat, hipStream_t st);
template <typename T2>
void launch_diag(T2 *s, uint64_t len, uint32_t k, const DiagTargets &t, uint64_t cmask, uint64_t cval,
                 const T2 *entries, hipStream_t st);
template <typename T2> size_t stage_lds_bytes(uint32_t L);
template <typename T2>
void launch_stage(T2 *s, const StageDesc &d, const StageOp *ops, const T2 *coefs, unsigned grid, hipStream_t st);
template <typename T2>
void launch_masked_norm(const T2 *s, uint64_t len, uint64_t mask, uint64_t val, double *partial, unsigned blocks,
                        double *out, hipStream_t st);
template <typename TA, typename TB>
void launch_compare(const TA *a, const TB *b, uint64_t len, double *partial, unsigned blocks, hipStream_t st);
uint32_t expect_segment_bits(uint32_t n);
// <Z_q> and <Z_q Z_{q+1}> of one state among several: segment sums over
// 2^B-amplitude segments, then one combine thread per output value.
struct ExpDesc {
    uint64_t state_off;  // first amplitude of the state
    uint64_t seg_off;    // first double of its segment sums (2 B per segment)
    uint64_t nsegs;      // segments
    uint32_t B;          // segment bits
    uint32_t n;          // qubits reported: n values of z, n - 1 of zz
    uint32_t z_off, zz_off;
};
ExpDesc expect_desc(uint32_t state_qubits, uint32_t n);
template <typename T2>
void launch_expect_z(const T2 *s, uint32_t n, double *seg_sums, double *z, double *zz, hipStream_t st);
template <typename T2>
void launch_expect_batch(const T2 *s, const ExpDesc *descs, uint32_t ndesc, const uint64_t *seg_start,
                         uint64_t total_segs, const uint32_t *out_start, uint32_t total_outs, double *seg_sums,
                         double *z, double *zz, hipStream_t st);
// Zero-fills n 32-bit words with a kernel.
void launch_zero_words(uint32_t *p, uint32_t n, hipStream_t st);
// Zero-fills `len` amplitudes and sets amplitude offs[i] to 1 for each state i.
template <typename T2>
void launch_batch_init(T2 *s, uint64_t len, const uint64_t *offs, uint32_t count, hipStream_t st);
template <typename T2>
void launch_collapse(T2 *s, uint64_t len, uint32_t qubit, uint32_t outcome, double keep, hipStream_t st);
size_t sort_temp_bytes(uint64_t shots);
// Two-pass sampling (src/measure.rs:104-174) in three parts: the chunk sums
// and their prefix; the uniforms' sort into w.targets (on any stream); then,
// after both, targets = sorted uniforms * total, the bounds and the walk.
template <typename T2>
void launch_sample_sums(const T2 *s, uint64_t len, uint64_t chunk, SampleScratch &w, hipStream_t st);
hipError_t launch_sort_uniforms(const double *uniforms, uint64_t shots, SampleScratch &w, hipStream_t st);
template <typename T2>
hipError_t launch_sample_walk(const T2 *s, uint64_t len, uint64_t chunk, uint64_t shots, SampleScratch &w,
                              uint64_t *out, hipStream_t st);

}  // namespace qvd

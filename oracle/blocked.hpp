// ORACLE — test infrastructure only (see matrix.hpp).
//
// A CPU restatement of the reference's blocked state-vector engine, used both
// as the parity checker and as the timed CPU baseline:
//   - blocks of LANES real then LANES imaginary values   (src/simd.rs:15-53, src/state.rs:121-124)
//   - pdep loop-index spreading                            (src/kernels.rs:22-43)
//   - Layout / DensePlan / DiagonalPlan / dense_sweep      (src/kernels.rs:78-473)
//   - Plan::new folding lane controls into the matrix      (src/kernels.rs:484-525)
//   - StateVector with logical->physical layout            (src/state.rs)
//   - swap_physical                                        (src/state.rs:212-268)
//   - measurement and two-pass sampling                    (src/measure.rs)
// The SIMD vectors are AVX2 (8 x f32 / 4 x f64) as in the reference; the
// arithmetic per lane is the reference's: fused multiply-adds for mul_add and
// neg_mul_add, plain products and sums elsewhere.
#pragma once

#include <immintrin.h>
#include <omp.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "matrix.hpp"

namespace orc {

inline size_t deposit(size_t value, size_t mask) {
#if defined(__BMI2__)
    return (size_t)_pdep_u64(value, mask);
#else
    size_t out = 0, bit = 1;
    while (mask) {
        size_t lowest = mask & (~mask + 1);
        if (value & bit) out |= lowest;
        mask &= mask - 1;
        bit <<= 1;
    }
    return out;
#endif
}

// ---------------------------------------------------------------------------
// SIMD vectors (simd.rs:26-53 semantics).
// ---------------------------------------------------------------------------
struct F32x8 {
    __m256 v;
    static F32x8 load(const float *p) { return {_mm256_loadu_ps(p)}; }
    void store(float *p) const { _mm256_storeu_ps(p, v); }
    static F32x8 zero() { return {_mm256_setzero_ps()}; }
    static F32x8 mul_add(F32x8 a, F32x8 b, F32x8 c) { return {_mm256_fmadd_ps(a.v, b.v, c.v)}; }
    static F32x8 neg_mul_add(F32x8 a, F32x8 b, F32x8 c) { return {_mm256_fnmadd_ps(a.v, b.v, c.v)}; }
    static F32x8 mul(F32x8 a, F32x8 b) { return {_mm256_mul_ps(a.v, b.v)}; }
    static F32x8 add(F32x8 a, F32x8 b) { return {_mm256_add_ps(a.v, b.v)}; }
    static F32x8 sub(F32x8 a, F32x8 b) { return {_mm256_sub_ps(a.v, b.v)}; }
    using Perm = __m256i;
    // Lane l takes the value of lane l ^ x.
    static Perm xor_perm(size_t x) {
        alignas(32) int idx[8];
        for (int l = 0; l < 8; l++) idx[l] = l ^ (int)x;
        return _mm256_load_si256((const __m256i *)idx);
    }
    F32x8 permute(Perm p) const { return {_mm256_permutevar8x32_ps(v, p)}; }
};

struct F64x4 {
    __m256d v;
    static F64x4 load(const double *p) { return {_mm256_loadu_pd(p)}; }
    void store(double *p) const { _mm256_storeu_pd(p, v); }
    static F64x4 zero() { return {_mm256_setzero_pd()}; }
    static F64x4 mul_add(F64x4 a, F64x4 b, F64x4 c) { return {_mm256_fmadd_pd(a.v, b.v, c.v)}; }
    static F64x4 neg_mul_add(F64x4 a, F64x4 b, F64x4 c) { return {_mm256_fnmadd_pd(a.v, b.v, c.v)}; }
    static F64x4 mul(F64x4 a, F64x4 b) { return {_mm256_mul_pd(a.v, b.v)}; }
    static F64x4 add(F64x4 a, F64x4 b) { return {_mm256_add_pd(a.v, b.v)}; }
    static F64x4 sub(F64x4 a, F64x4 b) { return {_mm256_sub_pd(a.v, b.v)}; }
    using Perm = __m256i;
    // A double lane is two 32-bit lanes: permute the pairs.
    static Perm xor_perm(size_t x) {
        alignas(32) int idx[8];
        for (int l = 0; l < 4; l++) {
            idx[2 * l] = 2 * (l ^ (int)x);
            idx[2 * l + 1] = 2 * (l ^ (int)x) + 1;
        }
        return _mm256_load_si256((const __m256i *)idx);
    }
    F64x4 permute(Perm p) const {
        return {_mm256_castps_pd(_mm256_permutevar8x32_ps(_mm256_castpd_ps(v), p))};
    }
};

template <typename T> struct Real;
template <> struct Real<float> {
    static constexpr size_t LANES = 8;
    static constexpr unsigned LANE_BITS = 3;
    using V = F32x8;
};
template <> struct Real<double> {
    static constexpr size_t LANES = 4;
    static constexpr unsigned LANE_BITS = 2;
    using V = F64x4;
};

// Iterations per parallel task (kernels.rs:61-63).
constexpr size_t TASK_ITERATIONS = size_t(1) << 11;

template <typename F>
inline void for_ranges(size_t count, bool parallel, F &&body) {
    if (!parallel || count <= TASK_ITERATIONS || omp_get_max_threads() == 1) {
        body(size_t(0), count);
        return;
    }
    const long tasks = (long)((count + TASK_ITERATIONS - 1) / TASK_ITERATIONS);
#pragma omp parallel for schedule(dynamic, 1)
    for (long task = 0; task < tasks; task++) {
        size_t start = (size_t)task * TASK_ITERATIONS;
        body(start, std::min(start + TASK_ITERATIONS, count));
    }
}

// ---------------------------------------------------------------------------
// Layout (kernels.rs:78-156).
// ---------------------------------------------------------------------------
struct Layout {
    std::vector<std::pair<size_t, unsigned>> high;  // (matrix bit, block bit), sorted by block bit
    std::vector<std::pair<size_t, unsigned>> low;   // (matrix bit, lane bit)
    std::vector<size_t> offsets;
    size_t free_mask = 0, control_value = 0;

    Layout(const std::vector<size_t> &targets, const std::vector<std::pair<size_t, bool>> &controls,
           unsigned region_bits, unsigned lane_bits) {
        for (size_t j = 0; j < targets.size(); j++) {
            size_t q = targets[j];
            if (q < lane_bits)
                low.push_back({j, (unsigned)q});
            else
                high.push_back({j, (unsigned)(q - lane_bits)});
        }
        std::stable_sort(high.begin(), high.end(), [](auto &a, auto &b) { return a.second < b.second; });
        size_t high_mask = 0;
        for (auto &h : high) high_mask |= size_t(1) << h.second;
        for (size_t h = 0; h < (size_t(1) << high.size()); h++) offsets.push_back(deposit(h, high_mask));
        size_t control_mask = 0;
        for (auto &c : controls) {
            if (c.first < lane_bits) throw std::logic_error("low control qubits must be folded into the matrix");
            control_mask |= size_t(1) << (c.first - lane_bits);
            if (c.second) control_value |= size_t(1) << (c.first - lane_bits);
        }
        size_t region_mask = (size_t(1) << region_bits) - 1;
        if ((high_mask | control_mask) & ~region_mask) throw std::logic_error("gate qubit outside the region");
        free_mask = region_mask & ~high_mask & ~control_mask;
    }
    size_t matrix_index(size_t h, size_t lane) const {
        size_t index = 0;
        for (size_t m = 0; m < high.size(); m++) index |= ((h >> m) & 1) << high[m].first;
        for (auto &l : low) index |= ((lane >> l.second) & 1) << l.first;
        return index;
    }
    size_t low_lane_mask() const {
        size_t m = 0;
        for (auto &l : low) m |= size_t(1) << l.second;
        return m;
    }
    size_t iterations() const { return size_t(1) << __builtin_popcountll(free_mask); }
};

template <typename T> struct AlignedVec {
    T *ptr = nullptr;
    size_t len = 0;
    AlignedVec() = default;
    explicit AlignedVec(size_t n) : len(n) {
        size_t bytes = ((n * sizeof(T) + 63) / 64) * 64;
        ptr = (T *)std::aligned_alloc(64, bytes ? bytes : 64);
        if (!ptr) throw std::bad_alloc();
    }
    AlignedVec(const AlignedVec &) = delete;
    AlignedVec &operator=(const AlignedVec &) = delete;
    AlignedVec(AlignedVec &&o) noexcept : ptr(o.ptr), len(o.len) { o.ptr = nullptr; o.len = 0; }
    AlignedVec &operator=(AlignedVec &&o) noexcept {
        std::swap(ptr, o.ptr);
        std::swap(len, o.len);
        return *this;
    }
    ~AlignedVec() { std::free(ptr); }
};

// ---------------------------------------------------------------------------
// DensePlan / DiagonalPlan (kernels.rs:158-473).
// ---------------------------------------------------------------------------
template <typename T> struct DensePlan {
    using R = Real<T>;
    using V = typename R::V;
    Layout layout;
    size_t nh, nt;
    std::vector<size_t> spreads;  // lane xor pattern of each low-target combination
    std::vector<T> coefficients;  // per (row, col, t): LANES re then LANES im

    DensePlan(const Matrix &m, const std::vector<size_t> &targets,
              const std::vector<std::pair<size_t, bool>> &controls, unsigned region_bits)
        : layout(targets, controls, region_bits, R::LANE_BITS) {
        nh = size_t(1) << layout.high.size();
        nt = size_t(1) << layout.low.size();
        size_t low_mask = layout.low_lane_mask();
        for (size_t t = 0; t < nt; t++) spreads.push_back(deposit(t, low_mask));
        for (size_t row = 0; row < nh; row++)
            for (size_t col = 0; col < nh; col++)
                for (size_t s : spreads) {
                    for (size_t lane = 0; lane < R::LANES; lane++)
                        coefficients.push_back((T)m.at(layout.matrix_index(row, lane), layout.matrix_index(col, lane ^ s)).real());
                    for (size_t lane = 0; lane < R::LANES; lane++)
                        coefficients.push_back((T)m.at(layout.matrix_index(row, lane), layout.matrix_index(col, lane ^ s)).imag());
                }
    }
    void apply(T *region, bool parallel) const;
};

template <typename T, size_t NH, size_t NT>
void dense_sweep(const DensePlan<T> &plan, T *region, bool parallel) {
    using R = Real<T>;
    using V = typename R::V;
    constexpr size_t L = R::LANES;
    const size_t free = plan.layout.free_mask, control_value = plan.layout.control_value;
    size_t offsets[NH];
    for (size_t h = 0; h < NH; h++) offsets[h] = plan.layout.offsets[h] * 2 * L;
    typename V::Perm perms[NT];
    for (size_t t = 0; t < NT; t++) perms[t] = V::xor_perm(plan.spreads[t]);
    const T *coef = plan.coefficients.data();
    constexpr size_t stride = NT * 2 * L;  // coefficients per (row, col)
    for_ranges(plan.layout.iterations(), parallel, [&](size_t start, size_t end) {
        size_t base = deposit(start, free) | control_value;
        for (size_t it = start; it < end; it++) {
            T *block = region + base * 2 * L;
            V re[NH][NT], im[NH][NT];
            for (size_t h = 0; h < NH; h++) {
                V xr = V::load(block + offsets[h]), xi = V::load(block + offsets[h] + L);
                for (size_t t = 0; t < NT; t++) {
                    re[h][t] = t == 0 ? xr : xr.permute(perms[t]);
                    im[h][t] = t == 0 ? xi : xi.permute(perms[t]);
                }
            }
            // Every output row from the loaded inputs: four accumulators
            // (rr, ir, ii, ri), one fused multiply-add per input term each,
            // then (rr - ii, ir + ri) (kernels.rs:311-386).
            for (size_t row = 0; row < NH; row++) {
                V rr = V::zero(), ir = V::zero(), ii = V::zero(), ri = V::zero();
                const T *c0 = coef + row * NH * stride;
                for (size_t col = 0; col < NH; col++)
                    for (size_t t = 0; t < NT; t++) {
                        size_t o = col * stride + t * 2 * L;
                        V wr = V::load(c0 + o), wi = V::load(c0 + o + L);
                        rr = V::mul_add(wr, re[col][t], rr);
                        ir = V::mul_add(wr, im[col][t], ir);
                        ii = V::mul_add(wi, im[col][t], ii);
                        ri = V::mul_add(wi, re[col][t], ri);
                    }
                T *p = block + offsets[row];
                V::sub(rr, ii).store(p);
                V::add(ir, ri).store(p + L);
            }
            base = (((base | ~free) + 1) & free) | control_value;
        }
    });
}

template <typename T> void DensePlan<T>::apply(T *region, bool parallel) const {
#define ORC_SHAPE(a, b) \
    if (nh == a && nt == b) return dense_sweep<T, a, b>(*this, region, parallel);
    ORC_SHAPE(1, 2) ORC_SHAPE(1, 4) ORC_SHAPE(1, 8)
    ORC_SHAPE(2, 1) ORC_SHAPE(2, 2) ORC_SHAPE(2, 4) ORC_SHAPE(2, 8)
    ORC_SHAPE(4, 1) ORC_SHAPE(4, 2) ORC_SHAPE(4, 4) ORC_SHAPE(4, 8)
    ORC_SHAPE(8, 1) ORC_SHAPE(8, 2) ORC_SHAPE(8, 4) ORC_SHAPE(8, 8)
    ORC_SHAPE(16, 1) ORC_SHAPE(16, 2) ORC_SHAPE(16, 4)
    ORC_SHAPE(32, 1) ORC_SHAPE(32, 2)
    ORC_SHAPE(64, 1)
#undef ORC_SHAPE
    throw std::logic_error("unsupported dense gate shape");
}

template <typename T> struct DiagonalPlan {
    using R = Real<T>;
    using V = typename R::V;
    Layout layout;
    size_t nh;
    std::vector<T> factors;  // per high combination: LANES re then LANES im

    DiagonalPlan(const std::vector<C64> &entries, const std::vector<size_t> &targets,
                 const std::vector<std::pair<size_t, bool>> &controls, unsigned region_bits)
        : layout(targets, controls, region_bits, R::LANE_BITS) {
        nh = size_t(1) << layout.high.size();
        for (size_t h = 0; h < nh; h++) {
            for (size_t lane = 0; lane < R::LANES; lane++) factors.push_back((T)entries[layout.matrix_index(h, lane)].real());
            for (size_t lane = 0; lane < R::LANES; lane++) factors.push_back((T)entries[layout.matrix_index(h, lane)].imag());
        }
    }
    void apply(T *region, bool parallel) const {
        using Vv = V;
        constexpr size_t L = R::LANES;
        const size_t free = layout.free_mask, control_value = layout.control_value;
        std::vector<size_t> offs(nh);
        for (size_t h = 0; h < nh; h++) offs[h] = layout.offsets[h] * 2 * L;
        const T *f = factors.data();
        const size_t NH = nh;
        for_ranges(layout.iterations(), parallel, [&](size_t start, size_t end) {
            size_t base = deposit(start, free) | control_value;
            for (size_t it = start; it < end; it++) {
                T *block = region + base * 2 * L;
                for (size_t h = 0; h < NH; h++) {
                    Vv fr = Vv::load(f + h * 2 * L), fi = Vv::load(f + h * 2 * L + L);
                    T *p = block + offs[h];
                    Vv re = Vv::load(p), im = Vv::load(p + L);
                    Vv out_re = Vv::neg_mul_add(fi, im, Vv::mul(fr, re));
                    Vv out_im = Vv::mul_add(fi, re, Vv::mul(fr, im));
                    out_re.store(p);
                    out_im.store(p + L);
                }
                base = (((base | ~free) + 1) & free) | control_value;
            }
        });
    }
};

constexpr size_t MAX_GATE_QUBITS = 6;  // kernels.rs:475-476

template <typename T> struct Plan {
    std::unique_ptr<DensePlan<T>> dense;
    std::unique_ptr<DiagonalPlan<T>> diag;

    // Plan::new (kernels.rs:488-525).
    Plan(const Matrix &matrix, const std::vector<size_t> &qubits,
         const std::vector<std::pair<size_t, bool>> &controls, unsigned region_bits) {
        const size_t lane_bits = Real<T>::LANE_BITS;
        std::vector<std::pair<size_t, bool>> low, high;
        for (auto &c : controls) (c.first < lane_bits ? low : high).push_back(c);
        Matrix m;
        std::vector<size_t> targets;
        if (low.empty()) {
            m = matrix;
            targets = qubits;
        } else {
            size_t state = 0;
            for (size_t i = 0; i < low.size(); i++) state |= (size_t)low[i].second << i;
            for (auto &c : low) targets.push_back(c.first);
            for (size_t q : qubits) targets.push_back(q);
            m = matrix.controlled((unsigned)low.size(), state);
        }
        if (targets.size() > MAX_GATE_QUBITS)
            throw std::invalid_argument("gates on more than 6 qubits (including controls on lane qubits) are not supported");
        if (m.is_diagonal())
            diag.reset(new DiagonalPlan<T>(m.diagonal_entries(), targets, high, region_bits));
        else
            dense.reset(new DensePlan<T>(m, targets, high, region_bits));
    }
    void apply(T *region, bool parallel) const {
        if (dense) dense->apply(region, parallel);
        else diag->apply(region, parallel);
    }
};

// ---------------------------------------------------------------------------
// StateVector (state.rs) and measurement (measure.rs).
// ---------------------------------------------------------------------------
constexpr size_t CHUNK_BLOCKS = size_t(1) << 12;  // measure.rs:12-13

template <typename T> struct StateVector {
    using R = Real<T>;
    static constexpr size_t L = R::LANES;
    unsigned num_qubits;
    AlignedVec<T> buffer;
    std::vector<size_t> layout;  // layout[logical] = physical

    explicit StateVector(unsigned n) : num_qubits(n) {
        if (n >= 56) throw std::invalid_argument("too many qubits");
        size_t amplitudes = std::max(size_t(1) << n, L);
        buffer = AlignedVec<T>(2 * amplitudes);
        T *d = buffer.ptr;
        const long blocks = (long)(buffer.len / (2 * L));
#pragma omp parallel for schedule(static)
        for (long b = 0; b < blocks; b++) std::memset(d + (size_t)b * 2 * L, 0, 2 * L * sizeof(T));
        d[0] = (T)1.0;
        for (unsigned q = 0; q < n; q++) layout.push_back(q);
    }
    size_t len() const { return size_t(1) << num_qubits; }
    size_t blocks() const { return buffer.len / (2 * L); }
    unsigned block_bits() const { return (unsigned)__builtin_ctzll(blocks()); }
    T *raw() { return buffer.ptr; }
    const T *raw() const { return buffer.ptr; }

    size_t physical_index(size_t index) const {
        size_t p = 0;
        for (unsigned q = 0; q < num_qubits; q++) p |= ((index >> q) & 1) << layout[q];
        return p;
    }
    size_t logical_index(size_t index) const {
        size_t l = 0;
        for (unsigned q = 0; q < num_qubits; q++) l |= ((index >> layout[q]) & 1) << q;
        return l;
    }
    bool identity_layout() const {
        for (unsigned q = 0; q < num_qubits; q++)
            if (layout[q] != q) return false;
        return true;
    }
    static size_t re_offset(size_t physical) {
        return (physical >> R::LANE_BITS) * 2 * L + (physical & (L - 1));
    }
    C64 get(size_t index) const {
        size_t o = re_offset(physical_index(index));
        return C64((double)buffer.ptr[o], (double)buffer.ptr[o + L]);
    }
    void set(size_t index, C64 v) {
        size_t o = re_offset(physical_index(index));
        buffer.ptr[o] = (T)v.real();
        buffer.ptr[o + L] = (T)v.imag();
    }
    void to_vec(C64 *out) const {
        const long n = (long)len();
        const bool ident = identity_layout();
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) {
            size_t o = re_offset(ident ? (size_t)i : physical_index((size_t)i));
            out[i] = C64((double)buffer.ptr[o], (double)buffer.ptr[o + L]);
        }
    }
    // state.rs:172-177 (the association of the final sum differs from rayon's;
    // the reference's own result is not association-stable either).
    double norm_sqr() const {
        const size_t chunk = size_t(1) << 14;
        const long chunks = (long)((buffer.len + chunk - 1) / chunk);
        std::vector<double> part(chunks);
#pragma omp parallel for schedule(static)
        for (long c = 0; c < chunks; c++) {
            double s = 0;
            size_t end = std::min(buffer.len, (size_t)(c + 1) * chunk);
            for (size_t i = (size_t)c * chunk; i < end; i++) s += (double)buffer.ptr[i] * (double)buffer.ptr[i];
            part[c] = s;
        }
        double total = 0;
        for (double s : part) total += s;
        return total;
    }
    // state.rs:181-198
    void apply(const Matrix &matrix, const std::vector<size_t> &qubits,
               const std::vector<std::pair<size_t, bool>> &controls) {
        for (size_t q : qubits)
            if (q >= num_qubits) throw std::invalid_argument("qubit out of range");
        for (auto &c : controls)
            if (c.first >= num_qubits) throw std::invalid_argument("qubit out of range");
        if (qubits.empty()) return;
        std::vector<size_t> targets;
        for (size_t q : qubits) targets.push_back(layout[q]);
        std::vector<std::pair<size_t, bool>> pc;
        for (auto &c : controls) pc.push_back({layout[c.first], c.second});
        Plan<T> plan(matrix, targets, pc, block_bits());
        plan.apply(buffer.ptr, true);
    }
    void scale(double factor) {
        const long n = (long)buffer.len;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) buffer.ptr[i] = (T)((double)buffer.ptr[i] * factor);
    }
    // state.rs:212-268
    void swap_physical(const std::vector<std::pair<size_t, size_t>> &pairs) {
        if (pairs.empty()) return;
        const size_t lb = R::LANE_BITS;
        size_t seen = 0;
        std::vector<std::pair<size_t, size_t>> shifts;
        for (auto &pr : pairs) {
            if (pr.first < lb || pr.second < lb || pr.first == pr.second) throw std::invalid_argument("swap of a lane qubit");
            size_t ba = size_t(1) << (pr.first - lb), bb = size_t(1) << (pr.second - lb);
            if (seen & (ba | bb)) throw std::invalid_argument("qubit swapped twice in one pass");
            seen |= ba | bb;
            shifts.push_back({pr.first - lb, pr.second - lb});
        }
        auto permute = [&](size_t i) {
            size_t j = i & ~seen;
            for (auto &s : shifts) j |= ((i >> s.first) & 1) << s.second | ((i >> s.second) & 1) << s.first;
            return j;
        };
        const size_t bl = 2 * L;
        T *d = buffer.ptr;
        const long nb = (long)blocks();
#pragma omp parallel for schedule(static)
        for (long i = 0; i < nb; i++) {
            size_t j = permute((size_t)i);
            if ((size_t)i < j)
                for (size_t e = 0; e < bl; e++) std::swap(d[(size_t)i * bl + e], d[j * bl + e]);
        }
        for (auto &q : layout)
            for (auto &pr : pairs) {
                if (q == pr.first) q = pr.second;
                else if (q == pr.second) q = pr.first;
            }
    }

    // measure.rs:25-40
    void probabilities(double *out) const {
        const long n = (long)len();
        const bool ident = identity_layout();
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) {
            size_t o = re_offset(ident ? (size_t)i : physical_index((size_t)i));
            double re = (double)buffer.ptr[o], im = (double)buffer.ptr[o + L];
            out[i] = re * re + im * im;
        }
    }
    // measure.rs:43-65 (per-chunk sequential sums; the final association is
    // sequential over chunks).
    double probability_one(size_t qubit) const {
        if (qubit >= num_qubits) throw std::invalid_argument("qubit out of range");
        const size_t pq = layout[qubit];
        const size_t chunk_len = 2 * L * CHUNK_BLOCKS;
        const long chunks = (long)((buffer.len + chunk_len - 1) / chunk_len);
        std::vector<double> part(chunks);
#pragma omp parallel for schedule(static)
        for (long c = 0; c < chunks; c++) {
            double s = 0;
            size_t nb = std::min(buffer.len - (size_t)c * chunk_len, chunk_len) / (2 * L);
            for (size_t b = 0; b < nb; b++) {
                const T *blk = buffer.ptr + (size_t)c * chunk_len + b * 2 * L;
                size_t index = ((size_t)c * CHUNK_BLOCKS + b) << R::LANE_BITS;
                double bs = 0;
                for (size_t l = 0; l < L; l++) {
                    double re = (double)blk[l], im = (double)blk[l + L];
                    if ((((index | l) >> pq) & 1) == 1) bs += re * re + im * im;
                }
                s += bs;
            }
            part[c] = s;
        }
        double total = 0;
        for (double s : part) total += s;
        return total;
    }
    // measure.rs:68-80 with the uniform supplied by the caller.
    bool measure(size_t qubit, double uniform) {
        double p1 = std::min(std::max(probability_one(qubit), 0.0), 1.0);
        bool outcome = uniform < p1;
        double p = outcome ? p1 : 1.0 - p1;
        C64 keep(1.0 / std::sqrt(p), 0.0), zero(0.0, 0.0);
        std::vector<C64> entries = outcome ? std::vector<C64>{zero, keep} : std::vector<C64>{keep, zero};
        DiagonalPlan<T> plan(entries, {layout[qubit]}, {}, block_bits());
        plan.apply(buffer.ptr, true);
        return outcome;
    }
    // measure.rs:104-174: two-pass sampling. `uniforms` are the draws the
    // reference takes from its generator (measure.rs:125); the result is in
    // the order of the sorted targets (logical indices).
    std::vector<size_t> sample_unordered(const double *uniforms, size_t shots) const {
        std::vector<size_t> outcomes;
        if (shots == 0) return outcomes;
        const size_t chunk_len = 2 * L * CHUNK_BLOCKS;
        const size_t chunks = (buffer.len + chunk_len - 1) / chunk_len;
        std::vector<double> sums(chunks);
#pragma omp parallel for schedule(static)
        for (long c = 0; c < (long)chunks; c++) {
            size_t nb = std::min(buffer.len - (size_t)c * chunk_len, chunk_len) / (2 * L);
            double s = 0;  // sequential in physical order
            for (size_t b = 0; b < nb; b++) {
                const T *blk = buffer.ptr + (size_t)c * chunk_len + b * 2 * L;
                for (size_t l = 0; l < L; l++) {
                    double re = (double)blk[l], im = (double)blk[l + L];
                    s += re * re + im * im;
                }
            }
            sums[c] = s;
        }
        std::vector<double> prefix(chunks + 1);
        prefix[0] = 0.0;
        for (size_t c = 0; c < chunks; c++) prefix[c + 1] = prefix[c] + sums[c];
        const double total = prefix[chunks];
        std::vector<double> targets(shots);
        for (size_t i = 0; i < shots; i++) targets[i] = uniforms[i] * total;
        std::sort(targets.begin(), targets.end());
        std::vector<size_t> bounds(chunks + 1);
        for (size_t c = 0; c <= chunks; c++)
            bounds[c] = (size_t)(std::lower_bound(targets.begin(), targets.end(), prefix[c]) - targets.begin());
        const size_t n = len();
        outcomes.assign(shots, 0);
#pragma omp parallel for schedule(dynamic, 1)
        for (long c = 0; c < (long)chunks; c++) {
            size_t lo = bounds[c], hi = std::min(bounds[c + 1], shots);
            if (lo >= hi) continue;
            double cumulative = prefix[c];
            size_t shot = lo;
            size_t last_nonzero = ((size_t)c * CHUNK_BLOCKS) << R::LANE_BITS;
            size_t nb = std::min(buffer.len - (size_t)c * chunk_len, chunk_len) / (2 * L);
            bool done = false;
            for (size_t b = 0; b < nb && !done; b++) {
                const T *blk = buffer.ptr + (size_t)c * chunk_len + b * 2 * L;
                for (size_t l = 0; l < L; l++) {
                    double re = (double)blk[l], im = (double)blk[l + L];
                    double p = re * re + im * im;
                    if (p == 0.0) continue;
                    size_t index = (((size_t)c * CHUNK_BLOCKS + b) << R::LANE_BITS) | l;
                    last_nonzero = index;
                    cumulative += p;
                    while (shot < hi && targets[shot] < cumulative) outcomes[shot++] = index;
                    if (shot == hi) { done = true; break; }
                }
            }
            while (shot < hi) outcomes[shot++] = last_nonzero;  // rounding leftovers
            for (size_t s = lo; s < hi; s++) outcomes[s] = std::min(outcomes[s], n - 1);
        }
        if (!identity_layout())
            for (auto &o : outcomes) o = logical_index(o);
        return outcomes;
    }
};

}  // namespace orc

// ORACLE — test infrastructure only (see matrix.hpp).
//
// Gate fusion (src/fusion.rs), cache-blocked stages (src/blocking.rs), the
// state-vector driver (src/simulator.rs:104-238) and the naive FP64 reference
// simulator (src/reference.rs), restated in C++ behind a C API that the
// tests load with ctypes. The blocked engine it drives is blocked.hpp.
#include "qvd_oracle.h"

#include <chrono>
#include <cstdio>
#include <map>
#include <string>

#include "blocked.hpp"

namespace orc {

// ---------------------------------------------------------------------------
// Fusion (fusion.rs:16-171).
// ---------------------------------------------------------------------------
struct Op {
    std::vector<size_t> targets;
    std::vector<std::pair<size_t, bool>> controls;
    Matrix matrix;

    std::vector<size_t> qubits() const {
        std::vector<size_t> q = targets;
        for (auto &c : controls) q.push_back(c.first);
        return q;
    }
    size_t width() const { return targets.size() + controls.size(); }
    // The op as a plain matrix on its controls followed by its targets.
    void full(std::vector<size_t> &qs, Matrix &m) const {
        if (controls.empty()) {
            qs = targets;
            m = matrix;
            return;
        }
        size_t state = 0;
        for (size_t i = 0; i < controls.size(); i++) state |= (size_t)controls[i].second << i;
        qs.clear();
        for (auto &c : controls) qs.push_back(c.first);
        for (size_t q : targets) qs.push_back(q);
        m = matrix.controlled((unsigned)controls.size(), state);
    }
};

static bool contains(const std::vector<size_t> &v, size_t x) { return std::find(v.begin(), v.end(), x) != v.end(); }

struct Group {
    size_t seed;
    std::vector<size_t> members;
    std::vector<size_t> qubits;

    template <typename W>
    void grow(const std::vector<Op> &ops, std::vector<bool> &done, size_t max_qubits, bool connected_only, const W &width) {
        std::vector<size_t> blocked;
        for (size_t i = seed + 1; i < ops.size(); i++) {
            if (done[i]) continue;
            if (connected_only) {
                bool all_blocked = true;
                for (size_t q : qubits)
                    if (!contains(blocked, q)) { all_blocked = false; break; }
                if (all_blocked) break;
            }
            std::vector<size_t> oq = ops[i].qubits();
            bool touches_blocked = false;
            for (size_t q : oq)
                if (contains(blocked, q)) touches_blocked = true;
            std::vector<size_t> fresh;
            for (size_t q : oq)
                if (!contains(qubits, q)) fresh.push_back(q);
            bool connected = fresh.size() < oq.size();
            bool fits = width(ops[i]) <= max_qubits && qubits.size() + fresh.size() <= max_qubits;
            if (!touches_blocked && fits && (connected || !connected_only)) {
                done[i] = true;
                members.push_back(i);
                for (size_t q : fresh) qubits.push_back(q);
            } else {
                for (size_t q : oq)
                    if (!contains(blocked, q)) blocked.push_back(q);
            }
        }
        std::sort(members.begin(), members.end());
    }
};

template <typename F>
std::vector<Op> fuse_where(const std::vector<Op> &ops, size_t max_qubits, const F &fusable) {
    auto width = [&](const Op &op) { return fusable(op) ? op.width() : SIZE_MAX; };
    std::vector<bool> done(ops.size(), false);
    std::vector<Op> fused;
    for (size_t seed = 0; seed < ops.size(); seed++) {
        if (done[seed]) continue;
        done[seed] = true;
        Group g{seed, {seed}, ops[seed].qubits()};
        if (width(ops[seed]) <= max_qubits)
            for (bool connected_only : {true, false}) g.grow(ops, done, max_qubits, connected_only, width);
        if (g.members.size() == 1) {
            fused.push_back(ops[seed]);
            continue;
        }
        std::sort(g.qubits.begin(), g.qubits.end());
        Matrix m = Matrix::identity((unsigned)g.qubits.size());
        for (size_t i : g.members) {
            std::vector<size_t> oq;
            Matrix om;
            ops[i].full(oq, om);
            m = om.embed(oq, g.qubits).mul(m);
        }
        fused.push_back(Op{g.qubits, {}, m});
    }
    return fused;
}

static std::vector<Op> fuse(const std::vector<Op> &ops, size_t max_qubits) {
    return fuse_where(ops, max_qubits, [](const Op &) { return true; });
}

// ---------------------------------------------------------------------------
// Cache blocking (blocking.rs).
// ---------------------------------------------------------------------------
struct BlockingStats {
    size_t stages = 0, swaps = 0, kernels = 0;
};

template <typename T> struct RegionOp {
    size_t control_mask = 0, control_value = 0;
    std::vector<size_t> diagonal_global;
    std::vector<std::unique_ptr<Plan<T>>> plans;

    RegionOp(const Op &op, const std::vector<size_t> &layout, size_t local_limit, unsigned region_bits) {
        std::vector<std::pair<size_t, bool>> local_controls;
        for (auto &c : op.controls) {
            size_t p = layout[c.first];
            if (p >= local_limit) {
                control_mask |= size_t(1) << (p - local_limit);
                control_value |= (size_t)c.second << (p - local_limit);
            } else {
                local_controls.push_back({p, c.second});
            }
        }
        std::vector<size_t> physical;
        for (size_t q : op.targets) physical.push_back(layout[q]);
        bool all_local = true;
        for (size_t p : physical)
            if (p >= local_limit) all_local = false;
        if (all_local) {
            plans.emplace_back(new Plan<T>(op.matrix, physical, local_controls, region_bits));
            return;
        }
        // A diagonal gate with global targets: one local diagonal per
        // assignment of the global target bits (blocking.rs:91-126).
        std::vector<C64> diagonal = op.matrix.diagonal_entries();
        std::vector<size_t> global, local;
        for (size_t j = 0; j < physical.size(); j++) (physical[j] >= local_limit ? global : local).push_back(j);
        std::vector<size_t> local_targets;
        for (size_t j : local) local_targets.push_back(physical[j]);
        for (size_t j : global) diagonal_global.push_back(physical[j] - local_limit);
        for (size_t a = 0; a < (size_t(1) << global.size()); a++) {
            std::vector<C64> entries;
            for (size_t sub = 0; sub < (size_t(1) << local.size()); sub++) {
                size_t index = 0;
                for (size_t m = 0; m < global.size(); m++) index |= ((a >> m) & 1) << global[m];
                for (size_t m = 0; m < local.size(); m++) index |= ((sub >> m) & 1) << local[m];
                entries.push_back(diagonal[index]);
            }
            plans.emplace_back(new Plan<T>(Matrix::diagonal(entries), local_targets, local_controls, region_bits));
        }
    }
    void apply(size_t region, T *ptr) const {
        if ((region & control_mask) != control_value) return;
        size_t assignment = 0;
        for (size_t m = 0; m < diagonal_global.size(); m++) assignment |= ((region >> diagonal_global[m]) & 1) << m;
        plans[diagonal_global.empty() ? 0 : assignment]->apply(ptr, false);
    }
};

static bool runs_locally(const Op &op, const std::vector<size_t> &layout, size_t local_limit) {
    if (op.matrix.is_diagonal()) return true;
    for (size_t q : op.targets)
        if (layout[q] >= local_limit) return false;
    return true;
}

template <typename F>
static std::vector<size_t> extract(const std::vector<Op> &ops, const std::vector<bool> &done, F &&can_run) {
    std::vector<size_t> taken, blocked;
    for (size_t i = 0; i < ops.size(); i++) {
        if (done[i]) continue;
        std::vector<size_t> oq = ops[i].qubits();
        bool touches = false;
        for (size_t q : oq)
            if (contains(blocked, q)) touches = true;
        if (!touches && can_run(ops[i])) {
            taken.push_back(i);
        } else {
            for (size_t q : oq)
                if (!contains(blocked, q)) blocked.push_back(q);
        }
    }
    return taken;
}

static std::vector<std::pair<size_t, size_t>> plan_swaps(const std::vector<Op> &ops, const std::vector<bool> &done,
                                                         const std::vector<size_t> &layout, size_t lane_bits,
                                                         size_t local_limit) {
    const size_t capacity = local_limit - lane_bits;
    std::vector<size_t> wanted;
    extract(ops, done, [&](const Op &op) {
        if (op.matrix.is_diagonal()) return true;
        std::vector<size_t> fresh;
        for (size_t q : op.targets)
            if (layout[q] >= lane_bits && !contains(wanted, q)) fresh.push_back(q);
        if (wanted.size() + fresh.size() > capacity) return false;
        for (size_t q : fresh) wanted.push_back(q);
        return true;
    });
    std::vector<std::pair<size_t, size_t>> pairs;
    if (wanted.empty()) return pairs;
    std::vector<size_t> incoming, wanted_pos;
    for (size_t q : wanted) {
        wanted_pos.push_back(layout[q]);
        if (layout[q] >= local_limit) incoming.push_back(layout[q]);
    }
    size_t k = 0;
    for (size_t p = local_limit; p-- > lane_bits && k < incoming.size();) {
        if (contains(wanted_pos, p)) continue;
        pairs.push_back({p, incoming[k++]});
    }
    return pairs;
}

template <typename T>
BlockingStats apply_blocked(StateVector<T> &state, const std::vector<Op> &ops, unsigned region_bits, size_t max_fused) {
    const size_t lane_bits = Real<T>::LANE_BITS;
    const unsigned block_bits = state.block_bits();
    BlockingStats stats;
    if (block_bits <= region_bits + 1) {
        for (const Op &op : fuse(ops, max_fused)) {
            state.apply(op.matrix, op.targets, op.controls);
            stats.stages++;
            stats.kernels++;
        }
        return stats;
    }
    const size_t local_limit = lane_bits + region_bits;
    const size_t regions = size_t(1) << (block_bits - region_bits);
    const size_t region_len = (size_t(1) << region_bits) * 2 * Real<T>::LANES;
    std::vector<bool> done(ops.size(), false);
    size_t remaining = ops.size();
    while (remaining > 0) {
        std::vector<size_t> layout = state.layout;
        std::vector<size_t> taken = extract(ops, done, [&](const Op &op) { return runs_locally(op, layout, local_limit); });
        if (!taken.empty()) {
            std::vector<Op> stage_ops;
            for (size_t i : taken) {
                done[i] = true;
                stage_ops.push_back(ops[i]);
            }
            remaining -= taken.size();
            std::vector<Op> fused = fuse_where(stage_ops, max_fused, [&](const Op &op) {
                for (size_t q : op.qubits())
                    if (layout[q] >= local_limit) return false;
                return true;
            });
            stats.kernels += fused.size();
            std::vector<std::unique_ptr<RegionOp<T>>> stage;
            for (const Op &op : fused) stage.emplace_back(new RegionOp<T>(op, layout, local_limit, region_bits));
            T *base = state.raw();
#pragma omp parallel for schedule(dynamic, 1)
            for (long region = 0; region < (long)regions; region++) {
                T *ptr = base + (size_t)region * region_len;
                for (auto &op : stage) op->apply((size_t)region, ptr);
            }
            stats.stages++;
        }
        if (remaining == 0) break;
        auto pairs = plan_swaps(ops, done, layout, lane_bits, local_limit);
        if (pairs.empty()) {
            size_t i = 0;
            while (done[i]) i++;
            done[i] = true;
            remaining--;
            state.apply(ops[i].matrix, ops[i].targets, ops[i].controls);
            stats.stages++;
            stats.kernels++;
        } else {
            state.swap_physical(pairs);
            stats.swaps++;
        }
    }
    return stats;
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C API.
// ---------------------------------------------------------------------------
using namespace orc;

namespace {
thread_local std::string g_error;

struct Handle {
    int precision;  // 0 = f32, 1 = f64
    std::unique_ptr<StateVector<float>> f;
    std::unique_ptr<StateVector<double>> d;
};

template <typename Fn>
int guarded(Fn &&fn) {
    try {
        fn();
        return 0;
    } catch (const std::invalid_argument &e) {
        g_error = e.what();
        return -1;
    } catch (const std::bad_alloc &) {
        g_error = "out of memory";
        return -2;
    } catch (const std::exception &e) {
        g_error = e.what();
        return -3;
    }
}

Matrix read_matrix(const double *m, unsigned k) {
    size_t dim = size_t(1) << k;
    std::vector<C64> e(dim * dim);
    for (size_t i = 0; i < dim * dim; i++) e[i] = C64(m[2 * i], m[2 * i + 1]);
    return Matrix::from(std::move(e));
}

Op read_op(const orc_op &o) {
    Op op;
    for (uint32_t j = 0; j < o.num_targets; j++) op.targets.push_back(o.targets[j]);
    for (uint32_t j = 0; j < o.num_controls; j++)
        op.controls.push_back({o.controls[j], o.control_values ? o.control_values[j] != 0 : true});
    op.matrix = read_matrix(o.matrix, o.num_targets);
    return op;
}

template <typename F>
auto dispatch(void *h, F &&f) {
    Handle *s = (Handle *)h;
    if (!s) throw std::invalid_argument("null state");
    if (s->precision == 0) return f(*s->f);
    return f(*s->d);
}
}  // namespace

extern "C" {

const char *orc_last_error(void) { return g_error.c_str(); }

void orc_set_threads(int threads) { omp_set_num_threads(threads > 0 ? threads : 1); }
int orc_max_threads(void) { return omp_get_max_threads(); }

uint64_t orc_deposit(uint64_t value, uint64_t mask) { return deposit(value, mask); }

int orc_gate_matrix(const char *name, const double *params, uint32_t num_params, double *out, uint32_t capacity,
                    uint32_t *num_qubits) {
    return guarded([&] {
        Matrix m = gate_matrix(name, std::vector<double>(params, params + num_params));
        if (capacity < m.data.size()) throw std::invalid_argument("output too small");
        for (size_t i = 0; i < m.data.size(); i++) {
            out[2 * i] = m.data[i].real();
            out[2 * i + 1] = m.data[i].imag();
        }
        *num_qubits = m.qubits();
    });
}

// gate_op (simulator.rs:106-119): controls kept separate.
int orc_gate_controls(const char *name, uint32_t *num_controls) {
    return guarded([&] {
        unsigned c = 0;
        std::string b;
        if (!controlled_form(name, c, b)) c = 0;
        *num_controls = c;
    });
}

int orc_matrix_embed(const double *m, uint32_t k, const uint32_t *gate_qubits, const uint32_t *qubits, uint32_t n,
                     double *out) {
    return guarded([&] {
        Matrix a = read_matrix(m, k);
        Matrix e = a.embed(std::vector<size_t>(gate_qubits, gate_qubits + k), std::vector<size_t>(qubits, qubits + n));
        for (size_t i = 0; i < e.data.size(); i++) {
            out[2 * i] = e.data[i].real();
            out[2 * i + 1] = e.data[i].imag();
        }
    });
}

int orc_matrix_controlled(const double *m, uint32_t k, uint32_t controls, uint64_t state, double *out) {
    return guarded([&] {
        Matrix e = read_matrix(m, k).controlled(controls, state);
        for (size_t i = 0; i < e.data.size(); i++) {
            out[2 * i] = e.data[i].real();
            out[2 * i + 1] = e.data[i].imag();
        }
    });
}

// reference.rs:8-28 — gather, matrix-vector product, scatter; FP64.
int orc_reference_apply(double *state, uint32_t num_qubits, const double *matrix, const uint32_t *qubits, uint32_t k) {
    return guarded([&] {
        Matrix m = read_matrix(matrix, k);
        size_t mask = 0;
        for (uint32_t j = 0; j < k; j++) {
            if (qubits[j] >= num_qubits) throw std::invalid_argument("qubit out of range");
            mask |= size_t(1) << qubits[j];
        }
        C64 *s = (C64 *)state;
        size_t dim = size_t(1) << k;
        std::vector<size_t> off(dim);
        for (size_t sub = 0; sub < dim; sub++) {
            size_t o = 0;
            for (uint32_t j = 0; j < k; j++) o |= ((sub >> j) & 1) << qubits[j];
            off[sub] = o;
        }
        std::vector<C64> input(dim);
        const size_t len = size_t(1) << num_qubits;
        for (size_t base = 0; base < len; base++) {
            if (base & mask) continue;
            for (size_t sub = 0; sub < dim; sub++) input[sub] = s[base | off[sub]];
            for (size_t row = 0; row < dim; row++) {
                C64 acc(0.0, 0.0);
                for (size_t col = 0; col < dim; col++) acc += cmul(m.at(row, col), input[col]);
                s[base | off[row]] = acc;
            }
        }
    });
}

// Fusion exposed for its known-answer tests: the fused ops' targets (each
// at most 16 entries) and controls.
int orc_fuse(const orc_op *ops, uint64_t n, uint32_t max_qubits, uint32_t *out_count, uint32_t *out_targets,
             uint32_t *out_num_targets, uint32_t *out_num_controls, double *out_matrices, uint64_t matrix_capacity) {
    return guarded([&] {
        std::vector<Op> v;
        for (uint64_t i = 0; i < n; i++) v.push_back(read_op(ops[i]));
        std::vector<Op> f = fuse(v, max_qubits);
        *out_count = (uint32_t)f.size();
        uint64_t mo = 0;
        for (size_t i = 0; i < f.size(); i++) {
            if (f[i].targets.size() > 16) throw std::invalid_argument("fused op too wide for the output");
            out_num_targets[i] = (uint32_t)f[i].targets.size();
            out_num_controls[i] = (uint32_t)f[i].controls.size();
            for (size_t j = 0; j < f[i].targets.size(); j++) out_targets[16 * i + j] = (uint32_t)f[i].targets[j];
            if (out_matrices) {
                if (mo + 2 * f[i].matrix.data.size() > matrix_capacity) throw std::invalid_argument("matrix output too small");
                for (auto &e : f[i].matrix.data) {
                    out_matrices[mo++] = e.real();
                    out_matrices[mo++] = e.imag();
                }
            }
        }
    });
}

void *orc_sv_create(uint32_t num_qubits, int precision) {
    Handle *h = nullptr;
    int rc = guarded([&] {
        if (precision != 0 && precision != 1) throw std::invalid_argument("precision must be 0 (f32) or 1 (f64)");
        h = new Handle{precision, nullptr, nullptr};
        if (precision == 0) h->f.reset(new StateVector<float>(num_qubits));
        else h->d.reset(new StateVector<double>(num_qubits));
    });
    if (rc != 0) {
        delete h;
        return nullptr;
    }
    return h;
}

void orc_sv_destroy(void *h) { delete (Handle *)h; }

int orc_sv_set(void *h, const double *amps, uint64_t len) {
    return guarded([&] {
        dispatch(h, [&](auto &s) {
            if (len != s.len()) throw std::invalid_argument("amplitude count does not match the state");
            const C64 *a = (const C64 *)amps;
            const long n = (long)len;
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n; i++) s.set((size_t)i, a[i]);
            return 0;
        });
    });
}

int orc_sv_get(void *h, double *amps, uint64_t len) {
    return guarded([&] {
        dispatch(h, [&](auto &s) {
            if (len != s.len()) throw std::invalid_argument("amplitude count does not match the state");
            s.to_vec((C64 *)amps);
            return 0;
        });
    });
}

int orc_sv_apply(void *h, const double *matrix, const uint32_t *targets, uint32_t k, const uint32_t *controls,
                 const uint8_t *values, uint32_t c) {
    return guarded([&] {
        Matrix m = read_matrix(matrix, k);
        std::vector<size_t> t(targets, targets + k);
        std::vector<std::pair<size_t, bool>> cs;
        for (uint32_t j = 0; j < c; j++) cs.push_back({controls[j], values ? values[j] != 0 : true});
        dispatch(h, [&](auto &s) {
            s.apply(m, t, cs);
            return 0;
        });
    });
}

// apply_ops (simulator.rs:122-147); region_bits < 0 disables blocking.
int orc_sv_apply_ops(void *h, const orc_op *ops, uint64_t n, uint32_t max_fused_qubits, int32_t region_bits,
                     orc_stats *stats) {
    return guarded([&] {
        std::vector<Op> v;
        for (uint64_t i = 0; i < n; i++) v.push_back(read_op(ops[i]));
        auto t0 = std::chrono::steady_clock::now();
        size_t max_fused = std::max<uint32_t>(max_fused_qubits, 1);
        orc_stats st{};
        dispatch(h, [&](auto &s) {
            if (region_bits >= 0) {
                BlockingStats b = apply_blocked(s, v, (unsigned)region_bits, max_fused);
                st.passes = b.stages + b.swaps;
                st.kernels = b.kernels;
                st.swaps = b.swaps;
            } else {
                std::vector<Op> f = fuse(v, max_fused);
                for (auto &op : f) s.apply(op.matrix, op.targets, op.controls);
                st.passes = f.size();
                st.kernels = f.size();
            }
            return 0;
        });
        st.gates = n;
        st.gate_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (stats) *stats = st;
    });
}

int orc_sv_norm_sqr(void *h, double *out) {
    return guarded([&] { *out = dispatch(h, [&](auto &s) { return s.norm_sqr(); }); });
}

int orc_sv_scale(void *h, double factor) {
    return guarded([&] { dispatch(h, [&](auto &s) { s.scale(factor); return 0; }); });
}

int orc_sv_probability_one(void *h, uint32_t qubit, double *out) {
    return guarded([&] { *out = dispatch(h, [&](auto &s) { return s.probability_one(qubit); }); });
}

int orc_sv_probabilities(void *h, double *out, uint64_t len) {
    return guarded([&] {
        dispatch(h, [&](auto &s) {
            if (len != s.len()) throw std::invalid_argument("length does not match the state");
            s.probabilities(out);
            return 0;
        });
    });
}

int orc_sv_measure(void *h, uint32_t qubit, double uniform, uint8_t *outcome) {
    return guarded([&] { *outcome = dispatch(h, [&](auto &s) { return s.measure(qubit, uniform); }) ? 1 : 0; });
}

// measure.rs:83-87 — measure, then X on outcome 1.
int orc_sv_reset(void *h, uint32_t qubit, double uniform) {
    return guarded([&] {
        dispatch(h, [&](auto &s) {
            if (s.measure(qubit, uniform)) s.apply(gate_matrix("x", {}), {qubit}, {});
            return 0;
        });
    });
}

int orc_sv_sample(void *h, const double *uniforms, uint64_t shots, uint64_t *out) {
    return guarded([&] {
        dispatch(h, [&](auto &s) {
            std::vector<size_t> o = s.sample_unordered(uniforms, shots);
            for (size_t i = 0; i < o.size(); i++) out[i] = o[i];
            return 0;
        });
    });
}

int orc_sv_swap_physical(void *h, const uint32_t *pairs, uint32_t num_pairs) {
    return guarded([&] {
        std::vector<std::pair<size_t, size_t>> p;
        for (uint32_t i = 0; i < num_pairs; i++) p.push_back({pairs[2 * i], pairs[2 * i + 1]});
        dispatch(h, [&](auto &s) {
            s.swap_physical(p);
            return 0;
        });
    });
}

int orc_sv_layout(void *h, uint32_t *out) {
    return guarded([&] {
        dispatch(h, [&](auto &s) {
            for (size_t q = 0; q < s.layout.size(); q++) out[q] = (uint32_t)s.layout[q];
            return 0;
        });
    });
}

// Counts from outcomes (simulator.rs:187-238): integer keys over the
// measured (qubit, clbit) pairs, sorted; the caller formats the bitstrings
// (classical bit 0 on the right). Writes the distinct keys (low 64 bits,
// num_clbits <= 64) and their counts; returns the number of distinct keys
// through out_distinct.
int orc_count_outcomes(const uint64_t *outcomes, uint64_t shots, const uint32_t *measure_qubits,
                       const uint32_t *measure_clbits, uint32_t num_measures, uint64_t *out_keys,
                       uint64_t *out_counts, uint64_t *out_distinct) {
    return guarded([&] {
        std::vector<uint64_t> keys(shots);
        for (uint64_t s = 0; s < shots; s++) {
            uint64_t k = 0;
            for (uint32_t m = 0; m < num_measures; m++) {
                if (measure_clbits[m] >= 64) throw std::invalid_argument("more than 64 classical bits");
                k |= ((outcomes[s] >> measure_qubits[m]) & 1) << measure_clbits[m];
            }
            keys[s] = k;
        }
        std::sort(keys.begin(), keys.end());
        uint64_t d = 0;
        for (uint64_t s = 0; s < shots;) {
            uint64_t e = s;
            while (e < shots && keys[e] == keys[s]) e++;
            out_keys[d] = keys[s];
            out_counts[d] = e - s;
            d++;
            s = e;
        }
        *out_distinct = d;
    });
}

}  // extern "C"

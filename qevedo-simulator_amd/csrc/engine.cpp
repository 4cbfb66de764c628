// The C ABI of the state-vector engine (include/qvd_hip.h): state handles in
// HBM, single ops, planned op sequences, reductions, measurement and
// sampling. Host-side validation turns every assert of the reference into a
// status code.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <type_traits>
#include <unordered_map>

#include "common.h"
#include "dense_mfma.h"
#include "kernels.h"
#include "mlp.h"
#include "plan.h"
#include "vprog.h"
#include "vstage.h"

// Tiles per item of a batch's stage queue (at most).
#ifndef QVD_QUEUE_CHUNK
#define QVD_QUEUE_CHUNK 16
#endif

using namespace qvd;

namespace qvd {

thread_local std::string g_last_error;
void set_last_error(const std::string &m) { g_last_error = m; }

namespace {

// An owned device allocation.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    explicit DBuf(size_t b) { alloc(b); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t b) {
        if (p) {
            (void)hipFree(p);
            p = nullptr;
            bytes = 0;
        }
        if (b == 0) return;
        QVD_HIP(hipMalloc(&p, b));
        bytes = b;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

template <typename Fn>
int guard(Fn &&fn) {
    try {
        fn();
        return QVD_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("out of host memory");
        return QVD_ENOMEM;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return QVD_EINVAL;
    }
}

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace
}  // namespace qvd

struct qvd_sv {
    uint32_t n = 0;
    int precision = QVD_C64;
    int device = 0;
    uint64_t len = 0;
    DBuf amps;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;      // sampling: the uniforms' copy and sort (made by the first qvd_sv_sample)
    hipEvent_t side_done = nullptr;  // recorded after them
    void *sample_host = nullptr;     // pinned host staging of the outcomes, grown to the largest call and kept
    size_t sample_host_bytes = 0;
    DBuf scratch;    // small reduction buffers
    DBuf sample_ws;  // sampling workspace, grown to the largest call and kept
    DBuf plan_ws;    // a plan's uploaded programs, tables and matrices, grown to the largest plan and kept
    double last_sample_seconds = 0;
    std::vector<double> stage_seconds;  // register-tile stage kernels of the last timed apply_ops
    bool zero_pending = false;          // the state is |0...0> but the amplitudes are not written yet
    uint64_t zero_qubits = 0;           // qubits known to be |0> (every amplitude with the bit set is zero)
    uint64_t stale_qubits = 0;          // of those, qubits whose amplitudes have not been written since
                                        // |0...0> (memory stale); every other amplitude is written

    size_t elem() const { return precision == QVD_C64 ? sizeof(float2) : sizeof(double2); }
    void bind() const { QVD_HIP(hipSetDevice(device)); }
    void sync() const { QVD_HIP(hipStreamSynchronize(stream)); }
    double *scratch_doubles(size_t count) {
        if (scratch.bytes < count * sizeof(double)) scratch.alloc(std::max<size_t>(count * sizeof(double), 1 << 16));
        return scratch.as<double>();
    }
};

// A batch: every circuit's state back to back in one allocation, the
// register-tile programs of all stages, one launch table per stage level and
// the HIP graph that replays reset, levels and expectation values.
struct qvd_batch {
    int precision = QVD_C64;
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t count = 0;
    uint64_t len = 0;            // amplitudes of all states
    uint64_t num_z = 0, num_zz = 0;
    uint64_t gates = 0, kernels = 0, stages = 0;
    DBuf amps, ins, coefs, ldsops, ints, descs, items, sync, basis;
    DBuf edescs, seg_start, out_start, seg_sums, z, zz;
    uint32_t levels = 0, nitems = 0;
    unsigned grid = 0;
    std::vector<uint64_t> state_off;  // first amplitude of each circuit's state
    std::vector<uint32_t> qubits;     // each circuit's qubit count
    std::vector<uint32_t> fixed;      // each circuit's register-tile geometry (fixed lane bits)
    std::vector<uint32_t> circuit_stages;  // each circuit's number of stages
    bool init_states = false;         // the replay writes every state first (a circuit without stages, or one
                                      // whose stages leave stale qubits); otherwise the stages' zeros are lazy
    uint64_t stage_bytes = 0, other_bytes = 0, skipped_tiles = 0;  // per replay
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double *host_out = nullptr;  // pinned: a run's <Z>, <ZZ> and the queue's error word, copied back together
    ~qvd_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (stream) (void)hipStreamDestroy(stream);
        if (host_out) (void)hipHostFree(host_out);
    }
};

struct qvd_mlp {
    qvd::Mlp net;
    qvd_mlp(uint32_t in, uint32_t hidden) : net(in, hidden) {}
};

namespace qvd {
namespace {

template <typename F>
auto with_type(const qvd_sv *sv, F &&f) {
    if (sv->precision == QVD_C64) return f((float2 *)sv->amps.p);
    return f((double2 *)sv->amps.p);
}

// |0...0> is written lazily: qvd_sv_create and qvd_sv_set_zero only mark the
// state, the next qvd_sv_apply_ops whose plan starts with a register-tile
// stage makes that stage's tiles from |0...0> in registers (no init pass and
// no tile loads), and a plan of register-tile stages only may leave the
// amplitudes of stale qubits unwritten (run_plan). Any other use of the
// state writes the missing zeros first.
void materialize(qvd_sv *sv) {
    if (!sv->zero_pending && !sv->stale_qubits) return;
    sv->bind();
    with_type(sv, [&](auto *s) {
        if (sv->zero_pending) launch_init_zero(s, sv->len, sv->stream);
        else launch_zero_stale(s, sv->len, sv->stale_qubits, sv->stream);
        return 0;
    });
    QVD_HIP(hipGetLastError());
    sv->sync();
    sv->zero_pending = false;
    sv->stale_qubits = 0;
}

// Every entry point that uses a state's amplitudes starts here. It may change
// them, so the qubits known to be |0> are forgotten (qvd_sv_apply_ops keeps
// track of them itself).
void check_handle(const qvd_sv *sv) {
    require(sv != nullptr, "null state handle");
    materialize(const_cast<qvd_sv *>(sv));
    const_cast<qvd_sv *>(sv)->zero_qubits = 0;
}

uint64_t qubit_mask(uint32_t n) { return n >= 64 ? ~uint64_t(0) : (uint64_t(1) << n) - 1; }

// The qubits still known to be |0> after an op: a non-diagonal op may set
// its targets; a diagonal op, and a control, leave every bit as it is.
uint64_t after_op(uint64_t zero, const Op &op) {
    if (!op.diagonal)
        for (uint32_t q : op.targets) zero &= ~(uint64_t(1) << q);
    return zero;
}

// Tiles' worth of amplitudes a register-tile stage loads: none when it starts
// from |0...0>; each stale qubit among the bits its visited tiles span halves
// them (the tiles of skipped qubits are not visited).
uint64_t stage_read_tiles(const VStageDesc &d) {
    if (d.from_zero) return 0;
    uint64_t span = 0;
    for (uint32_t j = 0; j < d.L; j++) span |= uint64_t(1) << d.local_pos[j];
    for (uint32_t j = 0; j < d.nglob; j++) span |= uint64_t(1) << d.glob_pos[j];
    return d.num_tiles >> __builtin_popcountll(d.zero_mem & span);
}

// Drops from a stage's tile index the non-tile qubits known to be |0>: a tile
// with such a bit set holds only zeros, and the stage (dense ops on tile
// qubits only; diagonals and controls on the others) keeps them zero, so the
// stage neither reads nor writes it. Returns the number of tiles dropped.
uint64_t drop_zero_tiles(VStageDesc &d, uint64_t zero) {
    uint32_t k = 0;
    for (uint32_t j = 0; j < d.nglob; j++)
        if (!((zero >> d.glob_pos[j]) & 1)) d.glob_pos[k++] = d.glob_pos[j];
    const uint64_t before = d.num_tiles;
    d.nglob = k;
    d.num_tiles = uint64_t(1) << k;
    return before - d.num_tiles;
}

// Validates qubits of an op the way the reference asserts them
// (src/state.rs:182-188; python/src/lib.rs:29-43).
void check_op_qubits(uint32_t n, const std::vector<uint32_t> &targets,
                     const std::vector<std::pair<uint32_t, bool>> &controls) {
    std::vector<uint32_t> all = targets;
    for (auto &c : controls) all.push_back(c.first);
    for (uint32_t q : all) require(q < n, "qubit " + std::to_string(q) + " out of range for " + std::to_string(n) + " qubits");
    std::vector<uint32_t> s = all;
    std::sort(s.begin(), s.end());
    require(std::adjacent_find(s.begin(), s.end()) == s.end(), "an op cannot act on the same qubit twice");
}

template <typename T2> T2 to_t2(C64 v) {
    T2 r;
    r.x = (decltype(r.x))v.real();
    r.y = (decltype(r.y))v.imag();
    return r;
}

// Copies each 1-qubit gate's four coefficients (already in the state's
// precision) into the instruction, where the kernel's look-ahead fetch reads
// them with the rest of the instruction.
template <typename T2> void inline_coefficients(std::vector<VInstr> &ins, const std::vector<T2> &coefs) {
    static_assert(4 * sizeof(T2) <= sizeof(VInstr::inl), "inline coefficient block too small");
    for (VInstr &v : ins) {
        const uint32_t op = v.op & VOP_MASK;
        if ((op != V_U1R && op != V_U1L) || (size_t)v.coef + 4 > coefs.size()) continue;
        std::memcpy(v.inl, &coefs[v.coef], 4 * sizeof(T2));
    }
}

// Pair-phase runs (V_PHASES) read (a, b, w) per pair; the program keeps
// (a, b, coefficient index) triples. This appends each run's pairs as
// records (a, b, then w's words in the state's precision) to the integer
// table and points the instruction at them (inl[0]), so the kernel loads a
// pair in one scalar fetch with no dependent coefficient load.
template <typename T2>
void inline_phase_records(std::vector<VInstr> &ins, std::vector<uint32_t> &ints, const std::vector<T2> &coefs) {
    constexpr uint32_t W = sizeof(T2) / 4;
    for (VInstr &v : ins) {
        if ((v.op & VOP_MASK) != V_PHASES) continue;
        // Pairs with register bit r (always the first of the two) come
        // grouped by r as (b, w) records, inl[1 + r] of them (r <
        // VREG_SLOTS); the other pairs follow as (a, b, w) records,
        // inl[1 + VREG_SLOTS] of them, so the kernel walks each register
        // bit's pairs with that bit fixed at compile time.
        const uint32_t off = (uint32_t)ints.size();
        uint32_t counts[VREG_SLOTS + 1] = {};
        for (uint32_t pass = 0; pass <= VREG_SLOTS; pass++) {
            for (uint32_t i = 0; i < v.k; i++) {
                const uint32_t a = ints[v.aux + 3 * i], b = ints[v.aux + 3 * i + 1], ci = ints[v.aux + 3 * i + 2];
                require(ci < coefs.size(), "internal: pair phase index out of range");
                const bool reg = (a & VT_KIND) == VT_REG;
                require(!reg || (a & 0xffffu) < VREG_SLOTS, "internal: register bit out of range in a phase run");
                if (reg ? (a & 0xffffu) != pass : pass != VREG_SLOTS) continue;
                uint32_t w[W];
                std::memcpy(w, &coefs[ci], sizeof(T2));
                if (!reg) ints.push_back(a);
                ints.push_back(b);
                ints.insert(ints.end(), w, w + W);
                counts[pass]++;
            }
        }
        v.inl[0] = off;
        for (uint32_t r = 0; r <= VREG_SLOTS; r++) v.inl[1 + r] = counts[r];
    }
}

// One op as its own HBM pass: the dense kernel over the controlled subspace,
// or the diagonal kernel.
int g_num_cus = 0;
int num_cus(int device) {
    if (g_num_cus == 0) {
        hipDeviceProp_t p;
        QVD_HIP(hipGetDeviceProperties(&p, device));
        g_num_cus = p.multiProcessorCount;
    }
    return g_num_cus;
}

// The coefficients a full-state op's kernel reads: the diagonal of a
// diagonal op, else the row-major matrix, in the state's precision.
template <typename T2>
std::vector<T2> full_coefficients(const Op &op) {
    std::vector<T2> c;
    if (op.diagonal) {
        c.resize(op.m.dim());
        for (size_t i = 0; i < c.size(); i++) c[i] = to_t2<T2>(op.m.at(i, i));
    } else {
        c.resize(op.m.a.size());
        for (size_t i = 0; i < c.size(); i++) c[i] = to_t2<T2>(op.m.a[i]);
    }
    return c;
}

// One op as its own HBM pass, its coefficients (full_coefficients) already
// in device memory at `coef`; enqueued on the handle's stream, not waited for.
template <typename T2>
void apply_full(qvd_sv *sv, T2 *s, const Op &op, const T2 *coef, bool valu_only = false) {
    const uint32_t k = (uint32_t)op.targets.size();
    if (k == 0) return;
    if (k > MAX_GATE_QUBITS)
        fail(QVD_EUNSUPPORTED, "ops on more than " + std::to_string(MAX_GATE_QUBITS) + " target qubits are not supported");
    uint64_t cmask = 0, cval = 0;
    for (auto &c : op.controls) {
        cmask |= uint64_t(1) << c.first;
        if (c.second) cval |= uint64_t(1) << c.first;
    }
    if (op.diagonal) {
        DiagTargets t{};
        for (uint32_t j = 0; j < k; j++) t.q[j] = op.targets[j];
        launch_diag<T2>(s, sv->len, k, t, cmask, cval, coef, sv->stream);
        QVD_HIP(hipGetLastError());
        return;
    }
    GateDesc d{};
    d.k = k;
    std::vector<uint32_t> ins = op.targets;
    for (auto &c : op.controls) ins.push_back(c.first);
    std::sort(ins.begin(), ins.end());
    d.nins = (uint32_t)ins.size();
    for (size_t j = 0; j < ins.size(); j++) d.ins[j] = ins[j];
    for (uint32_t j = 0; j < (1u << k); j++) {
        uint64_t o = 0;
        for (uint32_t b = 0; b < k; b++) o |= uint64_t((j >> b) & 1) << op.targets[b];
        d.offsets[j] = o;
    }
    d.ctrl_val = cval;
    d.iterations = uint64_t(1) << (sv->n - d.nins);
    if constexpr (sizeof(T2) == sizeof(float2)) {
        // 5- and 6-qubit gates in complex64: FP32 matrix cores.
        if ((k == 5 || k == 6) && !valu_only) {
            QVD_HIP(launch_dense_mfma(s, d, coef, (unsigned)num_cus(sv->device), sv->stream));
            return;
        }
    }
    launch_dense<T2>(s, d, coef, sv->stream);
    QVD_HIP(hipGetLastError());
}

// Host data packed for one upload: each part 256-byte aligned.
struct Upload {
    std::vector<unsigned char> bytes;
    size_t put(const void *p, size_t n) {
        const size_t at = (bytes.size() + 255) & ~size_t(255);
        bytes.resize(at + std::max<size_t>(n, 1));
        if (n) std::memcpy(bytes.data() + at, p, n);
        return at;
    }
    template <typename T> size_t put(const std::vector<T> &v) { return put(v.data(), v.size() * sizeof(T)); }
    // Copies everything into the handle's plan arena (grown when needed) on
    // its stream; returns the arena's base.
    char *to_device(qvd_sv *sv) {
        if (sv->plan_ws.bytes < bytes.size()) {
            sv->sync();  // the arena may still feed a kernel of the previous call
            sv->plan_ws.alloc(std::max(bytes.size(), sv->plan_ws.bytes * 2));
        }
        if (!bytes.empty())
            QVD_HIP(hipMemcpyAsync(sv->plan_ws.p, bytes.data(), bytes.size(), hipMemcpyHostToDevice, sv->stream));
        return sv->plan_ws.as<char>();
    }
};

Op make_op(uint32_t n, const double *matrix, const uint32_t *targets, uint32_t k, const uint32_t *controls,
           const uint8_t *values, uint32_t c) {
    require(k == 0 || (targets && matrix), "null targets or matrix");
    require(c == 0 || controls, "null controls");
    Op op;
    op.targets.assign(targets, targets + k);
    for (uint32_t j = 0; j < c; j++) op.controls.push_back({controls[j], values ? values[j] != 0 : true});
    check_op_qubits(n, op.targets, op.controls);
    require(k <= 20, "op too wide");
    if (k > 0) op.m = HMatrix::from_interleaved(matrix, k);
    else op.m = HMatrix::identity(0);
    op.diagonal = op.m.is_diagonal();
    return op;
}

// Tile sizes: 2^13 amplitudes (64 KiB of LDS) in complex64, 2^12 in
// complex128; the lowest 5 / 4 qubits always in the tile (256-byte runs).
struct TileConfig {
    uint32_t L, contig;
};
TileConfig default_tiles(int precision) { return precision == QVD_C64 ? TileConfig{13, 5} : TileConfig{12, 4}; }

// Lowers a planned stage into device descriptors.
template <typename T2>
void lower_stage(const Stage &st, uint32_t n, std::vector<StageOp> &ops, std::vector<T2> &coefs, StageDesc &d) {
    std::vector<int> tile_bit(n, -1);
    for (size_t j = 0; j < st.local.size(); j++) tile_bit[st.local[j]] = (int)j;
    d = StageDesc{};
    d.L = (uint32_t)st.local.size();
    for (uint32_t j = 0; j < d.L; j++) d.local_pos[j] = st.local[j];
    d.nglob = 0;
    for (uint32_t q = 0; q < n; q++)
        if (tile_bit[q] < 0) d.glob_pos[d.nglob++] = q;
    d.num_tiles = uint64_t(1) << d.nglob;
    d.op_off = (uint32_t)ops.size();
    d.num_ops = 0;
    for (const Op &op : st.ops) {
        if (op.targets.empty()) continue;
        StageOp so{};
        so.k = (uint32_t)op.targets.size();
        std::vector<uint32_t> ins;
        for (auto &c : op.controls) {
            if (tile_bit[c.first] >= 0) {
                so.lc_mask |= 1u << tile_bit[c.first];
                if (c.second) so.lc_val |= 1u << tile_bit[c.first];
                ins.push_back((uint32_t)tile_bit[c.first]);
            } else {
                so.gc_mask |= uint64_t(1) << c.first;
                if (c.second) so.gc_val |= uint64_t(1) << c.first;
            }
        }
        so.coef_off = (uint32_t)coefs.size();
        if (op.diagonal) {
            so.kind = OP_DIAG;
            for (uint32_t j = 0; j < so.k; j++) {
                const uint32_t q = op.targets[j];
                so.tgt[j] = tile_bit[q] >= 0 ? (uint32_t)tile_bit[q] : (GLOBAL_BIT | q);
            }
            for (size_t i = 0; i < op.m.dim(); i++) coefs.push_back(to_t2<T2>(op.m.at(i, i)));
        } else {
            so.kind = OP_DENSE;
            for (uint32_t j = 0; j < so.k; j++) {
                const int b = tile_bit[op.targets[j]];
                require(b >= 0, "internal: dense stage op on a non-tile qubit");
                so.tgt[j] = (uint32_t)b;
                ins.push_back((uint32_t)b);
            }
            for (const C64 &v : op.m.a) coefs.push_back(to_t2<T2>(v));
        }
        std::sort(ins.begin(), ins.end());
        so.nins = (uint32_t)ins.size();
        for (size_t j = 0; j < ins.size(); j++) so.ins[j] = ins[j];
        ops.push_back(so);
        d.num_ops++;
    }
}


// Events around kernel launches on the handle's stream (QVD_FLAG_TIMING).
struct LaunchTimer {
    bool on = false;
    hipStream_t stream = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> stage, other;
    void begin(std::vector<std::pair<hipEvent_t, hipEvent_t>> &v) {
        if (!on) return;
        hipEvent_t a, b;
        QVD_HIP(hipEventCreate(&a));
        QVD_HIP(hipEventCreate(&b));
        v.push_back({a, b});
        QVD_HIP(hipEventRecord(a, stream));
    }
    void end(std::vector<std::pair<hipEvent_t, hipEvent_t>> &v) {
        if (on) QVD_HIP(hipEventRecord(v.back().second, stream));
    }
    static double total(std::vector<std::pair<hipEvent_t, hipEvent_t>> &v) {
        double t = 0;
        for (auto &e : v) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) t += ms * 1e-3;
        }
        return t;
    }
    ~LaunchTimer() {
        for (auto *v : {&stage, &other})
            for (auto &e : *v) {
                (void)hipEventDestroy(e.first);
                (void)hipEventDestroy(e.second);
            }
    }
};

// Runs a plan on the state. zero: the qubits known to be |0> before it (all
// of them when the state is a pending |0...0>); returns those after it.
template <typename T2>
uint64_t run_plan(qvd_sv *sv, T2 *s, const std::vector<Step> &steps, qvd_stats &st, uint32_t flags, uint64_t zero,
                  uint32_t fixed_lanes) {
    const bool timing = (flags & QVD_FLAG_TIMING) != 0;
    LaunchTimer timer;
    timer.on = timing;
    timer.stream = sv->stream;
    // Lower every stage first so one upload serves the whole plan.
    std::vector<StageOp> sops;
    std::vector<T2> coefs;
    std::vector<StageDesc> descs;
    // Register-tile stages: programs, their coefficients (complex128 until
    // upload), LDS ops and integer tables.
    std::vector<VInstr> vins;
    std::vector<C64> vco;
    std::vector<StageOp> vlds;
    std::vector<uint32_t> vints;
    std::vector<VStageDesc> vdescs;
    PendingDiag pend(sv->n);
    VGeom geom = vstage_geom<T2>();
    if (fixed_lanes) geom.f_bits = fixed_lanes;
    const bool skip_zero = (flags & QVD_FLAG_ALL_TILES) == 0;
    const uint64_t state_bytes = sv->len * sizeof(T2);
    // Amplitudes not written yet: all of them (a pending |0...0>), or those
    // with a bit of a stale qubit set. A plan of register-tile stages only
    // carries them: its first stage makes a pending |0...0> in registers,
    // each stage sets the amplitudes of the stale qubits to zero instead of
    // loading them and writes every tile it visits, and the qubits whose
    // tiles it skips (known to be |0>) are the stale ones after it. Another
    // plan makes a pending |0...0> in its first stage when that is a
    // register-tile stage visiting every tile; otherwise, and for stale
    // qubits, the zeros are written first.
    bool all_reg = true;
    for (const Step &step : steps) all_reg = all_reg && step.is_stage && step.stage.reg_tile;
    bool pending = sv->zero_pending;
    uint64_t stale = pending ? 0 : sv->stale_qubits;
    bool write_zeros = false;
    if ((pending && !(steps[0].is_stage && steps[0].stage.reg_tile)) || (!all_reg && stale)) {
        write_zeros = true;
        pending = false;
        stale = 0;
    }
    for (const Step &step : steps) {
        if (!step.is_stage) {
            zero = after_op(zero, step.op);
            st.other_bytes += 2 * state_bytes;
            continue;
        }
        if (step.stage.reg_tile) {
            VProgram p = build_vprogram(step.stage, sv->n, geom, pend, vco, vlds, vints, true);
            require(p.desc.ccount * sizeof(T2) <= VCONST_BYTES, "internal: stage coefficients exceed the LDS budget");
            p.desc.instr_off = (uint32_t)vins.size();
            uint64_t glob = 0;  // the stage's non-tile qubits
            for (uint32_t j = 0; j < p.desc.nglob; j++) glob |= uint64_t(1) << p.desc.glob_pos[j];
            const uint64_t dropped = skip_zero ? drop_zero_tiles(p.desc, zero) : 0;
            st.skipped_tiles += dropped;
            if (pending) {
                // the first step
                pending = false;
                if (all_reg || dropped == 0) p.desc.from_zero = 1;
                else write_zeros = true;
            } else {
                p.desc.zero_mem = stale;
            }
            if (all_reg) stale = skip_zero ? glob & zero : 0;
            st.stage_bytes += (sizeof(T2) << p.desc.L) * (p.desc.num_tiles + stage_read_tiles(p.desc));
            vins.insert(vins.end(), p.instr.begin(), p.instr.end());
            vdescs.push_back(p.desc);
        } else {
            StageDesc d;
            lower_stage<T2>(step.stage, sv->n, sops, coefs, d);
            descs.push_back(d);
            st.stage_bytes += 2 * state_bytes;
        }
        for (const Op &op : step.stage.ops) zero = after_op(zero, op);
    }
    if (write_zeros) {
        const uint64_t unwritten = sv->zero_pending ? sv->len : sv->len - (sv->len >> __builtin_popcountll(sv->stale_qubits));
        materialize(sv);
        st.other_bytes += unwritten * sizeof(T2);
    }
    std::vector<T2> vco_t(vco.size());
    for (size_t i = 0; i < vco.size(); i++) vco_t[i] = to_t2<T2>(vco[i]);
    inline_coefficients(vins, vco_t);
    inline_phase_records(vins, vints, vco_t);
    // Everything the launches read goes up in one copy into the handle's
    // plan arena (kept between calls), full-state ops' coefficients included.
    Upload up;
    const size_t o_ops = up.put(sops), o_coef = up.put(coefs), o_vins = up.put(vins), o_vco = up.put(vco_t);
    const size_t o_vlds = up.put(vlds), o_vints = up.put(vints), o_vdescs = up.put(vdescs);
    std::vector<size_t> o_full;
    for (const Step &step : steps)
        if (!step.is_stage) o_full.push_back(up.put(full_coefficients<T2>(step.op)));
    char *base = up.to_device(sv);
    sv->zero_pending = false;  // written above, or made by the first stage below
    sv->stale_qubits = stale;
    const StageOp *dops = reinterpret_cast<const StageOp *>(base + o_ops);
    const T2 *dcoef = reinterpret_cast<const T2 *>(base + o_coef);
    const VInstr *dvins = reinterpret_cast<const VInstr *>(base + o_vins);
    const T2 *dvco = reinterpret_cast<const T2 *>(base + o_vco);
    const StageOp *dvlds = reinterpret_cast<const StageOp *>(base + o_vlds);
    const uint32_t *dvints = reinterpret_cast<const uint32_t *>(base + o_vints);
    const VStageDesc *dvdescs = reinterpret_cast<const VStageDesc *>(base + o_vdescs);
    size_t fi = 0;
    const int cus = num_cus(sv->device);
    size_t si = 0, vi = 0;
    std::vector<char> reg_launch;  // per timed stage launch: register-tile or not
    for (const Step &step : steps) {
        if (step.is_stage && step.stage.reg_tile) {
            const VStageDesc &d = vdescs[vi];
            st.kernels += step.stage.ops.size();
            const unsigned grid = (unsigned)std::min<uint64_t>(d.num_tiles, (uint64_t)cus);
            timer.begin(timer.stage);
            reg_launch.push_back(1);
            QVD_HIP(launch_vstage<T2>(s, dvdescs + vi, 1, nullptr, dvins, dvco, dvlds, dvints, grid, sv->stream,
                                      d.from_zero != 0 || d.zero_mem != 0, geom.f_bits));
            vi++;
            timer.end(timer.stage);
            st.stages++;
            st.passes++;
        } else if (step.is_stage) {
            const StageDesc &d = descs[si++];
            st.kernels += d.num_ops;
            if (d.num_ops == 0) continue;
            const size_t lds = stage_lds_bytes<T2>(d.L);
            const unsigned per_cu = lds <= 80 * 1024 ? 2 : 1;
            const uint64_t want = (uint64_t)cus * per_cu * 2;
            const unsigned grid = (unsigned)std::min<uint64_t>(d.num_tiles, want);
            timer.begin(timer.stage);
            reg_launch.push_back(0);
            launch_stage<T2>(s, d, dops, dcoef, grid, sv->stream);
            QVD_HIP(hipGetLastError());
            timer.end(timer.stage);
            st.stages++;
            st.passes++;
        } else {
            timer.begin(timer.other);
            apply_full<T2>(sv, s, step.op, reinterpret_cast<const T2 *>(base + o_full[fi++]),
                           (flags & QVD_FLAG_DENSE_VALU) != 0);
            timer.end(timer.other);
            st.kernels++;
            st.passes++;
        }
    }
    sv->sync();
    if (timing) {
        st.stage_kernel_seconds = LaunchTimer::total(timer.stage);
        st.other_kernel_seconds = LaunchTimer::total(timer.other);
        sv->stage_seconds.clear();
        for (size_t i = 0; i < timer.stage.size(); i++) {
            if (!reg_launch[i]) continue;
            float ms = 0;
            QVD_HIP(hipEventElapsedTime(&ms, timer.stage[i].first, timer.stage[i].second));
            sv->stage_seconds.push_back(ms * 1e-3);
        }
    }
    return zero;
}

uint64_t state_cap_bytes() {
    const char *e = std::getenv("QVD_MAX_STATE_BYTES");
    if (!e || !*e) return 0;
    return std::strtoull(e, nullptr, 10);
}

// Coefficient bytes the register-tile program of one stage needs in LDS
// (every stage starts with no pending diagonals: the previous one ends with
// a scale that applies them).
size_t stage_coef_bytes(const Stage &st, uint32_t n, int precision, const VGeom &vg) {
    PendingDiag pend(n);
    std::vector<C64> co;
    std::vector<StageOp> lds;
    std::vector<uint32_t> ints;
    const VProgram p = build_vprogram(st, n, vg, pend, co, lds, ints, true);
    return (size_t)p.desc.ccount * (precision == QVD_C64 ? sizeof(float2) : sizeof(double2));
}

// Splits register-tile stages whose programs would need more coefficients
// than the LDS block of one stage holds. Each stage is built as the kernel
// will run it; one that does not fit is cut after the longest run of its
// ops whose program fits (found by bisection), and the rest is split the
// same way (each part is a pass over the state).
std::vector<Step> split_for_coefficients(std::vector<Step> steps, uint32_t n, int precision, const VGeom &vg) {
    std::vector<Step> out;
    for (Step &s : steps) {
        if (!s.is_stage || !s.stage.reg_tile || stage_coef_bytes(s.stage, n, precision, vg) <= VCONST_BYTES) {
            out.push_back(std::move(s));
            continue;
        }
        std::vector<Op> rest = std::move(s.stage.ops);
        while (!rest.empty()) {
            Step part;
            part.stage.local = s.stage.local;
            part.stage.reg_tile = true;
            part.stage.role_policy = s.stage.role_policy;
            auto fits = [&](size_t k) {
                part.stage.ops.assign(rest.begin(), rest.begin() + (ptrdiff_t)k);
                return stage_coef_bytes(part.stage, n, precision, vg) <= VCONST_BYTES;
            };
            size_t lo = 1, hi = rest.size();
            require(fits(1), "internal: one op's coefficients exceed the LDS budget of a stage");
            while (lo < hi) {
                const size_t mid = (lo + hi + 1) / 2;
                if (fits(mid)) lo = mid;
                else hi = mid - 1;
            }
            part.stage.ops.assign(rest.begin(), rest.begin() + (ptrdiff_t)lo);
            rest.erase(rest.begin(), rest.begin() + (ptrdiff_t)lo);
            out.push_back(std::move(part));
        }
    }
    return out;
}

// Everything the device computes with when it runs a plan, as bytes: each
// register-tile stage's program (instructions, stage descriptor,
// coefficients, integer tables and LDS ops) and each full-state op (targets,
// controls, matrix). The kernels are deterministic, so two plans with equal
// images give bit-identical amplitudes; plans that apply two gates on
// different qubits in a different order round differently and have
// different images.
std::vector<uint8_t> plan_image(const std::vector<Step> &steps, uint32_t n, int precision, uint32_t fixed_lanes = 0) {
    VGeom vg = precision == QVD_C64 ? vstage_geom<float2>() : vstage_geom<double2>();
    if (fixed_lanes) vg.f_bits = fixed_lanes;
    std::vector<uint8_t> img;
    auto put = [&](const void *p, size_t bytes) {
        const uint8_t *b = static_cast<const uint8_t *>(p);
        img.insert(img.end(), b, b + bytes);
    };
    auto put32 = [&](uint32_t x) { put(&x, sizeof x); };
    PendingDiag pend(n);
    for (const Step &step : steps) {
        if (!step.is_stage) {
            put32(0xffffffffu);
            put32((uint32_t)step.op.targets.size());
            put(step.op.targets.data(), step.op.targets.size() * sizeof(uint32_t));
            put32((uint32_t)step.op.controls.size());
            for (auto &c : step.op.controls) {
                put32(c.first);
                put32(c.second ? 1u : 0u);
            }
            put(step.op.m.a.data(), step.op.m.a.size() * sizeof(C64));
            continue;
        }
        require(step.stage.reg_tile, "internal: plan images cover register-tile plans");
        std::vector<C64> co;
        std::vector<StageOp> lds;
        std::vector<uint32_t> ints;
        const VProgram p = build_vprogram(step.stage, n, vg, pend, co, lds, ints, true);
        put32((uint32_t)p.instr.size());
        // the fields up to the last role map (no padding inside that range)
        for (const VInstr &v : p.instr) put(&v, offsetof(VInstr, rmap) + sizeof v.rmap);
        const VStageDesc &d = p.desc;
        put(&d, offsetof(VStageDesc, instr_off));
        put(&d.num_tiles, offsetof(VStageDesc, cbegin) - offsetof(VStageDesc, num_tiles));
        put32(d.ccount);
        put(co.data(), co.size() * sizeof(C64));
        put32((uint32_t)ints.size());
        put(ints.data(), ints.size() * sizeof(uint32_t));
        put32((uint32_t)lds.size());
        for (const StageOp &o : lds) {
            put(&o, offsetof(StageOp, gc_mask));  // kind .. lc_val
            put(&o.gc_mask, sizeof o.gc_mask + sizeof o.gc_val + sizeof o.coef_off);
        }
    }
    return img;
}

// Features of each register-tile stage of a plan, in launch order: what the
// stage's program does per tile (counts of each instruction kind, in
// log(1 + count)), the state size and the precision.
std::vector<float> plan_features(const std::vector<Step> &steps, uint32_t n, int precision, uint32_t fixed_lanes = 0) {
    VGeom vg = precision == QVD_C64 ? vstage_geom<float2>() : vstage_geom<double2>();
    if (fixed_lanes) vg.f_bits = fixed_lanes;
    std::vector<float> rows;
    PendingDiag pend(n);
    std::vector<C64> co;
    std::vector<StageOp> lds;
    std::vector<uint32_t> ints;
    for (const Step &step : steps) {
        if (!step.is_stage || !step.stage.reg_tile) continue;
        const size_t lds0 = lds.size();
        VProgram p = build_vprogram(step.stage, n, vg, pend, co, lds, ints, true);
        double f[QVD_PLAN_FEATURES] = {};
        f[0] = n / 32.0;
        f[1] = precision == QVD_C128 ? 1.0 : 0.0;
        for (const VInstr &v : p.instr) {
            const uint32_t op = v.op & VOP_MASK;
            const bool ctrl = (v.op & VF_CTRL) != 0;
            if (v.op & VF_GCTRL) f[22]++;
            f[23]++;
            switch (op) {
                case V_U1R:
                    if (ctrl || v.form == FORM_FULL) f[5]++;
                    else if (v.form == FORM_UNIT_ROT || v.form == FORM_UNIT_IMAG || v.form == FORM_UNIT_REAL) f[2]++;
                    else if (v.form == FORM_UNIT_DIAG) f[3]++;
                    else f[4]++;
                    break;
                case V_U1L:
                    if (ctrl) f[9]++;
                    else if (v.slot < 4) f[6]++;
                    else if (v.form == FORM_UNIT_ROT || v.form == FORM_UNIT_IMAG || v.form == FORM_UNIT_REAL) f[7]++;
                    else f[8]++;
                    break;
                case V_DIAG:
                    f[10]++;
                    if (v.threaded) f[11]++;
                    break;
                case V_SIGNS:
                    f[12]++;
                    // pairs through the generic loop, plus the masks evaluated
                    f[13] += v.k + ((v.form & 1) ? 1 : 0);
                    for (int t = 0; t < 10; t++) f[13] += v.inl[4 + t] != 0 ? 1 : 0;
                    if (v.neg) f[14]++;
                    break;
                case V_PHASES:
                    f[15]++;
                    f[16] += v.k;
                    if (v.form == 1) f[17]++;
                    break;
                case V_XCHG: f[18]++; break;
                case V_LDSOP:
                    f[19]++;
                    f[20] += double(1u << (2 * lds[v.aux].k));
                    break;
                case V_SCALE:
                    if (v.form != 1) f[21]++;
                    break;
                default: break;
            }
        }
        (void)lds0;
        for (int k = 2; k < QVD_PLAN_FEATURES; k++) f[k] = std::log1p(f[k]);
        for (int k = 0; k < QVD_PLAN_FEATURES; k++) rows.push_back((float)f[k]);
    }
    return rows;
}

// The planner's cost model, when one is installed (qvd_planner_set_model):
// predicted log(seconds per amplitude) of a register-tile stage from its
// features, with the feature and target normalisation of its training set.
struct PlannerModel {
    std::mutex mu;
    std::unique_ptr<Mlp> net;
    std::vector<float> mean, scale;
    float ymean = 0, yscale = 1;
    // The variant chosen for each op sequence seen since the model was
    // installed (hash of the state size, precision, options and ops).
    std::unordered_map<uint64_t, uint32_t> chosen;
};
PlannerModel g_model;

uint64_t fnv(uint64_t h, const void *p, size_t bytes) {
    const unsigned char *c = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < bytes; i++) h = (h ^ c[i]) * 1099511628211ull;
    return h;
}

uint64_t ops_hash(uint32_t n, int precision, const qvd_options &opt, const qvd_op *ops, uint64_t num_ops,
                  uint32_t fixed_lanes) {
    uint64_t h = 1469598103934665603ull;
    h = fnv(h, &n, sizeof n);
    h = fnv(h, &fixed_lanes, sizeof fixed_lanes);
    h = fnv(h, &precision, sizeof precision);
    h = fnv(h, &opt.max_fused_qubits, sizeof opt.max_fused_qubits);
    h = fnv(h, &opt.tile_bits, sizeof opt.tile_bits);
    // the candidate set: every variant, or those equal to the base variant's plan
    h = fnv(h, &opt.planner, sizeof opt.planner);
    h = fnv(h, &opt.variant, sizeof opt.variant);
    for (uint64_t i = 0; i < num_ops; i++) {
        const qvd_op &o = ops[i];
        h = fnv(h, &o.num_targets, sizeof o.num_targets);
        h = fnv(h, o.targets, o.num_targets * sizeof(uint32_t));
        h = fnv(h, &o.num_controls, sizeof o.num_controls);
        if (o.num_controls) {
            h = fnv(h, o.controls, o.num_controls * sizeof(uint32_t));
            h = fnv(h, o.control_values, o.num_controls);
        }
        h = fnv(h, o.matrix, (size_t(2) << (2 * o.num_targets)) * sizeof(double));
    }
    return h;
}

// A plan's identity for de-duplicating variants: its steps' tile qubits,
// op counts and role policies.
std::vector<uint32_t> plan_signature(const std::vector<Step> &steps) {
    std::vector<uint32_t> sig;
    for (const Step &st : steps) {
        sig.push_back(st.is_stage ? (uint32_t)st.stage.ops.size() : 0xffffffffu);
        if (!st.is_stage) continue;
        sig.push_back(st.stage.role_policy);
        sig.insert(sig.end(), st.stage.local.begin(), st.stage.local.end());
    }
    return sig;
}

// The variants whose plans have the same image as variant `base` (itself
// included): the plans that reproduce its amplitudes bit for bit.
std::vector<uint32_t> equivalent_variants(const std::vector<std::vector<Step>> &plans, uint32_t base, uint32_t n,
                                          int precision, uint32_t fixed_lanes = 0) {
    std::vector<std::vector<uint8_t>> img(plans.size());
    std::vector<std::thread> workers;
    std::vector<std::exception_ptr> errors(plans.size());
    for (size_t var = 0; var < plans.size(); var++)
        workers.emplace_back([&, var] {
            try {
                img[var] = plan_image(plans[var], n, precision, fixed_lanes);
            } catch (...) {
                errors[var] = std::current_exception();
            }
        });
    for (auto &w : workers) w.join();
    for (auto &e : errors)
        if (e) std::rethrow_exception(e);
    std::vector<uint32_t> out;
    for (uint32_t var = 0; var < plans.size(); var++)
        if (img[var] == img[base]) out.push_back(var);
    return out;
}

// Validates the ops and plans them for a state of n qubits in `precision`:
// the heuristic (options.variant picks a variant of it), or, with
// QVD_PLANNER_LEARNED and a cost model installed, the variant whose stages
// the model predicts to run fastest (the heuristic when no model is set).
std::vector<Step> make_plan(uint32_t n, int precision, const qvd_op *ops, uint64_t num_ops, const qvd_options &opt,
                            uint32_t *chosen = nullptr, uint32_t fixed_lanes = 0) {
    require(num_ops == 0 || ops != nullptr, "null ops");
    std::vector<Op> v;
    v.reserve(num_ops);
    for (uint64_t i = 0; i < num_ops; i++) {
        const qvd_op &o = ops[i];
        Op op = make_op(n, o.matrix, o.targets, o.num_targets, o.controls, o.control_values, o.num_controls);
        if (op.targets.size() > MAX_GATE_QUBITS)
            fail(QVD_EUNSUPPORTED, "ops on more than " + std::to_string(MAX_GATE_QUBITS) + " target qubits are not supported");
        v.push_back(std::move(op));
    }
    if (chosen) *chosen = 0;
    if (opt.tile_bits < 0) return plan_unblocked(v, opt.max_fused_qubits);
    TileConfig tc = default_tiles(precision);
    VGeom vg = precision == QVD_C64 ? vstage_geom<float2>() : vstage_geom<double2>();
    if (!fixed_lanes) fixed_lanes = opt.fixed_lanes;  // a batch's per-circuit geometry, or the caller's
    if (fixed_lanes) {
        require(fixed_lanes + 1 >= vg.f_bits && fixed_lanes <= vg.f_bits + 1,
                "fixed_lanes must be 0 or " + std::to_string(vg.f_bits - 1) + ".." + std::to_string(vg.f_bits + 1));
        vg.f_bits = fixed_lanes;
    }
    PlanParams p;
    p.num_qubits = n;
    p.max_fused = std::max<uint32_t>(opt.max_fused_qubits, 1);
    p.max_stage_op = 4;
    if (opt.tile_bits == 0 && n >= vg.L()) {
        // Register-tile stages: the fixed lane bits are the lowest qubits.
        p.reg_tile = true;
        p.tile_bits = vg.L();
        p.contig = vg.f_bits;
    } else {
        p.tile_bits = opt.tile_bits > 0 ? (uint32_t)opt.tile_bits : tc.L;
        require(p.tile_bits <= 14 - (precision == QVD_C128 ? 1 : 0) && p.tile_bits >= 1,
                "tile_bits out of range (1..14 for complex64, 1..13 for complex128)");
        p.contig = std::min(tc.contig, p.tile_bits);
    }
    require(opt.variant < PLAN_VARIANTS, "plan variant out of range");
    p.variant = opt.variant;
    if (!p.reg_tile) return plan_heuristic(v, p);
    require(opt.planner <= QVD_PLANNER_LEARNED_ANY, "unknown planner");
    std::unique_lock<std::mutex> lock(g_model.mu);
    if (opt.planner == QVD_PLANNER_HEURISTIC || !g_model.net) {
        lock.unlock();
        if (chosen) *chosen = p.variant;
        return split_for_coefficients(plan_heuristic(v, p), n, precision, vg);
    }
    const bool same_amplitudes = opt.planner == QVD_PLANNER_LEARNED;
    // Learned: a choice made for the same ops before is reused.
    const uint64_t key = ops_hash(n, precision, opt, ops, num_ops, vg.f_bits);
    auto hit = g_model.chosen.find(key);
    if (hit != g_model.chosen.end()) {
        p.variant = hit->second;
        if (chosen) *chosen = p.variant;
        return split_for_coefficients(plan_heuristic(v, p), n, precision, vg);
    }
    // Otherwise every variant is planned (in parallel; identical plans are
    // scored once). QVD_PLANNER_LEARNED keeps the variants whose plan image
    // equals the image of the heuristic plan (options.variant): they run the
    // same arithmetic, so whichever the model picks, the amplitudes are
    // bit-identical to the heuristic's. QVD_PLANNER_LEARNED_ANY keeps them
    // all. The kept plans' stages are scored by the cost model in one forward
    // pass on the device.
    const uint32_t F = QVD_PLAN_FEATURES;
    const uint32_t base = p.variant;
    std::vector<std::vector<Step>> plans(PLAN_VARIANTS);
    std::vector<std::vector<float>> feats(PLAN_VARIANTS);
    std::vector<char> dup(PLAN_VARIANTS, 0);
    {
        std::vector<std::thread> workers;
        std::vector<std::exception_ptr> errors(PLAN_VARIANTS);
        for (uint32_t var = 0; var < PLAN_VARIANTS; var++)
            workers.emplace_back([&, var] {
                try {
                    PlanParams pv = p;
                    pv.variant = var;
                    plans[var] = split_for_coefficients(plan_heuristic(v, pv), n, precision, vg);
                } catch (...) {
                    errors[var] = std::current_exception();
                }
            });
        for (auto &w : workers) w.join();
        for (auto &e : errors)
            if (e) std::rethrow_exception(e);
        std::vector<std::vector<uint32_t>> sigs(PLAN_VARIANTS);
        for (uint32_t var = 0; var < PLAN_VARIANTS; var++) {
            sigs[var] = plan_signature(plans[var]);
            for (uint32_t u = 0; u < var && !dup[var]; u++)
                if (!dup[u] && sigs[u] == sigs[var]) dup[var] = 1;
        }
        if (same_amplitudes) {
            std::vector<uint32_t> keep = equivalent_variants(plans, base, n, precision, vg.f_bits);
            for (uint32_t var = 0; var < PLAN_VARIANTS; var++)
                if (std::find(keep.begin(), keep.end(), var) == keep.end()) dup[var] = 1;
        }
        workers.clear();
        for (uint32_t var = 0; var < PLAN_VARIANTS; var++) {
            if (dup[var]) continue;
            workers.emplace_back([&, var] {
                try {
                    feats[var] = plan_features(plans[var], n, precision, vg.f_bits);
                } catch (...) {
                    errors[var] = std::current_exception();
                }
            });
        }
        for (auto &w : workers) w.join();
        for (auto &e : errors)
            if (e) std::rethrow_exception(e);
    }
    std::vector<float> rows;
    std::vector<size_t> first(PLAN_VARIANTS + 1, 0);
    std::vector<double> fixed(PLAN_VARIANTS, 0.0);
    const double amps = std::ldexp(1.0, (int)n);
    const double elem = precision == QVD_C64 ? 8.0 : 16.0;
    for (uint32_t var = 0; var < PLAN_VARIANTS; var++) {
        std::vector<float> &r = feats[var];
        for (size_t k = 0; k < r.size(); k++) {
            const size_t j = k % F;
            r[k] = (r[k] - g_model.mean[j]) / g_model.scale[j];
        }
        rows.insert(rows.end(), r.begin(), r.end());
        first[var + 1] = rows.size() / F;
        // A full-state pass outside a stage: one read and one write at the
        // HBM copy rate.
        for (const Step &s : plans[var])
            if (!s.is_stage) fixed[var] += 2 * amps * elem / 4.5e12;
    }
    std::vector<float> pred(rows.size() / F);
    if (!pred.empty()) g_model.net->forward(rows.data(), pred.size(), pred.data());
    uint32_t best = base;
    double best_t = 0;
    bool found = false;
    for (uint32_t var = 0; var < PLAN_VARIANTS; var++) {
        if (dup[var]) continue;
        double t = fixed[var];
        for (size_t k = first[var]; k < first[var + 1]; k++)
            t += std::exp(g_model.ymean + g_model.yscale * pred[k]) * amps;
        if (!found || t < best_t) {
            best = var;
            best_t = t;
            found = true;
        }
    }
    if (g_model.chosen.size() >= 256) g_model.chosen.clear();
    g_model.chosen[key] = best;
    if (chosen) *chosen = best;
    return std::move(plans[best]);
}

void count_plan(const std::vector<Step> &steps, qvd_stats &st) {
    for (const Step &s : steps) {
        st.passes++;
        if (s.is_stage) {
            st.stages++;
            st.kernels += s.stage.ops.size();
        } else {
            st.kernels++;
        }
    }
}

// Fails when a wait of the batch's stage queue ran out (after a replay).
const char *const QUEUE_TIMEOUT = "a batch stage waited more than 10 s for its previous stage; results are invalid";

void check_queue(const qvd_batch *b) {
    uint32_t err = 0;
    QVD_HIP(hipMemcpy(&err, b->sync.as<uint32_t>() + VQUEUE_ERROR, sizeof(err), hipMemcpyDeviceToHost));
    if (err) fail(QVD_EHIP, QUEUE_TIMEOUT);
}

// Builds a batch: plans each circuit on max(n, L) qubits (L = the register
// tile), lowers its stages, lays out the level launch tables and the
// expectation descriptors, uploads them and captures the graph.
template <typename T2>
void build_batch(qvd_batch *b, const qvd_circuit *circuits, uint32_t count, const qvd_options &opt_in,
                 const std::vector<uint32_t> &fixed) {
    const VGeom vg = vstage_geom<T2>();
    const uint32_t L = vg.L();
    b->fixed = fixed;
    qvd_options opt = opt_in;
    opt.tile_bits = 0;  // register tiles
    std::vector<uint64_t> state_off(count);
    std::vector<std::vector<VStageDesc>> stage_descs(count);
    std::vector<VInstr> vins;
    std::vector<C64> vco;
    std::vector<StageOp> vlds;
    std::vector<uint32_t> vints;
    std::vector<ExpDesc> ed(count);
    std::vector<uint64_t> seg_start(count + 1, 0);
    std::vector<uint32_t> out_start(count + 1, 0);
    uint64_t off = 0, seg_off = 0;
    for (uint32_t c = 0; c < count; c++) {
        const qvd_circuit &ci = circuits[c];
        require(ci.num_qubits >= 1 && ci.num_qubits <= 30, "batched circuits have 1 to 30 qubits");
        require(ci.num_ops == 0 || ci.ops != nullptr, "null ops");
        // Ops must stay on the circuit's own qubits, not the ancillas.
        for (uint64_t i = 0; i < ci.num_ops; i++) {
            const qvd_op &o = ci.ops[i];
            for (uint32_t j = 0; j < o.num_targets; j++)
                require(o.targets && o.targets[j] < ci.num_qubits, "circuit " + std::to_string(c) + ": qubit out of range");
            for (uint32_t j = 0; j < o.num_controls; j++)
                require(o.controls && o.controls[j] < ci.num_qubits, "circuit " + std::to_string(c) + ": qubit out of range");
        }
        const uint32_t npad = std::max(ci.num_qubits, L);
        VGeom gc = vg;
        gc.f_bits = fixed[c];
        std::vector<Step> steps = make_plan(npad, b->precision, ci.ops, ci.num_ops, opt, nullptr, gc.f_bits);
        PendingDiag pend(npad);
        state_off[c] = off;
        // Every circuit starts from |0...0>: its stages skip the tiles known
        // to be zero. Its first stage makes its tiles in registers; the
        // tiles it skips are left unwritten (stale), and each later stage
        // sets the amplitudes of the stale qubits to zero instead of loading
        // them and leaves the qubits it skips stale. A circuit with stale
        // qubits after its last stage (or no stages) needs its state written
        // first; the replay then writes every state and no stage is lazy.
        uint64_t zero = (opt.flags & QVD_FLAG_ALL_TILES) ? 0 : qubit_mask(npad);
        uint64_t stale = 0;
        for (const Step &st : steps) {
            if (!st.is_stage || !st.stage.reg_tile)
                fail(QVD_EUNSUPPORTED, "circuit " + std::to_string(c) +
                                           " needs a full-state pass; run it with qvd_sv_apply_ops");
            VProgram p = build_vprogram(st.stage, npad, gc, pend, vco, vlds, vints, true);
            require(p.desc.ccount * sizeof(T2) <= VCONST_BYTES, "internal: stage coefficients exceed the LDS budget");
            p.desc.instr_off = (uint32_t)vins.size();
            p.desc.state_off = off;
            uint64_t glob = 0;  // the stage's non-tile qubits
            for (uint32_t j = 0; j < p.desc.nglob; j++) glob |= uint64_t(1) << p.desc.glob_pos[j];
            b->skipped_tiles += drop_zero_tiles(p.desc, zero);
            p.desc.zero_mem = stale;
            stale = glob & zero;
            for (const Op &op : st.stage.ops) zero = after_op(zero, op);
            vins.insert(vins.end(), p.instr.begin(), p.instr.end());
            stage_descs[c].push_back(p.desc);
            b->kernels += st.stage.ops.size();
        }
        if (steps.empty() || stale) b->init_states = true;
        b->gates += ci.num_ops;
        ExpDesc &e = ed[c];
        e = expect_desc(npad, ci.num_qubits);
        e.state_off = off;
        e.seg_off = seg_off;
        e.z_off = (uint32_t)b->num_z;
        e.zz_off = (uint32_t)b->num_zz;
        seg_off += e.nsegs * 2 * e.B;
        seg_start[c + 1] = seg_start[c] + e.nsegs;
        out_start[c + 1] = out_start[c] + 2 * ci.num_qubits - 1;
        b->num_z += ci.num_qubits;
        b->num_zz += ci.num_qubits - 1;
        off += uint64_t(1) << npad;
    }
    b->len = off;
    b->state_off = state_off;
    for (uint32_t c = 0; c < count; c++) {
        b->qubits.push_back(circuits[c].num_qubits);
        b->circuit_stages.push_back((uint32_t)stage_descs[c].size());
    }
    // The stage queue: stage r of every circuit (level r) before stage r + 1
    // of any, circuits in the same order on every level, so the tiles an item
    // waits for were queued a whole level earlier. An item holds up to
    // QVD_QUEUE_CHUNK tiles (fewer for states of few tiles, so that their
    // stages still spread over several workgroups, and near the end of the
    // queue), interleaved: item j of a
    // stage of m items takes tiles j, j + m, j + 2m, ..., so the workgroups
    // working on a stage at one time touch neighbouring tiles, as they do
    // when each takes one tile. Its state's counter must first reach the
    // tiles of the state's earlier stages.
    std::vector<VStageDesc> all_descs;
    std::vector<VQueueItem> items;
    std::vector<uint64_t> queued(count, 0);
    uint64_t remaining = 0;  // tiles of the queue from the current stage on
    for (uint32_t c = 0; c < count; c++)
        for (const VStageDesc &d : stage_descs[c]) remaining += d.num_tiles;
    unsigned resident = 0;
    QVD_HIP(vstage_resident_groups<T2>(&resident));
    for (uint32_t r = 0;; r++) {
        bool any = false;
        for (uint32_t c = 0; c < count; c++) {
            if (r >= stage_descs[c].size()) continue;
            any = true;
            VStageDesc d = stage_descs[c][r];
            // Lazy zeros unless the states are written first.
            d.from_zero = r == 0 && !b->init_states;
            if (b->init_states) d.zero_mem = 0;
            b->stage_bytes += (sizeof(T2) << d.L) * (d.num_tiles + stage_read_tiles(d));
            const uint32_t di = (uint32_t)all_descs.size();
            all_descs.push_back(d);
            // Large items while much work remains (each item costs a fetch, a
            // wait and a release), small ones towards the end of the queue,
            // so that the workgroups finish together.
            const uint64_t chunk = std::max<uint64_t>(
                1, std::min<uint64_t>({QVD_QUEUE_CHUNK, d.num_tiles / 8, remaining / (4 * (uint64_t)resident)}));
            remaining -= d.num_tiles;
            const uint64_t m = (d.num_tiles + chunk - 1) / chunk;  // items of this stage
            for (uint64_t j = 0; j < m; j++)
                items.push_back(VQueueItem{di, c, (uint32_t)queued[c], (uint32_t)((d.num_tiles - j + m - 1) / m), j, m});
            queued[c] += d.num_tiles;
            require(queued[c] < (uint64_t(1) << 32), "circuit " + std::to_string(c) + " has too many stage tiles");
            b->stages++;
        }
        if (!any) break;
        b->levels++;
    }
    require(items.size() < (uint64_t(1) << 32), "batch too large for one stage queue");
    // Initialisation (when it runs) writes every state; the expectation
    // values read every state.
    b->other_bytes = (b->init_states ? 2 : 1) * off * sizeof(T2);
    b->nitems = (uint32_t)items.size();
    b->grid = (unsigned)std::min<uint64_t>(resident, std::max<uint64_t>(items.size(), 1));
    // Upload.
    std::vector<T2> vco_t(vco.size());
    for (size_t i = 0; i < vco.size(); i++) vco_t[i] = to_t2<T2>(vco[i]);
    inline_coefficients(vins, vco_t);
    inline_phase_records(vins, vints, vco_t);
    auto upload = [&](DBuf &buf, const void *src, size_t bytes) {
        buf.alloc(std::max<size_t>(bytes, 16));
        if (bytes) QVD_HIP(hipMemcpy(buf.p, src, bytes, hipMemcpyHostToDevice));
    };
    upload(b->ins, vins.data(), vins.size() * sizeof(VInstr));
    upload(b->coefs, vco_t.data(), vco_t.size() * sizeof(T2));
    upload(b->ldsops, vlds.data(), vlds.size() * sizeof(StageOp));
    upload(b->ints, vints.data(), vints.size() * sizeof(uint32_t));
    upload(b->descs, all_descs.data(), all_descs.size() * sizeof(VStageDesc));
    upload(b->items, items.data(), items.size() * sizeof(VQueueItem));
    b->sync.alloc((VQUEUE_COUNTERS + count) * sizeof(uint32_t));
    upload(b->basis, state_off.data(), state_off.size() * sizeof(uint64_t));
    upload(b->edescs, ed.data(), ed.size() * sizeof(ExpDesc));
    upload(b->seg_start, seg_start.data(), seg_start.size() * sizeof(uint64_t));
    upload(b->out_start, out_start.data(), out_start.size() * sizeof(uint32_t));
    b->seg_sums.alloc(std::max<uint64_t>(seg_off, 1) * sizeof(double));
    b->z.alloc(std::max<uint64_t>(b->num_z, 1) * sizeof(double));
    b->zz.alloc(std::max<uint64_t>(b->num_zz, 1) * sizeof(double));
    if (b->host_out) QVD_HIP(hipHostFree(b->host_out));
    b->host_out = nullptr;
    QVD_HIP(hipHostMalloc((void **)&b->host_out, (b->num_z + b->num_zz + 1) * sizeof(double), hipHostMallocDefault));
    const uint64_t bytes = b->len * sizeof(T2);
    const uint64_t cap = state_cap_bytes();
    if (cap && bytes > cap) fail(QVD_ENOMEM, "batch states of " + std::to_string(bytes) + " bytes exceed QVD_MAX_STATE_BYTES");
    b->amps.alloc(bytes);
    // One eager run first (also sets the kernels' LDS attributes outside the
    // capture), then the capture of the same sequence.
    T2 *s = b->amps.as<T2>();
    auto enqueue = [&]() {
        // Every circuit's first stage starts from |0...0> without loading,
        // unless a circuit ends with stale qubits or has no stage: then every
        // state is written first.
        if (b->init_states) launch_batch_init<T2>(s, b->len, b->basis.as<uint64_t>(), count, b->stream);
        launch_zero_words(b->sync.as<uint32_t>(), VQUEUE_COUNTERS + count, b->stream);
        QVD_HIP(hipGetLastError());
        if (b->nitems)
            QVD_HIP(launch_vstage_queue<T2>(s, b->descs.as<VStageDesc>(), b->items.as<VQueueItem>(), b->nitems,
                                            b->sync.as<uint32_t>(), b->ins.as<VInstr>(), b->coefs.as<T2>(),
                                            b->ldsops.as<StageOp>(), b->ints.as<uint32_t>(), b->grid, b->stream));
        launch_expect_batch<T2>(s, b->edescs.as<ExpDesc>(), count, b->seg_start.as<uint64_t>(), seg_start[count],
                                b->out_start.as<uint32_t>(), out_start[count], b->seg_sums.as<double>(),
                                b->z.as<double>(), b->zz.as<double>(), b->stream);
        QVD_HIP(hipGetLastError());
    };
    enqueue();
    QVD_HIP(hipStreamSynchronize(b->stream));
    check_queue(b);
    QVD_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    try {
        enqueue();
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(b->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    QVD_HIP(hipStreamEndCapture(b->stream, &b->graph));
    QVD_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
}

// A batch handle with its own stream on the current device.
std::unique_ptr<qvd_batch> new_batch(int precision, uint32_t count) {
    std::unique_ptr<qvd_batch> b(new qvd_batch());
    b->precision = precision;
    b->count = count;
    QVD_HIP(hipGetDevice(&b->device));
    QVD_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    return b;
}

// The autotuner's choices for this process: class key -> fixed lane bits.
struct TuneCache {
    std::mutex mu;
    std::unordered_map<uint64_t, uint32_t> choice;
} g_tune;

// A circuit's class for the autotuner: precision, size, and each op's
// targets, controls and control values (not its matrix), FNV-1a.
uint64_t circuit_class(const qvd_circuit &c, int precision, const qvd_options &opt) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint64_t v) {
        for (int i = 0; i < 8; i++) {
            h ^= (v >> (8 * i)) & 0xff;
            h *= 1099511628211ull;
        }
    };
    mix((uint64_t)precision);
    mix(c.num_qubits);
    mix(opt.max_fused_qubits);
    mix(opt.variant);
    mix(c.num_ops);
    for (uint64_t i = 0; i < c.num_ops; i++) {
        const qvd_op &o = c.ops[i];
        mix(((uint64_t)o.num_targets << 32) | o.num_controls);
        for (uint32_t j = 0; j < o.num_targets; j++) mix(o.targets[j]);
        for (uint32_t j = 0; j < o.num_controls; j++) mix(((uint64_t)o.controls[j] << 1) | (o.control_values ? o.control_values[j] & 1 : 1));
    }
    return h;
}

// Seconds of the fastest of `reps` replays of a built batch's graph (after
// one unmeasured replay).
double replay_seconds(qvd_batch *b, int reps) {
    QVD_HIP(hipGraphLaunch(b->exec, b->stream));
    QVD_HIP(hipStreamSynchronize(b->stream));
    double best = 1e30;
    for (int r = 0; r < reps; r++) {
        const double t0 = now();
        QVD_HIP(hipGraphLaunch(b->exec, b->stream));
        QVD_HIP(hipStreamSynchronize(b->stream));
        best = std::min(best, now() - t0);
    }
    check_queue(b);
    return best;
}

template <typename T2>
void build_batch(qvd_batch *b, const qvd_circuit *circuits, uint32_t count, const qvd_options &opt_in,
                 const std::vector<uint32_t> &fixed);

// Each circuit's register-tile geometry: its own request, else the
// autotuner's (QVD_FLAG_AUTOTUNE), else the default. The autotuner runs each
// class not seen before as a batch of its own with every candidate geometry
// (the default and one fixed lane bit fewer or more), on up to enough of the
// class's circuits for 8 tiles per resident workgroup, and keeps the
// fastest; a candidate must beat the default by 1 % to be kept.
template <typename T2>
std::vector<uint32_t> batch_geometry(const qvd_circuit *circuits, uint32_t count, int precision,
                                     const qvd_options &opt) {
    const VGeom vg = vstage_geom<T2>();
    const uint32_t def = vg.f_bits, lo = def - 1, hi = def + 1;
    std::vector<uint32_t> fixed(count, 0);
    for (uint32_t c = 0; c < count; c++) {
        const uint32_t f = circuits[c].fixed_lanes;
        require(f == 0 || (f >= lo && f <= hi),
                "circuit " + std::to_string(c) + ": fixed_lanes must be 0 or " + std::to_string(lo) + ".." + std::to_string(hi));
        fixed[c] = f;
    }
    if (!(opt.flags & QVD_FLAG_AUTOTUNE)) {
        for (uint32_t &f : fixed)
            if (!f) f = def;
        return fixed;
    }
    std::vector<uint64_t> keys(count, 0);
    std::vector<uint64_t> order;  // class keys in order of first appearance
    std::unordered_map<uint64_t, std::vector<uint32_t>> members;
    for (uint32_t c = 0; c < count; c++) {
        if (fixed[c]) continue;
        keys[c] = circuit_class(circuits[c], precision, opt);
        auto &m = members[keys[c]];
        if (m.empty()) order.push_back(keys[c]);
        m.push_back(c);
    }
    unsigned resident = 0;
    QVD_HIP(vstage_resident_groups<T2>(&resident));
    qvd_options popt = opt;
    popt.flags &= ~QVD_FLAG_AUTOTUNE;
    for (uint64_t key : order) {
        {
            std::lock_guard<std::mutex> lock(g_tune.mu);
            auto hit = g_tune.choice.find(key);
            if (hit != g_tune.choice.end()) continue;
        }
        // The probe: the class's first circuits, up to 8 tiles per resident
        // workgroup per stage.
        std::vector<qvd_circuit> probe;
        uint64_t tiles = 0;
        for (uint32_t c : members[key]) {
            if (tiles >= 8ull * resident) break;
            probe.push_back(circuits[c]);
            tiles += uint64_t(1) << (std::max(circuits[c].num_qubits, vg.L()) - vg.L());
        }
        uint32_t best = def;
        double best_s = 0, def_s = 0;
        for (uint32_t f = lo; f <= hi; f++) {
            std::unique_ptr<qvd_batch> pb = new_batch(precision, (uint32_t)probe.size());
            build_batch<T2>(pb.get(), probe.data(), (uint32_t)probe.size(), popt,
                            std::vector<uint32_t>(probe.size(), f));
            const double t = replay_seconds(pb.get(), 3);
            if (f == def) def_s = t;
            if (f == lo || t < best_s) {
                best_s = t;
                best = f;
            }
        }
        if (best != def && best_s > 0.99 * def_s) best = def;
        std::lock_guard<std::mutex> lock(g_tune.mu);
        g_tune.choice[key] = best;
    }
    std::lock_guard<std::mutex> lock(g_tune.mu);
    for (uint32_t c = 0; c < count; c++)
        if (!fixed[c]) fixed[c] = g_tune.choice.at(keys[c]);
    return fixed;
}

// The walk's marks: the running sum every 2^6 amplitudes (1/64 of the
// amplitude count, in doubles), so a shot's resumed sum is at most 64 long.
constexpr uint32_t SAMPLE_MARK_BITS = 6;

// Chunks of the reference's CHUNK_BLOCKS blocks (src/measure.rs:12-13):
// 4096 x 8 amplitudes in complex64, 4096 x 4 in complex128.
uint64_t sample_chunk(const qvd_sv *sv) { return sv->precision == QVD_C64 ? 32768 : 16384; }

// The sampler's buffers carved from one workspace (256-byte aligned parts),
// with the walk's marks every 2^mark_bits amplitudes; returns the bytes
// needed, and fills `w`, the uniforms' and the outputs' device addresses when
// `base` is not null.
size_t carve_sample(const qvd_sv *sv, uint64_t shots, uint32_t mark_bits, char *base, SampleScratch &w, double **du,
                    uint64_t **dout) {
    const uint64_t chunk = sample_chunk(sv);
    const uint64_t nc = (sv->len + chunk - 1) / chunk;
    const uint64_t nmarks = (sv->len + (uint64_t(1) << mark_bits) - 1) >> mark_bits;
    w.mark_bits = mark_bits;
    w.sort_temp_bytes = sort_temp_bytes(shots);
    require(w.sort_temp_bytes > 0, "radix sort temporary size query failed");
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) & ~size_t(255);
        return at;
    };
    const size_t o_sums = carve(nc * sizeof(double)), o_prefix = carve((nc + 1) * sizeof(double));
    const size_t o_bounds = carve((nc + 1) * sizeof(uint64_t));
    const size_t o_ts = carve(shots * sizeof(double)), o_du = carve(shots * sizeof(double));
    const size_t o_out = carve(shots * sizeof(uint64_t)), o_temp = carve(w.sort_temp_bytes);
    const size_t o_marks = carve(nmarks * sizeof(double)), o_walked = carve(nc * sizeof(uint32_t));
    if (base) {
        w.marks = reinterpret_cast<double *>(base + o_marks);
        w.walked = reinterpret_cast<uint32_t *>(base + o_walked);
        w.sums = reinterpret_cast<double *>(base + o_sums);
        w.prefix = reinterpret_cast<double *>(base + o_prefix);
        w.bounds = reinterpret_cast<uint64_t *>(base + o_bounds);
        w.targets = reinterpret_cast<double *>(base + o_ts);
        w.sort_temp = base + o_temp;
        *du = reinterpret_cast<double *>(base + o_du);
        *dout = reinterpret_cast<uint64_t *>(base + o_out);
    }
    return off;
}

// Two-pass sampling with the buffers carved from `base` (at least
// carve_sample bytes of device memory): copies the uniforms in, runs the
// sampler on the handle's stream, copies the outcomes out and waits. With
// `side` (a second stream and an event of the handle), the uniforms' copy and
// sort run there while the chunk sums and their prefix run on the handle's
// stream; they depend on nothing of each other until the targets are scaled.
void sample_with(qvd_sv *sv, const double *uniforms, uint64_t shots, uint32_t mark_bits, uint64_t *out, char *base,
                 bool side, void *host_stage) {
    const uint64_t chunk = sample_chunk(sv);
    SampleScratch w;
    double *du = nullptr;
    uint64_t *dout = nullptr;
    carve_sample(sv, shots, mark_bits, base, w, &du, &dout);
    with_type(sv, [&](auto *s) {
        if (side) {
            // the sums first: the copy from pageable memory may hold the host
            launch_sample_sums(s, sv->len, chunk, w, sv->stream);
            QVD_HIP(hipMemcpyAsync(du, uniforms, shots * sizeof(double), hipMemcpyHostToDevice, sv->side));
            QVD_HIP(launch_sort_uniforms(du, shots, w, sv->side));
            QVD_HIP(hipEventRecord(sv->side_done, sv->side));
            QVD_HIP(hipStreamWaitEvent(sv->stream, sv->side_done, 0));
        } else {
            QVD_HIP(hipMemcpyAsync(du, uniforms, shots * sizeof(double), hipMemcpyHostToDevice, sv->stream));
            launch_sample_sums(s, sv->len, chunk, w, sv->stream);
            QVD_HIP(launch_sort_uniforms(du, shots, w, sv->stream));
        }
        QVD_HIP(launch_sample_walk(s, sv->len, chunk, shots, w, dout, sv->stream));
        return 0;
    });
    if (host_stage) {
        // One DMA into pinned memory and one host copy: a copy straight into
        // pageable memory goes through the runtime's staging in pieces with
        // host work between them (about 1.8 ms for 8 MB against 0.2 ms).
        QVD_HIP(hipMemcpyAsync(host_stage, dout, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, sv->stream));
        sv->sync();
        std::memcpy(out, host_stage, shots * sizeof(uint64_t));
    } else {
        QVD_HIP(hipMemcpyAsync(out, dout, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, sv->stream));
        sv->sync();
    }
}

}  // namespace
}  // namespace qvd

extern "C" {

const char *qvd_last_error(void) { return qvd::g_last_error.c_str(); }

int qvd_device_info(int *num_devices, char *arch, uint32_t arch_capacity) {
    return guard([&] {
        require(num_devices != nullptr, "null output");
        int count = 0;
        hipError_t e = hipGetDeviceCount(&count);
        if (e != hipSuccess) count = 0;
        *num_devices = count;
        if (arch && arch_capacity) {
            arch[0] = 0;
            if (count > 0) {
                hipDeviceProp_t p;
                QVD_HIP(hipGetDeviceProperties(&p, 0));
                std::strncpy(arch, p.gcnArchName, arch_capacity - 1);
                arch[arch_capacity - 1] = 0;
            }
        }
    });
}

int qvd_state_vector_fits(uint32_t num_qubits, int precision, int *fits, uint64_t *state_bytes,
                          uint64_t *device_bytes) {
    return guard([&] {
        require(fits != nullptr, "null output");
        require(precision == QVD_C64 || precision == QVD_C128, "precision must be QVD_C64 or QVD_C128");
        require(num_qubits <= 63, "num_qubits out of range");
        int dev = 0;
        QVD_HIP(hipGetDevice(&dev));
        hipDeviceProp_t p;
        QVD_HIP(hipGetDeviceProperties(&p, dev));
        const double bytes = std::ldexp(1.0, (int)num_qubits) * (precision == QVD_C64 ? 8.0 : 16.0);
        *fits = bytes <= 0.8 * (double)p.totalGlobalMem ? 1 : 0;
        if (state_bytes) *state_bytes = num_qubits < 60 ? (uint64_t(1) << num_qubits) * (precision == QVD_C64 ? 8 : 16) : ~uint64_t(0);
        if (device_bytes) *device_bytes = (uint64_t)p.totalGlobalMem;
    });
}

int qvd_sv_create(uint32_t num_qubits, int precision, qvd_sv **out) {
    return guard([&] {
        require(out != nullptr, "null output handle");
        *out = nullptr;
        require(precision == QVD_C64 || precision == QVD_C128, "precision must be QVD_C64 or QVD_C128");
        require(num_qubits <= 34, "a state vector of more than 34 qubits does not fit in memory");
        std::unique_ptr<qvd_sv> sv(new qvd_sv());
        sv->n = num_qubits;
        sv->precision = precision;
        sv->len = uint64_t(1) << num_qubits;
        const uint64_t bytes = sv->len * sv->elem();
        const uint64_t cap = state_cap_bytes();
        if (cap && bytes > cap)
            fail(QVD_ENOMEM, "state of " + std::to_string(bytes) + " bytes exceeds QVD_MAX_STATE_BYTES");
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) fail(QVD_EHIP, "no HIP device is visible");
        QVD_HIP(hipGetDevice(&sv->device));
        sv->bind();
        QVD_HIP(hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking));
        sv->amps.alloc(bytes);
        sv->zero_pending = true;
        sv->zero_qubits = qubit_mask(num_qubits);
        *out = sv.release();
    });
}

int qvd_sv_destroy(qvd_sv *sv) {
    return guard([&] {
        if (!sv) return;
        sv->bind();
        if (sv->stream) (void)hipStreamSynchronize(sv->stream);
        sv->amps.alloc(0);
        sv->scratch.alloc(0);
        if (sv->side) (void)hipStreamSynchronize(sv->side);
        if (sv->sample_host) (void)hipHostFree(sv->sample_host);
        if (sv->side_done) (void)hipEventDestroy(sv->side_done);
        if (sv->side) (void)hipStreamDestroy(sv->side);
        if (sv->stream) (void)hipStreamDestroy(sv->stream);
        delete sv;
    });
}

int qvd_sv_num_qubits(const qvd_sv *sv, uint32_t *out) {
    return guard([&] {
        check_handle(sv);
        *out = sv->n;
    });
}

int qvd_sv_precision(const qvd_sv *sv, int *out) {
    return guard([&] {
        check_handle(sv);
        *out = sv->precision;
    });
}

int qvd_sv_set_zero(qvd_sv *sv) {
    return guard([&] {
        require(sv != nullptr, "null state handle");
        sv->zero_pending = true;
        sv->zero_qubits = qubit_mask(sv->n);
        sv->stale_qubits = 0;
    });
}

int qvd_sv_copy(qvd_sv *dst, const qvd_sv *src) {
    return guard([&] {
        require(dst != nullptr && src != nullptr, "null state handle");
        require(dst->n == src->n && dst->precision == src->precision, "states of different sizes or precisions");
        require(dst->device == src->device, "states on different devices");
        if (dst == src) return;
        // dst is overwritten whole: a pending |0...0> in src is copied as pending.
        dst->zero_pending = src->zero_pending;
        dst->zero_qubits = src->zero_qubits;
        dst->stale_qubits = src->stale_qubits;  // stale memory is copied as stale
        if (src->zero_pending) return;
        src->sync();
        dst->bind();
        QVD_HIP(hipMemcpyAsync(dst->amps.p, src->amps.p, src->len * src->elem(), hipMemcpyDeviceToDevice,
                               dst->stream));
        dst->sync();
    });
}

int qvd_sv_set_amplitudes(qvd_sv *sv, const double *amplitudes, uint64_t len) {
    return guard([&] {
        require(sv != nullptr, "null state handle");
        require(amplitudes != nullptr, "null amplitudes");
        require(len == sv->len, "expected " + std::to_string(sv->len) + " amplitudes, got " + std::to_string(len));
        sv->zero_pending = false;  // every amplitude is written below
        sv->zero_qubits = 0;
        sv->stale_qubits = 0;
        sv->bind();
        const uint64_t chunk = std::min<uint64_t>(len, uint64_t(1) << 24);
        DBuf stage(chunk * sizeof(double2));
        for (uint64_t off = 0; off < len; off += chunk) {
            const uint64_t cnt = std::min(chunk, len - off);
            QVD_HIP(hipMemcpyAsync(stage.p, amplitudes + 2 * off, cnt * sizeof(double2), hipMemcpyHostToDevice, sv->stream));
            with_type(sv, [&](auto *s) {
                launch_narrow(stage.as<double2>(), s + off, cnt, sv->stream);
                return 0;
            });
            QVD_HIP(hipGetLastError());
            sv->sync();
        }
    });
}

int qvd_sv_get_amplitudes(const qvd_sv *sv, double *amplitudes, uint64_t len) {
    return guard([&] {
        check_handle(sv);
        require(amplitudes != nullptr, "null amplitudes");
        require(len == sv->len, "expected " + std::to_string(sv->len) + " amplitudes, got " + std::to_string(len));
        sv->bind();
        const uint64_t chunk = std::min<uint64_t>(len, uint64_t(1) << 24);
        DBuf stage(chunk * sizeof(double2));
        for (uint64_t off = 0; off < len; off += chunk) {
            const uint64_t cnt = std::min(chunk, len - off);
            with_type(sv, [&](auto *s) {
                launch_widen(s + off, stage.as<double2>(), cnt, sv->stream);
                return 0;
            });
            QVD_HIP(hipGetLastError());
            QVD_HIP(hipMemcpyAsync(amplitudes + 2 * off, stage.p, cnt * sizeof(double2), hipMemcpyDeviceToHost, sv->stream));
            sv->sync();
        }
    });
}

int qvd_sv_apply(qvd_sv *sv, const double *matrix, const uint32_t *targets, uint32_t num_targets,
                 const uint32_t *controls, const uint8_t *control_values, uint32_t num_controls) {
    return guard([&] {
        check_handle(sv);
        Op op = make_op(sv->n, matrix, targets, num_targets, controls, control_values, num_controls);
        sv->bind();
        with_type(sv, [&](auto *s) {
            using T2 = std::remove_pointer_t<decltype(s)>;
            Upload up;
            const size_t o = up.put(full_coefficients<T2>(op));
            apply_full(sv, s, op, reinterpret_cast<const T2 *>(up.to_device(sv) + o));
            sv->sync();
            return 0;
        });
    });
}


int qvd_plan_ops(uint32_t num_qubits, int precision, const qvd_op *ops, uint64_t num_ops,
                 const qvd_options *options, qvd_stats *stats) {
    return guard([&] {
        require(precision == QVD_C64 || precision == QVD_C128, "precision must be QVD_C64 or QVD_C128");
        require(num_qubits <= 34, "a state vector of more than 34 qubits does not fit in memory");
        qvd_options opt{4, 0, QVD_PLANNER_HEURISTIC, 0};
        if (options) opt = *options;
        const double t0 = now();
        uint32_t variant = 0;
        std::vector<Step> steps = make_plan(num_qubits, precision, ops, num_ops, opt, &variant);
        qvd_stats st{};
        st.gates = num_ops;
        st.plan_seconds = now() - t0;
        st.plan_variant = variant;
        st.fixed_lanes = opt.fixed_lanes ? opt.fixed_lanes
                                         : (precision == QVD_C64 ? vstage_geom<float2>() : vstage_geom<double2>()).f_bits;
        count_plan(steps, st);
        if (stats) *stats = st;
    });
}

namespace qvd {
namespace {

// Plans `ops` and runs the plan on `sv` (qvd_sv_apply_ops after the
// geometry is settled).
void apply_plan(qvd_sv *sv, const qvd_op *ops, uint64_t num_ops, const qvd_options &opt, qvd_stats &st) {
    const double t0 = now();
    uint32_t variant = 0;
    std::vector<Step> steps = make_plan(sv->n, sv->precision, ops, num_ops, opt, &variant);
    const double t1 = now();
    st = qvd_stats{};
    st.gates = num_ops;
    st.plan_variant = variant;
    st.fixed_lanes = opt.fixed_lanes ? opt.fixed_lanes
                                     : (sv->precision == QVD_C64 ? vstage_geom<float2>() : vstage_geom<double2>()).f_bits;
    // An empty plan leaves a pending |0...0> pending.
    sv->bind();
    if (!steps.empty()) {
        const uint64_t zero = sv->zero_pending ? qubit_mask(sv->n) : sv->zero_qubits;
        sv->zero_qubits = 0;  // until the plan has run
        sv->zero_qubits = with_type(sv, [&](auto *s) {
            return run_plan(sv, s, steps, st, opt.flags, zero, opt.fixed_lanes);
        });
    }
    const double t2 = now();
    st.plan_seconds = t1 - t0;
    st.gate_seconds = t2 - t0;
}

// The single-state geometry autotuner's choices for this process.
TuneCache g_tune_sv;

// The register-tile geometry for `ops` on `sv` under QVD_FLAG_AUTOTUNE: the
// fastest of the default and one fixed lane bit fewer or more (kept only if
// 1 % faster than the default), timed once per circuit class by running the
// plan twice with each on a scratch state of the same size that knows the
// same qubits to be zero (the amplitudes do not change the time; the tiles
// known to be zero do). 0 (the default) when the plan has no register-tile
// stages or the scratch state cannot be allocated.
uint32_t tune_geometry(qvd_sv *sv, const qvd_op *ops, uint64_t num_ops, const qvd_options &opt) {
    const VGeom vg = sv->precision == QVD_C64 ? vstage_geom<float2>() : vstage_geom<double2>();
    if (opt.tile_bits != 0 || sv->n < vg.L()) return 0;
    qvd_circuit c{};
    c.num_qubits = sv->n;
    c.ops = ops;
    c.num_ops = num_ops;
    const uint64_t key = circuit_class(c, sv->precision, opt) ^
                         (0x9e3779b97f4a7c15ull * (1 + opt.planner + 8 * uint64_t(opt.flags & QVD_FLAG_ALL_TILES)));
    {
        std::lock_guard<std::mutex> lock(g_tune_sv.mu);
        auto hit = g_tune_sv.choice.find(key);
        if (hit != g_tune_sv.choice.end()) return hit->second;
    }
    uint32_t best = 0;
    qvd_sv *probe = nullptr;
    if (qvd_sv_create(sv->n, sv->precision, &probe) == QVD_OK) {
        qvd_options o = opt;
        o.flags = (opt.flags & ~QVD_FLAG_AUTOTUNE) | QVD_FLAG_TIMING;
        double best_t = 0;
        for (uint32_t f : {vg.f_bits, vg.f_bits - 1, vg.f_bits + 1}) {
            o.fixed_lanes = f;
            double t = 1e300;
            try {
                for (int rep = 0; rep < 2; rep++) {
                    probe->zero_pending = sv->zero_pending;
                    probe->zero_qubits = sv->zero_qubits;
                    probe->stale_qubits = sv->stale_qubits;
                    qvd_stats st;
                    apply_plan(probe, ops, num_ops, o, st);
                    t = std::min(t, st.stage_kernel_seconds + st.other_kernel_seconds);
                }
            } catch (...) {
                // a candidate that cannot run is skipped; the default's
                // failure is the call's own
                if (f != vg.f_bits) continue;
                (void)qvd_sv_destroy(probe);
                throw;
            }
            if (f == vg.f_bits) best_t = t;
            else if (t < 0.99 * best_t) {
                best = f;
                best_t = t;
            }
        }
        (void)qvd_sv_destroy(probe);
    }
    std::lock_guard<std::mutex> lock(g_tune_sv.mu);
    g_tune_sv.choice[key] = best;
    return best;
}

}  // namespace
}  // namespace qvd

int qvd_sv_apply_ops(qvd_sv *sv, const qvd_op *ops, uint64_t num_ops, const qvd_options *options, qvd_stats *stats) {
    return guard([&] {
        require(sv != nullptr, "null state handle");
        const double t0 = now();
        qvd_options opt{4, 0, QVD_PLANNER_HEURISTIC, 0};
        if (options) opt = *options;
        if (!opt.fixed_lanes && (opt.flags & QVD_FLAG_AUTOTUNE)) opt.fixed_lanes = tune_geometry(sv, ops, num_ops, opt);
        qvd_stats st{};
        apply_plan(sv, ops, num_ops, opt, st);
        st.gate_seconds = now() - t0;  // with the autotuner's probes when they ran
        if (stats) *stats = st;
    });
}

int qvd_sv_norm_sqr(const qvd_sv *sv_c, double *out) {
    return guard([&] {
        check_handle(sv_c);
        require(out != nullptr, "null output");
        qvd_sv *sv = const_cast<qvd_sv *>(sv_c);
        sv->bind();
        const unsigned blocks = 1024;
        double *buf = sv->scratch_doubles(blocks + 1);
        with_type(sv, [&](auto *s) {
            launch_masked_norm(s, sv->len, 0, 0, buf, blocks, buf + blocks, sv->stream);
            return 0;
        });
        QVD_HIP(hipGetLastError());
        QVD_HIP(hipMemcpyAsync(out, buf + blocks, sizeof(double), hipMemcpyDeviceToHost, sv->stream));
        sv->sync();
    });
}

int qvd_sv_scale(qvd_sv *sv, double factor) {
    return guard([&] {
        check_handle(sv);
        sv->bind();
        with_type(sv, [&](auto *s) {
            launch_scale(s, sv->len, factor, sv->stream);
            return 0;
        });
        QVD_HIP(hipGetLastError());
        sv->sync();
    });
}

static double probability_one_impl(qvd_sv *sv, uint32_t qubit) {
    require(qubit < sv->n, "qubit " + std::to_string(qubit) + " out of range for " + std::to_string(sv->n) + " qubits");
    const unsigned blocks = 1024;
    double *buf = sv->scratch_doubles(blocks + 1);
    const uint64_t m = uint64_t(1) << qubit;
    with_type(sv, [&](auto *s) {
        launch_masked_norm(s, sv->len, m, m, buf, blocks, buf + blocks, sv->stream);
        return 0;
    });
    QVD_HIP(hipGetLastError());
    double r = 0;
    QVD_HIP(hipMemcpyAsync(&r, buf + blocks, sizeof(double), hipMemcpyDeviceToHost, sv->stream));
    sv->sync();
    return r;
}

int qvd_sv_probability_one(const qvd_sv *sv_c, uint32_t qubit, double *out) {
    return guard([&] {
        check_handle(sv_c);
        require(out != nullptr, "null output");
        qvd_sv *sv = const_cast<qvd_sv *>(sv_c);
        sv->bind();
        *out = probability_one_impl(sv, qubit);
    });
}

int qvd_sv_probabilities(const qvd_sv *sv_c, double *out, uint64_t len) {
    return guard([&] {
        check_handle(sv_c);
        require(out != nullptr, "null output");
        require(len == sv_c->len, "length does not match the state");
        qvd_sv *sv = const_cast<qvd_sv *>(sv_c);
        sv->bind();
        // |a|^2 in FP64 on the device, copied out in pieces of at most 2^24
        // values (128 MiB of staging).
        const uint64_t piece = std::min<uint64_t>(len, uint64_t(1) << 24);
        DBuf buf(piece * sizeof(double));
        for (uint64_t off = 0; off < len; off += piece) {
            const uint64_t cnt = std::min(piece, len - off);
            with_type(sv, [&](auto *s) {
                launch_probabilities(s + off, cnt, buf.as<double>(), sv->stream);
                return 0;
            });
            QVD_HIP(hipGetLastError());
            QVD_HIP(hipMemcpyAsync(out + off, buf.p, cnt * sizeof(double), hipMemcpyDeviceToHost, sv->stream));
            sv->sync();
        }
    });
}

int qvd_sv_expect_pauli(const qvd_sv *sv_c, const uint64_t *x_masks, const uint64_t *z_masks, uint64_t num_terms,
                        double *out) {
    return guard([&] {
        check_handle(sv_c);
        if (num_terms == 0) return;
        require(x_masks != nullptr && z_masks != nullptr && out != nullptr, "null masks or output");
        qvd_sv *sv = const_cast<qvd_sv *>(sv_c);
        const uint64_t all = sv->n == 64 ? ~uint64_t(0) : (uint64_t(1) << sv->n) - 1;
        for (uint64_t t = 0; t < num_terms; t++)
            require(((x_masks[t] | z_masks[t]) & ~all) == 0, "Pauli string acts on a qubit out of range");
        sv->bind();
        // Terms grouped by their X part: each group of up to PAULI_TERMS
        // terms is one pass over the state.
        std::vector<uint64_t> order(num_terms);
        for (uint64_t t = 0; t < num_terms; t++) order[t] = t;
        std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return x_masks[a] < x_masks[b]; });
        const unsigned blocks = 1024;
        const uint32_t W = 2 * PAULI_TERMS;
        double *buf = sv->scratch_doubles((size_t)blocks * W + W + PAULI_TERMS);
        double *partial = buf, *sums = buf + (size_t)blocks * W;
        uint64_t *dz = reinterpret_cast<uint64_t *>(sums + W);
        std::vector<uint64_t> zs;
        std::vector<double> h(W);
        for (uint64_t g = 0; g < num_terms;) {
            const uint64_t x = x_masks[order[g]];
            zs.clear();
            uint64_t e = g;
            while (e < num_terms && x_masks[order[e]] == x && zs.size() < PAULI_TERMS) zs.push_back(z_masks[order[e++]]);
            const uint32_t k = (uint32_t)zs.size();
            QVD_HIP(hipMemcpyAsync(dz, zs.data(), k * sizeof(uint64_t), hipMemcpyHostToDevice, sv->stream));
            with_type(sv, [&](auto *s) {
                launch_expect_pauli(s, sv->len, x, dz, k, partial, blocks, sums, sv->stream);
                return 0;
            });
            QVD_HIP(hipGetLastError());
            QVD_HIP(hipMemcpyAsync(h.data(), sums, 2 * k * sizeof(double), hipMemcpyDeviceToHost, sv->stream));
            sv->sync();
            for (uint32_t t = 0; t < k; t++) {
                // P = i^popcount(x & z) X^x Z^z; its expectation is real for
                // a Hermitian string, so only the real part is kept.
                const uint64_t idx = order[g + t];
                const int ny = __builtin_popcountll(x & z_masks[idx]) & 3;
                const double re = h[2 * t], im = h[2 * t + 1];
                out[idx] = ny == 0 ? re : ny == 1 ? -im : ny == 2 ? -re : im;
            }
            g = e;
        }
    });
}

int qvd_sv_expect_z(const qvd_sv *sv_c, double *z, double *zz) {
    return guard([&] {
        check_handle(sv_c);
        require(z != nullptr, "null output");
        qvd_sv *sv = const_cast<qvd_sv *>(sv_c);
        const uint32_t n = sv->n;
        if (n == 0) return;
        sv->bind();
        const uint32_t B = expect_segment_bits(n);
        const uint64_t segs = uint64_t(1) << (n - B);
        // segment sums, then the results, in the handle's reduction scratch
        double *seg = sv->scratch_doubles(segs * 2 * B + 2 * n);
        double *res = seg + segs * 2 * B;
        with_type(sv, [&](auto *s) {
            launch_expect_z(s, n, seg, res, res + n, sv->stream);
            return 0;
        });
        QVD_HIP(hipGetLastError());
        std::vector<double> h(2 * n);
        QVD_HIP(hipMemcpyAsync(h.data(), res, (2 * n - 1) * sizeof(double), hipMemcpyDeviceToHost, sv->stream));
        sv->sync();
        std::copy(h.begin(), h.begin() + n, z);
        if (zz) std::copy(h.begin() + n, h.begin() + 2 * n - 1, zz);
    });
}

static uint8_t measure_impl(qvd_sv *sv, uint32_t qubit, double uniform) {
    double p1 = probability_one_impl(sv, qubit);
    p1 = std::min(std::max(p1, 0.0), 1.0);
    const bool outcome = uniform < p1;
    const double p = outcome ? p1 : 1.0 - p1;
    const double keep = 1.0 / std::sqrt(p);
    with_type(sv, [&](auto *s) {
        launch_collapse(s, sv->len, qubit, outcome ? 1u : 0u, keep, sv->stream);
        return 0;
    });
    QVD_HIP(hipGetLastError());
    sv->sync();
    return outcome ? 1 : 0;
}

int qvd_sv_measure(qvd_sv *sv, uint32_t qubit, double uniform, uint8_t *outcome) {
    return guard([&] {
        check_handle(sv);
        require(outcome != nullptr, "null output");
        sv->bind();
        *outcome = measure_impl(sv, qubit, uniform);
    });
}

int qvd_sv_reset(qvd_sv *sv, uint32_t qubit, double uniform) {
    return guard([&] {
        check_handle(sv);
        sv->bind();
        if (measure_impl(sv, qubit, uniform)) {
            Op x;
            x.targets = {qubit};
            x.m = HMatrix(1);
            x.m.at(0, 1) = C64(1, 0);
            x.m.at(1, 0) = C64(1, 0);
            x.diagonal = false;
            with_type(sv, [&](auto *s) {
                using T2 = std::remove_pointer_t<decltype(s)>;
                Upload up;
                const size_t o = up.put(full_coefficients<T2>(x));
                apply_full(sv, s, x, reinterpret_cast<const T2 *>(up.to_device(sv) + o));
                sv->sync();
                return 0;
            });
        }
    });
}

int qvd_sv_sample(const qvd_sv *sv_c, const double *uniforms, uint64_t shots, uint64_t *out) {
    return guard([&] {
        check_handle(sv_c);
        qvd_sv *sv = const_cast<qvd_sv *>(sv_c);
        if (shots == 0) return;
        require(uniforms != nullptr && out != nullptr, "null uniforms or output");
        require(shots < (uint64_t(1) << 31), "at most 2^31 - 1 shots per call");
        sv->bind();
        const double t0 = now();
        // One workspace per handle, grown to the largest call and kept: no
        // allocation on repeated calls.
        SampleScratch w;
        const size_t need = carve_sample(sv, shots, SAMPLE_MARK_BITS, nullptr, w, nullptr, nullptr);
        if (sv->sample_ws.bytes < need) sv->sample_ws.alloc(need);
        if (!sv->side) {
            QVD_HIP(hipStreamCreateWithFlags(&sv->side, hipStreamNonBlocking));
            QVD_HIP(hipEventCreateWithFlags(&sv->side_done, hipEventDisableTiming));
        }
        if (sv->sample_host_bytes < shots * sizeof(uint64_t)) {
            if (sv->sample_host) QVD_HIP(hipHostFree(sv->sample_host));
            sv->sample_host = nullptr;
            sv->sample_host_bytes = 0;
            QVD_HIP(hipHostMalloc(&sv->sample_host, shots * sizeof(uint64_t), hipHostMallocDefault));
            sv->sample_host_bytes = shots * sizeof(uint64_t);
        }
        sample_with(sv, uniforms, shots, SAMPLE_MARK_BITS, out, sv->sample_ws.as<char>(), true, sv->sample_host);
        sv->last_sample_seconds = now() - t0;
    });
}

int qvd_sv_sample_noalloc(const qvd_sv *sv_c, const double *uniforms, uint64_t shots, uint64_t *out) {
    return guard([&] {
        check_handle(sv_c);
        qvd_sv *sv = const_cast<qvd_sv *>(sv_c);
        if (shots == 0) return;
        require(uniforms != nullptr && out != nullptr, "null uniforms or output");
        require(shots < (uint64_t(1) << 31), "at most 2^31 - 1 shots per call");
        sv->bind();
        const double t0 = now();
        // The finest marks that fit the fixed workspace: coarser marks only
        // lengthen each shot's resumed sum (at most one chunk).
        SampleScratch w;
        uint32_t mark_bits = SAMPLE_MARK_BITS;
        size_t need = carve_sample(sv, shots, mark_bits, nullptr, w, nullptr, nullptr);
        while (need > NOALLOC_SCRATCH_BYTES && (uint64_t(2) << mark_bits) <= sample_chunk(sv))
            need = carve_sample(sv, shots, ++mark_bits, nullptr, w, nullptr, nullptr);
        if (need > NOALLOC_SCRATCH_BYTES)
            fail(QVD_EINVAL, "qvd_sv_sample_noalloc: " + std::to_string(shots) + " shots on " + std::to_string(sv->n) +
                                 " qubits need " + std::to_string(need) + " bytes of workspace, more than its fixed " +
                                 std::to_string(NOALLOC_SCRATCH_BYTES) + " (use qvd_sv_sample)");
        void *ptr = nullptr;
        QVD_HIP(noalloc_scratch(&ptr));
        // Every call on this device shares the one workspace: calls from this
        // process are serialised here (the call returns after its GPU work).
        static std::mutex mu;
        std::lock_guard<std::mutex> lock(mu);
        sample_with(sv, uniforms, shots, mark_bits, out, static_cast<char *>(ptr), false, nullptr);
        sv->last_sample_seconds = now() - t0;
    });
}

int qvd_sv_compare(const qvd_sv *a_c, const qvd_sv *b, double *max_abs, double *diff_sq, double *ref_sq) {
    return guard([&] {
        check_handle(a_c);
        check_handle(b);
        require(a_c->n == b->n, "states of different sizes");
        require(a_c->device == b->device, "states on different devices");
        qvd_sv *a = const_cast<qvd_sv *>(a_c);
        a->bind();
        QVD_HIP(hipStreamSynchronize(b->stream));
        const unsigned blocks = 1024;
        double *buf = a->scratch_doubles(3 * blocks);
        with_type(a, [&](auto *x) {
            with_type(b, [&](auto *y) {
                launch_compare(x, y, a->len, buf, blocks, a->stream);
                return 0;
            });
            return 0;
        });
        QVD_HIP(hipGetLastError());
        std::vector<double> h(3 * blocks);
        QVD_HIP(hipMemcpyAsync(h.data(), buf, h.size() * sizeof(double), hipMemcpyDeviceToHost, a->stream));
        a->sync();
        double m = 0, d = 0, r = 0;
        for (unsigned i = 0; i < blocks; i++) {
            m = std::max(m, h[3 * i]);
            d += h[3 * i + 1];
            r += h[3 * i + 2];
        }
        if (max_abs) *max_abs = std::sqrt(m);
        if (diff_sq) *diff_sq = d;
        if (ref_sq) *ref_sq = r;
    });
}

int qvd_sv_last_sample_seconds(const qvd_sv *sv, double *out) {
    return guard([&] {
        check_handle(sv);
        *out = sv->last_sample_seconds;
    });
}


int qvd_batch_create(const qvd_circuit *circuits, uint32_t count, int precision, const qvd_options *options,
                     qvd_batch **out) {
    return guard([&] {
        require(out != nullptr, "null output handle");
        *out = nullptr;
        require(precision == QVD_C64 || precision == QVD_C128, "precision must be QVD_C64 or QVD_C128");
        require(count >= 1 && circuits != nullptr, "a batch needs at least one circuit");
        qvd_options opt{4, 0, QVD_PLANNER_HEURISTIC, 0};
        if (options) opt = *options;
        int devices = 0;
        if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) fail(QVD_EHIP, "no HIP device is visible");
        std::vector<uint32_t> fixed = precision == QVD_C64 ? batch_geometry<float2>(circuits, count, precision, opt)
                                                            : batch_geometry<double2>(circuits, count, precision, opt);
        std::unique_ptr<qvd_batch> b = new_batch(precision, count);
        if (precision == QVD_C64) build_batch<float2>(b.get(), circuits, count, opt, fixed);
        else build_batch<double2>(b.get(), circuits, count, opt, fixed);
        *out = b.release();
    });
}

int qvd_batch_destroy(qvd_batch *batch) {
    return guard([&] {
        if (!batch) return;
        (void)hipSetDevice(batch->device);
        (void)hipStreamSynchronize(batch->stream);
        delete batch;
    });
}

int qvd_batch_amplitudes(const qvd_batch *batch, uint32_t index, void *out, uint64_t capacity) {
    return guard([&] {
        require(batch != nullptr, "null batch handle");
        require(index < batch->count, "circuit index out of range");
        require(out != nullptr, "null output");
        const uint64_t len = uint64_t(1) << batch->qubits[index];
        require(capacity >= len, "output holds fewer than 2^num_qubits amplitudes");
        const size_t amp = batch->precision == QVD_C64 ? sizeof(float2) : sizeof(double2);
        QVD_HIP(hipSetDevice(batch->device));
        // The ancillas are the state's top qubits, in |0>: the circuit's state
        // is the first 2^n amplitudes.
        QVD_HIP(hipMemcpy(out, batch->amps.as<char>() + batch->state_off[index] * amp, len * amp,
                          hipMemcpyDeviceToHost));
    });
}

int qvd_batch_geometry(const qvd_batch *batch, uint32_t *fixed_lanes, uint32_t *stages, uint32_t count) {
    return guard([&] {
        require(batch != nullptr, "null batch handle");
        require(count <= batch->count, "count exceeds the batch's circuits");
        for (uint32_t c = 0; c < count; c++) {
            if (fixed_lanes) fixed_lanes[c] = batch->fixed[c];
            if (stages) stages[c] = batch->circuit_stages[c];
        }
    });
}

int qvd_batch_outputs(const qvd_batch *batch, uint64_t *num_z, uint64_t *num_zz) {
    return guard([&] {
        require(batch != nullptr, "null batch handle");
        if (num_z) *num_z = batch->num_z;
        if (num_zz) *num_zz = batch->num_zz;
    });
}

int qvd_batch_run(qvd_batch *batch, double *z, double *zz, qvd_stats *stats) {
    return guard([&] {
        require(batch != nullptr, "null batch handle");
        QVD_HIP(hipSetDevice(batch->device));
        const double t0 = now();
        QVD_HIP(hipGraphLaunch(batch->exec, batch->stream));
        // The outputs and the queue's error word come back in one pinned
        // buffer, queued behind the replay, so one wait covers everything.
        double *hz = batch->host_out, *hzz = hz + batch->num_z;
        uint32_t *herr = reinterpret_cast<uint32_t *>(hzz + batch->num_zz);
        if (batch->num_z)
            QVD_HIP(hipMemcpyAsync(hz, batch->z.p, batch->num_z * sizeof(double), hipMemcpyDeviceToHost,
                                   batch->stream));
        if (batch->num_zz)
            QVD_HIP(hipMemcpyAsync(hzz, batch->zz.p, batch->num_zz * sizeof(double), hipMemcpyDeviceToHost,
                                   batch->stream));
        QVD_HIP(hipMemcpyAsync(herr, batch->sync.as<uint32_t>() + VQUEUE_ERROR, sizeof(uint32_t),
                               hipMemcpyDeviceToHost, batch->stream));
        QVD_HIP(hipStreamSynchronize(batch->stream));
        const double t1 = now();
        if (*herr) fail(QVD_EHIP, QUEUE_TIMEOUT);
        if (z && batch->num_z) std::memcpy(z, hz, batch->num_z * sizeof(double));
        if (zz && batch->num_zz) std::memcpy(zz, hzz, batch->num_zz * sizeof(double));
        if (stats) {
            qvd_stats st{};
            st.gates = batch->gates;
            st.kernels = batch->kernels;
            st.passes = batch->levels;
            st.stages = batch->stages;
            st.gate_seconds = t1 - t0;
            st.stage_bytes = batch->stage_bytes;
            st.other_bytes = batch->other_bytes;
            st.skipped_tiles = batch->skipped_tiles;
            *stats = st;
        }
    });
}

int qvd_plan_features(uint32_t num_qubits, int precision, const qvd_op *ops, uint64_t num_ops,
                      const qvd_options *options, float *features, uint64_t capacity, uint64_t *rows) {
    return guard([&] {
        require(precision == QVD_C64 || precision == QVD_C128, "precision must be QVD_C64 or QVD_C128");
        require(num_qubits <= 34, "a state vector of more than 34 qubits does not fit in memory");
        require(rows != nullptr, "null output");
        qvd_options opt{4, 0, QVD_PLANNER_HEURISTIC, 0, 0};
        if (options) opt = *options;
        std::vector<Step> steps = make_plan(num_qubits, precision, ops, num_ops, opt);
        std::vector<float> f = plan_features(steps, num_qubits, precision, opt.fixed_lanes);
        const uint64_t r = f.size() / QVD_PLAN_FEATURES;
        *rows = r;
        require(r <= capacity || features == nullptr, "feature buffer too small");
        if (features) std::copy(f.begin(), f.end(), features);
    });
}

int qvd_plan_equivalent_variants(uint32_t num_qubits, int precision, const qvd_op *ops, uint64_t num_ops,
                                 const qvd_options *options, uint32_t *mask) {
    return guard([&] {
        require(precision == QVD_C64 || precision == QVD_C128, "precision must be QVD_C64 or QVD_C128");
        require(num_qubits <= 34, "a state vector of more than 34 qubits does not fit in memory");
        require(mask != nullptr, "null output");
        qvd_options opt{4, 0, QVD_PLANNER_HEURISTIC, 0, 0};
        if (options) opt = *options;
        require(opt.variant < PLAN_VARIANTS, "plan variant out of range");
        opt.planner = QVD_PLANNER_HEURISTIC;
        std::vector<std::vector<Step>> plans(PLAN_VARIANTS);
        for (uint32_t var = 0; var < PLAN_VARIANTS; var++) {
            qvd_options o = opt;
            o.variant = var;
            plans[var] = make_plan(num_qubits, precision, ops, num_ops, o);
        }
        *mask = 0;
        bool reg = true;
        for (const Step &st : plans[opt.variant]) reg &= !st.is_stage || st.stage.reg_tile;
        if (!reg) {
            *mask = 1u << opt.variant;  // LDS-tile plans: only the variant itself
            return;
        }
        for (uint32_t var : equivalent_variants(plans, opt.variant, num_qubits, precision, opt.fixed_lanes))
            *mask |= 1u << var;
    });
}

int qvd_sv_stage_seconds(const qvd_sv *sv, double *seconds, uint64_t capacity, uint64_t *count) {
    return guard([&] {
        check_handle(sv);
        require(count != nullptr, "null output");
        *count = sv->stage_seconds.size();
        if (!seconds) return;
        require(sv->stage_seconds.size() <= capacity, "buffer too small");
        std::copy(sv->stage_seconds.begin(), sv->stage_seconds.end(), seconds);
    });
}

int qvd_mlp_create(uint32_t inputs, uint32_t hidden, qvd_mlp **out) {
    return guard([&] {
        require(out != nullptr, "null output handle");
        *out = nullptr;
        *out = new qvd_mlp(inputs, hidden);
    });
}

int qvd_mlp_destroy(qvd_mlp *mlp) {
    return guard([&] { delete mlp; });
}

int qvd_mlp_num_params(const qvd_mlp *mlp, uint64_t *count) {
    return guard([&] {
        require(mlp != nullptr && count != nullptr, "null argument");
        *count = mlp->net.shape().params();
    });
}

int qvd_mlp_set_params(qvd_mlp *mlp, const float *params, uint64_t count) {
    return guard([&] {
        require(mlp != nullptr && params != nullptr, "null argument");
        require(count == mlp->net.shape().params(), "parameter count does not match the network");
        mlp->net.set_params(params);
    });
}

int qvd_mlp_get_params(const qvd_mlp *mlp, float *params, uint64_t count) {
    return guard([&] {
        require(mlp != nullptr && params != nullptr, "null argument");
        require(count == mlp->net.shape().params(), "parameter count does not match the network");
        mlp->net.get_params(params);
    });
}

int qvd_mlp_forward(qvd_mlp *mlp, const float *x, uint64_t rows, float *y) {
    return guard([&] {
        require(mlp != nullptr && (rows == 0 || (x != nullptr && y != nullptr)), "null argument");
        require(rows < (uint64_t(1) << 31), "too many rows");
        mlp->net.forward(x, rows, y);
    });
}

int qvd_mlp_gradient(qvd_mlp *mlp, const float *x, const float *t, uint64_t rows, float *grad, double *loss) {
    return guard([&] {
        require(mlp != nullptr && x != nullptr && t != nullptr && grad != nullptr, "null argument");
        require(rows >= 1 && rows < (uint64_t(1) << 31), "rows out of range");
        const double l = mlp->net.gradient(x, t, rows, grad);
        if (loss) *loss = l;
    });
}

int qvd_mlp_train(qvd_mlp *mlp, const float *x, const float *t, uint64_t rows, uint32_t epochs, uint32_t batch,
                  float lr, uint64_t seed, double *loss) {
    return guard([&] {
        require(mlp != nullptr && x != nullptr && t != nullptr, "null argument");
        require(rows >= 1 && rows < (uint64_t(1) << 31) && batch >= 1, "rows or batch out of range");
        require(lr > 0, "learning rate must be positive");
        const double l = mlp->net.train(x, t, rows, epochs, batch, lr, seed);
        if (loss) *loss = l;
    });
}

int qvd_planner_set_model(const qvd_mlp *mlp, const float *mean, const float *scale, float y_mean, float y_scale) {
    return guard([&] {
        std::lock_guard<std::mutex> lock(g_model.mu);
        if (!mlp) {
            g_model.net.reset();
            g_model.chosen.clear();
            return;
        }
        require(mean != nullptr && scale != nullptr, "null normalisation");
        require(mlp->net.shape().inputs == QVD_PLAN_FEATURES, "the cost model must take QVD_PLAN_FEATURES inputs");
        for (int k = 0; k < QVD_PLAN_FEATURES; k++) require(scale[k] > 0, "feature scales must be positive");
        require(y_scale > 0, "target scale must be positive");
        std::vector<float> p(mlp->net.shape().params());
        mlp->net.get_params(p.data());
        std::unique_ptr<Mlp> net(new Mlp(mlp->net.shape().inputs, mlp->net.shape().hidden));
        net->set_params(p.data());
        g_model.net = std::move(net);
        g_model.chosen.clear();
        g_model.mean.assign(mean, mean + QVD_PLAN_FEATURES);
        g_model.scale.assign(scale, scale + QVD_PLAN_FEATURES);
        g_model.ymean = y_mean;
        g_model.yscale = y_scale;
    });
}
}  // extern "C"

// Host planning (see plan.h).
#include "plan.h"

#include <algorithm>

namespace qvd {

namespace {

// A group being grown from a seed op (src/fusion.rs:119-172).
struct Group {
    size_t seed;
    std::vector<size_t> members;
    std::vector<uint32_t> qubits;

    template <typename W>
    void grow(const std::vector<Op> &ops, std::vector<char> &done, size_t max_qubits, bool connected_only,
              const W &width) {
        std::vector<uint32_t> blocked;
        for (size_t i = seed + 1; i < ops.size(); i++) {
            if (done[i]) continue;
            if (connected_only) {
                bool all = true;
                for (uint32_t q : qubits)
                    if (!contains(blocked, q)) {
                        all = false;
                        break;
                    }
                if (all) break;
            }
            const std::vector<uint32_t> oq = ops[i].qubits();
            bool touches = false;
            for (uint32_t q : oq) touches |= contains(blocked, q);
            std::vector<uint32_t> fresh;
            for (uint32_t q : oq)
                if (!contains(qubits, q)) fresh.push_back(q);
            const bool connected = fresh.size() < oq.size();
            const bool fits = width(ops[i]) <= max_qubits && qubits.size() + fresh.size() <= max_qubits;
            if (!touches && fits && (connected || !connected_only)) {
                done[i] = 1;
                members.push_back(i);
                qubits.insert(qubits.end(), fresh.begin(), fresh.end());
            } else {
                for (uint32_t q : oq)
                    if (!contains(blocked, q)) blocked.push_back(q);
            }
        }
        std::sort(members.begin(), members.end());
    }
};

}  // namespace

template <typename F>
std::vector<Op> fuse_where(const std::vector<Op> &ops, size_t max_qubits, const F &fusable) {
    auto width = [&](const Op &op) { return fusable(op) ? op.width() : SIZE_MAX; };
    std::vector<char> done(ops.size(), 0);
    std::vector<Op> out;
    for (size_t seed = 0; seed < ops.size(); seed++) {
        if (done[seed]) continue;
        done[seed] = 1;
        Group g{seed, {seed}, ops[seed].qubits()};
        if (width(ops[seed]) <= max_qubits) {
            g.grow(ops, done, max_qubits, true, width);
            g.grow(ops, done, max_qubits, false, width);
        }
        if (g.members.size() == 1) {
            out.push_back(ops[seed]);
            continue;
        }
        std::sort(g.qubits.begin(), g.qubits.end());
        HMatrix m = HMatrix::identity((uint32_t)g.qubits.size());
        for (size_t i : g.members) {
            std::vector<uint32_t> oq;
            HMatrix om;
            ops[i].full(oq, om);
            m = om.embed(oq, g.qubits).mul(m);
        }
        Op f;
        f.targets = g.qubits;
        f.m = std::move(m);
        f.diagonal = f.m.is_diagonal();
        out.push_back(std::move(f));
    }
    return out;
}

std::vector<Op> fuse(const std::vector<Op> &ops, size_t max_qubits) {
    return fuse_where(ops, max_qubits, [](const Op &) { return true; });
}

namespace {

// Take every pending op that can run now and whose qubits no earlier
// pending op blocks (src/blocking.rs:157-179).
template <typename F>
std::vector<size_t> extract(const std::vector<Op> &ops, const std::vector<char> &done, F &&can_run) {
    std::vector<size_t> taken;
    std::vector<uint32_t> blocked;
    for (size_t i = 0; i < ops.size(); i++) {
        if (done[i]) continue;
        const std::vector<uint32_t> oq = ops[i].qubits();
        bool touches = false;
        for (uint32_t q : oq) touches |= contains(blocked, q);
        if (!touches && can_run(ops[i])) {
            taken.push_back(i);
        } else {
            for (uint32_t q : oq)
                if (!contains(blocked, q)) blocked.push_back(q);
        }
    }
    return taken;
}

// extract over qubit masks (bit q = qubit q): the same ops in the same
// order, without building each op's qubit list on every pass. can_run takes
// the op's index.
template <typename F>
std::vector<size_t> extract_masked(const std::vector<uint64_t> &qmask, const std::vector<char> &done, F &&can_run) {
    std::vector<size_t> taken;
    uint64_t blocked = 0;
    for (size_t i = 0; i < qmask.size(); i++) {
        if (done[i]) continue;
        if (!(qmask[i] & blocked) && can_run(i)) taken.push_back(i);
        else blocked |= qmask[i];
    }
    return taken;
}

uint64_t mask_of(const std::vector<uint32_t> &qs) {
    uint64_t m = 0;
    for (uint32_t q : qs) m |= uint64_t(1) << q;
    return m;
}

}  // namespace

std::vector<Step> plan_unblocked(const std::vector<Op> &ops, uint32_t max_fused) {
    std::vector<Step> steps;
    for (Op &op : fuse(ops, std::max<uint32_t>(max_fused, 1))) {
        Step s;
        s.is_stage = false;
        s.op = std::move(op);
        steps.push_back(std::move(s));
    }
    return steps;
}

std::vector<Step> plan_heuristic(const std::vector<Op> &ops, const PlanParams &p) {
    const uint32_t n = p.num_qubits;
    const size_t max_fused = std::max<uint32_t>(p.max_fused, 1);
    std::vector<Step> steps;
    if (ops.empty()) return steps;
    const uint32_t L = std::min(p.tile_bits, n);
    const uint32_t contig = std::min(p.contig, L);
    require(n <= 64, "internal: the planner's qubit masks hold 64 qubits");
    std::vector<char> is_local(n, 0);
    std::vector<uint32_t> local;
    for (uint32_t q = 0; q < L; q++) {
        local.push_back(q);
        is_local[q] = 1;
    }
    // Every op's qubits and targets as masks, and whether it may run in a
    // tile at all (diagonal, or dense and narrow enough for the in-tile
    // kernel).
    std::vector<uint64_t> qmask(ops.size()), tmask(ops.size());
    std::vector<char> stageable(ops.size());
    for (size_t i = 0; i < ops.size(); i++) {
        qmask[i] = mask_of(ops[i].qubits());
        tmask[i] = mask_of(ops[i].targets);
        stageable[i] = ops[i].diagonal || ops[i].targets.size() <= p.max_stage_op;
    }
    const uint64_t low = contig >= 64 ? ~uint64_t(0) : (uint64_t(1) << contig) - 1;
    // A dense op runs inside a tile when its targets are tile qubits and it
    // is narrow enough for the in-tile kernel; a diagonal op always runs
    // (its non-tile targets and every non-tile control are fixed per tile).
    auto runs_under = [&](size_t i, uint64_t tile) {
        return ops[i].diagonal || (stageable[i] && (tmask[i] & ~tile) == 0);
    };
    // Coupling graph of the circuit (qubits that share an op), for the
    // region candidates (variant bit 4 clear).
    const bool regions = (p.variant & 16) == 0;
    std::vector<std::vector<uint32_t>> adj;
    if (regions) {
        adj.assign(n, {});
        for (const Op &op : ops) {
            const std::vector<uint32_t> qs = op.qubits();
            for (uint32_t a : qs)
                for (uint32_t b : qs)
                    if (a != b && !contains(adj[a], b)) adj[a].push_back(b);
        }
        for (auto &nb : adj) std::sort(nb.begin(), nb.end());
    }
    std::vector<char> done(ops.size(), 0);
    size_t remaining = ops.size();
    while (remaining > 0) {
        const uint64_t local_mask = mask_of(local);
        std::vector<size_t> taken = extract_masked(qmask, done, [&](size_t i) { return runs_under(i, local_mask); });
        if (!taken.empty()) {
            std::vector<Op> stage_ops;
            for (size_t i : taken) {
                done[i] = 1;
                stage_ops.push_back(ops[i]);
            }
            remaining -= taken.size();
            Step s;
            s.stage.local = local;
            s.stage.reg_tile = p.reg_tile;
            s.stage.role_policy = (p.variant & 8) ? 0 : 1;
            if (p.reg_tile) {
                // The register-tile program merges 1-qubit runs itself and
                // keeps gates in unit form; wider fusion would only add
                // arithmetic to a compute-bound tile.
                s.stage.ops = std::move(stage_ops);
            } else {
                // Fuse only ops on tile qubits: merging a non-tile control or
                // diagonal target into a dense matrix would need it in the tile.
                s.stage.ops = fuse_where(stage_ops, std::min<size_t>(max_fused, p.max_stage_op), [&](const Op &op) {
                    for (uint32_t q : op.qubits())
                        if (!is_local[q]) return false;
                    return true;
                });
            }
            steps.push_back(std::move(s));
        }
        if (remaining == 0) break;
        // Lookahead (src/blocking.rs:181-226): the qubits the coming ops
        // need, within the tile's free capacity.
        const uint32_t full_cap = L - contig;
        auto wanted_set = [&](uint32_t capacity) {
            uint64_t wanted = 0;
            extract_masked(qmask, done, [&](size_t i) {
                if (ops[i].diagonal) return true;
                if (!stageable[i]) return false;
                const uint64_t fresh = tmask[i] & ~low & ~wanted;
                if ((uint32_t)(__builtin_popcountll(wanted) + __builtin_popcountll(fresh)) > capacity) return false;
                wanted |= fresh;
                return true;
            });
            std::vector<uint32_t> qs;
            for (uint32_t q = 0; q < n; q++)
                if ((wanted >> q) & 1) qs.push_back(q);
            return qs;
        };
        auto tile_of = [&](const std::vector<uint32_t> &wanted) {
            std::vector<uint32_t> next;
            for (uint32_t q = 0; q < contig; q++) next.push_back(q);
            for (uint32_t q : wanted) next.push_back(q);
            for (uint32_t q = contig; q < n && next.size() < L; q++)
                if (!contains(next, q)) next.push_back(q);
            std::sort(next.begin(), next.end());
            return next;
        };
        const uint32_t cut = std::min<uint32_t>(p.variant & 3, full_cap > 1 ? full_cap - 1 : 0);
        std::vector<uint32_t> next = tile_of(wanted_set(full_cap - cut));
        if (p.variant & 4) {
            // The candidate tile that runs the most pending ops next.
            auto runnable_count = [&](const std::vector<uint32_t> &tile) {
                const uint64_t tm = mask_of(tile);
                return extract_masked(qmask, done, [&](size_t i) { return runs_under(i, tm); }).size();
            };
            std::vector<std::vector<uint32_t>> cands;
            for (uint32_t c = full_cap; c >= 1 && c + 4 > full_cap; c--) cands.push_back(tile_of(wanted_set(c)));
            // the most-targeted qubits among the next pending dense ops
            std::vector<uint32_t> count(n, 0), first(n, UINT32_MAX);
            uint32_t seen = 0;
            for (size_t i = 0; i < ops.size() && seen < 8 * L; i++) {
                if (done[i] || ops[i].diagonal) continue;
                for (uint32_t q : ops[i].targets)
                    if (q >= contig) {
                        count[q]++;
                        first[q] = std::min(first[q], seen);
                    }
                seen++;
            }
            std::vector<uint32_t> byuse;
            for (uint32_t q = contig; q < n; q++)
                if (count[q]) byuse.push_back(q);
            std::stable_sort(byuse.begin(), byuse.end(), [&](uint32_t a, uint32_t b) {
                return count[a] != count[b] ? count[a] > count[b] : first[a] < first[b];
            });
            if (byuse.size() > full_cap) byuse.resize(full_cap);
            cands.push_back(tile_of(byuse));
            size_t best = next == local ? 0 : runnable_count(next);
            for (const auto &cnd : cands) {
                if (cnd == local) continue;
                const size_t k = runnable_count(cnd);
                if (k > best) {
                    best = k;
                    next = cnd;
                }
            }
        }
        if (regions) {
            // Region candidates: for every seed qubit, the breadth-first
            // region of the coupling graph that fills the free capacity (a
            // window along a line, a patch of a lattice). The next tile is the
            // candidate (with the lookahead's) under which the most pending
            // ops run now, so the tiles sweep the circuit's light cones
            // instead of chasing its earliest pending op.
            auto runnable = [&](const std::vector<uint32_t> &tile) {
                const uint64_t tm = mask_of(tile);
                return extract_masked(qmask, done, [&](size_t i) { return runs_under(i, tm); }).size();
            };
            std::vector<std::vector<uint32_t>> cands{next};
            for (uint32_t seed = contig; seed < n; seed++) {
                std::vector<uint32_t> region{seed};
                for (size_t i = 0; i < region.size() && region.size() < full_cap; i++)
                    for (uint32_t w : adj[region[i]]) {
                        if (region.size() >= full_cap) break;
                        if (w >= contig && !contains(region, w)) region.push_back(w);
                    }
                std::vector<uint32_t> cand = tile_of(region);
                if (std::find(cands.begin(), cands.end(), cand) == cands.end()) cands.push_back(std::move(cand));
            }
            // A region replaces the lookahead's tile only when more pending
            // ops run under it; one under which nothing runs (an op wider
            // than a stage is next) never does, so such an op still gets its
            // own full-state pass below.
            size_t best = next == local ? 0 : runnable(next);
            for (const auto &cnd : cands) {
                if (cnd == local) continue;
                const size_t k = runnable(cnd);
                if (k > best) {
                    best = k;
                    next = cnd;
                }
            }
        }
        bool same = next == local;
        if (same) {
            // Nothing new can be brought in (an op wider than the tile's
            // free capacity): one HBM pass for the first pending op.
            size_t i = 0;
            while (done[i]) i++;
            done[i] = 1;
            remaining--;
            Step s;
            s.is_stage = false;
            s.op = ops[i];
            steps.push_back(std::move(s));
            continue;
        }
        local = next;
        std::fill(is_local.begin(), is_local.end(), 0);
        for (uint32_t q : local) is_local[q] = 1;
    }
    return steps;
}

}  // namespace qvd

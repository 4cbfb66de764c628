// Host planning: gate fusion and LDS-tile stage extraction.
//
// The heuristic planner restates the reference's greedy fusion
// (src/fusion.rs:63-171) and its dependency-aware stage extraction with
// lookahead (src/blocking.rs:150-302). The difference is what a stage's
// "local" qubits are: in the reference they are the low physical bits of a
// 1 MiB region and a relabeling pass swaps new qubits in; here a stage picks
// any set of tile qubits (the lowest `contig` bits always included so every
// HBM access is a contiguous run of 2^contig amplitudes) and the tile loader
// gathers them, so no relabeling pass exists and the layout stays logical.
#pragma once

#include <cstdint>
#include <vector>

#include "common.h"

namespace qvd {

// Fuse ops into groups on at most max_qubits qubits; ops for which fusable
// is false never merge (src/fusion.rs:69-115).
template <typename F>
std::vector<Op> fuse_where(const std::vector<Op> &ops, size_t max_qubits, const F &fusable);
std::vector<Op> fuse(const std::vector<Op> &ops, size_t max_qubits);

// A stage: every op acts on tile qubits only (or is diagonal / controlled
// on non-tile qubits, resolved per tile).
struct Stage {
    std::vector<uint32_t> local;  // tile bit j -> physical qubit local[j], ascending
    std::vector<Op> ops;          // program order (fused for LDS-tile stages)
    bool reg_tile = false;        // run by the register-tile kernel (ops unfused)
    uint32_t role_policy = 0;     // register-tile programs: 0 = bits in order of need, 1 = most-used bits first
};

// One step of a plan: an LDS-tile stage, or a single op applied as its own
// HBM pass (an op with more non-tile targets than a tile holds).
struct Step {
    bool is_stage = true;
    Stage stage;
    Op op;
};

struct PlanParams {
    uint32_t num_qubits = 0;
    uint32_t tile_bits = 13;  // L
    uint32_t contig = 5;      // lowest qubits always in the tile
    uint32_t max_fused = 4;
    uint32_t max_stage_op = 4;  // widest dense op applied inside a tile
    bool reg_tile = false;      // stages for the register-tile kernel: ops kept unfused
    // Plan variant (0 = the heuristic). Bits 0-1: the lookahead's capacity is
    // reduced by that many qubits; bit 2: the next tile is the candidate (four
    // lookahead capacities and the most-targeted qubits) that runs the most
    // ops; bit 3: register-tile programs give register slots to bits in order
    // of need (clear: the most-used bits first); bit 4: the next tile is the
    // lookahead's alone, as src/blocking.rs:181-226 chooses it (clear: the
    // lookahead's or, if more pending ops run under it, the region of the
    // circuit's coupling graph grown from one seed qubit). Every variant
    // applies the same ops in a dependency-respecting order, so the amplitudes
    // agree up to rounding; two variants give bit-identical amplitudes only
    // when they run the same programs, since a different order of two gates
    // on different qubits rounds differently.
    uint32_t variant = 0;
};

constexpr uint32_t PLAN_VARIANTS = 32;

// The heuristic plan: src/blocking.rs's extract / plan_swaps loop with the
// lookahead's wanted set becoming the next stage's tile qubits.
std::vector<Step> plan_heuristic(const std::vector<Op> &ops, const PlanParams &p);

// Plan without blocking: fused ops, each its own HBM pass (options.region_bits = None).
std::vector<Step> plan_unblocked(const std::vector<Op> &ops, uint32_t max_fused);

}  // namespace qvd

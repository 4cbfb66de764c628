// Descriptors shared by the host launcher and the HIP kernels.
#pragma once

#include <cstdint>

namespace qvd {

constexpr uint32_t MAX_QUBITS = 36;
constexpr uint32_t MAX_GATE_QUBITS = 6;  // widest op the kernels apply (src/kernels.rs:475-476)
constexpr uint32_t GLOBAL_BIT = 0x80000000u;

enum StageOpKind : uint32_t { OP_DENSE = 0, OP_DIAG = 1 };

// One op of an LDS-tile stage.
//  - dense: tgt[j] = tile bit of matrix bit j; ins[] = the tile bits of the
//    targets and local controls, ascending (the group index is spread over
//    the remaining tile bits); lc_* = local controls in tile bits.
//  - diagonal: tgt[j] = tile bit of matrix bit j, or GLOBAL_BIT | physical
//    qubit for a non-tile target (fixed per tile).
//  - gc_* = non-tile controls as physical bits, checked once per tile.
struct StageOp {
    uint32_t kind;
    uint32_t k;
    uint32_t tgt[MAX_GATE_QUBITS];
    uint32_t nins;
    uint32_t ins[2 * MAX_GATE_QUBITS];
    uint32_t lc_mask, lc_val;
    uint64_t gc_mask, gc_val;
    uint32_t coef_off;  // first coefficient (complex) of this op
    uint32_t pad;
};

// One stage: the tile's qubits and the ops applied to every tile.
struct StageDesc {
    uint32_t L;        // tile bits
    uint32_t nglob;    // non-tile bits
    uint32_t num_ops;
    uint32_t op_off;   // first StageOp of the stage in the op array
    uint64_t num_tiles;
    uint32_t local_pos[MAX_QUBITS];  // tile bit j -> physical bit, ascending
    uint32_t glob_pos[MAX_QUBITS];   // tile-index bit j -> physical bit, ascending
};

// A full-state op (one HBM pass).
struct GateDesc {
    uint32_t k;
    uint32_t nins;                    // sorted targets + controls
    uint32_t ins[MAX_QUBITS];         // ascending physical bits
    uint64_t offsets[1 << MAX_GATE_QUBITS];  // amplitude offset of matrix index j
    uint64_t ctrl_val;                // control bits that must be set
    uint64_t iterations;              // 2^(n - k - controls)
};

}  // namespace qvd

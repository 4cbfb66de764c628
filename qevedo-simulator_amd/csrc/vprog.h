// Register-tile stage programs.
//
// A register-tile stage keeps its tile of 2^L amplitudes in VGPRs: thread t
// of a workgroup of W = 64 * 2^w threads holds R = 2^r amplitudes, and
// L = 6 + w + r. Tile bits play three roles:
//   - lane bits (the 6 low bits of the thread index): the lowest F of them
//     are tile bits 0..F-1, physical qubits 0..F-1, so every access is a run
//     of 2^F consecutive amplitudes (128 bytes); the other 6 - F lane bits
//     hold any tile bits. Gates on lane bits use cross-lane shuffles;
//   - wave bits: the remaining bits of the thread index; a gate on one needs
//     an exchange through LDS that makes it a lane or register bit;
//   - register bits: the index into the thread's registers; gates on them
//     are thread-local arithmetic.
// A program is a list of VInstr the kernel interprets in order; every
// workgroup runs the same program on each of its tiles.
#pragma once

#include <cstddef>

#include <cstdint>
#include <vector>

#include "common.h"
#include "device.h"
#include "plan.h"

namespace qvd {

enum VOp : uint32_t {
    V_U1R = 0,   // 1-qubit gate on a register bit
    V_U1L = 1,   // 1-qubit gate on a lane bit
    V_DIAG = 2,  // diagonal on up to 6 bits (thread, register or non-tile)
    V_XCHG = 3,  // exchange wave bits and register bits through LDS
    V_LDSOP = 4, // any op of up to 4 targets, applied with the tile in LDS
    V_SCALE = 5, // multiply by per-thread x per-register x per-tile factors
    V_SIGNS = 6, // a run of commuting CZ-like gates: sign = parity of a quadratic form over bits
    V_PHASES = 7, // a run of commuting pair phases: factor = product of w over pairs with both bits 1
};

// VInstr::op holds the VOp in its low byte, these flags above it and, in
// bits 16-23, the kernel code path (VHandler) that runs the instruction, so
// the kernel reads one word and branches once to dispatch, and loads other
// fields only where used.
constexpr uint32_t VOP_MASK = 0xffu;
constexpr uint32_t VF_CTRL = 1u << 8;   // controls on thread or register bits (tmask / rmask nonzero)
constexpr uint32_t VF_GCTRL = 1u << 31;  // controls on non-tile qubits (gmask nonzero): the sign bit, one compare
constexpr uint32_t VH_SHIFT = 16;

// Register bits a handler can name (the geometries use 4 or 5).
constexpr uint32_t VREG_SLOTS = 5;

// Kernel code paths. Each fixes at compile time what the VOp, form, slot and
// rotation leave open, so an instruction costs one switch in the kernel.
enum VHandler : uint32_t {
    H_U1R_ROT = 0,    // + 4 * register bit + k          (FORM_UNIT_ROT)
    H_U1R_DIAG = 20,  // + register bit
    H_U1R_ANTI = 25,  // + register bit
    H_U1R_FULL = 30,  // + register bit
    H_U1R_CTRL = 35,  // + register bit                  (full form, controls)
    H_U1L_ROT = 40,   // + 4 * lane bit + k              (FORM_UNIT_ROT)
    H_U1L_DIAG = 64,  // + lane bit
    H_U1L_ANTI = 70,  // + lane bit
    H_U1L_FULL = 76,  // + lane bit
    H_U1L_CTRL = 82,  // + lane bit                      (full form, controls)
    H_DIAG = 88,      // + 2 * threaded + (form == 1)
    H_SIGNS = 92,
    H_PHASES = 93,
    H_SCALE1 = 94,    // SCALE form 1: one uniform factor
    H_SCALE = 95,
    H_XCHG = 96,
    H_LDSOP = 97,
    H_U1R_IMAG = 98,  // + register bit                  (FORM_UNIT_IMAG)
    H_U1R_REAL = 103, // + register bit                  (FORM_UNIT_REAL)
    H_U1L_IMAG = 108, // + lane bit
    H_U1L_REAL = 114, // + lane bit
    H_COUNT = 120,
};
static_assert(H_U1R_DIAG == 4 * VREG_SLOTS && H_U1R_IMAG + 2 * VREG_SLOTS == H_U1L_IMAG, "register-bit handler blocks");

// A V_SIGNS instruction's masks in VInstr::inl: inl[r] (r < VSIGN_TT) the
// thread bits that flip register bit r's linear term, inl[VSIGN_TT + t]
// (t < 10) the thread bits u > t whose product with bit t enters the parity.
constexpr uint32_t VSIGN_TT = 6;

// Forms of a 1-qubit gate's coefficients c[0..3]:
//   FORM_FULL      [[c0, c1], [c2, c3]]
//   FORM_UNIT_DIAG [[1, c1], [c2, 1]]   (4 fused multiply-adds per amplitude)
//   FORM_UNIT_ANTI [[c0, 1], [1, c3]]
//   FORM_UNIT_ROT  [[1, i^k], [i^(2-k), 1]]  (an add and an FMA by 2 per part; k = aux & 3)
//   FORM_UNIT_IMAG [[1, i s1], [i s2, 1]]    (c1 = (0, s1), c2 = (0, s2): one FMA per part, as for
//                                             RX-like gates)
//   FORM_UNIT_REAL [[1, s1], [s2, 1]]        (c1 = (s1, 0), c2 = (s2, 0): one FMA per part, as for
//                                             RY-like gates)
enum VForm : uint32_t {
    FORM_FULL = 0,
    FORM_UNIT_DIAG = 1,
    FORM_UNIT_ANTI = 2,
    FORM_UNIT_ROT = 3,
    FORM_UNIT_IMAG = 4,
    FORM_UNIT_REAL = 5,
};

constexpr uint32_t VT_THREAD = 0u << 16;  // target bit is bit x of the thread index
constexpr uint32_t VT_REG = 1u << 16;     // bit x of the register index
constexpr uint32_t VT_GLOBAL = 2u << 16;  // physical qubit x outside the tile
constexpr uint32_t VT_KIND = 3u << 16;

constexpr uint32_t VMAX_W_BITS = 4;
constexpr uint32_t VMAX_R_BITS = 6;
constexpr uint32_t VMAX_L_FLEX = 4;  // flexible lane bits
constexpr uint32_t VMAX_REGS = 32;

// The fields the common instructions read (the op word, counts, offsets, a
// sign run's mask and a complex64 gate's coefficients) fill the first
// 64 bytes, and the struct is 64-byte aligned, so interpreting such an
// instruction touches one scalar-cache line: a long stage program keeps its
// hot lines in the scalar cache instead of streaming whole records.
struct alignas(64) VInstr {
    uint32_t op;    // VOp | VF_* flags
    uint32_t k;     // DIAG: number of targets; SCALE: number of non-tile factors; SIGNS / PHASES: pairs
    uint32_t aux;   // LDSOP: index of its StageOp; SCALE: first non-tile qubit in the int array
    uint32_t coef;  // first coefficient (complex) of the op
    uint32_t form;
    uint32_t slot;  // U1R: register bit; U1L: lane bit
    uint64_t neg;   // SIGNS: per-register signs; DIAG, form 1: entries equal to -1 (the others are +1)
    // U1R / U1L: the four matrix coefficients in the state's precision (4 x
    // float2 or 4 x double2), written when the program is uploaded, so the
    // kernel reads them from the instruction instead of following coef.
    // PHASES: inl[0] is the offset of its pair records in the int array.
    uint32_t inl[16];
    uint32_t tmask, tval;  // controls on thread-index bits
    uint32_t rmask, rval;  // controls on register-index bits
    uint64_t gmask, gval;  // controls on non-tile qubits (checked per tile)
    uint32_t tgt[MAX_GATE_QUBITS];            // DIAG targets (VT_* | bit)
    uint32_t threaded;                        // DIAG: some target or control is a thread-index bit
    uint8_t rtab[VMAX_REGS];                  // DIAG: register-index part of the entry index, per register
    uint8_t lmap[VMAX_L_FLEX];                // XCHG: tile bit of each flexible lane bit
    uint8_t wmap[VMAX_W_BITS];                // XCHG: tile bit of each wave bit
    uint8_t rmap[VMAX_R_BITS];                // XCHG: tile bit of each register bit
};

// The handler of an instruction (VF_* flags and fields already set); the
// register-bit count must be at most VREG_SLOTS.
uint32_t vhandler(const VInstr &v);

static_assert(offsetof(VInstr, neg) == 24 && offsetof(VInstr, inl) == 32 && sizeof(VInstr) % 64 == 0,
              "a complex64 gate's fields and a sign run's mask sit in an instruction's first 64 bytes");

// One register-tile stage as launched. A launch runs one or more of them
// (distinct states of a batch), each on its own range of workgroups.
struct VStageDesc {
    uint32_t L;
    uint32_t nglob;
    uint32_t num_instr;
    uint32_t instr_off;
    uint64_t num_tiles;
    uint64_t state_off;              // first amplitude of the stage's state in the launch's buffer
    uint32_t local_pos[MAX_QUBITS];  // tile bit -> physical qubit (ascending)
    uint32_t glob_pos[MAX_QUBITS];   // tile-index bit -> physical qubit
    uint8_t lmap0[VMAX_L_FLEX];      // role map at load
    uint8_t wmap0[VMAX_W_BITS];
    uint8_t rmap0[VMAX_R_BITS];
    uint8_t lmapF[VMAX_L_FLEX];      // role map at store
    uint8_t wmapF[VMAX_W_BITS];
    uint8_t rmapF[VMAX_R_BITS];
    uint32_t cbegin, ccount;         // the stage's coefficients, copied to LDS once per workgroup
    uint32_t from_zero;              // nonzero: the state is |0...0> before this stage (its tiles are not loaded)
    uint32_t fixed_lanes;            // F of the geometry the program was built for
    uint64_t zero_mem;               // stale qubits: an amplitude with one of these bits set is zero and its memory
                                     // has not been written since |0...0> (it is not loaded)
};

// LDS budget of a stage's coefficient block, in bytes.
constexpr uint32_t VCONST_BYTES = 24 * 1024;

// Geometry of the register-tile kernel for one precision.
struct VGeom {
    uint32_t w_bits;  // wave bits (W = 64 << w_bits threads)
    uint32_t r_bits;  // register bits
    uint32_t f_bits;  // fixed lane bits (tile bits 0..F-1 = physical qubits 0..F-1)
    uint32_t L() const { return 6 + w_bits + r_bits; }
    uint32_t flex_lanes() const { return 6 - f_bits; }
};

// Per-qubit diagonal factors not yet applied to the amplitudes: the engine
// keeps 1-qubit gates in unit form (D * M with M's diagonal or off-diagonal
// equal to 1) and carries D forward to the next gate on that qubit, where it
// is absorbed, or to the end of the stage, where a SCALE applies all of them.
struct PendingDiag {
    std::vector<C64> d0, d1;  // per physical qubit
    explicit PendingDiag(uint32_t n = 0) : d0(n, C64(1, 0)), d1(n, C64(1, 0)) {}
    bool identity(uint32_t q) const { return d0[q] == C64(1, 0) && d1[q] == C64(1, 0); }
};

// Builds the program of one stage. `coefs` (complex128) and `ldsops` are
// appended to; the caller converts coefficients to the state's precision.
struct VProgram {
    std::vector<VInstr> instr;
    VStageDesc desc;
};

VProgram build_vprogram(const Stage &stage, uint32_t n, const VGeom &g, PendingDiag &pending,
                        std::vector<C64> &coefs, std::vector<StageOp> &ldsops, std::vector<uint32_t> &ints,
                        bool flush_all);

}  // namespace qvd

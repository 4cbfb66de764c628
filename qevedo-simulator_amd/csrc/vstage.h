// Launcher of the register-tile stage kernel (vstage.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"
#include "vprog.h"

namespace qvd {

template <typename T2> VGeom vstage_geom();
template <typename T2> size_t vstage_lds_bytes();
// Runs ndesc register-tile stages in one launch of `grid` workgroups:
// descriptor i (device memory) owns workgroups block_start[i] ..
// block_start[i + 1] - 1 (device memory, ndesc + 1 entries; unused and may
// be null when ndesc == 1). All pointers are device pointers. lazy_zero: a
// descriptor may have from_zero set (its tiles are made, not loaded) or
// stale qubits in zero_mem (their amplitudes are set to zero, not loaded).
template <typename T2>
hipError_t launch_vstage(T2 *s, const VStageDesc *descs, uint32_t ndesc, const uint32_t *block_start,
                         const VInstr *prog, const T2 *coefs, const StageOp *ldsops, const uint32_t *ints, unsigned grid,
                         hipStream_t st, bool lazy_zero = false, uint32_t fixed_lanes = 0);

// One item of a stage queue: tiles tile0, tile0 + step, ... (count of them)
// of stage `desc`, which belongs to state `state` and may start once that
// state's completed-tile counter has reached `need` (the tiles of its earlier
// stages).
struct VQueueItem {
    uint32_t desc;
    uint32_t state;
    uint32_t need;
    uint32_t count;
    uint64_t tile0;
    uint64_t step;
};

// Words of a stage queue's synchronization buffer before the per-state
// counters: the next item to take, and an error flag.
constexpr uint32_t VQUEUE_HEAD = 0, VQUEUE_ERROR = 1, VQUEUE_COUNTERS = 2;
// A wait longer than this (in 100 MHz real-time ticks, 10 s) sets the error
// flag instead of waiting on.
constexpr uint64_t VQUEUE_WAIT_TICKS = 1000000000ull;

// Runs every stage of several states as one launch: `grid` workgroups take
// items (device memory, `nitems` of them, each item's dependencies earlier
// in the array) in order from a queue, wait until the item's state has
// finished its earlier stages, run the item's tiles and count them done.
// `sync` (device memory) holds VQUEUE_COUNTERS words and one counter per
// state, all zero at launch; the error word is nonzero afterwards when a wait
// ran out.
template <typename T2>
hipError_t launch_vstage_queue(T2 *s, const VStageDesc *descs, const VQueueItem *items, uint32_t nitems,
                               uint32_t *sync, const VInstr *prog, const T2 *coefs, const StageOp *ldsops,
                               const uint32_t *ints, unsigned grid, hipStream_t st);

// Workgroups of the stage kernel the device holds at once (per-CU
// occupancy times the CU count).
template <typename T2>
hipError_t vstage_resident_groups(unsigned *groups);

}  // namespace qvd

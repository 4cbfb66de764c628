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
// be null when ndesc == 1). All pointers are device pointers.
template <typename T2>
hipError_t launch_vstage(T2 *s, const VStageDesc *descs, uint32_t ndesc, const uint32_t *block_start,
                         const VInstr *prog, const T2 *coefs, const StageOp *ldsops, const uint32_t *ints, unsigned grid,
                         hipStream_t st);

}  // namespace qvd

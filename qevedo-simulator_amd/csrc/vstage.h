// Launcher of the register-tile stage kernel (vstage.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"
#include "vprog.h"

namespace qvd {

template <typename T2> VGeom vstage_geom();
template <typename T2> size_t vstage_lds_bytes();
template <typename T2>
hipError_t launch_vstage(T2 *s, const VStageDesc &d, const VInstr *prog, const T2 *coefs, const StageOp *ldsops,
                         const uint32_t *ints, unsigned grid, hipStream_t st);

}  // namespace qvd

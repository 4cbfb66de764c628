// Launcher of the FP32 matrix-core dense kernel for fused 5- and 6-qubit
// gates in complex64 (dense_mfma.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace qvd {

typedef float floatx4 __attribute__((ext_vector_type(4)));

// One pass of a dense op with d.k = 5 or 6 targets over the controlled
// subspace (same descriptor as launch_dense); mat is the 2^k x 2^k matrix,
// row-major, in device memory. cus = compute units (grid size).
hipError_t launch_dense_mfma(float2 *s, const GateDesc &d, const float2 *mat, unsigned cus, hipStream_t st);

}  // namespace qvd

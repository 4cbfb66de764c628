// Fused dense gates of K = 5 or 6 targets in complex64, applied as real
// GEMMs on FP32 matrix cores (v_mfma_f32_16x16x4f32, FP32 accumulation).
//
// A column is one group of D = 2^K amplitudes the gate mixes (the iteration
// index spread over the non-target, non-control bits, as in k_dense). With
// U = Ur + i Ui and a column a = ar + i ai, the output is
//     [or; oi] = R [ar; ai],   R = [[Ur, -Ui], [Ui, Ur]]   (2D x 2D real),
// so a block of NC columns is the real product (2D x 2D) x (2D x NC).
// A workgroup of 4 waves keeps R in LDS for all of its column blocks, stages
// each block's amplitudes through LDS (coalesced loads and stores), and each
// wave computes the 2D x 16 output of its 16 columns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

#include "dense_mfma.h"

namespace qvd {

namespace {

constexpr int NC = 64;         // columns per block
constexpr int WAVES = 4;       // 16 columns each
constexpr int PAD = NC + 1;    // row pitch of the staged block (floats)

__device__ __forceinline__ uint64_t spread(uint64_t x, const uint32_t *pos, uint32_t m) {
    for (uint32_t j = 0; j < m; j++) {
        const uint32_t p = pos[j];
        const uint64_t low = x & ((uint64_t(1) << p) - 1);
        x = ((x >> p) << (p + 1)) | low;
    }
    return x;
}

template <int K>
__global__ __launch_bounds__(64 * WAVES) void k_dense_mfma(float2 *__restrict__ s, const GateDesc d,
                                                           const float2 *__restrict__ mat) {
    constexpr int D = 1 << K, M = 2 * D;  // R is M x M
    extern __shared__ float lds[];
    float *R = lds;              // M x M, row-major
    float *Br = R + M * M;       // D x PAD: real parts, row j = target index, column = block column
    float *Bi = Br + D * PAD;    // D x PAD: imaginary parts
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    for (uint32_t e = tid; e < (uint32_t)(M * M); e += 64 * WAVES) {
        const uint32_t r = e / M, k = e % M;
        const uint32_t rr = r % D, kk = k % D;
        const float2 u = mat[rr * D + kk];
        const bool rim = r >= (uint32_t)D, kim = k >= (uint32_t)D;
        R[e] = rim == kim ? u.x : (rim ? u.y : -u.y);
    }

    // Load order: target index fastest when the lowest target is qubit 0
    // (then a column's amplitudes are contiguous), else column fastest.
    const bool jfast = d.offsets[1] == 1;
    const uint64_t blocks = (d.iterations + NC - 1) / NC;
    for (uint64_t blk = blockIdx.x; blk < blocks; blk += gridDim.x) {
        const uint64_t c0 = blk * NC;
        __syncthreads();  // R written / previous block's output read
        for (uint32_t e = tid; e < (uint32_t)(D * NC); e += 64 * WAVES) {
            const uint32_t j = jfast ? e % D : e / NC, c = jfast ? e / D : e % NC;
            float2 a = make_float2(0.0f, 0.0f);
            if (c0 + c < d.iterations) a = s[(spread(c0 + c, d.ins, d.nins) | d.ctrl_val) + d.offsets[j]];
            Br[j * PAD + c] = a.x;
            Bi[j * PAD + c] = a.y;
        }
        __syncthreads();

        // acc[mt] = rows mt*16 .. mt*16+15 of this wave's 16 columns.
        floatx4 acc[M / 16];
#pragma unroll
        for (int mt = 0; mt < M / 16; mt++) acc[mt] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        const uint32_t col = wave * 16 + (lane & 15), ksub = lane >> 4;
#pragma unroll 4
        for (int ks = 0; ks < M / 4; ks++) {
            const uint32_t k = ks * 4 + ksub;
            const float b = k < (uint32_t)D ? Br[k * PAD + col] : Bi[(k - D) * PAD + col];
#pragma unroll
            for (int mt = 0; mt < M / 16; mt++) {
                const float a = R[(mt * 16 + (lane & 15)) * M + k];
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[mt], 0, 0, 0);
            }
        }
        __syncthreads();  // every wave has read Br / Bi
        // Accumulator element i of lane l: row 4 * (l / 16) + i, column l % 16.
#pragma unroll
        for (int mt = 0; mt < M / 16; mt++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t row = mt * 16 + 4 * (lane >> 4) + i;
                const uint32_t c = wave * 16 + (lane & 15);
                if (row < (uint32_t)D) Br[row * PAD + c] = acc[mt][i];
                else Bi[(row - D) * PAD + c] = acc[mt][i];
            }
        }
        __syncthreads();
        for (uint32_t e = tid; e < (uint32_t)(D * NC); e += 64 * WAVES) {
            const uint32_t j = jfast ? e % D : e / NC, c = jfast ? e / D : e % NC;
            if (c0 + c < d.iterations)
                s[(spread(c0 + c, d.ins, d.nins) | d.ctrl_val) + d.offsets[j]] =
                    make_float2(Br[j * PAD + c], Bi[j * PAD + c]);
        }
    }
}

template <int K>
hipError_t launch_k(float2 *s, const GateDesc &d, const float2 *mat, unsigned grid, hipStream_t st) {
    constexpr int D = 1 << K, M = 2 * D;
    const size_t lds = (size_t)(M * M + 2 * D * PAD) * sizeof(float);
    static std::atomic<uint64_t> attr_set{0};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const uint64_t bit = uint64_t(1) << (dev & 63);
    if (!(attr_set.load() & bit)) {
        e = hipFuncSetAttribute((const void *)k_dense_mfma<K>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr_set.fetch_or(bit);
    }
    k_dense_mfma<K><<<grid, 64 * WAVES, lds, st>>>(s, d, mat);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_dense_mfma(float2 *s, const GateDesc &d, const float2 *mat, unsigned cus, hipStream_t st) {
    const uint64_t blocks = (d.iterations + NC - 1) / NC;
    const unsigned grid = (unsigned)std::min<uint64_t>(blocks, (uint64_t)cus);
    if (d.k == 5) return launch_k<5>(s, d, mat, grid, st);
    if (d.k == 6) return launch_k<6>(s, d, mat, grid, st);
    return hipErrorInvalidValue;
}

}  // namespace qvd

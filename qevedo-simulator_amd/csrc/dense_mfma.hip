// Fused dense gates of K = 5 or 6 targets in complex64, applied as real
// GEMMs on FP32 matrix cores (v_mfma_f32_16x16x4f32, FP32 accumulation).
//
// A column is one group of D = 2^K amplitudes the gate mixes (the iteration
// index spread over the non-target, non-control bits, as in k_dense). With
// U = Ur + i Ui and a column a = ar + i ai, the output is
//     [or; oi] = R [ar; ai],   R = [[Ur, -Ui], [Ui, Ur]]   (2D x 2D real),
// so a block of NC columns is the real product (2D x 2D) x (2D x NC).
// A workgroup of WAVES waves (4 for K = 5, 8 for K = 6: one 16-row tile of R
// each) stages each block's amplitudes through LDS (coalesced loads and
// stores, the next block's loads in flight during the products); each wave
// keeps its rows of R in registers for the whole launch and computes those
// rows of the output for all NC columns. Blocks of 32 columns keep a 6-qubit
// workgroup at 78 VGPRs, so three share a CU (six waves per SIMD); the grid
// is 12 workgroups per CU. (Measured on the 32-qubit precision form: 64
// columns at 4 workgroups per CU 0.855 s, 32 at 12 0.826 s, 128 at 4 0.987 s;
// an XOR-swizzled LDS layout without bank conflicts 1.13 s, its per-lane
// operand addresses costing more than the conflicts.)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

#include "dense_mfma.h"

namespace qvd {

namespace {

constexpr int NC = 32;         // columns per block
constexpr int PAD = NC + 1;    // row pitch of the staged block (floats)

__device__ __forceinline__ uint64_t spread(uint64_t x, const uint32_t *pos, uint32_t m) {
    for (uint32_t j = 0; j < m; j++) {
        const uint32_t p = pos[j];
        const uint64_t low = x & ((uint64_t(1) << p) - 1);
        x = ((x >> p) << (p + 1)) | low;
    }
    return x;
}

template <int K, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_dense_mfma(float2 *__restrict__ s, const GateDesc d,
                                                           const float2 *__restrict__ mat) {
    constexpr int D = 1 << K, M = 2 * D;       // R is M x M
    constexpr int MPW = M / 16 / WAVES;        // 16-row tiles of R per wave
    constexpr int KS = M / 4;                  // k steps of 4
    constexpr int PER = D * NC / (64 * WAVES); // amplitudes staged per thread
    extern __shared__ float lds[];
    float *Br = lds;                                 // D x PAD: real parts (row j, column c)
    float *Bi = Br + D * PAD;                        // D x PAD: imaginary parts
    uint64_t *colbase = (uint64_t *)(Bi + D * PAD);  // 2 x NC: first amplitude of each column (double-buffered)
    uint64_t *offs = colbase + 2 * NC;               // D: offset of target index j
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t blocks = (d.iterations + NC - 1) / NC;
    // Target index fastest when the lowest target is qubit 0 (a column's
    // amplitudes are then contiguous), else column fastest.
    const bool jfast = d.offsets[1] == 1;
    auto elem = [&](uint32_t e, uint32_t &j, uint32_t &c) {
        j = jfast ? e % D : e / NC;
        c = jfast ? e / D : e % NC;
    };
    auto bases = [&](uint64_t blk, uint64_t *out) {
        if (tid < (uint32_t)NC) {
            const uint64_t col = blk * NC + tid;
            out[tid] = col < d.iterations ? (spread(col, d.ins, d.nins) | d.ctrl_val) : ~uint64_t(0);
        }
    };

    // This wave's rows of R stay in registers for the whole launch: lane l
    // holds R[row][4 ks + l / 16] of row = 16 mt + l % 16.
    float ra[MPW][KS];
#pragma unroll
    for (int i = 0; i < MPW; i++) {
        const uint32_t row = (wave * MPW + i) * 16 + (lane & 15);
        const uint32_t rr = row % D;
        const bool rim = row >= (uint32_t)D;
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const uint32_t k = ks * 4 + (lane >> 4), kk = k % D;
            const bool kim = k >= (uint32_t)D;
            const float2 u = mat[rr * D + kk];
            ra[i][ks] = rim == kim ? u.x : (rim ? u.y : -u.y);
        }
    }
    if (tid < (uint32_t)D) offs[tid] = d.offsets[tid];
    uint64_t blk = blockIdx.x;
    if (blk >= blocks) return;
    uint32_t cur = 0;
    bases(blk, colbase);
    __syncthreads();
    float2 pre[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        uint32_t j, c;
        elem(tid + i * 64 * WAVES, j, c);
        const uint64_t b = colbase[c];
        pre[i] = b != ~uint64_t(0) ? s[b + offs[j]] : make_float2(0.0f, 0.0f);
    }
    for (;;) {
        __syncthreads();  // the previous block's output has been stored
#pragma unroll
        for (int i = 0; i < PER; i++) {
            uint32_t j, c;
            elem(tid + i * 64 * WAVES, j, c);
            Br[j * PAD + c] = pre[i].x;
            Bi[j * PAD + c] = pre[i].y;
        }
        const uint64_t next = blk + gridDim.x;
        if (next < blocks) bases(next, colbase + (cur ^ 1) * NC);
        __syncthreads();
        // The next block's loads fly while the matrix cores work.
        if (next < blocks) {
            const uint64_t *nb = colbase + (cur ^ 1) * NC;
#pragma unroll
            for (int i = 0; i < PER; i++) {
                uint32_t j, c;
                elem(tid + i * 64 * WAVES, j, c);
                const uint64_t b = nb[c];
                pre[i] = b != ~uint64_t(0) ? s[b + offs[j]] : make_float2(0.0f, 0.0f);
            }
        }
        floatx4 acc[MPW][NC / 16];
#pragma unroll
        for (int i = 0; i < MPW; i++)
#pragma unroll
            for (int g = 0; g < NC / 16; g++) acc[i][g] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const uint32_t k = ks * 4 + (lane >> 4);
            const float *row = k < (uint32_t)D ? Br + k * PAD : Bi + (k - D) * PAD;
            float b[NC / 16];
#pragma unroll
            for (int g = 0; g < NC / 16; g++) b[g] = row[g * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < MPW; i++)
#pragma unroll
                for (int g = 0; g < NC / 16; g++)
                    acc[i][g] = __builtin_amdgcn_mfma_f32_16x16x4f32(ra[i][ks], b[g], acc[i][g], 0, 0, 0);
        }
        __syncthreads();  // every wave has read Br / Bi
        // Accumulator element e of lane l: row 4 * (l / 16) + e, column l % 16.
#pragma unroll
        for (int i = 0; i < MPW; i++)
#pragma unroll
            for (int g = 0; g < NC / 16; g++)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const uint32_t row = (wave * MPW + i) * 16 + 4 * (lane >> 4) + e;
                    const uint32_t c = g * 16 + (lane & 15);
                    if (row < (uint32_t)D) Br[row * PAD + c] = acc[i][g][e];
                    else Bi[(row - D) * PAD + c] = acc[i][g][e];
                }
        __syncthreads();
        const uint64_t *cb = colbase + cur * NC;
        for (uint32_t e = tid; e < (uint32_t)(D * NC); e += 64 * WAVES) {
            uint32_t j, c;
            elem(e, j, c);
            const uint64_t b = cb[c];
            if (b != ~uint64_t(0)) s[b + offs[j]] = make_float2(Br[j * PAD + c], Bi[j * PAD + c]);
        }
        if (next >= blocks) break;
        blk = next;
        cur ^= 1;
    }
}

template <int K, int WAVES>
hipError_t launch_k(float2 *s, const GateDesc &d, const float2 *mat, unsigned grid, hipStream_t st) {
    constexpr int D = 1 << K;
    const size_t lds = (size_t)(2 * D * PAD) * sizeof(float) + (2 * NC + D) * sizeof(uint64_t);
    static std::atomic<uint64_t> attr_set{0};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const uint64_t bit = uint64_t(1) << (dev & 63);
    if (!(attr_set.load() & bit)) {
        e = hipFuncSetAttribute((const void *)k_dense_mfma<K, WAVES>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr_set.fetch_or(bit);
    }
    k_dense_mfma<K, WAVES><<<grid, 64 * WAVES, lds, st>>>(s, d, mat);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_dense_mfma(float2 *s, const GateDesc &d, const float2 *mat, unsigned cus, hipStream_t st) {
    const uint64_t blocks = (d.iterations + NC - 1) / NC;
    const unsigned grid = (unsigned)std::min<uint64_t>(blocks, (uint64_t)cus * 12);
    if (d.k == 5) return launch_k<5, 4>(s, d, mat, grid, st);
    if (d.k == 6) return launch_k<6, 8>(s, d, mat, grid, st);
    return hipErrorInvalidValue;
}

}  // namespace qvd

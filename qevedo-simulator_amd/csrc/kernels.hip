// HIP kernels of the MI355X state-vector engine (gfx950, wave64).
//
// State layout in HBM: 2^n interleaved complex amplitudes (float2 for
// complex64, double2 for complex128) in logical index order; qubit q is bit q
// of the index. Every kernel here is HBM-bound except the in-tile gate
// arithmetic of deep stages.
//
// Compiled with -ffp-contract=off: fused multiply-adds appear only where
// written (fma / fmaf), so the arithmetic is the same on every build.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <mutex>

#include "device.h"
#include "kernels.h"

namespace qvd {

// ---------------------------------------------------------------------------
// Complex helpers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float fmul_add(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmul_add(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T2> struct Real;
template <> struct Real<float2> { using T = float; };
template <> struct Real<double2> { using T = double; };

// acc += w * x, as four fused multiply-adds.
template <typename T2>
__device__ __forceinline__ void cmac(T2 &acc, const T2 w, const T2 x) {
    acc.x = fmul_add(w.x, x.x, acc.x);
    acc.x = fmul_add(-w.y, x.y, acc.x);
    acc.y = fmul_add(w.x, x.y, acc.y);
    acc.y = fmul_add(w.y, x.x, acc.y);
}

// f * a, with the rounding of the reference's diagonal kernel
// (re = f.re*a.re - f.im*a.im with the second product fused; im likewise,
// src/kernels.rs:464-465).
template <typename T2>
__device__ __forceinline__ T2 cmul_diag(const T2 f, const T2 a) {
    T2 r;
    r.x = fmul_add(-f.y, a.y, f.x * a.x);
    r.y = fmul_add(f.y, a.x, f.x * a.y);
    return r;
}

// Insert a zero bit at each of the ascending positions pos[0..m).
__device__ __forceinline__ uint64_t insert_zeros(uint64_t x, const uint32_t *pos, uint32_t m) {
    for (uint32_t j = 0; j < m; j++) {
        const uint32_t p = pos[j];
        const uint64_t low = x & ((uint64_t(1) << p) - 1);
        x = ((x >> p) << (p + 1)) | low;
    }
    return x;
}

// ---------------------------------------------------------------------------
// Initialisation and conversion.
// ---------------------------------------------------------------------------
template <typename T2>
__global__ void k_init_zero(T2 *__restrict__ s, uint64_t len) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
        T2 v;
        v.x = i == 0 ? 1 : 0;
        v.y = 0;
        s[i] = v;
    }
}

// Writes zero over every amplitude whose index has a bit of `stale` set.
template <typename T2>
__global__ void k_zero_stale(T2 *__restrict__ s, uint64_t len, uint64_t stale) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
        if (i & stale) {
            T2 v;
            v.x = 0;
            v.y = 0;
            s[i] = v;
        }
    }
}

template <typename T2>
__global__ void k_widen(const T2 *__restrict__ s, double2 *__restrict__ out, uint64_t count) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const T2 v = s[i];
        out[i] = make_double2((double)v.x, (double)v.y);
    }
}

template <typename T2>
__global__ void k_narrow(const double2 *__restrict__ in, T2 *__restrict__ s, uint64_t count) {
    using T = typename Real<T2>::T;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const double2 v = in[i];
        T2 o;
        o.x = (T)v.x;
        o.y = (T)v.y;
        s[i] = o;
    }
}

template <typename T2>
__global__ void k_scale(T2 *__restrict__ s, uint64_t len, double factor) {
    using T = typename Real<T2>::T;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
        T2 v = s[i];
        v.x = (T)((double)v.x * factor);
        v.y = (T)((double)v.y * factor);
        s[i] = v;
    }
}

// ---------------------------------------------------------------------------
// Full-state ops: one HBM pass each.
// ---------------------------------------------------------------------------

// A dense op on K targets: each thread owns one group of 2^K amplitudes
// (the iteration index spread over the non-target, non-control bits).
template <typename T2, int K>
__global__ __launch_bounds__(256) void k_dense(T2 *__restrict__ s, const GateDesc d, const T2 *__restrict__ mat) {
    constexpr int D = 1 << K;
    extern __shared__ unsigned char smem_raw[];
    T2 *m = reinterpret_cast<T2 *>(smem_raw);
    for (int i = threadIdx.x; i < D * D; i += blockDim.x) m[i] = mat[i];
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t it = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; it < d.iterations; it += stride) {
        const uint64_t base = insert_zeros(it, d.ins, d.nins) | d.ctrl_val;
        T2 a[D];
#pragma unroll
        for (int j = 0; j < D; j++) a[j] = s[base + d.offsets[j]];
#pragma unroll
        for (int r = 0; r < D; r++) {
            T2 acc;
            acc.x = 0;
            acc.y = 0;
#pragma unroll
            for (int c = 0; c < D; c++) cmac(acc, m[r * D + c], a[c]);
            s[base + d.offsets[r]] = acc;
        }
    }
}

// A diagonal op: amplitude i is multiplied by entries[index of its target
// bits] where every control bit holds.
template <typename T2>
__global__ __launch_bounds__(256) void k_diag(T2 *__restrict__ s, uint64_t len, const uint32_t k,
                                              const DiagTargets t, uint64_t cmask, uint64_t cval,
                                              const T2 *__restrict__ entries) {
    __shared__ T2 e[1 << MAX_GATE_QUBITS];
    for (uint32_t i = threadIdx.x; i < (1u << k); i += blockDim.x) e[i] = entries[i];
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
        if ((i & cmask) != cval) continue;
        uint32_t idx = 0;
        for (uint32_t j = 0; j < k; j++) idx |= (uint32_t)((i >> t.q[j]) & 1) << j;
        s[i] = cmul_diag(e[idx], s[i]);
    }
}

// ---------------------------------------------------------------------------
// LDS-tile stages: one HBM pass applies every op of a stage.
//
// A workgroup loads a tile of 2^L amplitudes whose index bits are the
// stage's tile qubits (the lowest `contig` qubits among them, so each HBM
// access is a contiguous run), applies the stage's ops in LDS, and writes
// the tile back. Non-tile qubits are fixed per tile: non-tile controls
// select whether an op runs on the tile at all, and non-tile diagonal targets
// select the entries used.
// ---------------------------------------------------------------------------
constexpr int STAGE_THREADS = 256;
constexpr int STAGE_LOG_THREADS = 8;

template <typename T2, int K>
__device__ __forceinline__ void tile_dense(T2 *__restrict__ tile, const T2 *__restrict__ m, const StageOp &op,
                                           uint32_t L) {
    constexpr int D = 1 << K;
    uint32_t off[D];
#pragma unroll
    for (int j = 0; j < D; j++) {
        uint32_t o = 0;
#pragma unroll
        for (int b = 0; b < K; b++) o |= (uint32_t)((j >> b) & 1) << op.tgt[b];
        off[j] = o;
    }
    const uint32_t groups = 1u << (L - op.nins);
    for (uint32_t g = threadIdx.x; g < groups; g += STAGE_THREADS) {
        uint32_t i0 = g;
        for (uint32_t j = 0; j < op.nins; j++) {
            const uint32_t p = op.ins[j];
            i0 = ((i0 >> p) << (p + 1)) | (i0 & ((1u << p) - 1));
        }
        i0 |= op.lc_val;
        T2 a[D];
#pragma unroll
        for (int j = 0; j < D; j++) a[j] = tile[i0 | off[j]];
#pragma unroll
        for (int r = 0; r < D; r++) {
            T2 acc;
            acc.x = 0;
            acc.y = 0;
#pragma unroll
            for (int c = 0; c < D; c++) cmac(acc, m[r * D + c], a[c]);
            tile[i0 | off[r]] = acc;
        }
    }
}

template <typename T2>
__device__ __forceinline__ void tile_diag(T2 *__restrict__ tile, const T2 *__restrict__ e, const StageOp &op,
                                          uint32_t L, uint64_t base) {
    uint32_t gidx = 0, lmask[MAX_GATE_QUBITS];
    for (uint32_t j = 0; j < op.k; j++) {
        const uint32_t t = op.tgt[j];
        if (t & GLOBAL_BIT) {
            gidx |= (uint32_t)((base >> (t & ~GLOBAL_BIT)) & 1) << j;
            lmask[j] = 0;
        } else {
            lmask[j] = 1u << t;
        }
    }
    const uint32_t tsize = 1u << L;
    for (uint32_t i = threadIdx.x; i < tsize; i += STAGE_THREADS) {
        if ((i & op.lc_mask) != op.lc_val) continue;
        uint32_t idx = gidx;
        for (uint32_t j = 0; j < op.k; j++) idx |= (uint32_t)((i & lmask[j]) != 0) << j;
        tile[i] = cmul_diag(e[idx], tile[i]);
    }
}

template <typename T2>
__global__ __launch_bounds__(STAGE_THREADS) void k_stage(T2 *__restrict__ s, const StageDesc d,
                                                         const StageOp *__restrict__ ops,
                                                         const T2 *__restrict__ coefs) {
    extern __shared__ unsigned char smem_raw[];
    T2 *tile = reinterpret_cast<T2 *>(smem_raw);
    const uint32_t L = d.L;
    const uint32_t tsize = 1u << L;
    T2 *mat = tile + tsize;  // coefficients of the current op (<= 16 x 16)
    const uint32_t tid = threadIdx.x;
    const uint32_t lowbits = L < STAGE_LOG_THREADS ? L : STAGE_LOG_THREADS;
    uint64_t toff = 0;
    for (uint32_t j = 0; j < lowbits; j++) toff |= (uint64_t)((tid >> j) & 1) << d.local_pos[j];
    const bool active = tid < tsize;
    for (uint64_t t = blockIdx.x; t < d.num_tiles; t += gridDim.x) {
        uint64_t base = 0;
        for (uint32_t j = 0; j < d.nglob; j++) base |= ((t >> j) & 1) << d.glob_pos[j];
        for (uint32_t m = 0; m < tsize; m += STAGE_THREADS) {
            uint64_t moff = 0;
            for (uint32_t j = lowbits; j < L; j++) moff |= (uint64_t)((m >> j) & 1) << d.local_pos[j];
            if (active) tile[m + tid] = s[base + moff + toff];
        }
        __syncthreads();
        for (uint32_t o = 0; o < d.num_ops; o++) {
            const StageOp &op = ops[d.op_off + o];
            if ((base & op.gc_mask) != op.gc_val) continue;
            const uint32_t ncoef = op.kind == OP_DENSE ? (1u << (2 * op.k)) : (1u << op.k);
            for (uint32_t c = tid; c < ncoef; c += STAGE_THREADS) mat[c] = coefs[op.coef_off + c];
            __syncthreads();
            if (op.kind == OP_DENSE) {
                switch (op.k) {
                    case 1: tile_dense<T2, 1>(tile, mat, op, L); break;
                    case 2: tile_dense<T2, 2>(tile, mat, op, L); break;
                    case 3: tile_dense<T2, 3>(tile, mat, op, L); break;
                    case 4: tile_dense<T2, 4>(tile, mat, op, L); break;
                    default: break;
                }
            } else {
                tile_diag(tile, mat, op, L, base);
            }
            __syncthreads();
        }
        for (uint32_t m = 0; m < tsize; m += STAGE_THREADS) {
            uint64_t moff = 0;
            for (uint32_t j = lowbits; j < L; j++) moff |= (uint64_t)((m >> j) & 1) << d.local_pos[j];
            if (active) s[base + moff + toff] = tile[m + tid];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// FP64 reductions.
// ---------------------------------------------------------------------------
constexpr int RED_THREADS = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum of v over the block; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *scratch) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += scratch[w];
    return r;
}

template <typename T2>
__device__ __forceinline__ double prob(const T2 a) {
    const double re = (double)a.x, im = (double)a.y;
    return re * re + im * im;
}

// Per-block partial sums of |a_i|^2 over amplitudes with (i & mask) == val.
template <typename T2>
__global__ __launch_bounds__(RED_THREADS) void k_masked_norm(const T2 *__restrict__ s, uint64_t len, uint64_t mask,
                                                             uint64_t val, double *__restrict__ partial) {
    __shared__ double scratch[RED_THREADS / 64];
    double acc = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // U loads in flight per thread, summed in index order. An amplitude
    // outside the mask adds +0.0, which leaves the sum (never -0.0) as it is;
    // its load re-reads the batch's first address (cached) instead, so the
    // loads need no branch, which would wait for each of them.
    constexpr int U = 8;
    for (; i + (U - 1) * stride < len; i += U * stride) {
        T2 a[U];
        bool in[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t k = i + u * stride;
            in[u] = (k & mask) == val;
            a[u] = s[in[u] ? k : i];
        }
#pragma unroll
        for (int u = 0; u < U; u++) acc += in[u] ? prob(a[u]) : 0.0;
    }
    for (; i < len; i += stride)
        if ((i & mask) == val) acc += prob(s[i]);
    const double r = block_sum(acc, scratch);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// |a_i|^2 in FP64, in index order.
template <typename T2>
__global__ void k_probabilities(const T2 *__restrict__ s, uint64_t len, double *__restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) out[i] = prob(s[i]);
}

// Pauli strings sharing one X part x (qubits with X or Y) and differing in
// their Z parts z[t] (qubits with Z or Y): per block, the sums over j of
// (-1)^popcount((j ^ x) & z[t]) conj(a_j) a_(j ^ x), real and imaginary
// parts in FP64, written to partial[block * 2 * nterms + 2 t (+1)]. With the
// phase i^popcount(x & z[t]) applied afterwards this is <P_t>. KT >= nterms
// slots are computed (a power of two, so a group of few terms does not pay
// for PAULI_TERMS).
template <typename T2, int KT>
__global__ __launch_bounds__(RED_THREADS) void k_expect_pauli(const T2 *__restrict__ s, uint64_t len, uint64_t x,
                                                              const uint64_t *__restrict__ zs, uint32_t nterms,
                                                              double *__restrict__ partial) {
    __shared__ double scratch[RED_THREADS / 64];
    double re[KT], im[KT];
    uint64_t z[KT];
#pragma unroll
    for (int t = 0; t < KT; t++) {
        re[t] = im[t] = 0;
        z[t] = t < (int)nterms ? zs[t] : 0;
    }
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    auto add = [&](uint64_t j, const T2 a, const T2 b) {
        const double ar = a.x, ai = a.y, br = b.x, bi = b.y;
        const double vr = ar * br + ai * bi, vi = ar * bi - ai * br;  // conj(a) b
        const uint64_t k = j ^ x;
#pragma unroll
        for (int t = 0; t < KT; t++) {  // every slot: unused ones have z = 0 and are not written
            const bool neg = __popcll(k & z[t]) & 1;
            re[t] += neg ? -vr : vr;
            im[t] += neg ? -vi : vi;
        }
    };
    // U iterations' loads in flight per thread, summed in index order.
    constexpr int U = 4;
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; j + (U - 1) * stride < len; j += U * stride) {
        T2 a[U], b[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            a[u] = s[j + u * stride];
            b[u] = s[(j + u * stride) ^ x];
        }
#pragma unroll
        for (int u = 0; u < U; u++) add(j + u * stride, a[u], b[u]);
    }
    for (; j < len; j += stride) add(j, s[j], s[j ^ x]);
#pragma unroll
    for (int t = 0; t < KT; t++) {
        const double r = block_sum(re[t], scratch);
        const double i = block_sum(im[t], scratch);
        if (threadIdx.x == 0 && t < (int)nterms) {
            partial[(uint64_t)blockIdx.x * 2 * nterms + 2 * t] = r;
            partial[(uint64_t)blockIdx.x * 2 * nterms + 2 * t + 1] = i;
        }
    }
}

// out[v] = sum over blocks b, in block order, of partial[b * width + v].
__global__ void k_sum_columns(const double *__restrict__ partial, uint32_t blocks, uint32_t width,
                              double *__restrict__ out) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= width) return;
    double r = 0;
    for (uint32_t b = 0; b < blocks; b++) r += partial[(uint64_t)b * width + v];
    out[v] = r;
}

// Sequential sum of partial[0..m) into out[0].
__global__ void k_sum_partials(const double *__restrict__ partial, uint32_t m, double *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double r = 0;
        for (uint32_t i = 0; i < m; i++) r += partial[i];
        out[0] = r;
    }
}

// Per-block max |a-b|, sum |a-b|^2 and sum |b|^2 over two states.
template <typename TA, typename TB>
__global__ __launch_bounds__(RED_THREADS) void k_compare(const TA *__restrict__ a, const TB *__restrict__ b,
                                                         uint64_t len, double *__restrict__ partial) {
    __shared__ double scratch[RED_THREADS / 64];
    __shared__ double mx[RED_THREADS / 64];
    double m = 0, d2 = 0, r2 = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
        const TA x = a[i];
        const TB y = b[i];
        const double dr = (double)x.x - (double)y.x, di = (double)x.y - (double)y.y;
        const double dd = dr * dr + di * di;
        d2 += dd;
        m = dd > m ? dd : m;
        r2 += (double)y.x * (double)y.x + (double)y.y * (double)y.y;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double t = __shfl_xor(m, o, 64);
        m = t > m ? t : m;
    }
    if ((threadIdx.x & 63) == 0) mx[threadIdx.x >> 6] = m;
    const double s1 = block_sum(d2, scratch);
    const double s2 = block_sum(r2, scratch);
    if (threadIdx.x == 0) {
        double mm = 0;
        for (int w = 0; w < RED_THREADS / 64; w++) mm = mx[w] > mm ? mx[w] : mm;
        partial[3 * blockIdx.x] = mm;
        partial[3 * blockIdx.x + 1] = s1;
        partial[3 * blockIdx.x + 2] = s2;
    }
}

// <Z_q> and <Z_q Z_q+1> in one pass. A block covers a segment of 2^B
// contiguous amplitudes; per segment it writes 2B values: the total T, the
// per-bit sums Q[q] = sum p [bit q] (q < B) and the per-pair sums
// R[q] = sum p [bit q xor bit q+1] (q < B-1). Bits above B are constant in a
// segment and are resolved by k_expect_combine.
constexpr int EZ_TID_BITS = 10;  // bits from the thread index (1024 threads)
constexpr int EZ_MAX_M_BITS = 6; // bits from the per-thread loop

// The descriptor of block b: `one` when there is a single state, else the
// entry i of `descs` with start[i] <= b < start[i + 1] (binary search).
template <typename S>
__device__ __forceinline__ uint32_t find_range(const S *__restrict__ start, uint32_t count, uint64_t b) {
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)start[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <typename T2>
__global__ __launch_bounds__(1 << EZ_TID_BITS) void k_expect_segments(const T2 *__restrict__ s_all, ExpDesc one,
                                                         const ExpDesc *__restrict__ descs, uint32_t ndesc,
                                                         const uint64_t *__restrict__ seg_start,
                                                         double *__restrict__ out_all) {
    constexpr int MAXV = 2 * (EZ_TID_BITS + EZ_MAX_M_BITS);  // values per segment, at most
    __shared__ double scratch[(1 << EZ_TID_BITS) / 64][MAXV];
    ExpDesc d = one;
    uint64_t seg = blockIdx.x;
    if (ndesc > 1) {
        const uint32_t di = find_range(seg_start, ndesc, blockIdx.x);
        d = descs[di];
        seg = blockIdx.x - seg_start[di];
    }
    const T2 *s = s_all + d.state_off;
    double *out = out_all + d.seg_off;
    const uint32_t B = d.B;
    const uint32_t tb = B < EZ_TID_BITS ? B : EZ_TID_BITS;
    const uint32_t mb = B - tb;
    const uint32_t tid = threadIdx.x;
    const bool active = tid < (1u << tb);
    double T = 0, S[EZ_MAX_M_BITS], P[EZ_MAX_M_BITS];
#pragma unroll
    for (int j = 0; j < EZ_MAX_M_BITS; j++) S[j] = P[j] = 0;
    if (active) {
        const uint64_t base = (seg << B) + tid;
        const uint32_t M = 1u << mb;
        auto add = [&](uint32_t m, double p) {
            T += p;
#pragma unroll
            for (int j = 0; j < EZ_MAX_M_BITS; j++) {
                if ((uint32_t)j < mb) {
                    const uint32_t bj = (m >> j) & 1;
                    if (bj) S[j] += p;
                    if ((uint32_t)j + 1 < mb && (bj ^ ((m >> (j + 1)) & 1))) P[j] += p;
                }
            }
        };
        // U loads in flight per thread (M is a power of two, so a multiple
        // of U once it reaches U: the loads need no bounds test, which would
        // put each in a branch of its own and wait for it), then the sums in
        // the order of m.
        constexpr uint32_t U = 8;
        if (M >= U) {
            for (uint32_t m0 = 0; m0 < M; m0 += U) {
                T2 a[U];
#pragma unroll
                for (uint32_t u = 0; u < U; u++) a[u] = s[base + ((uint64_t)(m0 + u) << tb)];
#pragma unroll
                for (uint32_t u = 0; u < U; u++) add(m0 + u, prob(a[u]));
            }
        } else {
            for (uint32_t m = 0; m < M; m++) add(m, prob(s[base + ((uint64_t)m << tb)]));
        }
    }
    // The segment's 2B values: T, Q[q] (q < B) and R[q] (q < B - 1), each
    // summed over the wave, then over the waves in order by one thread per
    // value after a single barrier.
    const uint32_t lane = tid & 63, wave = tid >> 6;
    auto value = [&](uint32_t k) -> double {
        if (k == 0) return T;
        if (k <= B) {  // Q[q]: T times the thread's bit q for q < tb, else S[q - tb]
            const uint32_t q = k - 1;
            if (q < tb) return ((tid >> q) & 1) ? T : 0.0;
            double v = 0;
#pragma unroll
            for (int j = 0; j < EZ_MAX_M_BITS; j++)
                if ((uint32_t)j == q - tb) v = S[j];
            return v;
        }
        const uint32_t q = k - 1 - B;  // R[q]
        if (q + 1 < tb) return (((tid >> q) ^ (tid >> (q + 1))) & 1) ? T : 0.0;
        if (q + 1 == tb) {
            // bit q from the thread, bit q+1 = loop bit 0.
            const double s0 = mb > 0 ? S[0] : 0.0;
            return ((tid >> q) & 1) ? T - s0 : s0;
        }
        double v = 0;
#pragma unroll
        for (int j = 0; j < EZ_MAX_M_BITS; j++)
            if ((uint32_t)j == q - tb) v = P[j];
        return v;
    };
    const uint32_t nv = 2 * B;  // T, B values Q, B - 1 values R
    for (uint32_t k = 0; k < nv; k++) {
        const double w = wave_sum(value(k));
        if (lane == 0) scratch[wave][k] = w;
    }
    __syncthreads();
    if (tid < nv) {
        double r = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += scratch[w][tid];
        out[seg * (2 * B) + tid] = r;
    }
}

// Combines the segment sums into z[q] = sum p (1 - 2 bit q) and
// zz[q] = sum p (1 - 2 (bit q xor bit q+1)). One thread per output, in
// blocks of 64 threads.
__global__ __launch_bounds__(64) void k_expect_combine(const double *__restrict__ seg_all, ExpDesc one, const ExpDesc *__restrict__ descs,
                                 uint32_t ndesc, const uint32_t *__restrict__ out_start, double *__restrict__ z,
                                 double *__restrict__ zz) {
    const uint32_t wg = threadIdx.x + blockIdx.x * blockDim.x;
    ExpDesc d = one;
    uint32_t w = wg;
    if (ndesc > 1) {
        if (wg >= out_start[ndesc]) return;
        const uint32_t di = find_range(out_start, ndesc, wg);
        d = descs[di];
        w = wg - out_start[di];
    }
    const uint32_t n = d.n, B = d.B;
    if (w >= 2 * n - 1) return;
    const double *seg_sums = seg_all + d.seg_off;
    // What the output takes from segment g besides its total T = o[0]:
    //   0: o[idx]                          (its bit, or both bits of its pair, inside the segment)
    //   1: T where bit sh of g is set      (<Z_q>, q at or above the segment bits)
    //   2: (g & 1) ? T - o[idx] : o[idx]   (<Z_q Z_q+1>, bit q the segment's top bit)
    //   3: T where bits sh and sh + 1 of g differ
    uint32_t mode, idx = 0, sh = 0;
    if (w < n) {
        const uint32_t q = w;
        if (q < B) mode = 0, idx = 1 + q;
        else mode = 1, sh = q - B;
    } else {
        const uint32_t q = w - n;
        if (q + 1 < B) mode = 0, idx = 1 + B + q;
        else if (q + 1 == B) mode = 2, idx = 1 + q;
        else mode = 3, sh = q - B;
    }
    // Branch-free: where an output takes nothing from a segment it adds
    // +0.0, which leaves the sum as it is (both sums start at +0.0 and add
    // non-negative values, so neither is ever -0.0).
    double tot = 0, part = 0;
    auto add = [&](uint64_t g, double T, double V) {
        tot += T;
        const bool b0 = (g >> sh) & 1, b1 = (g >> (sh + 1)) & 1;
        const double a = mode == 0 ? V : mode == 1 ? (b0 ? T : 0.0) : mode == 2 ? ((g & 1) ? T - V : V) : (b0 != b1 ? T : 0.0);
        part += a;
    };
    // The segments in order; G segments' values are loaded at a time, so a
    // state of many segments does not wait for one load per segment.
    // One pointer steps through a batch (precomputed per-segment addresses
    // would be loop invariants the compiler keeps, and spills).
    constexpr uint32_t G = 32;
    const uint64_t stride = 2 * (uint64_t)B;
    uint64_t g = 0;
    for (; g + G <= d.nsegs; g += G) {
        double T[G], V[G];
        const double *o = seg_sums + g * stride;
#pragma unroll
        for (uint32_t u = 0; u < G; u++, o += stride) {
            T[u] = o[0];
            V[u] = o[idx];
        }
#pragma unroll
        for (uint32_t u = 0; u < G; u++) add(g + u, T[u], V[u]);
    }
    for (; g < d.nsegs; g++) {
        const double *o = seg_sums + g * stride;
        add(g, o[0], o[idx]);
    }
    if (w < n) z[d.z_off + w] = tot - 2 * part;
    else zz[d.zz_off + w - n] = tot - 2 * part;
}

// Sets amplitude 0 of each state of a batch to 1 (after a zero fill).
template <typename T2>
__global__ void k_set_basis(T2 *__restrict__ s, const uint64_t *__restrict__ offs, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    T2 v;
    v.x = 1;
    v.y = 0;
    s[offs[i]] = v;
}

// Zero-fills n words (a kernel, not a memset: graph-captured memset nodes are
// avoided).
__global__ void k_zero_words(uint32_t *__restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Measurement collapse (src/measure.rs:68-80): the kept half scaled by
// `keep` in the state's precision, the other half zeroed.
// ---------------------------------------------------------------------------
template <typename T2>
__global__ void k_collapse(T2 *__restrict__ s, uint64_t len, uint32_t qubit, uint32_t outcome, double keep) {
    using T = typename Real<T2>::T;
    const T k = (T)keep;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
        T2 f;
        if (((i >> qubit) & 1) == outcome) {
            f.x = k;
            f.y = 0;
        } else {
            f.x = 0;
            f.y = 0;
        }
        s[i] = cmul_diag(f, s[i]);
    }
}

// ---------------------------------------------------------------------------
// Two-pass sampling (src/measure.rs:104-174). The reference's summation
// order is kept exactly: each chunk is summed sequentially in index order,
// the prefix over chunks is sequential, and the walk adds probabilities one
// by one; so one lane owns one chunk. Loads are 16-byte vectors.
// ---------------------------------------------------------------------------
// Amplitudes per batch of the one-lane-per-chunk passes: two batches live in
// registers (128 VGPRs of amplitudes).
template <typename T2>
constexpr int chunk_batch() {
    return sizeof(T2) == 8 ? 32 : 16;
}

// Streams one lane's amplitudes [i, end) through visit(batch, first index) in
// batches of U, two at a time in registers: batch b + 1's loads fly while
// batch b is visited. With one lane per chunk and few chunks (2^15 at 30
// qubits, half a wave per SIMD) nothing else covers a batch's memory round
// trip. visit returns false to stop (`stopped` is then set). issued() runs
// right after each batch's loads go out: a store placed there follows the
// loads the next wait is for, so that wait does not include the store (the
// memory counter of loads and stores runs in order). Returns the first index
// not visited; fewer than U amplitudes are left from there unless it stopped.
template <int U, typename T2, typename F, typename G>
__device__ __forceinline__ uint64_t stream_batches(const T2 *__restrict__ s, uint64_t i, uint64_t end, bool &stopped,
                                                   F &&visit, G &&issued) {
    stopped = false;
    if (i + U > end) return i;
    T2 a[U], b[U];
    uint64_t a_at = i;  // a holds [a_at, a_at + U)
#pragma unroll
    for (int j = 0; j < U; j++) a[j] = s[i + j];
    while (i + 2 * U <= end) {
#pragma unroll
        for (int j = 0; j < U; j++) b[j] = s[i + U + j];
        __builtin_amdgcn_sched_barrier(0);  // b's loads go out before a is used
        issued();
        if (!visit(a, i)) {
            stopped = true;
            return i + U;
        }
        i += U;
        // The next batch, or the last whole one (in bounds: i + U <= end):
        // an unconditional load keeps the waits for b exact.
        a_at = i + 2 * U <= end ? i + U : end - U;
#pragma unroll
        for (int j = 0; j < U; j++) a[j] = s[a_at + j];
        __builtin_amdgcn_sched_barrier(0);  // a's loads go out before b is used
        issued();
        if (!visit(b, i)) {
            stopped = true;
            return i + U;
        }
        i += U;
    }
    if (i + U <= end) {
        if (a_at != i) {
#pragma unroll
            for (int j = 0; j < U; j++) a[j] = s[i + j];
        }
        issued();  // between every two visits
        if (!visit(a, i)) stopped = true;
        i += U;
    }
    return i;
}

template <typename T2>
__global__ __launch_bounds__(64) void k_chunk_sums(const T2 *__restrict__ s, uint64_t len, uint64_t chunk,
                                                    uint64_t num_chunks, double *__restrict__ sums) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= num_chunks) return;
    const uint64_t start = c * chunk;
    const uint64_t end = start + chunk < len ? start + chunk : len;
    double acc = 0;
    constexpr int U = chunk_batch<T2>();
    bool stopped;
    uint64_t i = stream_batches<U>(s, start, end, stopped, [&](const T2(&v)[U], uint64_t) {
#pragma unroll
        for (int j = 0; j < U; j++) acc += prob(v[j]);
        return true;
    }, [] {});
    for (; i < end; i++) acc += prob(s[i]);
    sums[c] = acc;
}

// The sequential prefix over chunk sums (src/measure.rs:120-124): one lane
// adds them one by one. The sums pass through LDS in tiles: while lane 0
// adds up tile k in place, the workgroup's other lanes store tile k-1's
// prefixes and load tile k+1, all coalesced, so the dependent add chain
// waits on LDS only, never on a memory round trip.
constexpr uint32_t PREFIX_THREADS = 256;
constexpr uint32_t PREFIX_TILE = 2048;  // doubles per tile; two tiles of LDS

__global__ __launch_bounds__(PREFIX_THREADS) void k_prefix(const double *__restrict__ sums, uint64_t num_chunks,
                                                          double *__restrict__ prefix) {
    __shared__ double buf[2][PREFIX_TILE];
    constexpr uint32_t PER = PREFIX_TILE / PREFIX_THREADS;
    const uint32_t t = threadIdx.x;
    const uint64_t tiles = (num_chunks + PREFIX_TILE - 1) / PREFIX_TILE;
    if (t == 0) prefix[0] = 0;
#pragma unroll
    for (uint32_t j = 0; j < PER; j++) {
        const uint64_t c = j * PREFIX_THREADS + t;
        buf[0][j * PREFIX_THREADS + t] = c < num_chunks ? sums[c] : 0.0;
    }
    __syncthreads();
    double p = 0;
    for (uint64_t k = 0; k < tiles; k++) {
        double *cur = buf[k & 1], *other = buf[(k + 1) & 1];
        const uint64_t base = k * PREFIX_TILE;
        if (k > 0) {  // tile k-1's prefixes, finished in the last round
#pragma unroll
            for (uint32_t j = 0; j < PER; j++) {
                const uint64_t c = base - PREFIX_TILE + j * PREFIX_THREADS + t;
                prefix[c + 1] = other[j * PREFIX_THREADS + t];
            }
        }
        const bool more = k + 1 < tiles;
        double v[PER];
        if (more) {
#pragma unroll
            for (uint32_t j = 0; j < PER; j++) {
                const uint64_t c = base + PREFIX_TILE + j * PREFIX_THREADS + t;
                v[j] = c < num_chunks ? sums[c] : 0.0;
            }
        }
        if (t == 0) {
            const uint32_t n = num_chunks - base < PREFIX_TILE ? (uint32_t)(num_chunks - base) : PREFIX_TILE;
            uint32_t i = 0;
            for (; i + 16 <= n; i += 16) {
                double x[16];
#pragma unroll
                for (int j = 0; j < 16; j++) x[j] = cur[i + j];
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    p = p + x[j];
                    x[j] = p;
                }
#pragma unroll
                for (int j = 0; j < 16; j++) cur[i + j] = x[j];
            }
            for (; i < n; i++) {
                p = p + cur[i];
                cur[i] = p;
            }
        }
        if (more) {  // the same slots this thread read above
#pragma unroll
            for (uint32_t j = 0; j < PER; j++) other[j * PREFIX_THREADS + t] = v[j];
        }
        __syncthreads();
    }
    const double *last = buf[(tiles - 1) & 1];
    const uint64_t base = (tiles - 1) * PREFIX_TILE;
#pragma unroll
    for (uint32_t j = 0; j < PER; j++) {
        const uint64_t c = base + j * PREFIX_THREADS + t;
        if (c < num_chunks) prefix[c + 1] = last[j * PREFIX_THREADS + t];
    }
}

// targets = sorted uniforms * total, in place. The order survives: for a
// finite total >= 0, rounding u * total is monotone in u, so scaling the
// sorted uniforms gives sort(u * total) exactly.
__global__ void k_targets(double *__restrict__ targets, const double *__restrict__ total_ptr, uint64_t shots) {
    const double total = *total_ptr;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < shots; i += stride)
        targets[i] = targets[i] * total;
}

// bounds[c] = number of targets < prefix[c] (partition_point).
__global__ void k_bounds(const double *__restrict__ prefix, uint64_t num_chunks, const double *__restrict__ targets,
                         uint64_t shots, uint64_t *__restrict__ bounds) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > num_chunks) return;
    const double p = prefix[c];
    uint64_t lo = 0, hi = shots;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (targets[mid] < p) lo = mid + 1;
        else hi = mid;
    }
    bounds[c] = lo;
}

// The walk (src/measure.rs:143-166) in two kernels. Shot s of chunk c goes
// to the first index i of the chunk whose running sum prefix[c] + p_start +
// ... + p_i, added one by one in index order, exceeds targets[s]; a shot no
// index of the chunk reaches (rounding) goes to the chunk's last nonzero
// probability. Zero probabilities are never such an index: adding +0.0
// leaves the sum (never -0.0) unchanged, and every target of the chunk is at
// least prefix[c]. So the running sum may add them too.
//
// k_walk_marks runs each chunk's running sum, one lane per chunk, with no
// per-shot work, and stores it every 2^mark_bits amplitudes (marks[b] = the
// sum up to the end of the chunk's block b). It stops after the block where
// the sum passes the chunk's largest target, and stores how many blocks it
// walked. k_walk_place then gives every shot a thread of its own: the first
// mark above its target names its block, and the sum resumes from the mark
// before (bit for bit the value the sequential sum had there) over at most
// 2^mark_bits amplitudes. A shot that no index reaches was left over by a
// walk that went to the chunk's end; its thread scans back from the end for
// the last nonzero probability (rare, and cut short by the first nonzero).
template <typename T2>
__global__ __launch_bounds__(64) void k_walk_marks(const T2 *__restrict__ s, uint64_t len, uint64_t chunk,
                                                    uint64_t num_chunks, uint32_t mark_bits,
                                                    const double *__restrict__ prefix,
                                                    const uint64_t *__restrict__ bounds,
                                                    const double *__restrict__ targets, uint64_t shots,
                                                    double *__restrict__ marks, uint32_t *__restrict__ walked) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= num_chunks) return;
    const uint64_t lo = bounds[c];
    const uint64_t hi = bounds[c + 1] < shots ? bounds[c + 1] : shots;
    if (lo >= hi) return;  // no shot in this chunk: nothing is read
    const double t_last = targets[hi - 1];
    const uint64_t start = c * chunk;
    const uint64_t end = start + chunk < len ? start + chunk : len;
    const uint64_t block_mask = (uint64_t(1) << mark_bits) - 1;
    double *m = marks + (start >> mark_bits);
    double cumulative = prefix[c];
    uint32_t b = 0;
    bool pending = false;  // amplitudes added since the last mark
    // A mark is stored after the next batch's loads go out (see
    // stream_batches): a store before them would make the wait for them wait
    // for the store too, a memory round trip per mark. Marks gather in LDS
    // and leave 16 at a time, a whole 128-byte line per lane: one 8-byte
    // store per mark left lines partly written while the chunk streamed
    // through the caches.
    constexpr uint32_t MG = 16;
    __shared__ double mbuf[64][MG + 1];
    const uint32_t lane = threadIdx.x;
    uint32_t nbuf = 0, flushed = 0;
    bool mark_ready = false;
    double mark = 0;
    auto store_mark = [&] {
        if (mark_ready) {
            mbuf[lane][nbuf++] = mark;
            b++;
            if (nbuf == MG) {
#pragma unroll
                for (uint32_t k = 0; k < MG; k++) m[flushed + k] = mbuf[lane][k];
                flushed += MG;
                nbuf = 0;
            }
        }
        mark_ready = false;
    };
    constexpr int U = chunk_batch<T2>();  // divides 2^mark_bits
    bool stopped;
    uint64_t i = stream_batches<U>(s, start, end, stopped, [&](const T2(&v)[U], uint64_t at) {
#pragma unroll
        for (int j = 0; j < U; j++) cumulative += prob(v[j]);
        if (((at + U - start) & block_mask) == 0 || at + U == end) {
            mark = cumulative;
            mark_ready = true;
            pending = false;
            return !(t_last < cumulative);  // every shot of the chunk lands in a walked block
        }
        pending = true;
        return true;
    }, store_mark);
    store_mark();
    if (!stopped) {  // fewer than U amplitudes left: a state of fewer than U amplitudes
        for (; i < end; i++) {
            cumulative += prob(s[i]);
            pending = true;
        }
        if (pending) {
            mark = cumulative;
            mark_ready = true;
            store_mark();
        }
    }
    for (uint32_t k = 0; k < nbuf; k++) m[flushed + k] = mbuf[lane][k];
    walked[c] = b;
}

template <typename T2>
__global__ __launch_bounds__(256) void k_walk_place(const T2 *__restrict__ s, uint64_t len, uint64_t chunk,
                                                     uint64_t num_chunks, uint32_t mark_bits,
                                                     const double *__restrict__ prefix,
                                                     const uint64_t *__restrict__ bounds,
                                                     const double *__restrict__ targets, uint64_t shots,
                                                     const double *__restrict__ marks,
                                                     const uint32_t *__restrict__ walked,
                                                     uint64_t *__restrict__ out) {
    // Shots [bounds[0], bounds[num_chunks]) fall in some chunk: all of them
    // for uniforms in [0, 1).
    const uint64_t first = bounds[0];
    const uint64_t past = bounds[num_chunks] < shots ? bounds[num_chunks] : shots;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t shot = first + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; shot < past; shot += step) {
        // the chunk: bounds[c] <= shot < bounds[c + 1]
        uint64_t lo = 0, hi = num_chunks;
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) >> 1;
            if (bounds[mid] <= shot) lo = mid;
            else hi = mid;
        }
        const uint64_t c = lo;
        const double t = targets[shot];
        const double *m = marks + ((c * chunk) >> mark_bits);
        const uint32_t nb = walked[c];
        // the first walked block whose mark is not <= t (a NaN sum counts as
        // past every target, as no comparison with it succeeds)
        uint32_t a = 0, z = nb;
        while (a < z) {
            const uint32_t mid = (a + z) >> 1;
            if (m[mid] <= t) a = mid + 1;
            else z = mid;
        }
        const uint64_t cend = c * chunk + chunk < len ? c * chunk + chunk : len;
        uint64_t found = 0;
        bool done = false;
        if (a < nb) {
            double cumulative = a == 0 ? prefix[c] : m[a - 1];
            uint64_t i = c * chunk + ((uint64_t)a << mark_bits);
            const uint64_t stop = i + (uint64_t(1) << mark_bits) < cend ? i + (uint64_t(1) << mark_bits) : cend;
            constexpr int U = 8;
            for (; i + U <= stop && !done; i += U) {
                T2 v[U];
#pragma unroll
                for (int j = 0; j < U; j++) v[j] = s[i + j];
#pragma unroll
                for (int j = 0; j < U; j++) {
                    cumulative += prob(v[j]);
                    if (!done && t < cumulative) {
                        found = i + j;
                        done = true;
                    }
                }
            }
            for (; i < stop && !done; i++) {
                cumulative += prob(s[i]);
                if (t < cumulative) {
                    found = i;
                    done = true;
                }
            }
        }
        if (!done) {  // a rounding leftover: the chunk's last nonzero probability (a NaN counts), else its start
            found = c * chunk;
            for (uint64_t i = cend; i-- > c * chunk;)
                if (!(prob(s[i]) == 0.0)) {
                    found = i;
                    break;
                }
        }
        out[shot] = found;
    }
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
static inline unsigned grid_for(uint64_t work, unsigned threads, unsigned cap = 16384) {
    uint64_t g = (work + threads - 1) / threads;
    if (g == 0) g = 1;
    return (unsigned)(g < cap ? g : cap);
}

template <typename T2>
void launch_init_zero(T2 *s, uint64_t len, hipStream_t st) {
    k_init_zero<T2><<<grid_for(len, 256), 256, 0, st>>>(s, len);
}
template <typename T2>
void launch_zero_stale(T2 *s, uint64_t len, uint64_t stale, hipStream_t st) {
    k_zero_stale<T2><<<grid_for(len, 256), 256, 0, st>>>(s, len, stale);
}
template <typename T2>
void launch_widen(const T2 *s, double2 *out, uint64_t count, hipStream_t st) {
    k_widen<T2><<<grid_for(count, 256), 256, 0, st>>>(s, out, count);
}
template <typename T2>
void launch_narrow(const double2 *in, T2 *s, uint64_t count, hipStream_t st) {
    k_narrow<T2><<<grid_for(count, 256), 256, 0, st>>>(in, s, count);
}
template <typename T2>
void launch_scale(T2 *s, uint64_t len, double factor, hipStream_t st) {
    k_scale<T2><<<grid_for(len, 256), 256, 0, st>>>(s, len, factor);
}

template <typename T2>
void launch_dense(T2 *s, const GateDesc &d, const T2 *mat, hipStream_t st) {
    const size_t lds = sizeof(T2) << (2 * d.k);
    const unsigned grid = grid_for(d.iterations, 256, 65536);
    switch (d.k) {
        case 1: k_dense<T2, 1><<<grid, 256, lds, st>>>(s, d, mat); break;
        case 2: k_dense<T2, 2><<<grid, 256, lds, st>>>(s, d, mat); break;
        case 3: k_dense<T2, 3><<<grid, 256, lds, st>>>(s, d, mat); break;
        case 4: k_dense<T2, 4><<<grid, 256, lds, st>>>(s, d, mat); break;
        case 5: k_dense<T2, 5><<<grid, 256, lds, st>>>(s, d, mat); break;
        case 6: k_dense<T2, 6><<<grid, 256, lds, st>>>(s, d, mat); break;
        default: break;
    }
}

template <typename T2>
void launch_diag(T2 *s, uint64_t len, uint32_t k, const DiagTargets &t, uint64_t cmask, uint64_t cval,
                 const T2 *entries, hipStream_t st) {
    k_diag<T2><<<grid_for(len, 256, 65536), 256, 0, st>>>(s, len, k, t, cmask, cval, entries);
}

template <typename T2>
size_t stage_lds_bytes(uint32_t L) {
    return (sizeof(T2) << L) + sizeof(T2) * 256;
}

template <typename T2>
void launch_stage(T2 *s, const StageDesc &d, const StageOp *ops, const T2 *coefs, unsigned grid, hipStream_t st) {
    k_stage<T2><<<grid, STAGE_THREADS, stage_lds_bytes<T2>(d.L), st>>>(s, d, ops, coefs);
}

template <typename T2>
void launch_masked_norm(const T2 *s, uint64_t len, uint64_t mask, uint64_t val, double *partial, unsigned blocks,
                        double *out, hipStream_t st) {
    k_masked_norm<T2><<<blocks, RED_THREADS, 0, st>>>(s, len, mask, val, partial);
    k_sum_partials<<<1, 64, 0, st>>>(partial, blocks, out);
}

template <typename T2>
void launch_probabilities(const T2 *s, uint64_t len, double *out, hipStream_t st) {
    k_probabilities<T2><<<grid_for(len, 256), 256, 0, st>>>(s, len, out);
}

template <typename T2>
void launch_expect_pauli(const T2 *s, uint64_t len, uint64_t x, const uint64_t *zs, uint32_t nterms, double *partial,
                         unsigned blocks, double *out, hipStream_t st) {
    if (nterms <= 1) k_expect_pauli<T2, 1><<<blocks, RED_THREADS, 0, st>>>(s, len, x, zs, nterms, partial);
    else if (nterms <= 2) k_expect_pauli<T2, 2><<<blocks, RED_THREADS, 0, st>>>(s, len, x, zs, nterms, partial);
    else if (nterms <= 4) k_expect_pauli<T2, 4><<<blocks, RED_THREADS, 0, st>>>(s, len, x, zs, nterms, partial);
    else if (nterms <= 8) k_expect_pauli<T2, 8><<<blocks, RED_THREADS, 0, st>>>(s, len, x, zs, nterms, partial);
    else k_expect_pauli<T2, PAULI_TERMS><<<blocks, RED_THREADS, 0, st>>>(s, len, x, zs, nterms, partial);
    k_sum_columns<<<1, 64, 0, st>>>(partial, blocks, 2 * nterms, out);
}

ExpDesc expect_desc(uint32_t state_qubits, uint32_t n) {
    ExpDesc d{};
    d.B = expect_segment_bits(state_qubits);
    d.nsegs = uint64_t(1) << (state_qubits - d.B);
    d.n = n;
    return d;
}

template <typename T2>
void launch_expect_z(const T2 *s, uint32_t n, double *seg_sums, double *z, double *zz, hipStream_t st) {
    const ExpDesc d = expect_desc(n, n);
    k_expect_segments<T2><<<(unsigned)d.nsegs, 1 << EZ_TID_BITS, 0, st>>>(s, d, nullptr, 1, nullptr, seg_sums);
    const uint32_t outs = 2 * n - 1;
    k_expect_combine<<<(outs + 63) / 64, 64, 0, st>>>(seg_sums, d, nullptr, 1, nullptr, z, zz);
}

template <typename T2>
void launch_expect_batch(const T2 *s, const ExpDesc *descs, uint32_t ndesc, const uint64_t *seg_start,
                         uint64_t total_segs, const uint32_t *out_start, uint32_t total_outs, double *seg_sums,
                         double *z, double *zz, hipStream_t st) {
    const ExpDesc none{};
    k_expect_segments<T2><<<(unsigned)total_segs, 1 << EZ_TID_BITS, 0, st>>>(s, none, descs, ndesc, seg_start, seg_sums);
    k_expect_combine<<<(total_outs + 63) / 64, 64, 0, st>>>(seg_sums, none, descs, ndesc, out_start, z, zz);
}

void launch_zero_words(uint32_t *p, uint32_t n, hipStream_t st) {
    k_zero_words<<<(n + 255) / 256, 256, 0, st>>>(p, n);
}

template <typename T2>
void launch_batch_init(T2 *s, uint64_t len, const uint64_t *offs, uint32_t count, hipStream_t st) {
    k_init_zero<T2><<<grid_for(len, 256), 256, 0, st>>>(s, len);  // amplitude 0 of the first state is 1
    k_set_basis<T2><<<(count + 63) / 64, 64, 0, st>>>(s, offs, count);
}

template <typename TA, typename TB>
void launch_compare(const TA *a, const TB *b, uint64_t len, double *partial, unsigned blocks, hipStream_t st) {
    k_compare<TA, TB><<<blocks, RED_THREADS, 0, st>>>(a, b, len, partial);
}
template void launch_compare<float2, float2>(const float2 *, const float2 *, uint64_t, double *, unsigned, hipStream_t);
template void launch_compare<float2, double2>(const float2 *, const double2 *, uint64_t, double *, unsigned, hipStream_t);
template void launch_compare<double2, float2>(const double2 *, const float2 *, uint64_t, double *, unsigned, hipStream_t);
template void launch_compare<double2, double2>(const double2 *, const double2 *, uint64_t, double *, unsigned,
                                               hipStream_t);

uint32_t expect_segment_bits(uint32_t n) { return n < EZ_TID_BITS + EZ_MAX_M_BITS ? n : EZ_TID_BITS + EZ_MAX_M_BITS; }

template <typename T2>
void launch_collapse(T2 *s, uint64_t len, uint32_t qubit, uint32_t outcome, double keep, hipStream_t st) {
    k_collapse<T2><<<grid_for(len, 256, 65536), 256, 0, st>>>(s, len, qubit, outcome, keep);
}

// Workspace of the allocation-free sampler. In host code `scratch` names a
// host-side stand-in, not the device array: its device address comes from
// hipGetSymbolAddress, once per device, and only that address is used.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

hipError_t noalloc_scratch(void **ptr) {
    static std::mutex mu;
    static void *addr[64] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[dev] = p;
    }
    *ptr = addr[dev];
    return hipSuccess;
}

size_t sort_temp_bytes(uint64_t shots) {
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, (const double *)nullptr, (double *)nullptr, (int)shots) !=
        hipSuccess)
        return 0;
    return bytes;
}

template <typename T2>
void launch_sample_sums(const T2 *s, uint64_t len, uint64_t chunk, SampleScratch &w, hipStream_t st) {
    const uint64_t nc = (len + chunk - 1) / chunk;
    // 64-lane workgroups for the per-chunk kernels, so 2^15 chunks spread over every CU.
    k_chunk_sums<T2><<<grid_for(nc, 64, 1u << 30), 64, 0, st>>>(s, len, chunk, nc, w.sums);
    k_prefix<<<1, PREFIX_THREADS, 0, st>>>(w.sums, nc, w.prefix);
}

hipError_t launch_sort_uniforms(const double *uniforms, uint64_t shots, SampleScratch &w, hipStream_t st) {
    size_t bytes = w.sort_temp_bytes;
    return hipcub::DeviceRadixSort::SortKeys(w.sort_temp, bytes, uniforms, w.targets, (int)shots, 0,
                                             sizeof(double) * 8, st);
}

template <typename T2>
hipError_t launch_sample_walk(const T2 *s, uint64_t len, uint64_t chunk, uint64_t shots, SampleScratch &w,
                              uint64_t *out, hipStream_t st) {
    const uint64_t nc = (len + chunk - 1) / chunk;
    k_targets<<<grid_for(shots, 256), 256, 0, st>>>(w.targets, w.prefix + nc, shots);
    k_bounds<<<grid_for(nc + 1, 256, 1u << 30), 256, 0, st>>>(w.prefix, nc, w.targets, shots, w.bounds);
    k_walk_marks<T2><<<grid_for(nc, 64, 1u << 30), 64, 0, st>>>(s, len, chunk, nc, w.mark_bits, w.prefix, w.bounds,
                                                                w.targets, shots, w.marks, w.walked);
    k_walk_place<T2><<<grid_for(shots, 256), 256, 0, st>>>(s, len, chunk, nc, w.mark_bits, w.prefix, w.bounds,
                                                           w.targets, shots, w.marks, w.walked, out);
    return hipGetLastError();
}

#define QVD_INSTANTIATE(T2)                                                                                     \
    template void launch_init_zero<T2>(T2 *, uint64_t, hipStream_t);                                            \
    template void launch_zero_stale<T2>(T2 *, uint64_t, uint64_t, hipStream_t);                                 \
    template void launch_widen<T2>(const T2 *, double2 *, uint64_t, hipStream_t);                               \
    template void launch_narrow<T2>(const double2 *, T2 *, uint64_t, hipStream_t);                              \
    template void launch_scale<T2>(T2 *, uint64_t, double, hipStream_t);                                        \
    template void launch_dense<T2>(T2 *, const GateDesc &, const T2 *, hipStream_t);                            \
    template void launch_diag<T2>(T2 *, uint64_t, uint32_t, const DiagTargets &, uint64_t, uint64_t, const T2 *, \
                                  hipStream_t);                                                                 \
    template size_t stage_lds_bytes<T2>(uint32_t);                                                              \
    template void launch_stage<T2>(T2 *, const StageDesc &, const StageOp *, const T2 *, unsigned, hipStream_t); \
    template void launch_masked_norm<T2>(const T2 *, uint64_t, uint64_t, uint64_t, double *, unsigned, double *, \
                                         hipStream_t);                                                          \
    template void launch_expect_z<T2>(const T2 *, uint32_t, double *, double *, double *, hipStream_t);         \
    template void launch_probabilities<T2>(const T2 *, uint64_t, double *, hipStream_t);                        \
    template void launch_expect_pauli<T2>(const T2 *, uint64_t, uint64_t, const uint64_t *, uint32_t, double *, \
                                          unsigned, double *, hipStream_t);                                     \
    template void launch_expect_batch<T2>(const T2 *, const ExpDesc *, uint32_t, const uint64_t *, uint64_t,    \
                                          const uint32_t *, uint32_t, double *, double *, double *, hipStream_t); \
    template void launch_batch_init<T2>(T2 *, uint64_t, const uint64_t *, uint32_t, hipStream_t);              \
    template void launch_collapse<T2>(T2 *, uint64_t, uint32_t, uint32_t, double, hipStream_t);                 \
    template void launch_sample_sums<T2>(const T2 *, uint64_t, uint64_t, SampleScratch &, hipStream_t);         \
    template hipError_t launch_sample_walk<T2>(const T2 *, uint64_t, uint64_t, uint64_t, SampleScratch &,       \
                                               uint64_t *, hipStream_t);

QVD_INSTANTIATE(float2)
QVD_INSTANTIATE(double2)

}  // namespace qvd

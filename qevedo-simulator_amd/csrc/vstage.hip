// Register-tile stage kernel (gfx950).
//
// One workgroup of W = 64 << WB threads owns a tile of 2^L amplitudes,
// L = 6 + WB + RB, held in VGPRs: thread t keeps R = 2^RB amplitudes x[0..R).
// The lowest FB lane bits are tile bits 0..FB-1 (physical qubits 0..FB-1, so
// every load and store is a run of 2^FB consecutive amplitudes); the other
// lane bits, the wave bits and the register bits hold the other tile bits in
// a role map that LDS exchanges change. The workgroup interprets the stage's
// program (vprog.h) on each of its tiles: one switch per instruction on the
// handler the host resolved it to. The complex128 geometry (2 waves per
// SIMD) prefetches the next tile into a second register set; the complex64
// one (4 waves per SIMD) relies on the other waves to cover a tile's loads.
//
// Coefficients indexed per lane (diagonal tables, scale tables, LDS-op
// matrices) are copied to LDS once per workgroup; per-instruction uniform
// coefficients are read through the scalar cache (a 1-qubit gate's from its
// instruction). Inside the tile loop no vector memory access other than the
// tile's own loads and stores is issued, so no wait on a coefficient or an
// instruction field waits on the tile's memory traffic.
#include <hip/hip_runtime.h>

#include "device.h"
#include "vprog.h"
#include "vstage.h"

#include <atomic>
#include <type_traits>

// Constant address space: loads through these pointers are invariant, so
// uniform ones go through the scalar cache (s_load) and never wait behind
// the tile's vector loads.
#if defined(__HIP_DEVICE_COMPILE__)
#define QVD_CONST __attribute__((address_space(4)))
#else
#define QVD_CONST
#endif

namespace qvd {

namespace {

// Byte i of a byte array in constant memory, read as part of its aligned
// 32-bit word: the scalar unit loads whole words only, and a byte load would
// take the vector memory path, whose wait counter the tile's own loads and
// stores share.
__device__ __forceinline__ uint32_t cbyte(const QVD_CONST uint8_t *p, uint32_t i) {
    const uintptr_t a = (uintptr_t)(p + i);
    const QVD_CONST uint32_t *w = (const QVD_CONST uint32_t *)(a & ~uintptr_t(3));
    return (*w >> (8 * (uint32_t)(a & 3))) & 0xffu;
}

__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T2>
__device__ __forceinline__ T2 mk(decltype(T2::x) a, decltype(T2::x) b) {
    T2 r;
    r.x = a;
    r.y = b;
    return r;
}

// w * x
template <typename T2>
__device__ __forceinline__ T2 cm(const T2 w, const T2 x) {
    return mk<T2>(fma_(-w.y, x.y, w.x * x.x), fma_(w.y, x.x, w.x * x.y));
}
// a + w * x
template <typename T2>
__device__ __forceinline__ T2 cfma(const T2 w, const T2 x, const T2 a) {
    return mk<T2>(fma_(-w.y, x.y, fma_(w.x, x.x, a.x)), fma_(w.y, x.x, fma_(w.x, x.y, a.y)));
}
// a + i s x (s real)
template <typename T2>
__device__ __forceinline__ T2 imag_fma(const decltype(T2::x) s, const T2 x, const T2 a) {
    return mk<T2>(fma_(-s, x.y, a.x), fma_(s, x.x, a.y));
}
// a + s x (s real)
template <typename T2>
__device__ __forceinline__ T2 real_fma(const decltype(T2::x) s, const T2 x, const T2 a) {
    return mk<T2>(fma_(s, x.x, a.x), fma_(s, x.y, a.y));
}
// w0 * x0 + w1 * x1
template <typename T2>
__device__ __forceinline__ T2 cdot(const T2 w0, const T2 x0, const T2 w1, const T2 x1) {
    return cfma(w1, x1, cm(w0, x0));
}

// Value of v in lane ^ (1 << B): DPP quad permutes for bits 0 and 1, a
// half-row mirror followed by a quad reversal for bit 2, a row rotation by 8
// for bit 3 and a swizzle for bit 4 (bit 5 uses the row swap below). All 64
// lanes are active wherever this is called, and every source lane lies in
// the reading lane's row, so bound_ctrl changes no value; it lets the
// compiler fold the move into a following add or subtract as its DPP operand
// (the rotation form on lane bits 0-3 then costs one instruction per part).
template <int B>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v, uint32_t lane) {
    if constexpr (B == 0) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
    } else if constexpr (B == 1) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
    } else if constexpr (B == 2) {
        const int m = __builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, true);  // row_half_mirror: 7 - i
        return (uint32_t)__builtin_amdgcn_mov_dpp(m, 0x1B, 0xF, 0xF, true);      // quad_perm [3,2,1,0]
    } else if constexpr (B == 3) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, true);  // row_ror:8
    } else {
        // Bit 4 crosses DPP rows: the LDS crossbar's swizzle in bit mode
        // (lane ^ 16 within each half-wave; and-mask 0x1f, xor-mask 16) moves
        // it on the LDS pipe, off the VALU.
        static_assert(B == 4, "lane bit 5 uses the row swap");
        return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x401F);
    }
}

// Row swaps of gfx950 (B = 4: 16-lane rows, B = 5: 32-lane halves): the lanes
// with bit B set of a trade places with the lanes with bit B clear of b.
// Afterwards a holds (a's amplitude with bit B clear, b's amplitude with bit B
// clear) in its low / high lanes and b holds the two with bit B set, so a pair
// of registers carries two complete amplitude pairs; the same call undoes it.
template <int B>
__device__ __forceinline__ void row_swap(uint32_t &a, uint32_t &b) {
    if constexpr (B == 4) {
        const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
        a = r[0];
        b = r[1];
    } else {
        const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
        a = r[0];
        b = r[1];
    }
}
template <int B>
__device__ __forceinline__ void row_swap2(float2 &a, float2 &b) {
    uint32_t ax = __float_as_uint(a.x), ay = __float_as_uint(a.y), bx = __float_as_uint(b.x),
             by = __float_as_uint(b.y);
    row_swap<B>(ax, bx);
    row_swap<B>(ay, by);
    a = make_float2(__uint_as_float(ax), __uint_as_float(ay));
    b = make_float2(__uint_as_float(bx), __uint_as_float(by));
}
__device__ __forceinline__ void split64(double v, uint32_t &lo, uint32_t &hi) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    lo = (uint32_t)u;
    hi = (uint32_t)(u >> 32);
}
__device__ __forceinline__ double join64(uint32_t lo, uint32_t hi) {
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int B>
__device__ __forceinline__ void row_swap2(double2 &a, double2 &b) {
    uint32_t a0, a1, a2, a3, b0, b1, b2, b3;
    split64(a.x, a0, a1);
    split64(a.y, a2, a3);
    split64(b.x, b0, b1);
    split64(b.y, b2, b3);
    row_swap<B>(a0, b0);
    row_swap<B>(a1, b1);
    row_swap<B>(a2, b2);
    row_swap<B>(a3, b3);
    a = make_double2(join64(a0, a1), join64(a2, a3));
    b = make_double2(join64(b0, b1), join64(b2, b3));
}

template <int B>
__device__ __forceinline__ float2 lane_xor2(float2 v, uint32_t lane) {
    return make_float2(__uint_as_float(lane_xor<B>(__float_as_uint(v.x), lane)),
                       __uint_as_float(lane_xor<B>(__float_as_uint(v.y), lane)));
}
template <int B>
__device__ __forceinline__ double lane_xor_f64(double v, uint32_t lane) {
    uint32_t lo, hi;
    split64(v, lo, hi);
    return join64(lane_xor<B>(lo, lane), lane_xor<B>(hi, lane));
}
template <int B>
__device__ __forceinline__ double2 lane_xor2(double2 v, uint32_t lane) {
    return make_double2(lane_xor_f64<B>(v.x, lane), lane_xor_f64<B>(v.y, lane));
}

// Value of v in lane ^ (1 << B), B < 5, through the LDS crossbar's swizzle
// (bit mode: and-mask 0x1f, xor-mask 1 << B within each 32-lane half): it
// runs on the LDS pipe, so a gate whose FMAs keep the VALU busy does not also
// spend VALU cycles on the partner fetch.
template <int B>
__device__ __forceinline__ float2 lane_xor_swz2(float2 v) {
    static_assert(B < 5, "swizzle reaches lane bits 0-4");
    constexpr int PAT = 0x1F | ((1 << B) << 10);
    return make_float2(__int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v.x), PAT)),
                       __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v.y), PAT)));
}
template <int B>
__device__ __forceinline__ double2 lane_xor_swz2(double2 v) {
    return lane_xor2<B>(v, 0u);
}

template <typename T2>
__device__ __forceinline__ T2 sel(bool p, T2 a, T2 b) {
    return mk<T2>(p ? a.x : b.x, p ? a.y : b.y);
}

// Flips the sign of both parts of a where bit 31 of m is set.
__device__ __forceinline__ float2 xor_sign(float2 a, uint32_t m) {
    return make_float2(__uint_as_float(__float_as_uint(a.x) ^ m), __uint_as_float(__float_as_uint(a.y) ^ m));
}
__device__ __forceinline__ double2 xor_sign(double2 a, uint32_t m) {
    const unsigned long long w = (unsigned long long)m << 32;
    return make_double2(__longlong_as_double(__double_as_longlong(a.x) ^ w),
                        __longlong_as_double(__double_as_longlong(a.y) ^ w));
}

// A run of CZ-like gates: amplitude (thread, register j) changes sign when
// the quadratic form sum x_p x_q over the run's pairs is odd. Pairs of two
// register bits arrive as the per-register mask v.neg; pairs of a thread bit
// and a register bit r as the thread-bit mask inl[r] (register r's linear
// term is the parity of tid & inl[r]); pairs of thread bits t < u as the
// mask inl[4 + t] (the thread's constant term gains x_t * parity(tid &
// inl[4 + t])); pairs with a non-tile qubit are (encoded bit, encoded bit)
// in the int array. The parity of register j is built as t[j] = t[j with its
// lowest set bit cleared] ^ M[that bit], one exclusive-or per register, kept
// in bit 31.
template <typename T2, int RB>
__device__ __forceinline__ void signs(T2 (&x)[1 << RB], const QVD_CONST VInstr &v, const QVD_CONST uint32_t *ints,
                                      uint32_t tid, uint64_t base) {
    const uint64_t neg = v.neg;
    const uint32_t k = v.k, form = v.form;
    if (k == 0 && form == 0) {
#pragma unroll
        for (int j = 0; j < (1 << RB); j++) x[j] = xor_sign(x[j], (uint32_t)((neg >> j) & 1) << 31);
        return;
    }
    uint32_t al = 0, gam = 0, gm = 0;
    if (form & 1) {
#pragma unroll
        for (int r = 0; r < RB; r++) gam |= (__builtin_popcount(tid & v.inl[r]) & 1u) << r;
    }
    if (form & 2) {
#pragma unroll
        for (int t = 0; t < 10; t++) {
            const uint32_t m = v.inl[VSIGN_TT + t];
            if (m != 0) al ^= (tid >> t) & __builtin_popcount(tid & m);
        }
        al &= 1u;
    }
    for (uint32_t i = 0; i < k; i++) {
        const uint32_t a = ints[v.aux + 2 * i], b = ints[v.aux + 2 * i + 1];
        const uint32_t ka = a & VT_KIND, kb = b & VT_KIND, xa = a & 0xffffu, xb = b & 0xffffu;
        const uint32_t va = ka == VT_THREAD ? (tid >> xa) & 1 : (uint32_t)(base >> xa) & 1;
        const uint32_t vb = kb == VT_THREAD ? (tid >> xb) & 1 : (uint32_t)(base >> xb) & 1;
        if (ka != VT_REG && kb != VT_REG) al ^= va & vb;
        else if (ka == VT_REG) {
            if (kb == VT_THREAD) gam ^= vb << xa;
            else gm ^= vb << xa;
        } else {
            if (ka == VT_THREAD) gam ^= va << xb;
            else gm ^= va << xb;
        }
    }
    const uint32_t mask = gam ^ gm;
    uint32_t t[1 << RB];
    t[0] = al << 31;
#pragma unroll
    for (int j = 1; j < (1 << RB); j++) t[j] = t[j & (j - 1)] ^ (((mask >> __builtin_ctz(j)) & 1) << 31);
#pragma unroll
    for (int j = 0; j < (1 << RB); j++) x[j] = xor_sign(x[j], t[j] ^ ((uint32_t)((neg >> j) & 1) << 31));
}

// The unit-form pair (a, b) -> (a + i^K b, b + i^(2-K) a) (the second
// rotation is fixed by unitarity: M = [[1, m], [-conj(m), 1]]), computed in
// place: the second output is rebuilt from the first and b
// (b + i^(2-K) a = i^(2-K) o0 + 2 b, whose rotation only swaps and negates
// parts), so both outputs can take the registers of their inputs and no
// copy follows. One add and one FMA by 2 per part; the rounding differs from
// the direct form by at most one unit of the inputs' magnitude.
__device__ __forceinline__ float twice_plus(float b, float o) { return __builtin_fmaf(2.0f, b, o); }
__device__ __forceinline__ float twice_minus(float b, float o) { return __builtin_fmaf(2.0f, b, -o); }
__device__ __forceinline__ double twice_plus(double b, double o) { return __builtin_fma(2.0, b, o); }
__device__ __forceinline__ double twice_minus(double b, double o) { return __builtin_fma(2.0, b, -o); }
template <int K, typename T2>
__device__ __forceinline__ void rotpair(T2 &a, T2 &b) {
    if (K == 0) {  // (a + b, b - a)
        a.x = a.x + b.x;
        b.x = twice_minus(b.x, a.x);
        a.y = a.y + b.y;
        b.y = twice_minus(b.y, a.y);
    } else if (K == 2) {  // (a - b, b + a)
        a.x = a.x - b.x;
        b.x = twice_plus(b.x, a.x);
        a.y = a.y - b.y;
        b.y = twice_plus(b.y, a.y);
    } else if (K == 1) {  // (a + i b, b + i a)
        a.x = a.x - b.y;
        b.y = twice_plus(b.y, a.x);
        a.y = a.y + b.x;
        b.x = twice_minus(b.x, a.y);
    } else {  // (a - i b, b - i a)
        a.x = a.x + b.y;
        b.y = twice_minus(b.y, a.x);
        a.y = a.y - b.x;
        b.x = twice_plus(b.x, a.y);
    }
}

// 1-qubit gate on register bit B.
template <typename T2, int RB, int B, int FORM, bool CTRL, int K1 = 0>
__device__ __forceinline__ void u1r(T2 (&x)[1 << RB], const T2 *c, uint32_t tpred_ok, uint32_t rmask,
                                    uint32_t rval) {
    T2 c0, c1, c2, c3;
    if (FORM != FORM_UNIT_ROT) {
        c0 = c[0];
        c1 = c[1];
        c2 = c[2];
        c3 = c[3];
    }
    const bool tp = tpred_ok != 0;
#pragma unroll
    for (int j0 = 0; j0 < (1 << RB); j0++) {
        if (j0 & (1 << B)) continue;
        const int j1 = j0 | (1 << B);
        if (CTRL && ((uint32_t)j0 & rmask) != rval) continue;
        if (FORM == FORM_UNIT_ROT) {
            rotpair<K1>(x[j0], x[j1]);
            continue;
        }
        const T2 a = x[j0], b = x[j1];
        T2 o0, o1;
        if (FORM == FORM_UNIT_DIAG) {
            o0 = cfma(c1, b, a);
            o1 = cfma(c2, a, b);
        } else if (FORM == FORM_UNIT_IMAG) {
            o0 = imag_fma(c1.y, b, a);
            o1 = imag_fma(c2.y, a, b);
        } else if (FORM == FORM_UNIT_REAL) {
            o0 = real_fma(c1.x, b, a);
            o1 = real_fma(c2.x, a, b);
        } else if (FORM == FORM_UNIT_ANTI) {
            o0 = cfma(c0, a, b);
            o1 = cfma(c3, b, a);
        } else {
            o0 = cdot(c0, a, c1, b);
            o1 = cdot(c2, a, c3, b);
        }
        if (CTRL) {
            o0 = sel(tp, o0, a);
            o1 = sel(tp, o1, b);
        }
        x[j0] = o0;
        x[j1] = o1;
    }
}

// 1-qubit gate on lane bit B: the partner amplitude comes from lane ^ (1 << B).
template <typename T2, int RB, int FORM, bool CTRL, int B>
__device__ __forceinline__ void u1l_bit(T2 (&x)[1 << RB], const T2 *c, uint32_t lane, uint32_t tp,
                                        uint32_t rmask, uint32_t rval) {
    const bool hi = (lane >> B) & 1;
    const T2 cs = sel(hi, c[3], c[0]);  // coefficient of the lane's own amplitude
    const T2 cp = sel(hi, c[2], c[1]);  // coefficient of the partner's
#pragma unroll
    for (int j = 0; j < (1 << RB); j++) {
        if (CTRL && ((uint32_t)j & rmask) != rval) continue;
        const T2 a = x[j];
        // The FMA forms fetch the partner through the swizzle (LDS pipe),
        // the add-only rotation form (u1l_bit_rot) through DPP folded into
        // its adds.
        const T2 p = lane_xor_swz2<B>(a);
        T2 o;
        if (FORM == FORM_UNIT_DIAG) o = cfma(cp, p, a);
        else if (FORM == FORM_UNIT_IMAG) o = imag_fma(cp.y, p, a);
        else if (FORM == FORM_UNIT_REAL) o = real_fma(cp.x, p, a);
        else if (FORM == FORM_UNIT_ANTI) o = cfma(cs, a, p);
        else o = cdot(cs, a, cp, p);
        if (CTRL) o = sel(tp != 0, o, a);
        x[j] = o;
    }
}

// FORM_UNIT_ROT on lane bit B < 5: the lane with bit B clear computes
// a + i^K b and its partner b + i^(2-K) a, so each lane adds i^K (K odd) or
// +-1 (K even, the sign set by the lane's bit) times its partner's amplitude:
// two adds or FMAs per amplitude besides the partner fetch.
template <typename T2, int RB, int B, int K>
__device__ __forceinline__ void u1l_bit_rot(T2 (&x)[1 << RB], uint32_t lane) {
    using R = decltype(T2::x);
    const bool hi = (lane >> B) & 1;
    const R sgn = (K == 0) == hi ? R(-1) : R(1);  // K even: +1 where the own term leads, -1 where it trails
#pragma unroll
    for (int j = 0; j < (1 << RB); j++) {
        T2 a = x[j];
        const T2 p = lane_xor2<B>(a, lane);
        if constexpr (K == 1) {
            a.x = a.x - p.y;
            a.y = a.y + p.x;
        } else if constexpr (K == 3) {
            a.x = a.x + p.y;
            a.y = a.y - p.x;
        } else {
            a.x = fma_(sgn, p.x, a.x);
            a.y = fma_(sgn, p.y, a.y);
        }
        x[j] = a;
    }
}

// 1-qubit gate on lane bit B = 4 or 5: registers 2i and 2i + 1 are row-swapped
// so that each lane holds a whole amplitude pair (of register 2i in the lanes
// with bit B clear, of 2i + 1 in the others), the gate runs with uniform
// coefficients as on a register bit, and a second swap puts the results back.
template <typename T2, int RB, int FORM, bool CTRL, int B, int K1 = 0>
__device__ __forceinline__ void u1l_pair(T2 (&x)[1 << RB], const T2 *c, uint32_t lane, uint32_t tp,
                                         uint32_t rmask, uint32_t rval) {
    T2 c0, c1, c2, c3;
    if (FORM != FORM_UNIT_ROT) {
        c0 = c[0];
        c1 = c[1];
        c2 = c[2];
        c3 = c[3];
    }
    const bool hi = (lane >> B) & 1;
#pragma unroll
    for (int j = 0; j < (1 << RB); j += 2) {
        const uint32_t reg = hi ? (uint32_t)j + 1 : (uint32_t)j;
        if (CTRL && (rmask & 1) == 0 && (((uint32_t)j & rmask) != rval)) continue;  // neither register selected
        T2 a = x[j], b = x[j + 1];
        row_swap2<B>(a, b);
        if (FORM == FORM_UNIT_ROT) {
            rotpair<K1>(a, b);
            row_swap2<B>(a, b);
            x[j] = a;
            x[j + 1] = b;
            continue;
        }
        T2 o0, o1;
        if (FORM == FORM_UNIT_DIAG) {
            o0 = cfma(c1, b, a);
            o1 = cfma(c2, a, b);
        } else if (FORM == FORM_UNIT_IMAG) {
            o0 = imag_fma(c1.y, b, a);
            o1 = imag_fma(c2.y, a, b);
        } else if (FORM == FORM_UNIT_REAL) {
            o0 = real_fma(c1.x, b, a);
            o1 = real_fma(c2.x, a, b);
        } else if (FORM == FORM_UNIT_ANTI) {
            o0 = cfma(c0, a, b);
            o1 = cfma(c3, b, a);
        } else {
            o0 = cdot(c0, a, c1, b);
            o1 = cdot(c2, a, c3, b);
        }
        if (CTRL) {
            const bool on = tp != 0 && (reg & rmask) == rval;
            o0 = sel(on, o0, a);
            o1 = sel(on, o1, b);
        }
        row_swap2<B>(o0, o1);
        x[j] = o0;
        x[j + 1] = o1;
    }
}

// A run of pair phases: amplitude (thread, register j) gains w for every
// pair whose two bits are 1. Pairs with a register bit r scale the per-thread
// factor P[r] (w where the other bit is 1); pairs without one scale the
// per-thread factor T; pairs of two register bits arrive as the table R[j]
// (form 1). Register j then takes T * R[j] * prod over set bits b of j of
// P[b], built as F[j] = F[j with its lowest set bit cleared] * P[that bit].
template <typename T2, int RB>
__device__ __forceinline__ void phases(T2 (&x)[1 << RB], const QVD_CONST VInstr &v, const QVD_CONST T2 *coef,
                                       const QVD_CONST uint32_t *ints, uint32_t tid, uint64_t base) {
    const T2 one = mk<T2>(1, 0);
    T2 t = one;
    T2 pr[RB];
#pragma unroll
    for (int b = 0; b < RB; b++) pr[b] = one;
    // Pair records in the state's precision, built at upload: the pairs of
    // register bit r as (b, w), inl[1 + r] of them, then the others as
    // (a, b, w), inl[1 + VREG_SLOTS] of them. One scalar fetch per pair, none
    // depending on another load, and no dispatch on the register bit.
    constexpr uint32_t W = sizeof(T2) / 4;
    static_assert(RB <= (int)VREG_SLOTS, "the host writes one pair group per register bit below VREG_SLOTS");
    auto weight = [&](const QVD_CONST uint32_t *r) {
        T2 w;
        if constexpr (sizeof(T2) == 8) {
            w.x = __uint_as_float(r[0]);
            w.y = __uint_as_float(r[1]);
        } else {
            w.x = join64(r[0], r[1]);
            w.y = join64(r[2], r[3]);
        }
        return w;
    };
    auto bit = [&](uint32_t e) {
        const uint32_t xe = e & 0xffffu;
        return (e & VT_KIND) == VT_THREAD ? ((tid >> xe) & 1) != 0 : ((base >> xe) & 1) != 0;
    };
    const QVD_CONST uint32_t *rec = ints + v.inl[0];
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const uint32_t n = v.inl[1 + rb];
        for (uint32_t i = 0; i < n; i++, rec += 1 + W) pr[rb] = cm(sel(bit(rec[0]), weight(rec + 1), one), pr[rb]);
    }
    const uint32_t n_other = v.inl[1 + VREG_SLOTS];
    for (uint32_t i = 0; i < n_other; i++, rec += 2 + W) t = cm(sel(bit(rec[0]) && bit(rec[1]), weight(rec + 2), one), t);
    // f[j] is live only until its last child; the barriers keep the
    // scheduler from hoisting every product ahead of its use.
    const bool table = v.form == 1;
    const QVD_CONST T2 *rt = coef + v.coef;
    T2 f[1 << RB];
    f[0] = t;
#pragma unroll
    for (int j = 0; j < (1 << RB); j++) {
        if (j) f[j] = cm(pr[__builtin_ctz(j)], f[j & (j - 1)]);
        const T2 g = table ? cm(rt[j], f[j]) : f[j];
        x[j] = cm(g, x[j]);
        if ((j & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
}

// A diagonal op: entry index = (thread-index part, per lane) | (non-tile
// part, per tile) | rtab[j] (register part, precomputed per register).
// THREADED = false when no target or control is a thread-index bit: then
// every decision is uniform and sign flips are plain branches.
template <typename T2, int RB, bool SIGN, bool THREADED>
__device__ __forceinline__ void diag(T2 (&x)[1 << RB], const QVD_CONST VInstr &v, const T2 *__restrict__ cst, uint32_t tid,
                                     uint64_t base) {
    uint32_t it = 0, ig = 0;
#pragma unroll
    for (uint32_t j = 0; j < MAX_GATE_QUBITS; j++) {
        if (j >= v.k) continue;
        const uint32_t e = v.tgt[j], kind = e & VT_KIND, b = e & 0xffffu;
        if (THREADED && kind == VT_THREAD) it |= ((tid >> b) & 1) << j;
        else if (kind == VT_GLOBAL) ig |= (uint32_t)((base >> b) & 1) << j;
    }
    const bool tp = !THREADED || (tid & v.tmask) == v.tval;
#pragma unroll
    for (int j = 0; j < (1 << RB); j++) {
        if (((uint32_t)j & v.rmask) != v.rval) continue;
        const uint32_t ij = ig | cbyte(v.rtab, j);
        const T2 a = x[j];
        if (SIGN) {
            if (THREADED) {
                const bool neg = tp && ((v.neg >> (ij | it)) & 1);
                x[j] = neg ? mk<T2>(-a.x, -a.y) : a;
            } else if ((v.neg >> ij) & 1) {
                x[j] = mk<T2>(-a.x, -a.y);
            }
        } else {
            const T2 o = cm(cst[v.coef + (ij | it)], a);
            x[j] = tp ? o : a;
        }
    }
}

// In-LDS dense op (the general path): groups striped over the workgroup.
template <typename T2, int K, int W>
__device__ __forceinline__ void lds_dense(T2 *__restrict__ tile, const T2 *__restrict__ m, const QVD_CONST StageOp &op,
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
    for (uint32_t g = threadIdx.x; g < groups; g += W) {
        uint32_t i0 = g;
        for (uint32_t j = 0; j < op.nins; j++) {
            const uint32_t p = op.ins[j];
            i0 = ((i0 >> p) << (p + 1)) | (i0 & ((1u << p) - 1));
        }
        i0 |= op.lc_val;
        T2 a[D];
#pragma unroll
        for (int j = 0; j < D; j++) a[j] = tile[i0 | off[j]];
        // Rows one at a time (not unrolled): the row's offset is rebuilt
        // from its bits so no array is indexed at run time.
#pragma unroll 1
        for (int r = 0; r < D; r++) {
            T2 acc = mk<T2>(0, 0);
#pragma unroll
            for (int c = 0; c < D; c++) acc = cfma(m[r * D + c], a[c], acc);
            uint32_t orow = 0;
#pragma unroll
            for (int b = 0; b < K; b++) orow |= (uint32_t)((r >> b) & 1) << op.tgt[b];
            tile[i0 | orow] = acc;
        }
    }
}

// A 1-qubit gate's four coefficients, read from the instruction (inline
// copy in the state's precision).
// Workgroup barrier for the LDS tile: the wave's LDS accesses have completed
// (lgkmcnt 0), then every wave meets. Unlike __syncthreads it does not wait
// on global memory, so the tile's stores and prefetched loads stay in
// flight; the empty asm statements keep the compiler from moving memory
// accesses across.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // gfx9 encoding: lgkmcnt(0), vmcnt and expcnt not waited
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

template <typename T2>
__device__ __forceinline__ void inline_coefs(const QVD_CONST VInstr &v, T2 (&c)[4]) {
    const QVD_CONST T2 *p = (const QVD_CONST T2 *)v.inl;
#pragma unroll
    for (int j = 0; j < 4; j++) c[j] = p[j];
}

// Runs stage d on its tiles t_begin, t_begin + step, ... below t_end; s is
// the first amplitude of the stage's state. With MAY_ZERO the stage may be
// marked from_zero: its tiles are then made in registers (|0...0>), not
// loaded; or it may name stale qubits (zero_mem), whose amplitudes it sets to
// zero instead of loading them. FBC is the number of fixed lane bits, or -1 to read it from the
// descriptor (it enters only the offsets of the role maps, not the gate
// handlers).
template <typename T2, int WB, int RB, int FBC, bool MAY_ZERO>
__device__ __forceinline__ void stage_tiles(T2 *__restrict__ s, const QVD_CONST VStageDesc &d, uint64_t t_begin,
                                            uint64_t t_end, uint64_t step, const VInstr *__restrict__ prog,
                                            const T2 *__restrict__ coefs, const StageOp *__restrict__ ldsops,
                                            const uint32_t *__restrict__ ints) {
    constexpr int W = 64 << WB, R = 1 << RB;
    constexpr uint32_t L = 6 + WB + RB;
    // 16-byte aligned: dynamic LDS follows a kernel's static LDS (the queue
    // kernel's item slot), and a tile at a 4-byte offset would split every
    // 8- and 16-byte LDS access of the exchanges.
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    T2 *tile = reinterpret_cast<T2 *>(smem_raw);
    T2 *cst = tile + (1u << L);  // the stage's coefficients (index = coef - cbegin)
    const QVD_CONST VInstr *cprog = (const QVD_CONST VInstr *)prog;
    const QVD_CONST T2 *ccoef = (const QVD_CONST T2 *)coefs;
    const QVD_CONST StageOp *cops = (const QVD_CONST StageOp *)ldsops;
    const QVD_CONST uint32_t *cints = (const QVD_CONST uint32_t *)ints;
    const uint32_t tid = threadIdx.x, lane = tid & 63;

    for (uint32_t i = tid; i < d.ccount; i += W) cst[i] = coefs[d.cbegin + i];
    const T2 *cbase = cst - d.cbegin;  // so cbase[v.coef + i] is the LDS copy
    __syncthreads();

    // Physical offsets of the initial role map (loads) and the final one
    // (stores): thread part per lane, register part per register bit.
    const uint32_t FB = FBC >= 0 ? (uint32_t)FBC : d.fixed_lanes;
    const uint32_t FL = 6 - FB;  // flexible lane bits
    uint64_t t_off0 = lane & ((1u << FB) - 1), t_offF = t_off0;
#pragma unroll
    for (int s_ = 0; s_ < (int)VMAX_L_FLEX; s_++) {
        if (s_ < (int)FL) {
            t_off0 |= (uint64_t)((lane >> (FB + s_)) & 1) << d.local_pos[cbyte(d.lmap0, s_)];
            t_offF |= (uint64_t)((lane >> (FB + s_)) & 1) << d.local_pos[cbyte(d.lmapF, s_)];
        }
    }
#pragma unroll
    for (int s_ = 0; s_ < WB; s_++) {
        t_off0 |= (uint64_t)((tid >> (6 + s_)) & 1) << d.local_pos[cbyte(d.wmap0, s_)];
        t_offF |= (uint64_t)((tid >> (6 + s_)) & 1) << d.local_pos[cbyte(d.wmapF, s_)];
    }
    uint64_t p0[RB], pF[RB];
#pragma unroll
    for (int b = 0; b < RB; b++) {
        p0[b] = uint64_t(1) << d.local_pos[cbyte(d.rmap0, b)];
        pF[b] = uint64_t(1) << d.local_pos[cbyte(d.rmapF, b)];
    }

    // Role map at the start of every tile: the thread's part of the tile
    // index (fixed lane bits, flexible lane bits, wave bits) and the
    // register bits' tile bits (uniform).
    auto thread_index = [&](const QVD_CONST uint8_t *lm, const QVD_CONST uint8_t *wmp) {
        uint32_t ti = lane & ((1u << FB) - 1);
#pragma unroll
        for (int s_ = 0; s_ < (int)VMAX_L_FLEX; s_++)
            if (s_ < (int)FL) ti |= ((lane >> (FB + s_)) & 1) << cbyte(lm, s_);
#pragma unroll
        for (int s_ = 0; s_ < WB; s_++) ti |= ((tid >> (6 + s_)) & 1) << cbyte(wmp, s_);
        return ti;
    };
    const uint32_t ti0 = thread_index(d.lmap0, d.wmap0);
    uint32_t rm0[RB];
#pragma unroll
    for (int b = 0; b < RB; b++) rm0[b] = cbyte(d.rmap0, b);

    const QVD_CONST VInstr *const ip = cprog + d.instr_off;
    const QVD_CONST VInstr *const ip_end = ip + d.num_instr;

    // With 4 waves per SIMD (16-wave workgroups) the other waves cover a
    // tile's load latency; with 2 the next tile is prefetched into a second
    // register set.
    constexpr bool PREFETCH = WB <= 3;
    T2 x[R], y[PREFETCH ? R : 1];
    auto tile_base = [&](uint64_t t) {
        uint64_t base = 0;
        for (uint32_t j = 0; j < d.nglob; j++) base |= ((t >> j) & 1) << d.glob_pos[j];
        return base;
    };
    // With MAY_ZERO, amplitudes whose index has a bit of d.zero_mem set are
    // zero and their memory is stale: they are set, not loaded.
    auto load = [&](T2(&dst)[R], uint64_t base) {
        const T2 *src = s + base + t_off0;
        const uint64_t zm = MAY_ZERO ? d.zero_mem : 0;
        if (zm == 0) {
#pragma unroll
            for (int j = 0; j < R; j++) {
                uint64_t o = 0;
#pragma unroll
                for (int b = 0; b < RB; b++)
                    if ((j >> b) & 1) o += p0[b];
                dst[j] = src[o];
            }
        } else {
            const bool tz = ((base | t_off0) & zm) != 0;
#pragma unroll
            for (int j = 0; j < R; j++) {
                uint64_t o = 0;
#pragma unroll
                for (int b = 0; b < RB; b++)
                    if ((j >> b) & 1) o += p0[b];
                if (tz || (o & zm) != 0) dst[j] = mk<T2>(0, 0);
                else dst[j] = src[o];
            }
        }
    };

    // The first tile is loaded, and waited for, before the loop. With a
    // prefetch set the loop then copies the next tile into x at the end of
    // each iteration, after its stores; the wait for those loads leaves the
    // stores in flight, and no register entering the loop has a load pending
    // (the memory counter is shared by loads and stores, and a pending load
    // at the loop head would make every tile wait for its predecessor's
    // stores).
    // A stage that starts from |0...0>: amplitude 0 (tile 0, thread offset 0,
    // register 0) is 1, every other one 0.
    const bool zero = MAY_ZERO && d.from_zero != 0;
    auto make_zero = [&](T2(&dst)[R], uint64_t t) {
#pragma unroll
        for (int j = 0; j < R; j++) dst[j] = mk<T2>(j == 0 && t == 0 && t_off0 == 0 ? 1 : 0, 0);
    };
    uint64_t t = t_begin;
    if (t < t_end) {
        if (zero) make_zero(x, t);
        else load(x, tile_base(t));
    }
    if constexpr (PREFETCH) __builtin_amdgcn_s_waitcnt(0);
    for (; t < t_end; t += step) {
        const uint64_t base = tile_base(t);
        const uint64_t tn = t + step;
        if constexpr (PREFETCH)
            if (tn < t_end && !zero) load(y, tile_base(tn));

        // Current role map (an exchange changes it; each tile starts from
        // the stage's initial one).
        uint32_t ti = ti0;
        uint32_t rm[RB];
#pragma unroll
        for (int b = 0; b < RB; b++) rm[b] = rm0[b];

        // The program, one instruction at a time: the host has resolved each
        // instruction to a handler (VHandler), so dispatch is one switch and
        // every arm is straight-line code with its register bits, forms and
        // rotations fixed at compile time.
        auto handle = [&](auto hc, const QVD_CONST VInstr &v) {
            constexpr uint32_t H = decltype(hc)::value;
            // Controls on non-tile qubits come only with the full and
            // controlled 1-qubit forms, general diagonals and LDS ops (the
            // host checks this); those arms skip tiles where they do not
            // hold, and no other arm spends a compare on them.
            constexpr bool MAY_GCTRL = (H >= H_U1R_FULL && H < H_U1L_ROT) || (H >= H_U1L_FULL && H < H_SIGNS) ||
                                       H == H_LDSOP;
            if constexpr (MAY_GCTRL)
                if ((int32_t)v.op < 0 && (base & v.gmask) != v.gval) return;  // VF_GCTRL is the sign bit
            if constexpr (H >= H_U1R_IMAG) {
                // Unit forms with real or imaginary off-diagonals.
                T2 c[4];
                inline_coefs(v, c);
                if constexpr (H < H_U1L_IMAG) {
                    constexpr int F = H < H_U1R_REAL ? FORM_UNIT_IMAG : FORM_UNIT_REAL;
                    constexpr int B = (H - H_U1R_IMAG) % VREG_SLOTS;
                    if constexpr (B < RB) u1r<T2, RB, B, F, false>(x, c, 1, 0, 0);
                } else {
                    constexpr int F = H < H_U1L_REAL ? FORM_UNIT_IMAG : FORM_UNIT_REAL;
                    constexpr int B = (H - H_U1L_IMAG) % 6;
                    if constexpr (B < 5) u1l_bit<T2, RB, F, false, B>(x, c, lane, 1, 0, 0);
                    else u1l_pair<T2, RB, F, false, B>(x, c, lane, 1, 0, 0);
                }
            } else if constexpr (H < H_U1R_DIAG) {
                constexpr int B = (H - H_U1R_ROT) / 4, K = (H - H_U1R_ROT) % 4;
                if constexpr (B < RB) u1r<T2, RB, B, FORM_UNIT_ROT, false, K>(x, nullptr, 1, 0, 0);
            } else if constexpr (H < H_U1R_CTRL) {
                constexpr int F = H < H_U1R_ANTI ? FORM_UNIT_DIAG : H < H_U1R_FULL ? FORM_UNIT_ANTI : FORM_FULL;
                constexpr int B = (H - H_U1R_DIAG) % VREG_SLOTS;
                T2 c[4];
                inline_coefs(v, c);
                if constexpr (B < RB) u1r<T2, RB, B, F, false>(x, c, 1, 0, 0);
            } else if constexpr (H < H_U1L_ROT) {
                constexpr int B = H - H_U1R_CTRL;
                T2 c[4];
                inline_coefs(v, c);
                if constexpr (B < RB) u1r<T2, RB, B, FORM_FULL, true>(x, c, (tid & v.tmask) == v.tval, v.rmask, v.rval);
            } else if constexpr (H < H_U1L_DIAG) {
                constexpr int B = (H - H_U1L_ROT) / 4, K = (H - H_U1L_ROT) % 4;
                if constexpr (B < 5) u1l_bit_rot<T2, RB, B, K>(x, lane);
                else u1l_pair<T2, RB, FORM_UNIT_ROT, false, B, K>(x, nullptr, lane, 1, 0, 0);
            } else if constexpr (H < H_DIAG) {
                constexpr bool CTRL = H >= H_U1L_CTRL;
                constexpr int F = CTRL ? FORM_FULL
                                  : H < H_U1L_ANTI ? FORM_UNIT_DIAG : H < H_U1L_FULL ? FORM_UNIT_ANTI : FORM_FULL;
                constexpr int B = (H - H_U1L_DIAG) % 6;
                T2 c[4];
                inline_coefs(v, c);
                uint32_t tp = 1, rmask = 0, rval = 0;
                if constexpr (CTRL) {
                    tp = (tid & v.tmask) == v.tval;
                    rmask = v.rmask;
                    rval = v.rval;
                }
                if constexpr (B < 5) u1l_bit<T2, RB, F, CTRL, B>(x, c, lane, tp, rmask, rval);
                else u1l_pair<T2, RB, F, CTRL, B>(x, c, lane, tp, rmask, rval);
            } else if constexpr (H < H_SIGNS) {
                constexpr bool THREADED = ((H - H_DIAG) >> 1) != 0, SIGN = ((H - H_DIAG) & 1) != 0;
                diag<T2, RB, SIGN, THREADED>(x, v, cbase, tid, base);
            } else if constexpr (H == H_SIGNS) {
                signs<T2, RB>(x, v, cints, tid, base);
            } else if constexpr (H == H_PHASES) {
                phases<T2, RB>(x, v, ccoef, cints, tid, base);
            } else if constexpr (H == H_SCALE1) {
                const T2 f = ccoef[v.coef];
#pragma unroll
                for (int j = 0; j < R; j++) x[j] = cm(f, x[j]);
            } else if constexpr (H == H_SCALE) {
                T2 ft = cbase[v.coef + tid];
                const T2 *g = cbase + v.coef + W + R;
                for (uint32_t i = 0; i < v.k; i++) {
                    const uint32_t q = cints[v.aux + i];
                    ft = cm(g[2 * i + ((base >> q) & 1)], ft);
                }
#pragma unroll
                for (int j = 0; j < R; j++) x[j] = cm(cm(ccoef[v.coef + W + j], ft), x[j]);
            } else if constexpr (H == H_XCHG || H == H_LDSOP) {
                lds_barrier();
                // The thread part ti and the register part o of a tile index
                // hold disjoint tile bits, so ti | o = ti + o: one per-thread
                // base and a uniform offset per register (one add each).
                {
                    T2 *const tp = tile + ti;
#pragma unroll
                    for (int j = 0; j < R; j++) {
                        uint32_t o = 0;
#pragma unroll
                        for (int b = 0; b < RB; b++)
                            if ((j >> b) & 1) o |= 1u << rm[b];
                        tp[o] = x[j];
                    }
                }
                lds_barrier();
                if constexpr (H == H_LDSOP) {
                    const QVD_CONST StageOp &op = cops[v.aux];
                    const T2 *m = cbase + op.coef_off;
                    switch (op.k) {
                        case 2: lds_dense<T2, 2, W>(tile, m, op, L); break;
                        case 3: lds_dense<T2, 3, W>(tile, m, op, L); break;
                        case 4: lds_dense<T2, 4, W>(tile, m, op, L); break;
                        default: lds_dense<T2, 1, W>(tile, m, op, L); break;
                    }
                    lds_barrier();
                } else {
#pragma unroll
                    for (int b = 0; b < RB; b++) rm[b] = cbyte(v.rmap, b);
                    ti = thread_index(v.lmap, v.wmap);
                }
                const T2 *const tq = tile + ti;
#pragma unroll
                for (int j = 0; j < R; j++) {
                    uint32_t o = 0;
#pragma unroll
                    for (int b = 0; b < RB; b++)
                        if ((j >> b) & 1) o |= 1u << rm[b];
                    x[j] = tq[o];
                }
            }
        };
        static_assert(H_COUNT == 120, "the switch below lists every handler");
#define QVD_H(n) \
    case (n): handle(std::integral_constant<uint32_t, (n)>{}, v); break;
#define QVD_H4(n) QVD_H(n) QVD_H((n) + 1) QVD_H((n) + 2) QVD_H((n) + 3)
        for (const QVD_CONST VInstr *pv = ip; pv != ip_end; ++pv) {
            const QVD_CONST VInstr &v = *pv;
            const uint32_t w = v.op;
            switch ((w >> VH_SHIFT) & 0xffu) {
                QVD_H4(0) QVD_H4(4) QVD_H4(8) QVD_H4(12) QVD_H4(16) QVD_H4(20) QVD_H4(24) QVD_H4(28)
                QVD_H4(32) QVD_H4(36) QVD_H4(40) QVD_H4(44) QVD_H4(48) QVD_H4(52) QVD_H4(56) QVD_H4(60)
                QVD_H4(64) QVD_H4(68) QVD_H4(72) QVD_H4(76) QVD_H4(80) QVD_H4(84) QVD_H4(88) QVD_H4(92)
                QVD_H4(96) QVD_H4(100) QVD_H4(104) QVD_H4(108) QVD_H4(112) QVD_H4(116)
                default: break;
            }
        }
#undef QVD_H4
#undef QVD_H

        T2 *dst = s + base + t_offF;
#pragma unroll
        for (int j = 0; j < R; j++) {
            uint64_t o = 0;
#pragma unroll
            for (int b = 0; b < RB; b++)
                if ((j >> b) & 1) o += pF[b];
            dst[o] = x[j];
        }
        // The next tile's first exchange writes LDS only after a barrier, so
        // no barrier is needed here.
        if (zero) {
            if (tn < t_end) make_zero(x, tn);
        } else if constexpr (PREFETCH) {
#pragma unroll
            for (int j = 0; j < R; j++) x[j] = y[j];
        } else if (tn < t_end) {
            load(x, tile_base(tn));
        }
    }
}

// The complex64 geometry: 2^C64_WB waves of 2^C64_RB registers per thread
// (a tile of 2^(6 + C64_WB + C64_RB) = 2^14 amplitudes). 16 waves of 16
// registers (4 waves per SIMD, 110 VGPRs) beat 8 waves of 32 registers
// (2 waves per SIMD, 244 VGPRs) on every workload measured, although the
// latter decodes each instruction once for twice the amplitudes: 61.3 ms
// against 66.7 ms on the 30-qubit random circuit, 5.3 against 6.5 ms on
// QFT-30. QVD_C64_RB=5 builds the latter.
#ifndef QVD_C64_RB
#define QVD_C64_RB 4
#endif
constexpr int C64_RB = QVD_C64_RB, C64_WB = 8 - C64_RB;
// The complex128 geometry: a tile of 2^13 amplitudes, 2^C128_WB waves of
// 2^C128_RB registers. 16 waves of 8 registers (4 waves per SIMD) ran the
// 30-qubit random circuit and QFT-30 in the same time as 8 waves of 16
// (0.141 against 0.142 s, 12.3 against 12.0 ms): FP64 arithmetic, at half
// the FP32 rate, sets the pace either way. The default keeps 16 registers,
// with which the queue kernel needs no scratch.
#ifndef QVD_C128_RB
#define QVD_C128_RB 4
#endif
constexpr int C128_RB = QVD_C128_RB, C128_WB = 7 - C128_RB;

}  // namespace

template <typename T2, int WB, int RB, int FB, bool MAY_ZERO>
__global__ __launch_bounds__(64 << WB) void k_vstage(T2 *__restrict__ s, const VStageDesc *__restrict__ descs,
                                                     uint32_t ndesc, const uint32_t *__restrict__ block_start,
                                                     const VInstr *__restrict__ prog, const T2 *__restrict__ coefs,
                                                     const StageOp *__restrict__ ldsops,
                                                     const uint32_t *__restrict__ ints) {
    // The workgroup's stage: descriptor di owns workgroups
    // block_start[di] .. block_start[di + 1] - 1 (one descriptor: all).
    uint32_t di = 0, bl = blockIdx.x, nb = gridDim.x;
    if (ndesc > 1) {
        const QVD_CONST uint32_t *cbs = (const QVD_CONST uint32_t *)block_start;
        uint32_t lo = 0, hi = ndesc;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (cbs[mid] <= blockIdx.x) lo = mid;
            else hi = mid;
        }
        di = lo;
        bl = blockIdx.x - cbs[di];
        nb = cbs[di + 1] - cbs[di];
    }
    const QVD_CONST VStageDesc &d = ((const QVD_CONST VStageDesc *)descs)[di];
    stage_tiles<T2, WB, RB, FB, MAY_ZERO>(s + d.state_off, d, bl, d.num_tiles, nb, prog, coefs, ldsops, ints);
}

// Every stage of a batch of states in one launch. Items (runs of tiles of one
// stage) are taken in array order from a counter; an item's tiles read what
// every tile of its state's previous stage wrote, so the item waits until the
// state's counter of finished tiles has reached item.need. An item's
// dependencies all come earlier in the array and a workgroup only waits
// after it has started, so the earliest unfinished item taken always has its
// inputs and the queue drains with any number of resident workgroups.
//
// Tiles of one state are written and read on different XCDs, whose L2
// caches are not coherent with each other: every thread waits for its
// stores to reach L2 before the workgroup's count is released at agent scope
// (which writes L2 back), and a wait ends with an agent-scope acquire (which
// invalidates the stale lines) before any load of the item.
template <typename T2, int WB, int RB, int FB>
__global__ __launch_bounds__(64 << WB) void k_vstage_queue(T2 *__restrict__ s, const VStageDesc *__restrict__ descs,
                                                           const VQueueItem *__restrict__ items, uint32_t nitems,
                                                           uint32_t *__restrict__ sync,
                                                           const VInstr *__restrict__ prog,
                                                           const T2 *__restrict__ coefs,
                                                           const StageOp *__restrict__ ldsops,
                                                           const uint32_t *__restrict__ ints) {
    __shared__ uint32_t slot;
    uint32_t *const head = sync + VQUEUE_HEAD, *const error = sync + VQUEUE_ERROR;
    uint32_t *const done = sync + VQUEUE_COUNTERS;
    const QVD_CONST VQueueItem *citems = (const QVD_CONST VQueueItem *)items;
    // Wave 0 takes the items and waits. Every branch of this loop is wave-
    // uniform (scalar conditions; all 64 lanes of wave 0 run its atomics,
    // lane 0 adding the amount and the others 0): a branch on one lane
    // around the barriers lets the compiler rotate the loop so that the other
    // lanes of wave 0 pass the next barrier before lane 0 has taken an item.
    // Taking the next item before finishing this one keeps the queue
    // deadlock-free: the earliest unfinished item is always one whose
    // workgroup is running it or waiting only on finished items.
    const bool lead = __builtin_amdgcn_readfirstlane(threadIdx.x) == 0;
    const uint32_t first = (threadIdx.x & 63) == 0;
    // The first item is taken here, every later one while the previous
    // item's stores drain.
    uint32_t next = 0;
    if (lead)
        next = __builtin_amdgcn_readfirstlane(
            __hip_atomic_fetch_add(head, first, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    for (;;) {
        if (lead) {
            const uint32_t i = next;
            if (i < nitems && citems[i].need != 0) {
                const uint32_t st = citems[i].state, need = citems[i].need;
                const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
                for (;;) {
                    const uint32_t have = __builtin_amdgcn_readfirstlane(
                        __hip_atomic_load(done + st, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT));
                    if (have >= need) break;
                    if (__builtin_amdgcn_readfirstlane(
                            __hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0)
                        break;
                    if (__builtin_amdgcn_s_memrealtime() - t0 > VQUEUE_WAIT_TICKS) {
                        __hip_atomic_store(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(8);
                }
            }
            slot = i;
        }
        __syncthreads();
        const uint32_t i = __builtin_amdgcn_readfirstlane(slot);
        if (i >= nitems) break;
        const QVD_CONST VQueueItem &it = citems[i];
        const QVD_CONST VStageDesc &d = ((const QVD_CONST VStageDesc *)descs)[it.desc];
        stage_tiles<T2, WB, RB, -1, true>(s + d.state_off, d, it.tile0, it.tile0 + it.count * it.step, it.step, prog,
                                          coefs, ldsops, ints);
        uint32_t got = 0;
        if (lead) got = __hip_atomic_fetch_add(head, first, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(0);  // this thread's stores (and wave 0's fetch) have completed
        __syncthreads();
        if (lead) {
            __hip_atomic_fetch_add(done + it.state, first ? it.count : 0u, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
            next = __builtin_amdgcn_readfirstlane(got);
        }
    }
}

template <typename T2>
VGeom vstage_geom() {
    return sizeof(T2) == 8 ? VGeom{C64_WB, C64_RB, 4} : VGeom{C128_WB, C128_RB, 3};
}

template <typename T2>
size_t vstage_lds_bytes() {
    const VGeom g = vstage_geom<T2>();
    return (sizeof(T2) << g.L()) + VCONST_BYTES;
}

template <typename T2>
hipError_t launch_vstage(T2 *s, const VStageDesc *descs, uint32_t ndesc, const uint32_t *block_start,
                         const VInstr *prog, const T2 *coefs, const StageOp *ldsops, const uint32_t *ints, unsigned grid,
                         hipStream_t st, bool lazy_zero, uint32_t fixed_lanes) {
    const size_t lds = vstage_lds_bytes<T2>();
    using Kernel = void (*)(T2 *, const VStageDesc *, uint32_t, const uint32_t *, const VInstr *, const T2 *,
                            const StageOp *, const uint32_t *);
    // A stage that starts from |0...0> or names stale qubits has a kernel of
    // its own, so the others keep the code without those branches. The default geometry's
    // fixed lane bits are compiled in; any other count is read from the
    // descriptors.
    Kernel k, kz, kf, kfz;
    if constexpr (sizeof(T2) == 8) {
        k = k_vstage<T2, C64_WB, C64_RB, 4, false>;
        kz = k_vstage<T2, C64_WB, C64_RB, 4, true>;
        kf = k_vstage<T2, C64_WB, C64_RB, -1, false>;
        kfz = k_vstage<T2, C64_WB, C64_RB, -1, true>;
    } else {
        k = k_vstage<T2, C128_WB, C128_RB, 3, false>;
        kz = k_vstage<T2, C128_WB, C128_RB, 3, true>;
        kf = k_vstage<T2, C128_WB, C128_RB, -1, false>;
        kfz = k_vstage<T2, C128_WB, C128_RB, -1, true>;
    }
    // The LDS size attribute is set once per device, before any launch a
    // caller may capture into a graph.
    static std::atomic<uint64_t> attr_set{0};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const uint64_t bit = uint64_t(1) << (dev & 63);
    if (!(attr_set.load() & bit)) {
        e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute((const void *)kz, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute((const void *)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute((const void *)kfz, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr_set.fetch_or(bit);
    }
    if (fixed_lanes != 0 && fixed_lanes != vstage_geom<T2>().f_bits) {
        k = kf;
        kz = kfz;
    }
    if (lazy_zero) k = kz;
    k<<<grid, 64 << vstage_geom<T2>().w_bits, lds, st>>>(s, descs, ndesc, block_start, prog, coefs, ldsops, ints);
    return hipGetLastError();
}

namespace {

template <typename T2>
auto vstage_queue_kernel() {
    if constexpr (sizeof(T2) == 8) return k_vstage_queue<T2, C64_WB, C64_RB, 4>;
    else return k_vstage_queue<T2, C128_WB, C128_RB, 3>;
}

// Sets the queue kernel's dynamic LDS limit once per device (outside any
// graph capture: the first call comes from an eager run).
template <typename T2>
hipError_t queue_attributes() {
    static std::atomic<uint64_t> attr_set{0};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const uint64_t bit = uint64_t(1) << (dev & 63);
    if (!(attr_set.load() & bit)) {
        e = hipFuncSetAttribute((const void *)vstage_queue_kernel<T2>(), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)vstage_lds_bytes<T2>());
        if (e != hipSuccess) return e;
        attr_set.fetch_or(bit);
    }
    return hipSuccess;
}

}  // namespace

template <typename T2>
hipError_t vstage_resident_groups(unsigned *groups) {
    hipError_t e = queue_attributes<T2>();
    if (e != hipSuccess) return e;
    int dev = 0, per_cu = 0, cus = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)vstage_queue_kernel<T2>(),
                                                     64 << vstage_geom<T2>().w_bits, vstage_lds_bytes<T2>());
    if (e != hipSuccess) return e;
    *groups = (unsigned)std::max(1, per_cu) * (unsigned)std::max(1, cus);
    return hipSuccess;
}

template <typename T2>
hipError_t launch_vstage_queue(T2 *s, const VStageDesc *descs, const VQueueItem *items, uint32_t nitems,
                               uint32_t *sync, const VInstr *prog, const T2 *coefs, const StageOp *ldsops,
                               const uint32_t *ints, unsigned grid, hipStream_t st) {
    hipError_t e = queue_attributes<T2>();
    if (e != hipSuccess) return e;
    vstage_queue_kernel<T2>()<<<grid, 64 << vstage_geom<T2>().w_bits, vstage_lds_bytes<T2>(), st>>>(
        s, descs, items, nitems, sync, prog, coefs, ldsops, ints);
    return hipGetLastError();
}

template VGeom vstage_geom<float2>();
template VGeom vstage_geom<double2>();
template size_t vstage_lds_bytes<float2>();
template size_t vstage_lds_bytes<double2>();
template hipError_t launch_vstage<float2>(float2 *, const VStageDesc *, uint32_t, const uint32_t *, const VInstr *,
                                          const float2 *, const StageOp *, const uint32_t *, unsigned, hipStream_t,
                                          bool, uint32_t);
template hipError_t launch_vstage<double2>(double2 *, const VStageDesc *, uint32_t, const uint32_t *, const VInstr *,
                                           const double2 *, const StageOp *, const uint32_t *, unsigned, hipStream_t,
                                           bool, uint32_t);
template hipError_t launch_vstage_queue<float2>(float2 *, const VStageDesc *, const VQueueItem *, uint32_t, uint32_t *,
                                                const VInstr *, const float2 *, const StageOp *, const uint32_t *,
                                                unsigned, hipStream_t);
template hipError_t launch_vstage_queue<double2>(double2 *, const VStageDesc *, const VQueueItem *, uint32_t,
                                                 uint32_t *, const VInstr *, const double2 *, const StageOp *,
                                                 const uint32_t *, unsigned, hipStream_t);
template hipError_t vstage_resident_groups<float2>(unsigned *);
template hipError_t vstage_resident_groups<double2>(unsigned *);

}  // namespace qvd

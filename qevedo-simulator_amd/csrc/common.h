// Shared host-side definitions of the MI355X state-vector engine: status
// handling across the C ABI and the small f64 matrix type the planner works
// with.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qvd_hip.h"

namespace qvd {

// Errors thrown inside the library and turned into status codes at the ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] inline void fail(int code, const std::string &m) { throw Error(code, m); }
inline void require(bool ok, const std::string &m) {
    if (!ok) fail(QVD_EINVAL, m);
}

void set_last_error(const std::string &m);

inline void hip_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        fail(QVD_ENOMEM, std::string(what) + ": out of device memory");
    fail(QVD_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define QVD_HIP(call) ::qvd::hip_check((call), #call)

using C64 = std::complex<double>;

// Complex product without the C library's inf/nan recovery path.
inline C64 cmul(C64 a, C64 b) {
    return C64(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}
inline double nsq(C64 a) { return a.real() * a.real() + a.imag() * a.imag(); }

// Row-major 2^k x 2^k complex matrix; index bit j = qubit j of the op's
// target list (src/matrix.rs:1-18).
struct HMatrix {
    uint32_t k = 0;
    std::vector<C64> a;

    HMatrix() = default;
    explicit HMatrix(uint32_t k_) : k(k_), a(size_t(1) << (2 * k_), C64(0, 0)) {}
    size_t dim() const { return size_t(1) << k; }
    C64 &at(size_t r, size_t c) { return a[r * dim() + c]; }
    const C64 &at(size_t r, size_t c) const { return a[r * dim() + c]; }

    static HMatrix identity(uint32_t k) {
        HMatrix m(k);
        for (size_t i = 0; i < m.dim(); i++) m.at(i, i) = C64(1, 0);
        return m;
    }
    static HMatrix from_interleaved(const double *p, uint32_t k) {
        HMatrix m(k);
        for (size_t i = 0; i < m.a.size(); i++) m.a[i] = C64(p[2 * i], p[2 * i + 1]);
        return m;
    }
    // The exact test of src/matrix.rs:99-101.
    bool is_diagonal() const {
        for (size_t r = 0; r < dim(); r++)
            for (size_t c = 0; c < dim(); c++)
                if (r != c && nsq(at(r, c)) != 0.0) return false;
        return true;
    }
    // self * o, left entries with |a|^2 == 0 skipped (src/matrix.rs:108-124).
    HMatrix mul(const HMatrix &o) const {
        HMatrix out(k);
        const size_t d = dim();
        for (size_t r = 0; r < d; r++)
            for (size_t i = 0; i < d; i++) {
                C64 x = at(r, i);
                if (nsq(x) == 0.0) continue;
                for (size_t c = 0; c < d; c++) out.at(r, c) += cmul(x, o.at(i, c));
            }
        return out;
    }
    // The matrix on `qubits` (a superset of `gate_qubits`) acting as this
    // matrix on gate_qubits and identity elsewhere (src/matrix.rs:139-175).
    HMatrix embed(const std::vector<uint32_t> &gate_qubits, const std::vector<uint32_t> &qubits) const {
        std::vector<uint32_t> pos;
        for (uint32_t q : gate_qubits) {
            uint32_t p = 0;
            while (p < qubits.size() && qubits[p] != q) p++;
            require(p < qubits.size(), "qubit missing from embedding");
            pos.push_back(p);
        }
        size_t gmask = 0;
        for (uint32_t p : pos) gmask |= size_t(1) << p;
        HMatrix out((uint32_t)qubits.size());
        auto local = [&](size_t idx) {
            size_t v = 0;
            for (size_t j = 0; j < pos.size(); j++) v |= ((idx >> pos[j]) & 1) << j;
            return v;
        };
        for (size_t r = 0; r < out.dim(); r++)
            for (size_t c = 0; c < out.dim(); c++)
                if ((r & ~gmask) == (c & ~gmask)) out.at(r, c) = at(local(r), local(c));
        return out;
    }
    // Controls prepended on the low index bits, applied in `state`
    // (src/matrix.rs:180-195).
    HMatrix controlled(uint32_t controls, size_t state) const {
        HMatrix out = identity(controls + k);
        const size_t mask = (size_t(1) << controls) - 1;
        for (size_t r = 0; r < out.dim(); r++)
            for (size_t c = 0; c < out.dim(); c++) {
                bool rr = (r & mask) == state, cc = (c & mask) == state;
                if (rr && cc) out.at(r, c) = at(r >> controls, c >> controls);
                else if (rr || cc) out.at(r, c) = C64(0, 0);
            }
        return out;
    }
};

// A unitary op on logical qubits (src/fusion.rs:16-56).
struct Op {
    std::vector<uint32_t> targets;
    std::vector<std::pair<uint32_t, bool>> controls;
    HMatrix m;
    bool diagonal = false;  // cached m.is_diagonal()

    std::vector<uint32_t> qubits() const {
        std::vector<uint32_t> q = targets;
        for (auto &c : controls) q.push_back(c.first);
        return q;
    }
    size_t width() const { return targets.size() + controls.size(); }
    // The op as a plain matrix on controls followed by targets.
    void full(std::vector<uint32_t> &qs, HMatrix &mat) const {
        if (controls.empty()) {
            qs = targets;
            mat = m;
            return;
        }
        size_t state = 0;
        for (size_t i = 0; i < controls.size(); i++) state |= (size_t)controls[i].second << i;
        qs.clear();
        for (auto &c : controls) qs.push_back(c.first);
        for (uint32_t q : targets) qs.push_back(q);
        mat = m.controlled((uint32_t)controls.size(), state);
    }
};

inline bool contains(const std::vector<uint32_t> &v, uint32_t x) {
    for (uint32_t y : v)
        if (y == x) return true;
    return false;
}

}  // namespace qvd

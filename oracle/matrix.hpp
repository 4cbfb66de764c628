// ORACLE — test infrastructure only. Nothing in the product links or calls this
// file; tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it
// as the checker.
//
// Small dense complex matrices and the standard gate table, in double
// precision, restating the reference's conventions:
//   - row-major 2^k x 2^k, index bit j = qubit qubits[j]   (src/matrix.rs:1-18)
//   - exact-zero diagonal test                              (src/matrix.rs:99-101)
//   - product skipping zero left entries                    (src/matrix.rs:108-124)
//   - embed / controlled with controls on the low bits      (src/matrix.rs:139-195)
//   - Qiskit gate matrices                                  (src/circuit.rs:133-202)
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

using C64 = std::complex<double>;

// Complex product written out the way num-complex 0.4 does it
// (re = a.re*b.re - a.im*b.im, im = a.re*b.im + a.im*b.re), without the
// inf/nan recovery of the C library's complex multiply.
inline C64 cmul(C64 a, C64 b) {
    return C64(a.real() * b.real() - a.imag() * b.imag(),
               a.real() * b.imag() + a.imag() * b.real());
}
inline double norm_sqr(C64 a) { return a.real() * a.real() + a.imag() * a.imag(); }

struct Matrix {
    size_t dim = 0;
    std::vector<C64> data;

    Matrix() = default;
    explicit Matrix(size_t d) : dim(d), data(d * d, C64(0.0, 0.0)) {}
    static Matrix from(std::vector<C64> entries) {
        size_t d = (size_t)std::llround(std::sqrt((double)entries.size()));
        if (d * d != entries.size() || (d & (d - 1)) != 0)
            throw std::invalid_argument("not a 2^k x 2^k matrix");
        Matrix m;
        m.dim = d;
        m.data = std::move(entries);
        return m;
    }
    C64 &at(size_t r, size_t c) { return data[r * dim + c]; }
    const C64 &at(size_t r, size_t c) const { return data[r * dim + c]; }
    unsigned qubits() const {
        unsigned k = 0;
        while ((size_t(1) << k) < dim) k++;
        return k;
    }

    static Matrix identity(unsigned k) {
        Matrix m(size_t(1) << k);
        for (size_t i = 0; i < m.dim; i++) m.at(i, i) = C64(1.0, 0.0);
        return m;
    }
    static Matrix diagonal(const std::vector<C64> &entries) {
        Matrix m(entries.size());
        for (size_t i = 0; i < entries.size(); i++) m.at(i, i) = entries[i];
        return m;
    }
    // matrix.rs:99-101 — exact test, every off-diagonal |entry|^2 == 0.
    bool is_diagonal() const {
        for (size_t r = 0; r < dim; r++)
            for (size_t c = 0; c < dim; c++)
                if (r != c && norm_sqr(at(r, c)) != 0.0) return false;
        return true;
    }
    std::vector<C64> diagonal_entries() const {
        std::vector<C64> d(dim);
        for (size_t i = 0; i < dim; i++) d[i] = at(i, i);
        return d;
    }
    // matrix.rs:108-124 — self * other, left entries with |a|^2 == 0 skipped.
    Matrix mul(const Matrix &o) const {
        Matrix out(dim);
        for (size_t r = 0; r < dim; r++)
            for (size_t k = 0; k < dim; k++) {
                C64 a = at(r, k);
                if (norm_sqr(a) == 0.0) continue;
                for (size_t c = 0; c < dim; c++) out.at(r, c) += cmul(a, o.at(k, c));
            }
        return out;
    }
    Matrix adjoint() const {
        Matrix out(dim);
        for (size_t r = 0; r < dim; r++)
            for (size_t c = 0; c < dim; c++) out.at(c, r) = std::conj(at(r, c));
        return out;
    }
    // matrix.rs:139-175 — embed a gate on gate_qubits into the list qubits.
    Matrix embed(const std::vector<size_t> &gate_qubits, const std::vector<size_t> &qubits) const {
        if ((size_t(1) << gate_qubits.size()) != dim)
            throw std::invalid_argument("gate matrix does not match its qubits");
        std::vector<size_t> pos;
        for (size_t q : gate_qubits) {
            size_t p = 0;
            while (p < qubits.size() && qubits[p] != q) p++;
            if (p == qubits.size()) throw std::invalid_argument("qubit missing from embedding");
            pos.push_back(p);
        }
        size_t gate_mask = 0;
        for (size_t p : pos) gate_mask |= size_t(1) << p;
        size_t big = size_t(1) << qubits.size();
        auto local = [&](size_t index) {
            size_t v = 0;
            for (size_t j = 0; j < pos.size(); j++) v |= ((index >> pos[j]) & 1) << j;
            return v;
        };
        Matrix out(big);
        for (size_t r = 0; r < big; r++)
            for (size_t c = 0; c < big; c++)
                if ((r & ~gate_mask) == (c & ~gate_mask)) out.at(r, c) = at(local(r), local(c));
        return out;
    }
    // matrix.rs:180-195 — controls prepended on the low index bits.
    Matrix controlled(unsigned controls, size_t control_state) const {
        unsigned k = qubits();
        Matrix out = identity(controls + k);
        size_t mask = (size_t(1) << controls) - 1;
        for (size_t r = 0; r < out.dim; r++)
            for (size_t c = 0; c < out.dim; c++) {
                bool rr = (r & mask) == control_state, cc = (c & mask) == control_state;
                if (rr && cc)
                    out.at(r, c) = at(r >> controls, c >> controls);
                else if (rr || cc)
                    out.at(r, c) = C64(0.0, 0.0);
            }
        return out;
    }
    double distance(const Matrix &o) const {
        double d = 0;
        for (size_t i = 0; i < data.size(); i++) d = std::max(d, std::abs(data[i] - o.data[i]));
        return d;
    }
};

// ---------------------------------------------------------------------------
// Gate table (circuit.rs:14-202). Gates are named by their OpenQASM 2
// qelib1 names (circuit.rs:56-88) plus the few stdgates aliases the
// reference's QASM reader accepts for them.
// ---------------------------------------------------------------------------
inline C64 phase(double a) { return std::polar(1.0, a); }

// The matrix of an uncontrolled base gate (circuit.rs:143-200).
inline Matrix base_matrix(const std::string &n, const std::vector<double> &p) {
    const double h = 1.0 / std::sqrt(2.0);
    const C64 z(0, 0), one(1, 0);
    auto need = [&](size_t k) {
        if (p.size() != k) throw std::invalid_argument("gate " + n + " takes " + std::to_string(k) + " parameters");
    };
    if (n == "id" || n == "i") { need(0); return Matrix::identity(1); }
    if (n == "h") { need(0); return Matrix::from({C64(h, 0), C64(h, 0), C64(h, 0), C64(-h, 0)}); }
    if (n == "x") { need(0); return Matrix::from({z, one, one, z}); }
    if (n == "y") { need(0); return Matrix::from({z, C64(0, -1), C64(0, 1), z}); }
    if (n == "z") { need(0); return Matrix::diagonal({one, C64(-1, 0)}); }
    if (n == "s") { need(0); return Matrix::diagonal({one, C64(0, 1)}); }
    if (n == "sdg") { need(0); return Matrix::diagonal({one, C64(0, -1)}); }
    if (n == "t") { need(0); return Matrix::diagonal({one, phase(M_PI / 4)}); }
    if (n == "tdg") { need(0); return Matrix::diagonal({one, phase(-M_PI / 4)}); }
    if (n == "sx") { need(0); return Matrix::from({C64(.5, .5), C64(.5, -.5), C64(.5, -.5), C64(.5, .5)}); }
    if (n == "sxdg") { need(0); return Matrix::from({C64(.5, -.5), C64(.5, .5), C64(.5, .5), C64(.5, -.5)}); }
    if (n == "rx") {
        need(1);
        double s = std::sin(p[0] / 2), c = std::cos(p[0] / 2);
        return Matrix::from({C64(c, 0), C64(0, -s), C64(0, -s), C64(c, 0)});
    }
    if (n == "ry") {
        need(1);
        double s = std::sin(p[0] / 2), c = std::cos(p[0] / 2);
        return Matrix::from({C64(c, 0), C64(-s, 0), C64(s, 0), C64(c, 0)});
    }
    if (n == "rz") { need(1); return Matrix::diagonal({phase(-p[0] / 2), phase(p[0] / 2)}); }
    if (n == "p" || n == "u1" || n == "phase") { need(1); return Matrix::diagonal({one, phase(p[0])}); }
    if (n == "u" || n == "u3" || n == "U") {
        need(3);
        double s = std::sin(p[0] / 2), c = std::cos(p[0] / 2);
        return Matrix::from({C64(c, 0), -phase(p[2]) * s, phase(p[1]) * s, phase(p[1] + p[2]) * c});
    }
    // qasm.rs:1394-1425: u0 (identity, one ignored parameter), u2 = U(pi/2, .),
    // and the base of cu: e^(i gamma) U(theta, phi, lambda)
    if (n == "u0") { need(1); return Matrix::identity(1); }
    if (n == "u2") { need(2); return base_matrix("u", {M_PI / 2.0, p[0], p[1]}); }
    if (n == "u_gamma") {
        need(4);
        Matrix m = base_matrix("u", {p[0], p[1], p[2]});
        for (C64 &x : m.data) x = cmul(x, phase(p[3]));
        return m;
    }
    if (n == "swap") {
        need(0);
        return Matrix::from({one, z, z, z, z, z, one, z, z, one, z, z, z, z, z, one});
    }
    if (n == "rzz") {
        need(1);
        C64 a = phase(-p[0] / 2), b = phase(p[0] / 2);
        return Matrix::diagonal({a, b, b, a});
    }
    if (n == "rxx") {
        need(1);
        double s = std::sin(p[0] / 2), c = std::cos(p[0] / 2);
        C64 cc(c, 0), ms(0, -s);
        return Matrix::from({cc, z, z, ms, z, cc, ms, z, z, ms, cc, z, ms, z, z, cc});
    }
    throw std::invalid_argument("unknown gate `" + n + "`");
}

// controlled_form (circuit.rs:115-130): (leading controls, base gate name).
inline bool controlled_form(const std::string &n, unsigned &controls, std::string &base) {
    static const struct { const char *name; unsigned c; const char *base; } table[] = {
        {"cx", 1, "x"},   {"cy", 1, "y"},   {"cz", 1, "z"},     {"ch", 1, "h"},
        {"cp", 1, "p"},   {"cu1", 1, "p"},  {"crx", 1, "rx"},   {"cry", 1, "ry"},
        {"crz", 1, "rz"}, {"ccx", 2, "x"},  {"cswap", 1, "swap"},
        // aliases and the controlled gates of the reference's QASM reader
        // (qasm.rs:1394-1435), controls first
        {"CX", 1, "x"},   {"cnot", 1, "x"}, {"toffoli", 2, "x"}, {"fredkin", 1, "swap"},
        {"cphase", 1, "p"}, {"csx", 1, "sx"}, {"cu3", 1, "u3"}, {"cu", 1, "u_gamma"},
        {"c3x", 3, "x"},  {"c4x", 4, "x"},  {"c3sqrtx", 3, "sx"},
    };
    for (auto &e : table)
        if (n == e.name) { controls = e.c; base = e.base; return true; }
    return false;
}

// Gate::matrix (circuit.rs:133-202): controlled gates are their base gate
// with the controls prepended in state "all ones".
inline Matrix gate_matrix(const std::string &n, const std::vector<double> &p) {
    unsigned c; std::string b;
    if (controlled_form(n, c, b)) return base_matrix(b, p).controlled(c, (size_t(1) << c) - 1);
    return base_matrix(n, p);
}

}  // namespace orc

// The standard gate set in double precision (src/circuit.rs:14-202), named
// by OpenQASM 2 qelib1 / stdgates names, and its lowering to ops with the
// controls kept separate (gate_op, src/simulator.rs:106-119).
#include <cstring>
#include <string>

#include "common.h"

using namespace qvd;

namespace {

C64 polar1(double a) { return C64(std::cos(a), std::sin(a)); }

HMatrix m2(C64 a, C64 b, C64 c, C64 d) {
    HMatrix m(1);
    m.at(0, 0) = a;
    m.at(0, 1) = b;
    m.at(1, 0) = c;
    m.at(1, 1) = d;
    return m;
}

HMatrix diag(std::initializer_list<C64> e) {
    uint32_t k = 0;
    while ((size_t(1) << k) < e.size()) k++;
    HMatrix m(k);
    size_t i = 0;
    for (C64 v : e) {
        m.at(i, i) = v;
        i++;
    }
    return m;
}

struct Lowered {
    uint32_t controls = 0;
    HMatrix base;
};

// (number of leading controls, base gate) for every name; throws on an
// unknown name or a wrong parameter count.
Lowered lower(const std::string &name, const double *p, uint32_t np) {
    auto need = [&](uint32_t k) {
        require(np == k, "gate `" + name + "` takes " + std::to_string(k) + " parameter" + (k == 1 ? "" : "s"));
    };
    const double r = 1.0 / std::sqrt(2.0);
    const C64 z(0, 0), one(1, 0);
    auto base1 = [&](const std::string &n, const double *q) -> HMatrix {
        if (n == "id" || n == "i") return HMatrix::identity(1);
        if (n == "h") return m2(C64(r, 0), C64(r, 0), C64(r, 0), C64(-r, 0));
        if (n == "x") return m2(z, one, one, z);
        if (n == "y") return m2(z, C64(0, -1), C64(0, 1), z);
        if (n == "z") return diag({one, C64(-1, 0)});
        if (n == "s") return diag({one, C64(0, 1)});
        if (n == "sdg") return diag({one, C64(0, -1)});
        if (n == "t") return diag({one, polar1(M_PI / 4)});
        if (n == "tdg") return diag({one, polar1(-M_PI / 4)});
        if (n == "sx") return m2(C64(.5, .5), C64(.5, -.5), C64(.5, -.5), C64(.5, .5));
        if (n == "sxdg") return m2(C64(.5, -.5), C64(.5, .5), C64(.5, .5), C64(.5, -.5));
        if (n == "rx") {
            const double s = std::sin(q[0] / 2), c = std::cos(q[0] / 2);
            return m2(C64(c, 0), C64(0, -s), C64(0, -s), C64(c, 0));
        }
        if (n == "ry") {
            const double s = std::sin(q[0] / 2), c = std::cos(q[0] / 2);
            return m2(C64(c, 0), C64(-s, 0), C64(s, 0), C64(c, 0));
        }
        if (n == "rz") return diag({polar1(-q[0] / 2), polar1(q[0] / 2)});
        if (n == "p") return diag({one, polar1(q[0])});
        auto u3 = [&](double theta, double phi, double lam) {
            const double s = std::sin(theta / 2), c = std::cos(theta / 2);
            const C64 l = polar1(lam), f = polar1(phi), fl = polar1(phi + lam);
            return m2(C64(c, 0), C64(-l.real() * s, -l.imag() * s), C64(f.real() * s, f.imag() * s),
                      C64(fl.real() * c, fl.imag() * c));
        };
        if (n == "u") return u3(q[0], q[1], q[2]);
        // u2(phi, lambda) = U(pi/2, phi, lambda) (src/qasm.rs:1405)
        if (n == "u2") return u3(M_PI / 2.0, q[0], q[1]);
        // the base of cu(theta, phi, lambda, gamma): e^(i gamma) U(theta, phi, lambda)
        // (src/qasm.rs:1420-1425)
        if (n == "cu") {
            HMatrix m = u3(q[0], q[1], q[2]);
            const C64 g = polar1(q[3]);
            for (C64 &x : m.a) x = cmul(x, g);
            return m;
        }
        if (n == "swap") {
            HMatrix m(2);
            m.at(0, 0) = m.at(3, 3) = one;
            m.at(1, 2) = m.at(2, 1) = one;
            return m;
        }
        if (n == "rzz") {
            const C64 a = polar1(-q[0] / 2), b = polar1(q[0] / 2);
            return diag({a, b, b, a});
        }
        if (n == "rxx") {
            const double s = std::sin(q[0] / 2), c = std::cos(q[0] / 2);
            HMatrix m(2);
            for (int i = 0; i < 4; i++) m.at(i, i) = C64(c, 0);
            m.at(0, 3) = m.at(1, 2) = m.at(2, 1) = m.at(3, 0) = C64(0, -s);
            return m;
        }
        fail(QVD_EINVAL, "unknown gate `" + name + "`");
    };
    struct Entry {
        const char *name;
        uint32_t controls;
        const char *base;
        uint32_t params;
    };
    static const Entry table[] = {
        {"id", 0, "id", 0},   {"i", 0, "id", 0},     {"h", 0, "h", 0},       {"x", 0, "x", 0},
        {"y", 0, "y", 0},     {"z", 0, "z", 0},      {"s", 0, "s", 0},       {"sdg", 0, "sdg", 0},
        {"t", 0, "t", 0},     {"tdg", 0, "tdg", 0},  {"sx", 0, "sx", 0},     {"sxdg", 0, "sxdg", 0},
        {"rx", 0, "rx", 1},   {"ry", 0, "ry", 1},    {"rz", 0, "rz", 1},     {"p", 0, "p", 1},
        {"u1", 0, "p", 1},    {"phase", 0, "p", 1},  {"u", 0, "u", 3},       {"u3", 0, "u", 3},
        {"swap", 0, "swap", 0}, {"rzz", 0, "rzz", 1}, {"rxx", 0, "rxx", 1},
        {"cx", 1, "x", 0},    {"cnot", 1, "x", 0},   {"cy", 1, "y", 0},      {"cz", 1, "z", 0},
        {"ch", 1, "h", 0},    {"cp", 1, "p", 1},     {"cu1", 1, "p", 1},     {"cphase", 1, "p", 1},
        {"crx", 1, "rx", 1},  {"cry", 1, "ry", 1},   {"crz", 1, "rz", 1},    {"ccx", 2, "x", 0},
        {"toffoli", 2, "x", 0}, {"cswap", 1, "swap", 0}, {"fredkin", 1, "swap", 0},
        // the OpenQASM primitives and the qelib1 / stdgates names the
        // reference's reader maps (src/qasm.rs:1394-1435): u0 is the
        // identity (its parameter is a duration), the rest are their base
        // gates with controls first
        {"U", 0, "u", 3},     {"CX", 1, "x", 0},     {"u0", 0, "id", 1},     {"u2", 0, "u2", 2},
        {"csx", 1, "sx", 0},  {"cu3", 1, "u", 3},    {"cu", 1, "cu", 4},     {"c3x", 3, "x", 0},
        {"c4x", 4, "x", 0},   {"c3sqrtx", 3, "sx", 0},
    };
    for (const Entry &e : table)
        if (name == e.name) {
            need(e.params);
            Lowered l;
            l.controls = e.controls;
            l.base = base1(e.base, p);
            return l;
        }
    fail(QVD_EINVAL, "unknown gate `" + name + "`");
}

template <typename Fn>
int guard_call(Fn &&fn) {
    try {
        fn();
        return QVD_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return QVD_EINVAL;
    }
}

}  // namespace

extern "C" {

int qvd_gate_lower(const char *name, const double *params, uint32_t num_params, uint32_t *num_controls,
                   uint32_t *num_targets, double *matrix, uint32_t capacity) {
    return guard_call([&] {
        require(name && num_controls && num_targets, "null argument");
        require(num_params == 0 || params, "null parameters");
        Lowered l = lower(name, params, num_params);
        *num_controls = l.controls;
        *num_targets = l.base.k;
        if (matrix) {
            require(capacity >= l.base.a.size(), "matrix output too small");
            for (size_t i = 0; i < l.base.a.size(); i++) {
                matrix[2 * i] = l.base.a[i].real();
                matrix[2 * i + 1] = l.base.a[i].imag();
            }
        }
    });
}

int qvd_gate_matrix(const char *name, const double *params, uint32_t num_params, double *matrix, uint32_t capacity,
                    uint32_t *num_qubits) {
    return guard_call([&] {
        require(name && num_qubits, "null argument");
        require(num_params == 0 || params, "null parameters");
        Lowered l = lower(name, params, num_params);
        HMatrix full = l.controls ? l.base.controlled(l.controls, (size_t(1) << l.controls) - 1) : l.base;
        *num_qubits = full.k;
        if (matrix) {
            require(capacity >= full.a.size(), "matrix output too small");
            for (size_t i = 0; i < full.a.size(); i++) {
                matrix[2 * i] = full.a[i].real();
                matrix[2 * i + 1] = full.a[i].imag();
            }
        }
    });
}

}  // extern "C"

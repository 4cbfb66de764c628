// Register-tile stage programs (see vprog.h).
//
// The program generator orders a stage's ops by their true dependencies
// (the dependency-aware extraction of src/blocking.rs:157-179, applied to
// the lane and register bits of a tile instead of the local qubits of a
// region), inserts LDS exchanges when a gate needs a wave bit, merges runs
// of 1-qubit gates on one qubit, and keeps 1-qubit gates in unit form with
// their diagonal factors carried forward.
#include "vprog.h"

#include <algorithm>
#include <array>
#include <cmath>

namespace qvd {

namespace {

enum Role : uint8_t { ROLE_LANE, ROLE_WAVE, ROLE_REG, ROLE_NONE };

// Which tile bit each flexible lane / wave / register slot holds.
struct Roles {
    uint32_t F = 0;
    std::vector<uint8_t> lmap, wmap, rmap;
    std::vector<Role> role;     // tile bit -> role
    std::vector<uint8_t> slot;  // tile bit -> lane bit (lanes), wave slot, register slot

    void assign(uint32_t L, uint32_t f, const std::vector<uint8_t> &l, const std::vector<uint8_t> &w,
                const std::vector<uint8_t> &r) {
        F = f;
        lmap = l;
        wmap = w;
        rmap = r;
        role.assign(L, ROLE_NONE);
        slot.assign(L, 0);
        for (uint32_t b = 0; b < F; b++) {
            role[b] = ROLE_LANE;
            slot[b] = (uint8_t)b;
        }
        for (size_t s = 0; s < l.size(); s++) {
            role[l[s]] = ROLE_LANE;
            slot[l[s]] = (uint8_t)(F + s);
        }
        for (size_t s = 0; s < w.size(); s++) {
            role[w[s]] = ROLE_WAVE;
            slot[w[s]] = (uint8_t)s;
        }
        for (size_t s = 0; s < r.size(); s++) {
            role[r[s]] = ROLE_REG;
            slot[r[s]] = (uint8_t)s;
        }
    }
};

using M2 = std::array<C64, 4>;  // row-major 2x2

M2 mul2(const M2 &a, const M2 &b) {
    return {cmul(a[0], b[0]) + cmul(a[1], b[2]), cmul(a[0], b[1]) + cmul(a[1], b[3]),
            cmul(a[2], b[0]) + cmul(a[3], b[2]), cmul(a[2], b[1]) + cmul(a[3], b[3])};
}

bool has(const std::vector<uint8_t> &v, uint32_t x) { return std::find(v.begin(), v.end(), (uint8_t)x) != v.end(); }

// k when c is i^k to within 1e-14 (the rounding of sin and cos of pi/4 makes
// RX(pi/2) and RY(pi/2) miss the exact values by an ulp), else -1.
int quarter_turn(C64 c) {
    static const C64 rot[4] = {C64(1, 0), C64(0, 1), C64(-1, 0), C64(0, -1)};
    for (int k = 0; k < 4; k++)
        if (std::abs(c - rot[k]) < 1e-14) return k;
    return -1;
}

// Keeps each bit in its old slot of the same role where it can, so only the
// bits that move change slots.
std::vector<uint8_t> stabilize(const std::vector<uint8_t> &want, const std::vector<uint8_t> &old) {
    std::vector<uint8_t> out(want.size(), 0xff), rest;
    for (uint8_t b : want) {
        auto it = std::find(old.begin(), old.end(), b);
        if (it != old.end() && (size_t)(it - old.begin()) < out.size()) out[it - old.begin()] = b;
        else rest.push_back(b);
    }
    size_t k = 0;
    for (auto &s : out)
        if (s == 0xff) s = rest[k++];
    return out;
}

struct Gen {
    const Stage &st;
    uint32_t n;
    VGeom g;
    PendingDiag &pend;
    std::vector<C64> &coefs;
    std::vector<StageOp> &ldsops;
    std::vector<uint32_t> &ints;
    std::vector<int> tbit;  // physical -> tile bit (-1 if not in the tile)
    Roles roles;
    std::vector<VInstr> out;
    // Pending non-diagonal 1-qubit product per physical qubit (applied
    // after the pending diagonal).
    std::vector<M2> pmat;
    std::vector<char> has_pmat;
    // Pending run of commuting pair phases: the amplitude gains w when both
    // qubits p and q are 1. A run of w = -1 only is a sign run.
    struct PairPhase {
        uint32_t p, q;
        C64 w;
    };
    std::vector<PairPhase> prun;
    std::vector<char> in_run;

    Gen(const Stage &s, uint32_t n_, const VGeom &g_, PendingDiag &p, std::vector<C64> &c, std::vector<StageOp> &l,
        std::vector<uint32_t> &ii)
        : st(s), n(n_), g(g_), pend(p), coefs(c), ldsops(l), ints(ii), tbit(n_, -1), pmat(n_), has_pmat(n_, 0),
          in_run(n_, 0) {
        for (size_t j = 0; j < s.local.size(); j++) tbit[s.local[j]] = (int)j;
    }

    // Emits the pending run: as one V_SIGNS when every pair phase is -1
    // (pairs of two register bits become a per-register mask, the others go
    // to the int array as encoded bit pairs), else as one V_PHASES (pairs of
    // two register bits become a per-register factor table; the others go to
    // the int array as (encoded bit, encoded bit, coefficient index), a
    // register bit first when there is one).
    void flush_run() {
        if (prun.empty()) return;
        bool sign = true;
        for (auto &pp : prun)
            if (pp.w != C64(-1, 0)) sign = false;
        const uint32_t R = 1u << g.r_bits;
        VInstr v = blank(sign ? V_SIGNS : V_PHASES);
        v.aux = (uint32_t)ints.size();
        uint64_t rr = 0;
        std::vector<C64> rtable(R, C64(1, 0));
        bool has_rr = false;
        uint32_t tr_mask[VMAX_R_BITS] = {}, tt_mask[10] = {};
        bool has_tr = false, has_tt = false;
        for (auto &pp : prun) {
            uint32_t a = encode(pp.p), b = encode(pp.q);
            if (sign && (a & VT_KIND) != VT_GLOBAL && (b & VT_KIND) != VT_GLOBAL &&
                ((a & VT_KIND) == VT_THREAD || (b & VT_KIND) == VT_THREAD)) {
                // Sign pairs without a non-tile bit become bit masks (the sum
                // is mod 2, so a repeated pair cancels): a thread bit t and a
                // register bit r flip register r's linear term where thread
                // bit t is set (tr_mask[r]); two thread bits t < u add
                // x_t x_u to the thread's parity (tt_mask[t]).
                if ((a & VT_KIND) == VT_REG) std::swap(a, b);
                const uint32_t xa = a & 0xffffu, xb = b & 0xffffu;
                if ((b & VT_KIND) == VT_REG) {
                    tr_mask[xb] ^= 1u << xa;
                    has_tr = true;
                } else {
                    tt_mask[std::min(xa, xb)] ^= 1u << std::max(xa, xb);
                    has_tt = true;
                }
                continue;
            }
            if ((a & VT_KIND) == VT_REG && (b & VT_KIND) == VT_REG) {
                const uint32_t ra = a & 0xffffu, rb = b & 0xffffu;
                for (uint32_t j = 0; j < R; j++)
                    if (((j >> ra) & 1) && ((j >> rb) & 1)) {
                        rr ^= uint64_t(1) << j;
                        rtable[j] *= pp.w;
                    }
                has_rr = true;
                continue;
            }
            if ((b & VT_KIND) == VT_REG) std::swap(a, b);
            ints.push_back(a);
            ints.push_back(b);
            if (!sign) {
                ints.push_back((uint32_t)coefs.size());
                coefs.push_back(pp.w);
            }
            v.k++;
        }
        if (sign) {
            v.neg = rr;
            require(6 + g.w_bits <= 10 && g.r_bits <= VSIGN_TT && VSIGN_TT + 10 <= 16,
                    "internal: sign masks cover 10 thread and 6 register bits");
            v.form = (has_tr ? 1u : 0u) | (has_tt ? 2u : 0u);
            for (uint32_t r = 0; r < g.r_bits; r++) v.inl[r] = tr_mask[r];
            for (uint32_t t = 0; t < 10; t++) v.inl[VSIGN_TT + t] = tt_mask[t];
        } else if (has_rr) {
            v.form = 1;
            v.coef = (uint32_t)coefs.size();
            for (const C64 &c : rtable) coefs.push_back(c);
        }
        out.push_back(v);
        for (auto &pp : prun) in_run[pp.p] = in_run[pp.q] = 0;
        prun.clear();
    }

    bool touches_run(const Op &op) const {
        for (uint32_t q : op.qubits())
            if (in_run[q]) return true;
        return false;
    }

    // A diagonal op on at most two qubits (targets and controls together):
    // its entry for bits (x_a, x_b) factors as
    //   d00 * (d10 / d00)^x_a * (d01 / d00)^x_b * w^(x_a x_b),
    //   w = d11 d00 / (d10 d01),
    // so the constant and linear factors join the pending diagonals and w
    // joins the run of pair phases. Entries that are all +-1 keep exact signs.
    bool absorb_pair_diag(const Op &op) {
        const std::vector<uint32_t> qs = op.qubits();  // targets, then controls
        if (qs.size() > 2) return false;
        const size_t kt = op.targets.size();
        C64 e[4] = {C64(1, 0), C64(1, 0), C64(1, 0), C64(1, 0)};
        bool signs = true;
        for (uint32_t x = 0; x < (1u << qs.size()); x++) {
            bool ctrl = true;
            for (size_t c = 0; c < op.controls.size(); c++)
                if ((((x >> (kt + c)) & 1) != 0) != op.controls[c].second) ctrl = false;
            e[x] = ctrl ? op.m.at(x & ((1u << kt) - 1), x & ((1u << kt) - 1)) : C64(1, 0);
            if (e[x] != C64(1, 0) && e[x] != C64(-1, 0)) signs = false;
            if (e[x] == C64(0, 0)) return false;  // not unitary: leave it to V_DIAG
        }
        for (uint32_t q : qs) flush_pmat(q);
        const uint32_t a = qs[0];
        if (signs) {
            // (-1)^(c0 + c1 x_a + c2 x_b + c3 x_a x_b), kept exact
            int s[4];
            for (int x = 0; x < 4; x++) s[x] = e[x] == C64(-1, 0);
            const int c0 = s[0], c1 = s[1] ^ s[0];
            auto zflip = [&](uint32_t q) { pend.d1[q] = -pend.d1[q]; };
            if (c0) {
                pend.d0[a] = -pend.d0[a];
                pend.d1[a] = -pend.d1[a];
            }
            if (c1) zflip(a);
            if (qs.size() == 2) {
                const uint32_t b = qs[1];
                const int c2 = s[2] ^ s[0], c3 = s[3] ^ s[2] ^ s[1] ^ s[0];
                if (c2) zflip(b);
                if (c3) add_pair(a, b, C64(-1, 0));
            }
            return true;
        }
        pend.d0[a] *= e[0];
        pend.d1[a] *= e[1];
        if (qs.size() == 2) {
            const uint32_t b = qs[1];
            pend.d1[b] *= e[2] / e[0];
            C64 w = e[3] * e[0] / (e[1] * e[2]);
            // A pair phase within 1e-14 of -1 or 1 is that value: the rounding of
            // sin and cos of multiples of pi (RZZ(-pi/2), CP(pi)) leaves it an ulp
            // off, and an exact -1 joins a sign run (a sign flip) instead of a
            // phase run (a complex product).
            if (std::abs(w + 1.0) < 1e-14) w = C64(-1, 0);
            if (std::abs(w - 1.0) < 1e-14) w = C64(1, 0);
            if (w != C64(1, 0)) add_pair(a, b, w);
        }
        return true;
    }

    void add_pair(uint32_t a, uint32_t b, C64 w) {
        prun.push_back({a, b, w});
        in_run[a] = in_run[b] = 1;
    }

    // Where a physical qubit's bit lives now.
    uint32_t encode(uint32_t q) const {
        const int b = tbit[q];
        if (b < 0) return VT_GLOBAL | q;
        switch (roles.role[b]) {
            case ROLE_LANE: return VT_THREAD | roles.slot[b];
            case ROLE_WAVE: return VT_THREAD | (6u + roles.slot[b]);
            default: return VT_REG | roles.slot[b];
        }
    }

    void add_controls(VInstr &v, const std::vector<std::pair<uint32_t, bool>> &controls) const {
        for (auto &c : controls) {
            const uint32_t e = encode(c.first);
            const uint32_t x = e & 0xffffu;
            switch (e & VT_KIND) {
                case VT_THREAD:
                    v.tmask |= 1u << x;
                    if (c.second) v.tval |= 1u << x;
                    break;
                case VT_REG:
                    v.rmask |= 1u << x;
                    if (c.second) v.rval |= 1u << x;
                    break;
                default:
                    v.gmask |= uint64_t(1) << x;
                    if (c.second) v.gval |= uint64_t(1) << x;
            }
        }
    }

    VInstr blank(uint32_t op) const {
        VInstr v{};
        v.op = op;
        return v;
    }

    // Emits a 1-qubit gate m (row-major) on q.
    void emit_u1(uint32_t q, const M2 &m, uint32_t form, const std::vector<std::pair<uint32_t, bool>> &controls) {
        const int b = tbit[q];
        require(b >= 0 && roles.role[b] != ROLE_WAVE, "internal: 1-qubit gate on a wave bit");
        VInstr v = blank(roles.role[b] == ROLE_LANE ? V_U1L : V_U1R);
        v.slot = roles.slot[b];
        v.form = form;
        add_controls(v, controls);
        v.coef = (uint32_t)coefs.size();
        for (const C64 &c : m) coefs.push_back(c);
        out.push_back(v);
    }

    // Applies the pending product on q: U = P * D_q = D_new * M with M in
    // unit form; emits M and keeps D_new pending.
    void flush_pmat(uint32_t q) {
        if (!has_pmat[q]) return;
        has_pmat[q] = 0;
        const M2 &p = pmat[q];
        const M2 u = {cmul(p[0], pend.d0[q]), cmul(p[1], pend.d1[q]), cmul(p[2], pend.d0[q]), cmul(p[3], pend.d1[q])};
        if (nsq(u[1]) == 0.0 && nsq(u[2]) == 0.0) {  // diagonal after all
            pend.d0[q] = u[0];
            pend.d1[q] = u[3];
            return;
        }
        if (nsq(u[0]) >= nsq(u[1])) {
            // [[u0, u1], [u2, u3]] = diag(u0, u3) [[1, u1/u0], [u2/u3, 1]]
            const C64 a = u[1] / u[0], b = u[2] / u[3];
            const int ka = quarter_turn(a), kb = quarter_turn(b);
            if (ka >= 0 && kb == ((2 - ka) & 3)) {
                // Off-diagonals i^k and i^(2-k) = -conj(i^k) (to 1e-14): an
                // add and an FMA by 2 per part, in place.
                emit_u1(q, {C64(1, 0), a, b, C64(1, 0)}, FORM_UNIT_ROT, {});
                out.back().aux = (uint32_t)ka;
            } else if (std::abs(a.real()) <= 1e-15 * std::abs(a) && std::abs(b.real()) <= 1e-15 * std::abs(b)) {
                // Purely imaginary off-diagonals (RX-like; an exact zero real
                // part, or one below 1e-15 of the value, which is dropped).
                emit_u1(q, {C64(1, 0), C64(0, a.imag()), C64(0, b.imag()), C64(1, 0)}, FORM_UNIT_IMAG, {});
            } else if (std::abs(a.imag()) <= 1e-15 * std::abs(a) && std::abs(b.imag()) <= 1e-15 * std::abs(b)) {
                // Real off-diagonals (RY-like).
                emit_u1(q, {C64(1, 0), C64(a.real(), 0), C64(b.real(), 0), C64(1, 0)}, FORM_UNIT_REAL, {});
            } else {
                emit_u1(q, {C64(1, 0), a, b, C64(1, 0)}, FORM_UNIT_DIAG, {});
            }
            pend.d0[q] = u[0];
            pend.d1[q] = u[3];
        } else {
            // [[u0, u1], [u2, u3]] = diag(u1, u2) [[u0/u1, 1], [1, u3/u2]]
            emit_u1(q, {u[0] / u[1], C64(1, 0), C64(1, 0), u[3] / u[2]}, FORM_UNIT_ANTI, {});
            pend.d0[q] = u[1];
            pend.d1[q] = u[2];
        }
    }

    // A non-diagonal op on targets T (with controls) in the frame where the
    // pending diagonals D_T stay pending: D^-1 U D.
    HMatrix conjugate(const Op &op) const {
        HMatrix m = op.m;
        const size_t d = m.dim();
        for (size_t r = 0; r < d; r++)
            for (size_t c = 0; c < d; c++) {
                C64 fr(1, 0), fc(1, 0);
                for (size_t j = 0; j < op.targets.size(); j++) {
                    const uint32_t q = op.targets[j];
                    fr = cmul(fr, ((r >> j) & 1) ? pend.d1[q] : pend.d0[q]);
                    fc = cmul(fc, ((c >> j) & 1) ? pend.d1[q] : pend.d0[q]);
                }
                m.at(r, c) = cmul(m.at(r, c), fc) / fr;
            }
        return m;
    }

    void emit_diag(const Op &op) {
        for (uint32_t q : op.qubits()) flush_pmat(q);
        VInstr v = blank(V_DIAG);
        v.k = (uint32_t)op.targets.size();
        for (uint32_t j = 0; j < v.k; j++) {
            v.tgt[j] = encode(op.targets[j]);
            const uint32_t x = v.tgt[j] & 0xffffu;
            if ((v.tgt[j] & VT_KIND) == VT_THREAD) v.threaded = 1;
            if ((v.tgt[j] & VT_KIND) == VT_REG)
                for (uint32_t r = 0; r < (1u << g.r_bits); r++) v.rtab[r] |= (uint8_t)(((r >> x) & 1) << j);
        }
        add_controls(v, op.controls);
        if (v.tmask) v.threaded = 1;
        v.coef = (uint32_t)coefs.size();
        bool sign = true;
        uint64_t neg = 0;
        for (size_t i = 0; i < op.m.dim(); i++) {
            const C64 e = op.m.at(i, i);
            coefs.push_back(e);
            if (e == C64(-1, 0)) neg |= uint64_t(1) << i;
            else if (e != C64(1, 0)) sign = false;
        }
        if (sign) {  // entries +-1: a sign flip where the mask bit is set
            v.form = 1;
            v.neg = neg;
        }
        out.push_back(v);
    }

    void emit_ldsop(const Op &op) {
        for (uint32_t q : op.qubits()) flush_pmat(q);
        HMatrix m = conjugate(op);
        StageOp so{};
        so.kind = OP_DENSE;
        so.k = (uint32_t)op.targets.size();
        std::vector<uint32_t> ins;
        for (uint32_t j = 0; j < so.k; j++) {
            const int b = tbit[op.targets[j]];
            require(b >= 0, "internal: LDS op on a non-tile target");
            so.tgt[j] = (uint32_t)b;
            ins.push_back((uint32_t)b);
        }
        for (auto &c : op.controls) {
            const int b = tbit[c.first];
            if (b >= 0) {
                so.lc_mask |= 1u << b;
                if (c.second) so.lc_val |= 1u << b;
                ins.push_back((uint32_t)b);
            } else {
                so.gc_mask |= uint64_t(1) << c.first;
                if (c.second) so.gc_val |= uint64_t(1) << c.first;
            }
        }
        std::sort(ins.begin(), ins.end());
        so.nins = (uint32_t)ins.size();
        for (size_t j = 0; j < ins.size(); j++) so.ins[j] = ins[j];
        so.coef_off = (uint32_t)coefs.size();
        for (const C64 &c : m.a) coefs.push_back(c);
        VInstr v = blank(V_LDSOP);
        v.aux = (uint32_t)ldsops.size();
        v.gmask = so.gc_mask;
        v.gval = so.gc_val;
        ldsops.push_back(so);
        out.push_back(v);
    }

    // Applies every pending diagonal: one SCALE with per-thread,
    // per-register and per-tile (non-tile qubit) factors.
    void emit_scale() {
        const uint32_t W = 64u << g.w_bits, R = 1u << g.r_bits;
        std::vector<C64> ft(W, C64(1, 0)), fr(R, C64(1, 0));
        std::vector<uint32_t> globals;
        bool any = false;
        for (uint32_t q = 0; q < n; q++) {
            if (pend.identity(q)) continue;
            any = true;
            const uint32_t e = encode(q), x = e & 0xffffu;
            switch (e & VT_KIND) {
                case VT_THREAD:
                    for (uint32_t t = 0; t < W; t++) ft[t] = cmul(ft[t], ((t >> x) & 1) ? pend.d1[q] : pend.d0[q]);
                    break;
                case VT_REG:
                    for (uint32_t j = 0; j < R; j++) fr[j] = cmul(fr[j], ((j >> x) & 1) ? pend.d1[q] : pend.d0[q]);
                    break;
                default: globals.push_back(q);
            }
        }
        if (!any) return;
        VInstr v = blank(V_SCALE);
        v.coef = (uint32_t)coefs.size();
        coefs.insert(coefs.end(), ft.begin(), ft.end());
        coefs.insert(coefs.end(), fr.begin(), fr.end());
        v.k = (uint32_t)globals.size();
        for (uint32_t q : globals) {
            coefs.push_back(pend.d0[q]);
            coefs.push_back(pend.d1[q]);
        }
        v.aux = (uint32_t)ints.size();
        ints.insert(ints.end(), globals.begin(), globals.end());
        // Uniform factors: a plain complex scale.
        bool uniform = globals.empty();
        for (uint32_t t = 1; t < W && uniform; t++) uniform = ft[t] == ft[0];
        for (uint32_t j = 1; j < R && uniform; j++) uniform = fr[j] == fr[0];
        v.form = uniform ? 1 : 0;
        if (uniform) coefs[v.coef] = cmul(ft[0], fr[0]);
        out.push_back(v);
        for (uint32_t q = 0; q < n; q++) {
            pend.d0[q] = C64(1, 0);
            pend.d1[q] = C64(1, 0);
        }
    }

    // The flexible tile bit a 1-qubit gate needs as a lane or register bit
    // (none for diagonals and multi-qubit dense ops, which run anywhere).
    int need(const Op &op) const {
        if (op.diagonal || op.targets.size() != 1) return -1;
        const int b = tbit[op.targets[0]];
        return b >= (int)g.f_bits ? b : -1;
    }

    bool runnable(const Op &op) const {
        const int b = need(op);
        return b < 0 || roles.role[b] != ROLE_WAVE;
    }

    template <typename F>
    std::vector<size_t> extract(const std::vector<char> &done, F &&can_run) const {
        std::vector<size_t> taken;
        std::vector<uint32_t> blocked;
        for (size_t i = 0; i < st.ops.size(); i++) {
            if (done[i]) continue;
            const std::vector<uint32_t> oq = st.ops[i].qubits();
            bool touches = false;
            for (uint32_t q : oq) touches |= contains(blocked, q);
            if (!touches && can_run(st.ops[i])) {
                taken.push_back(i);
            } else {
                for (uint32_t q : oq)
                    if (!contains(blocked, q)) blocked.push_back(q);
            }
        }
        return taken;
    }

    // The flexible bits the coming gates want as lane or register bits, in
    // order of need, up to the number of such slots.
    std::vector<uint8_t> lookahead(const std::vector<char> &done) const {
        const size_t cap = g.r_bits + g.flex_lanes();
        std::vector<uint8_t> wanted;
        extract(done, [&](const Op &op) {
            const int b = need(op);
            if (b < 0 || has(wanted, (uint32_t)b)) return true;
            if (wanted.size() + 1 > cap) return false;
            wanted.push_back((uint8_t)b);
            return true;
        });
        if (st.role_policy == 1) {
            // Most-used first: the wanted bits by the number of pending ops
            // that need them in this stage (ties keep the order of need).
            std::vector<uint32_t> uses(st.local.size(), 0);
            for (size_t i = 0; i < st.ops.size(); i++) {
                if (done[i]) continue;
                const int b = need(st.ops[i]);
                if (b >= 0) uses[b]++;
            }
            std::stable_sort(wanted.begin(), wanted.end(), [&](uint8_t a, uint8_t b) { return uses[a] > uses[b]; });
        }
        return wanted;
    }

    // New role map: wanted bits first as register bits, then as lane bits;
    // current register and lane bits fill the rest of those slots; the
    // remaining flexible bits become wave bits.
    void choose_roles(const std::vector<uint8_t> &wanted, bool initial, std::vector<uint8_t> &l,
                      std::vector<uint8_t> &w, std::vector<uint8_t> &r) const {
        const uint32_t L = (uint32_t)st.local.size(), F = g.f_bits;
        std::vector<uint8_t> order = wanted;
        if (!initial) {
            for (uint8_t b : roles.rmap)
                if (!has(order, b)) order.push_back(b);
            for (uint8_t b : roles.lmap)
                if (!has(order, b)) order.push_back(b);
        }
        for (uint32_t b = F; b < L; b++)
            if (!has(order, b)) order.push_back((uint8_t)b);
        r.assign(order.begin(), order.begin() + g.r_bits);
        l.assign(order.begin() + g.r_bits, order.begin() + g.r_bits + g.flex_lanes());
        w.assign(order.begin() + g.r_bits + g.flex_lanes(), order.end());
        if (!initial) {
            // A wanted bit already in a lane slot can stay there.
            r = stabilize(r, roles.rmap);
            l = stabilize(l, roles.lmap);
            w = stabilize(w, roles.wmap);
        }
    }

    void emit_op(const Op &op) {
        if (op.diagonal) {
            // 1-qubit diagonal: folded into the pending diagonal (or the
            // pending product when one is open).
            if (op.controls.empty() && op.targets.size() == 1) {
                const uint32_t q = op.targets[0];
                const C64 a = op.m.at(0, 0), b = op.m.at(1, 1);
                if (has_pmat[q]) {
                    M2 &p = pmat[q];
                    p = {cmul(a, p[0]), cmul(a, p[1]), cmul(b, p[2]), cmul(b, p[3])};
                } else {
                    pend.d0[q] = cmul(a, pend.d0[q]);
                    pend.d1[q] = cmul(b, pend.d1[q]);
                }
                return;
            }
            if (absorb_pair_diag(op)) return;
            emit_diag(op);
            return;
        }
        if (touches_run(op)) flush_run();
        if (op.targets.size() == 1) {
            const uint32_t q = op.targets[0];
            for (auto &c : op.controls) flush_pmat(c.first);
            if (op.controls.empty()) {
                const M2 u = {op.m.at(0, 0), op.m.at(0, 1), op.m.at(1, 0), op.m.at(1, 1)};
                if (has_pmat[q]) pmat[q] = mul2(u, pmat[q]);
                else {
                    pmat[q] = u;
                    has_pmat[q] = 1;
                }
                return;
            }
            flush_pmat(q);
            HMatrix m = conjugate(op);
            emit_u1(q, {m.at(0, 0), m.at(0, 1), m.at(1, 0), m.at(1, 1)}, FORM_FULL, op.controls);
            return;
        }
        emit_ldsop(op);
    }

    VProgram run(bool flush_all) {
        const uint32_t L = (uint32_t)st.local.size();
        require(L == g.L(), "internal: stage tile size does not match the register tile");
        for (uint32_t b = 0; b < g.f_bits; b++) require(st.local[b] == b, "internal: fixed lane bits must be the lowest qubits");
        std::vector<char> done(st.ops.size(), 0);
        std::vector<uint8_t> l, w, r;
        choose_roles(lookahead(done), true, l, w, r);
        roles.assign(L, g.f_bits, l, w, r);
        VProgram p;
        p.desc = VStageDesc{};
        for (size_t s = 0; s < l.size(); s++) p.desc.lmap0[s] = l[s];
        for (size_t s = 0; s < w.size(); s++) p.desc.wmap0[s] = w[s];
        for (size_t s = 0; s < r.size(); s++) p.desc.rmap0[s] = r[s];
        size_t remaining = st.ops.size();
        while (remaining) {
            std::vector<size_t> taken = extract(done, [&](const Op &op) { return runnable(op); });
            for (size_t i : taken) {
                done[i] = 1;
                emit_op(st.ops[i]);
            }
            remaining -= taken.size();
            if (!remaining) break;
            // Exchange so the wanted bits become lane or register bits.
            // Pending products are emitted first, in the slots they have now.
            choose_roles(lookahead(done), false, l, w, r);
            for (uint32_t q = 0; q < n; q++) flush_pmat(q);
            flush_run();
            require(l != roles.lmap || w != roles.wmap || r != roles.rmap, "internal: exchange without progress");
            VInstr v = blank(V_XCHG);
            for (size_t s = 0; s < l.size(); s++) v.lmap[s] = l[s];
            for (size_t s = 0; s < w.size(); s++) v.wmap[s] = w[s];
            for (size_t s = 0; s < r.size(); s++) v.rmap[s] = r[s];
            roles.assign(L, g.f_bits, l, w, r);
            out.push_back(v);
        }
        for (uint32_t q = 0; q < n; q++) flush_pmat(q);
        flush_run();
        if (flush_all) emit_scale();
        for (size_t s = 0; s < roles.lmap.size(); s++) p.desc.lmapF[s] = roles.lmap[s];
        for (size_t s = 0; s < roles.wmap.size(); s++) p.desc.wmapF[s] = roles.wmap[s];
        for (size_t s = 0; s < roles.rmap.size(); s++) p.desc.rmapF[s] = roles.rmap[s];
        p.instr = std::move(out);
        return p;
    }
};

}  // namespace

VProgram build_vprogram(const Stage &stage, uint32_t n, const VGeom &g, PendingDiag &pending,
                        std::vector<C64> &coefs, std::vector<StageOp> &ldsops, std::vector<uint32_t> &ints,
                        bool flush_all) {
    Gen gen(stage, n, g, pending, coefs, ldsops, ints);
    const uint32_t cbegin = (uint32_t)coefs.size();
    VProgram p = gen.run(flush_all);
    VStageDesc &d = p.desc;
    d.cbegin = cbegin;
    d.ccount = (uint32_t)coefs.size() - cbegin;
    d.L = (uint32_t)stage.local.size();
    d.fixed_lanes = g.f_bits;
    for (uint32_t j = 0; j < d.L; j++) d.local_pos[j] = stage.local[j];
    d.nglob = 0;
    for (uint32_t q = 0; q < n; q++)
        if (gen.tbit[q] < 0) d.glob_pos[d.nglob++] = q;
    d.num_tiles = uint64_t(1) << d.nglob;
    d.num_instr = (uint32_t)p.instr.size();
    for (VInstr &v : p.instr) {
        if (v.tmask != 0 || v.rmask != 0) v.op |= VF_CTRL;
        if (v.gmask != 0) v.op |= VF_GCTRL;
        const uint32_t h = vhandler(v);
        require(v.gmask == 0 || (h >= H_U1R_FULL && h < H_U1L_ROT) || (h >= H_U1L_FULL && h < H_SIGNS) || h == H_LDSOP,
                "internal: non-tile controls on an instruction whose kernel path does not check them");
        v.op |= h << VH_SHIFT;
    }
    return p;
}

uint32_t vhandler(const VInstr &v) {
    const uint32_t op = v.op & VOP_MASK;
    const bool ctrl = (v.op & VF_CTRL) != 0;
    switch (op) {
        case V_U1R:
            require(v.slot < VREG_SLOTS, "register bit out of range");
            if (ctrl) return H_U1R_CTRL + v.slot;
            if (v.form == FORM_UNIT_ROT) return H_U1R_ROT + 4 * v.slot + (v.aux & 3);
            if (v.form == FORM_UNIT_DIAG) return H_U1R_DIAG + v.slot;
            if (v.form == FORM_UNIT_IMAG) return H_U1R_IMAG + v.slot;
            if (v.form == FORM_UNIT_REAL) return H_U1R_REAL + v.slot;
            if (v.form == FORM_UNIT_ANTI) return H_U1R_ANTI + v.slot;
            return H_U1R_FULL + v.slot;
        case V_U1L:
            require(v.slot < 6, "lane bit out of range");
            if (ctrl) return H_U1L_CTRL + v.slot;
            if (v.form == FORM_UNIT_ROT) return H_U1L_ROT + 4 * v.slot + (v.aux & 3);
            if (v.form == FORM_UNIT_DIAG) return H_U1L_DIAG + v.slot;
            if (v.form == FORM_UNIT_IMAG) return H_U1L_IMAG + v.slot;
            if (v.form == FORM_UNIT_REAL) return H_U1L_REAL + v.slot;
            if (v.form == FORM_UNIT_ANTI) return H_U1L_ANTI + v.slot;
            return H_U1L_FULL + v.slot;
        case V_DIAG: return H_DIAG + 2 * (v.threaded ? 1 : 0) + (v.form == 1 ? 1 : 0);
        case V_SIGNS: return H_SIGNS;
        case V_PHASES: return H_PHASES;
        case V_SCALE: return v.form == 1 ? H_SCALE1 : H_SCALE;
        case V_XCHG: return H_XCHG;
        case V_LDSOP: return H_LDSOP;
        default: require(false, "unknown register-tile instruction");
    }
    return H_COUNT;
}

}  // namespace qvd

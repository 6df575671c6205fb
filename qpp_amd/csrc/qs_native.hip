// qs_native.hip -- specialised pass kernels for a tile program, generated as
// HIP source and compiled at run time with hipRTC.
//
// k_tile (qs_fused.hip) interprets a pass: for every op it loads the op's
// head, dispatches on its code and writes all registers of the other of two
// register arrays.  A specialised kernel runs the same phases and LDS
// layout changes, one workgroup per tile, with the pass's ops fixed when the
// code is generated:
//  - matrix entries and diagonal factors come from a per-pass table at fixed
//    offsets, read with scalar loads where each op needs them (the address
//    is made opaque once per tile, so the compiler cannot hoist every
//    pass's constants out of the tile loop into registers);
//  - an X whose control lies in register bits only is a renaming of
//    registers at generation time and costs no instruction;
//  - a controlled op computes only the register groups its control admits;
//  - no dispatch and no copies between register arrays;
//  - tile-base runs and the global and LDS offsets of every register group
//    and thread bit are literals;
//  - a layout change whose two layouts share a tile bit (a thread bit at the
//    same position or a register bit) moves the tile in two halves through
//    half the LDS, so complex128 passes run 4 workgroups per CU instead of 2;
//  - the pass stores from a layout chosen next to its last phase.
// Before generation, the ops of each phase that act on one or two register
// bits are multiplied into fewer ops where that saves FP64 instructions
// (native_fuse).  The arithmetic of each op is written as in
// qs_tile_ops.hpp, so results agree with k_tile to rounding.  Programs use it on request (qs_program_compile_native); every
// other path keeps k_tile.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <sstream>
#include <string>
#include <vector>

#include "qs_fused.hpp"
#include "qs_internal.hpp"

namespace {

using qs::fail;
using C = std::complex<double>;

constexpr int kNR = 16;  // register groups (reg_bits 4)

struct Gen {
    std::ostringstream o;
    bool dbl = true;
    int perm[kNR];
    std::vector<double> tab;  // the pass's constants (re, im pairs)
    uint64_t stats[5] = {0, 0, 0, 0, 0};  // as FusedProgram::native_stats

    // a complex constant: its place in the pass's table
    std::string cx(double re, double im) {
        const size_t i = tab.size();
        tab.push_back(re);
        tab.push_back(im);
        return "mk(cm[" + std::to_string(i) + "], cm[" + std::to_string(i + 1) + "])";
    }
    std::string cx(const C& c) { return cx(c.real(), c.imag()); }
    std::string X(int rho) const { return "x[" + std::to_string(perm[rho]) + "]"; }
    void identity() {
        for (int i = 0; i < kNR; ++i) perm[i] = i;
    }
};

std::string u64(uint64_t v) { return std::to_string(v) + "ull"; }

// +1 or -1 for a diagonal entry that is exactly that, else 0
int sign_of(double re, double im) {
    if (im != 0.0) return 0;
    return re == 1.0 ? 1 : re == -1.0 ? -1 : 0;
}

// thread part of a phase's global index (single-bit offsets per thread bit)
std::string thr_g(const FPhase& ph) {
    std::string e = "(0ull";
    for (int j = 0; j < 8; ++j)
        e += " | ((unsigned long long)((t >> " + std::to_string(j) + ") & 1u) * " + u64(ph.g_thr[j]) + ")";
    return e + ")";
}

// thread part of a phase's swizzled LDS index, in bytes (es bytes a slot)
std::string thr_s(const FPhase& ph, uint32_t es) {
    std::string e = "(0u";
    for (int j = 0; j < 8; ++j)
        e += " ^ (((t >> " + std::to_string(j) + ") & 1u) * " + std::to_string(ph.s_thr[j] * es) + "u)";
    return e + ")";
}

uint64_t greg(const FPhase& ph, int rho) {
    uint64_t g = 0;
    for (int k = 0; k < 4; ++k)
        if ((rho >> k) & 1) g |= ph.g_rb[k];
    return g;
}

uint32_t sreg(const FPhase& ph, int rho) {
    uint32_t s = 0;
    for (int k = 0; k < 4; ++k)
        if ((rho >> k) & 1) s ^= ph.s_rb[k];
    return s;
}

std::string tile_base(const FPass& ps, const char* tile) {
    std::string e = "(0ull";
    for (int i = 0; i < ps.nruns; ++i) {
        const uint32_t r = ps.runs[i];
        const uint32_t src = r & 255u, dst = (r >> 8) & 255u, len = r >> 16;
        const uint64_t mask = len >= 64 ? ~uint64_t{0} : ((uint64_t{1} << len) - 1);
        e += " | (((" + std::string(tile) + " >> " + std::to_string(src) + ") & " + u64(mask) + ") << " +
             std::to_string(dst) + ")";
    }
    return e + ")";
}

// ---------------------------------------------------------------------------
// Pair fusion: within a phase, the ops that act only on register bits are
// gathered per bit (a 2x2 product) and per pair of bits (a 4x4 product, once
// an op joins the two bits: a CNOT or controlled gate between them, or a
// dense 2-bit gate).  Ops on other bits commute with a pending product and
// pass it; an op that involves a pending product's bits emits the product
// first.  A product goes out as one op when that is cheaper in FP64
// instructions than the ops it holds (U1 8, U2 16, diagonal 4, X 0 per
// register group), else its ops go out unchanged and in order.  On the
// random circuit a layer of 1-qubit gates, the CNOTs joining register bits
// and the next layer become 4x4 blocks: two layers of 8 U1 ops cost 1024
// instructions per tile-wave, the same as two 4x4 ops cost 512.
// ---------------------------------------------------------------------------
using M4 = std::array<C, 16>;  // column-major, index bit 0 <-> lower bit
using M2 = std::array<C, 4>;

M2 m2_mul(const M2& A, const M2& B) {
    M2 r{};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) r[i + 2 * j] = A[i] * B[2 * j] + A[i + 2] * B[1 + 2 * j];
    return r;
}
M4 m4_mul(const M4& A, const M4& B) {
    M4 r{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            C acc(0, 0);
            for (int k = 0; k < 4; ++k) acc += A[i + 4 * k] * B[k + 4 * j];
            r[i + 4 * j] = acc;
        }
    return r;
}
M4 m4_id() {
    M4 r{};
    for (int i = 0; i < 4; ++i) r[i + 4 * i] = C(1, 0);
    return r;
}
// U on index bit `which` (0: the lower bit of the pair) of a 4x4
M4 m4_on(const M2& U, int which) {
    M4 r{};
    for (int rr = 0; rr < 4; ++rr)
        for (int cc = 0; cc < 4; ++cc) {
            const int r0 = rr & 1, r1 = rr >> 1, c0 = cc & 1, c1 = cc >> 1;
            r[rr + 4 * cc] = which == 0 ? (r1 == c1 ? U[r0 + 2 * c0] : C(0, 0))
                                        : (r0 == c0 ? U[r1 + 2 * c1] : C(0, 0));
        }
    return r;
}
// U on index bit `which`, applied where the other bit equals v
M4 m4_ctrl(const M2& U, int which, int v) {
    M4 r = m4_id();
    for (int rr = 0; rr < 4; ++rr)
        for (int cc = 0; cc < 4; ++cc) {
            const int cr = which == 0 ? rr >> 1 : rr & 1, ccb = which == 0 ? cc >> 1 : cc & 1;
            if (cr != v || ccb != v) continue;
            const int tr = which == 0 ? rr & 1 : rr >> 1, tc = which == 0 ? cc & 1 : cc >> 1;
            r[rr + 4 * cc] = U[tr + 2 * tc];
        }
    return r;
}

constexpr int kCostU1 = 8, kCostU2 = 16, kCostDiag = 4;

struct Pending {
    bool pair = false;
    int a = -1, b = -1;       // bits (a < b for a pair)
    M2 s{};                   // single-bit product
    M4 m{};                   // pair product
    bool diag = true;         // single: every op diagonal
    int cost = 0;             // the held ops' FP64 cost
    std::vector<FOp> held;    // the held ops, in order
    // an X under a lane or tile condition taken in: s / m hold the product
    // where the condition fails, s1 / m1 where it holds
    bool cond = false;
    uint32_t tmask = 0, tval = 0;
    uint64_t omask = 0, oval = 0;
    M2 s1{};
    M4 m1{};
};

bool same_cond(const Pending& p, const FOp& x) {
    return p.tmask == x.tmask && p.tval == x.tval && p.omask == x.omask && p.oval == x.oval;
}
bool same_cond(const Pending& p, const Pending& q) {
    return p.tmask == q.tmask && p.tval == q.tval && p.omask == q.omask && p.oval == q.oval;
}

// the 2x2 an uncontrolled 1-bit op applies (U1, DIAG_R, X)
M2 op_m2(const FOp& o) {
    if (o.kind == OP_X) return {C(0, 0), C(1, 0), C(1, 0), C(0, 0)};
    if (o.kind == OP_DIAG_R) return {C(o.m[0], o.m[1]), C(0, 0), C(0, 0), C(o.m[2], o.m[3])};
    return {C(o.m[0], o.m[1]), C(o.m[2], o.m[3]), C(o.m[4], o.m[5]), C(o.m[6], o.m[7])};
}

int op_cost(const FOp& o) {
    switch (o.kind) {
        case OP_U1: return kCostU1;
        case OP_DIAG_R: case OP_DIAG_T: case OP_DIAG_O: return kCostDiag;
        case OP_U2: return kCostU2;
        case OP_U1B: return 4 * kCostU1;
        default: return 0;
    }
}

// register bits the op's register-group predicate depends on
uint32_t regok_bits(const FOp& o) {
    uint32_t m = 0;
    for (int b = 0; b < 4; ++b)
        for (int rho = 0; rho < kNR; ++rho)
            if (((o.regok >> rho) & 1u) != ((o.regok >> (rho ^ (1 << b))) & 1u)) {
                m |= 1u << b;
                break;
            }
    return m;
}

uint32_t op_bits(const FOp& o) {
    uint32_t m = regok_bits(o);
    switch (o.kind) {
        case OP_U1: case OP_X: case OP_DIAG_R: case OP_U1S: m |= 1u << o.j0; break;
        case OP_U2: case OP_U2S: m |= (1u << o.j0) | (1u << o.j1); break;
        case OP_U1B: m |= 15u; break;
        default: break;
    }
    return m;
}

struct Fuser {
    std::vector<C>& mats;
    std::vector<FOp> out;
    std::vector<Pending> pend;  // disjoint bit sets
    explicit Fuser(std::vector<C>& m) : mats(m) {}

    int find(int bit) const {
        for (size_t i = 0; i < pend.size(); ++i)
            if (pend[i].a == bit || pend[i].b == bit) return int(i);
        return -1;
    }
    void emit(size_t i) {
        Pending p = std::move(pend[i]);
        pend.erase(pend.begin() + long(i));
        FOp o{};
        o.regok = 0xFFFFu;
        if (p.cond) {
            // a product with an X under a condition: one select op when it
            // costs no more FP64 than the held ops (it also saves the X's
            // register swap); a lone X stays an X.  Pairs only take tile
            // conditions (16 entries selected per lane would take 64 more
            // VGPRs than a kernel at 4 waves per SIMD has)
            if (p.cost == 0 || (p.pair && p.cost < kCostU2)) {
                out.insert(out.end(), p.held.begin(), p.held.end());
                return;
            }
            o.tmask = p.tmask, o.tval = p.tval, o.omask = p.omask, o.oval = p.oval;
            o.j0 = p.a;
            o.moff = static_cast<uint32_t>(mats.size());
            if (!p.pair) {
                o.kind = OP_U1S;
                mats.insert(mats.end(), p.s.begin(), p.s.end());
                mats.insert(mats.end(), p.s1.begin(), p.s1.end());
            } else {
                o.kind = OP_U2S;
                o.j1 = p.b;
                mats.insert(mats.end(), p.m.begin(), p.m.end());
                mats.insert(mats.end(), p.m1.begin(), p.m1.end());
            }
            out.push_back(o);
            return;
        }
        if (!p.pair) {
            const int cost = p.diag ? kCostDiag : kCostU1;
            if (p.held.size() == 1 || cost >= p.cost) {
                out.insert(out.end(), p.held.begin(), p.held.end());
                return;
            }
            o.kind = p.diag ? OP_DIAG_R : OP_U1;
            o.j0 = p.a;
            if (p.diag) {
                o.m[0] = p.s[0].real(), o.m[1] = p.s[0].imag();
                o.m[2] = p.s[3].real(), o.m[3] = p.s[3].imag();
            } else {
                for (int t = 0; t < 4; ++t) o.m[2 * t] = p.s[t].real(), o.m[2 * t + 1] = p.s[t].imag();
            }
            out.push_back(o);
            return;
        }
        if (kCostU2 >= p.cost) {
            out.insert(out.end(), p.held.begin(), p.held.end());
            return;
        }
        o.kind = OP_U2;
        o.j0 = p.a;
        o.j1 = p.b;
        o.moff = static_cast<uint32_t>(mats.size());
        mats.insert(mats.end(), p.m.begin(), p.m.end());
        out.push_back(o);
    }
    void flush_bits(uint32_t bits) {
        for (size_t i = 0; i < pend.size();) {
            const Pending& p = pend[i];
            const bool hit = ((bits >> p.a) & 1u) || (p.pair && ((bits >> p.b) & 1u));
            if (hit)
                emit(i);
            else
                ++i;
        }
    }
    void flush_all() {
        while (!pend.empty()) emit(0);
    }
    // a 4x4 on bits (a < b) that the held op `o` applies
    void pair_op(int a, int b, const M4& G, const FOp& o) {
        int ia = find(a), ib = find(b);
        if (ia >= 0 && ia == ib && pend[ia].pair) {
            Pending& p = pend[ia];
            p.m = m4_mul(G, p.m);
            if (p.cond) p.m1 = m4_mul(G, p.m1);
            p.cost += op_cost(o);
            p.held.push_back(o);
            return;
        }
        // a pair with another bit is emitted; singles of a and b join
        if (ia >= 0 && pend[ia].pair) {
            emit(ia);
        }
        ib = find(b);
        if (ib >= 0 && pend[ib].pair) emit(ib);
        // singles under different conditions cannot share a product, and a
        // single under a lane condition goes out before a pair forms
        ia = find(a);
        if (ia >= 0 && pend[size_t(ia)].cond && pend[size_t(ia)].tmask != 0) emit(size_t(ia));
        ib = find(b);
        if (ib >= 0 && pend[size_t(ib)].cond && pend[size_t(ib)].tmask != 0) emit(size_t(ib));
        ia = find(a);
        ib = find(b);
        if (ia >= 0 && ib >= 0 && pend[ia].cond && pend[ib].cond && !same_cond(pend[ia], pend[ib])) {
            emit(size_t(ib));
            ia = find(a);
        }
        Pending np;
        np.pair = true;
        np.a = a;
        np.b = b;
        np.m = m4_id();
        np.m1 = m4_id();
        auto take = [&](int idx, int which) {
            Pending& q = pend[size_t(idx)];
            if (q.cond && !np.cond) {
                np.cond = true;
                np.tmask = q.tmask, np.tval = q.tval, np.omask = q.omask, np.oval = q.oval;
            }
            np.m = m4_mul(m4_on(q.s, which), np.m);
            np.m1 = m4_mul(m4_on(q.cond ? q.s1 : q.s, which), np.m1);
            np.cost += q.cost;
            np.held.insert(np.held.end(), q.held.begin(), q.held.end());
            pend.erase(pend.begin() + idx);
        };
        ia = find(a);
        if (ia >= 0) take(ia, 0);
        ib = find(b);
        if (ib >= 0) take(ib, 1);
        np.m = m4_mul(G, np.m);
        np.m1 = m4_mul(G, np.m1);
        np.cost += op_cost(o);
        np.held.push_back(o);
        pend.push_back(std::move(np));
    }
    void add(const FOp& o) {
        if (o.kind == OP_U1B) {  // a 2x2 on every register bit: four 1-bit ops
            for (int k = 0; k < 4; ++k) {
                FOp u{};
                u.kind = OP_U1;
                u.j0 = k;
                u.regok = 0xFFFFu;
                for (int t = 0; t < 4; ++t) {
                    u.m[2 * t] = mats[o.moff + 4 * k + t].real();
                    u.m[2 * t + 1] = mats[o.moff + 4 * k + t].imag();
                }
                add(u);
            }
            return;
        }
        const uint32_t rb = regok_bits(o);
        // an X whose only control is a lane or a tile condition joins the
        // product on its bit as a second product (where the condition holds)
        if (o.kind == OP_X && rb == 0 && (o.regok & 0xFFFFu) == 0xFFFFu && (o.tmask != 0) != (o.omask != 0)) {
            const int k = o.j0;
            int i = find(k);
            // (a pair under a lane condition would select 16 entries per
            // lane: the pair goes out first and the X starts a single)
            if (i >= 0 && ((pend[size_t(i)].cond && !same_cond(pend[size_t(i)], o)) ||
                           (pend[size_t(i)].pair && o.tmask != 0))) {
                emit(size_t(i));
                i = -1;
            }
            const M2 X = op_m2(o);
            if (i < 0) {
                Pending p;
                p.a = k;
                p.s = {C(1, 0), C(0, 0), C(0, 0), C(1, 0)};
                p.s1 = X;
                p.diag = false;
                p.cond = true;
                p.tmask = o.tmask, p.tval = o.tval, p.omask = o.omask, p.oval = o.oval;
                p.held.push_back(o);
                pend.push_back(std::move(p));
                return;
            }
            Pending& p = pend[size_t(i)];
            if (!p.pair) {
                p.s1 = m2_mul(X, p.cond ? p.s1 : p.s);
                p.diag = false;
            } else {
                p.m1 = m4_mul(m4_on(X, k == p.a ? 0 : 1), p.cond ? p.m1 : p.m);
            }
            if (!p.cond) {
                p.cond = true;
                p.tmask = o.tmask, p.tval = o.tval, p.omask = o.omask, p.oval = o.oval;
            }
            p.held.push_back(o);
            return;
        }
        const bool plain = o.tmask == 0 && o.omask == 0;
        const bool one_bit = o.kind == OP_U1 || o.kind == OP_X || o.kind == OP_DIAG_R;
        if (plain && one_bit && rb == 0 && (o.regok & 0xFFFFu) == 0xFFFFu) {
            const int j = o.j0;
            const int i = find(j);
            if (i < 0 && o.kind == OP_X) {  // a renaming, nothing to gain
                out.push_back(o);
                return;
            }
            const M2 U = op_m2(o);
            if (i >= 0 && pend[i].pair) {
                Pending& p = pend[i];
                p.m = m4_mul(m4_on(U, j == p.a ? 0 : 1), p.m);
                if (p.cond) p.m1 = m4_mul(m4_on(U, j == p.a ? 0 : 1), p.m1);
                p.cost += op_cost(o);
                p.held.push_back(o);
            } else if (i >= 0) {
                Pending& p = pend[i];
                p.s = m2_mul(U, p.s);
                if (p.cond) p.s1 = m2_mul(U, p.s1);
                p.diag = p.diag && o.kind == OP_DIAG_R && !p.cond;
                p.cost += op_cost(o);
                p.held.push_back(o);
            } else {
                Pending p;
                p.a = j;
                p.s = U;
                p.diag = o.kind == OP_DIAG_R;
                p.cost = op_cost(o);
                p.held.push_back(o);
                pend.push_back(std::move(p));
            }
            return;
        }
        // a 1-bit op controlled by exactly one other register bit, or an
        // uncontrolled dense 2-bit op: a 4x4 on the pair
        if (plain && one_bit && __builtin_popcount(rb) == 1 && !((rb >> o.j0) & 1u)) {
            const int c = __builtin_ctz(rb);
            int v = -1;
            uint32_t r1 = 0, r0 = 0;
            for (int rho = 0; rho < kNR; ++rho) ((rho >> c) & 1 ? r1 : r0) |= 1u << rho;
            if ((o.regok & 0xFFFFu) == r1) v = 1;
            if ((o.regok & 0xFFFFu) == r0) v = 0;
            if (v >= 0) {
                const int a = std::min(c, o.j0), b = std::max(c, o.j0);
                pair_op(a, b, m4_ctrl(op_m2(o), o.j0 == a ? 0 : 1, v), o);
                return;
            }
        }
        if (plain && o.kind == OP_U2 && rb == 0 && (o.regok & 0xFFFFu) == 0xFFFFu) {
            M4 G;
            for (int t = 0; t < 16; ++t) G[t] = mats[o.moff + t];
            pair_op(o.j0, o.j1, G, o);
            return;
        }
        flush_bits(op_bits(o));
        out.push_back(o);
    }
};

}  // namespace

namespace qs {

// The program with each phase's ops pair-fused (see Fuser); phases keep
// their layouts, their op ranges follow the new op list.
FusedProgram native_fuse(const FusedProgram& P) {
    if (P.reg_bits != 4) return P;
    FusedProgram Q;
    Q.nb = P.nb;
    Q.reg_bits = P.reg_bits;
    Q.passes = P.passes;
    Q.phases = P.phases;
    Q.mats = P.mats;
    Q.gates_per_pass = P.gates_per_pass;
    Q.dtype = P.dtype;
    for (FPhase& ph : Q.phases) {
        Fuser f(Q.mats);
        for (int i = ph.op_begin; i < ph.op_end; ++i) f.add(P.ops[i]);
        f.flush_all();
        ph.op_begin = static_cast<int>(Q.ops.size());
        for (FOp& o : f.out) {
            o.ctrl = (o.regok & 0xFFFFu) != 0xFFFFu || o.tmask != 0;
            Q.ops.push_back(o);
        }
        ph.op_end = static_cast<int>(Q.ops.size());
    }
    return Q;
}

}  // namespace qs

namespace {

// y = op(x) for one op, in place on the register array under the current
// renaming (qs_tile_ops.hpp run_op_h / run_op_ct / do_*)
// fx: an X on one register bit under a lane or tile condition, folded into
// this op (the next op of the phase on that bit): the op applies U X where
// the condition holds, U elsewhere (see fold_target).
void gen_op(Gen& g, const FusedProgram& P, const FOp& op, const FOp* fx = nullptr) {
    std::ostringstream& o = g.o;
    const uint32_t R = op.regok;
    const bool ct = op.ctrl != 0;
    // OP_U1S / OP_U2S: tmask / omask pick one of two products, not a control
    const bool selk = op.kind == OP_U1S || op.kind == OP_U2S;
    const bool has_t = op.tmask != 0 && !selk;
    const bool has_o = op.omask != 0 && !selk;
    auto rok = [&](int rho) { return !ct || ((R >> rho) & 1u); };
    o << "    {\n";
    if (has_o) o << "    if ((base & " << u64(op.omask) << ") == " << u64(op.oval) << ") {\n";
    if (has_t) o << "    const bool tk = (t & " << op.tmask << "u) == " << op.tval << "u;\n";
    auto put = [&](int rho, const std::string& v) {
        if (has_t)
            o << "    " << g.X(rho) << " = sel(tk, " << v << ", " << g.X(rho) << ");\n";
        else
            o << "    " << g.X(rho) << " = " << v << ";\n";
    };
    auto u1 = [&](int J, const std::string& u00, const std::string& u10, const std::string& u01,
                  const std::string& u11, bool controlled) {
        for (int rho = 0; rho < kNR; ++rho) {
            if ((rho >> J) & 1) continue;
            if (controlled && !rok(rho)) continue;
            const int r1 = rho | (1 << J);
            o << "    { const T a0 = " << g.X(rho) << ", b0 = " << g.X(r1) << ";\n";
            const std::string y0 = "mac2(" + u00 + ", a0, " + u01 + ", b0)";
            const std::string y1 = "mac2(" + u10 + ", a0, " + u11 + ", b0)";
            if (controlled && has_t) {
                o << "    " << g.X(rho) << " = sel(tk, " << y0 << ", a0);\n";
                o << "    " << g.X(r1) << " = sel(tk, " << y1 << ", b0); }\n";
            } else {
                o << "    " << g.X(rho) << " = " << y0 << ";\n";
                o << "    " << g.X(r1) << " = " << y1 << "; }\n";
            }
        }
    };
    switch (op.kind) {
        case OP_U1: {
            if (fx && fx->omask) {
                // U and U X (columns swapped) side by side in the table, the
                // tile's condition picks one: a uniform offset, no VALU work
                const size_t i0 = g.tab.size();
                const int src[8] = {0, 2, 4, 6, 4, 6, 0, 2};  // u00 u10 u01 u11, then u01 u11 u00 u10
                for (int k = 0; k < 8; ++k) (void)g.cx(op.m[src[k]], op.m[src[k] + 1]);
                o << "    const int so = (base & " << u64(fx->omask) << ") == " << u64(fx->oval) << " ? 8 : 0;\n";
                auto at = [&](int k) {
                    return "mk(cm[" + std::to_string(i0 + 2 * k) + " + so], cm[" + std::to_string(i0 + 2 * k + 1) + " + so])";
                };
                o << "    const T u00 = " << at(0) << ", u10 = " << at(1) << ", u01 = " << at(2) << ", u11 = " << at(3)
                  << ";\n";
                u1(op.j0, "u00", "u10", "u01", "u11", true);
                break;
            }
            o << "    const T u00 = " << g.cx(op.m[0], op.m[1]) << ", u10 = " << g.cx(op.m[2], op.m[3])
              << ", u01 = " << g.cx(op.m[4], op.m[5]) << ", u11 = " << g.cx(op.m[6], op.m[7]) << ";\n";
            if (fx && fx->tmask) {
                // U X on the lanes whose condition holds: the columns swapped
                // per lane (16 selects instead of 64 for the register swap)
                o << "    const bool xk = (t & " << fx->tmask << "u) == " << fx->tval << "u;\n";
                o << "    const T w00 = sel(xk, u01, u00), w10 = sel(xk, u11, u10), w01 = sel(xk, u00, u01), "
                     "w11 = sel(xk, u10, u11);\n";
                u1(op.j0, "w00", "w10", "w01", "w11", true);
                break;
            }
            u1(op.j0, "u00", "u10", "u01", "u11", true);
            break;
        }
        case OP_U1S: {
            // the product where the condition fails, then where it holds
            const C* M = P.mats.data() + op.moff;
            const size_t i0 = g.tab.size();
            for (int i = 0; i < 8; ++i) (void)g.cx(M[i]);
            if (op.omask) {  // a tile condition: a uniform offset into the table
                o << "    const int so = (base & " << u64(op.omask) << ") == " << u64(op.oval) << " ? 8 : 0;\n";
                auto at = [&](int k) {
                    return "mk(cm[" + std::to_string(i0 + 2 * k) + " + so], cm[" + std::to_string(i0 + 2 * k + 1) + " + so])";
                };
                o << "    const T u00 = " << at(0) << ", u10 = " << at(1) << ", u01 = " << at(2) << ", u11 = " << at(3)
                  << ";\n";
            } else {  // a lane condition: a select per entry
                auto at = [&](int k) {
                    return "mk(cm[" + std::to_string(i0 + 2 * k) + "], cm[" + std::to_string(i0 + 2 * k + 1) + "])";
                };
                o << "    const bool xk = (t & " << op.tmask << "u) == " << op.tval << "u;\n";
                o << "    const T u00 = sel(xk, " << at(4) << ", " << at(0) << "), u10 = sel(xk, " << at(5) << ", "
                  << at(1) << "), u01 = sel(xk, " << at(6) << ", " << at(2) << "), u11 = sel(xk, " << at(7) << ", "
                  << at(3) << ");\n";
            }
            u1(op.j0, "u00", "u10", "u01", "u11", false);
            break;
        }
        case OP_X: {
            const int J = op.j0;
            for (int rho = 0; rho < kNR; ++rho) {
                if ((rho >> J) & 1) continue;
                if (!rok(rho)) continue;
                const int r1 = rho | (1 << J);
                if (has_t) {
                    o << "    { const T a0 = " << g.X(rho) << ", b0 = " << g.X(r1) << ";\n";
                    o << "    " << g.X(rho) << " = sel(tk, b0, a0);\n";
                    o << "    " << g.X(r1) << " = sel(tk, a0, b0); }\n";
                } else if (op.omask) {
                    o << "    { const T a0 = " << g.X(rho) << "; " << g.X(rho) << " = " << g.X(r1) << "; "
                      << g.X(r1) << " = a0; }\n";
                } else {
                    std::swap(g.perm[rho], g.perm[r1]);  // a renaming: no code
                }
            }
            break;
        }
        case OP_DIAG_R: {
            const int J = op.j0;
            const int s0 = sign_of(op.m[0], op.m[1]), s1 = sign_of(op.m[2], op.m[3]);
            if (s0 && s1) {
                // entries +-1 (Z, CZ): a factor 1 is nothing, -1 flips the
                // sign bits (on the lanes the op's lane condition admits)
                if (has_t) o << "    const unsigned nm = tk ? 0x80000000u : 0u;\n";
                for (int rho = 0; rho < kNR; ++rho) {
                    if (!rok(rho) || ((rho >> J) & 1 ? s1 : s0) > 0) continue;
                    const std::string m = has_t ? "nm" : "0x80000000u";
                    o << "    " << g.X(rho) << " = mk(flip(" << g.X(rho) << ".x, " << m << "), flip(" << g.X(rho)
                      << ".y, " << m << "));\n";
                }
                break;
            }
            o << "    const T d0 = " << g.cx(op.m[0], op.m[1]) << ", d1 = " << g.cx(op.m[2], op.m[3]) << ";\n";
            for (int rho = 0; rho < kNR; ++rho) {
                if (!rok(rho)) continue;
                put(rho, std::string("cmul(") + (((rho >> J) & 1) ? "d1" : "d0") + ", " + g.X(rho) + ")");
            }
            break;
        }
        case OP_DIAG_T:
        case OP_DIAG_O: {
            const int s0 = sign_of(op.m[0], op.m[1]), s1 = sign_of(op.m[2], op.m[3]);
            if (s0 && s1) {
                // a factor +-1 per lane (thread bit) or per tile (bit outside
                // the tile): one sign mask, then the flips
                const std::string m0 = s0 < 0 ? "0x80000000u" : "0u", m1 = s1 < 0 ? "0x80000000u" : "0u";
                if (op.kind == OP_DIAG_T)
                    o << "    unsigned nm = (t & " << op.tbit << "u) != 0u ? " << m1 << " : " << m0 << ";\n";
                else
                    o << "    unsigned nm = (base & " << u64(op.dbit) << ") ? " << m1 << " : " << m0 << ";\n";
                if (has_t) o << "    nm = tk ? nm : 0u;\n";
                for (int rho = 0; rho < kNR; ++rho) {
                    if (!rok(rho)) continue;
                    o << "    " << g.X(rho) << " = mk(flip(" << g.X(rho) << ".x, nm), flip(" << g.X(rho) << ".y, nm));\n";
                }
                break;
            }
            o << "    const T d0 = " << g.cx(op.m[0], op.m[1]) << ", d1 = " << g.cx(op.m[2], op.m[3]) << ";\n";
            if (op.kind == OP_DIAG_T)
                o << "    const T dd = sel((t & " << op.tbit << "u) != 0u, d1, d0);\n";
            else
                o << "    const T dd = (base & " << u64(op.dbit) << ") ? d1 : d0;\n";
            for (int rho = 0; rho < kNR; ++rho) {
                if (!rok(rho)) continue;
                put(rho, "cmul(dd, " + g.X(rho) + ")");
            }
            break;
        }
        case OP_U2:
        case OP_U2S: {
            const int J0 = op.j0, J1 = op.j1;
            const C* M = P.mats.data() + op.moff;
            if (op.kind == OP_U2S) {
                // the product where the condition fails, then where it holds
                const size_t i0 = g.tab.size();
                for (int i = 0; i < 32; ++i) (void)g.cx(M[i]);
                auto at = [&](size_t k, const char* off) {
                    return "mk(cm[" + std::to_string(i0 + 2 * k) + off + "], cm[" + std::to_string(i0 + 2 * k + 1) +
                           off + "])";
                };
                if (op.omask) {
                    o << "    const int so = (base & " << u64(op.omask) << ") == " << u64(op.oval) << " ? 32 : 0;\n";
                    o << "    const T m[16] = {";
                    for (size_t i = 0; i < 16; ++i) o << (i ? ", " : "") << at(i, " + so");
                    o << "};\n";
                } else {
                    o << "    const bool xk = (t & " << op.tmask << "u) == " << op.tval << "u;\n";
                    o << "    const T m[16] = {";
                    for (size_t i = 0; i < 16; ++i)
                        o << (i ? ", " : "") << "sel(xk, " << at(i + 16, "") << ", " << at(i, "") << ")";
                    o << "};\n";
                }
            } else if (fx && fx->omask) {
                // M and M (X on the folded bit) side by side, picked by the
                // tile's condition
                const int bk = fx->j0 == J0 ? 1 : 2;
                const size_t i0 = g.tab.size();
                for (int i = 0; i < 16; ++i) (void)g.cx(M[i]);
                for (int i = 0; i < 16; ++i) (void)g.cx(M[(i & 3) + 4 * ((i >> 2) ^ bk)]);
                o << "    const int so = (base & " << u64(fx->omask) << ") == " << u64(fx->oval) << " ? 32 : 0;\n";
                o << "    const T m[16] = {";
                for (int i = 0; i < 16; ++i)
                    o << (i ? ", " : "") << "mk(cm[" << i0 + 2 * i << " + so], cm[" << i0 + 2 * i + 1 << " + so])";
                o << "};\n";
            } else {
                o << "    const T m[16] = {";
                for (int i = 0; i < 16; ++i) o << (i ? ", " : "") << g.cx(M[i]);
                o << "};\n";
            }
            for (int rho = 0; rho < kNR; ++rho) {
                if (((rho >> J0) & 1) || ((rho >> J1) & 1)) continue;
                if (!rok(rho)) continue;
                const int i[4] = {rho, rho | (1 << J0), rho | (1 << J1), rho | (1 << J0) | (1 << J1)};
                o << "    { const T v0 = " << g.X(i[0]) << ", v1 = " << g.X(i[1]) << ", v2 = " << g.X(i[2])
                  << ", v3 = " << g.X(i[3]) << ";\n";
                for (int r = 0; r < 4; ++r)
                    o << "    T o" << r << " = mac4(m[" << r << "], v0, m[" << r + 4 << "], v1, m[" << r + 8
                      << "], v2, m[" << r + 12 << "], v3);\n";
                for (int r = 0; r < 4; ++r) {
                    if (has_t)
                        o << "    " << g.X(i[r]) << " = sel(tk, o" << r << ", v" << r << ");\n";
                    else
                        o << "    " << g.X(i[r]) << " = o" << r << ";\n";
                }
                o << "    }\n";
            }
            break;
        }
        case OP_U1B: {
            const C* M = P.mats.data() + op.moff;
            for (int k = 0; k < 4; ++k) {
                o << "    { const T u00 = " << g.cx(M[4 * k]) << ", u10 = " << g.cx(M[4 * k + 1])
                  << ", u01 = " << g.cx(M[4 * k + 2]) << ", u11 = " << g.cx(M[4 * k + 3]) << ";\n";
                u1(k, "u00", "u10", "u01", "u11", false);
                o << "    }\n";
            }
            break;
        }
        default:
            break;
    }
    if (has_o) o << "    }\n";
    o << "    }\n";
    // each op's constants are loaded next to its arithmetic
    o << "    __builtin_amdgcn_sched_barrier(0);\n";
}

// How a layout change moves the tile through LDS.  Whole: every thread
// writes its 16 registers, then reads 16 (the tile's 2^12 amplitudes in LDS
// at once).  Split on a tile bit h: the amplitudes with bit h = 0 go first,
// then those with h = 1, each half through the same 2^11 slots: the
// swizzled index with bit h removed.  That is a bijection on each half (the
// swizzle is linear and maps the unit vector of h to itself for h < 4 and
// keeps bit h for h >= 4, so two indices of one half never meet); for
// h >= 4 it keeps the bank pattern, for h < 4 lanes may meet in a bank.  h is
//  - a thread bit at the same position j in both layouts: the threads with
//    bit j = k move all their registers in round k (whole waves for j >= 6,
//    half the lanes for j < 6), or
//  - a register bit of both layouts: every thread moves the 8 register
//    groups with h = k in round k, and the groups read land in the
//    registers just written (a renaming).
struct Split {
    int kind = 0;  // 0 whole, 1 thread bit, 2 register bit
    int h = 0;     // the tile bit
    int j = 0;     // thread bit (kind 1)
    int kf = 0, kt = 0;  // register bit of h in the old and new layout (kind 2)
};

Split find_split(const FPhase& A, const FPhase& B) {
    Split r;
    auto thread_bit = [&](int j) {
        if (A.tbits[j] != B.tbits[j]) return false;
        r.kind = 1;
        r.h = A.tbits[j];
        r.j = j;
        return true;
    };
    if (thread_bit(7) || thread_bit(6)) return r;  // whole waves skip a round
    for (int kf = 3; kf >= 0; --kf)
        for (int kt = 0; kt < 4; ++kt)
            if (A.regbits[kf] == B.regbits[kt]) {
                r.kind = 2;
                r.h = A.regbits[kf];
                r.kf = kf;
                r.kt = kt;
                return r;
            }
    for (int j = 5; j >= 0; --j)  // half the lanes of every wave
        if (thread_bit(j)) return r;
    return r;
}

// a swizzled LDS index with bit h removed (as a slot index; the run-time
// form on a byte index, bit h at byte bit h + sh)
uint32_t drop_bit(uint32_t s, int h) { return ((s >> (h + 1)) << h) | (s & ((1u << h) - 1)); }
std::string drop_bit(const std::string& s, int h, int sh) {
    const int b = h + sh;
    return "(((" + s + " >> " + std::to_string(b + 1) + ") << " + std::to_string(b) + ") | (" + s + " & " +
           std::to_string((1u << b) - 1) + "u))";
}

// The LDS slot of a thread part (bytes, run time) XOR a register part
// (slot index, a constant).  The two parts share only the low nibble of the
// slot index (the swizzle), so the rest of the register part adds and goes
// into the instruction's offset: one XOR per access.
std::string lds_at(const std::string& tb, uint32_t reg, uint32_t es) {
    const uint32_t lo = (reg & 15u) * es, hi = (reg & ~15u) * es;
    std::string a = "smem_raw + ";
    a += lo ? "(" + tb + " ^ " + std::to_string(lo) + "u)" : tb;
    if (hi) a += " + " + std::to_string(hi) + "u";
    return "*reinterpret_cast<T*>(" + a + ")";
}

// An X on register bit k whose only control is a lane condition or a tile
// condition (bits outside the tile) commutes with the phase's ops on other
// register bits, so it can wait for the next op on k and fold into it: a
// plain 1-bit op on k becomes U X where the condition holds (a per-lane
// select of its 4 entries, or a uniform choice between two tables for a
// tile condition), and so does a plain 2-bit op for a tile condition.  The
// X then costs no register swap.  Returns, per op of the phase, the index
// of the op it folds into (-1: emitted as it is).
bool plain_op(const FOp& o) {
    return o.tmask == 0 && o.omask == 0 && (o.regok & 0xFFFFu) == 0xFFFFu;
}

std::vector<int> fold_targets(const FusedProgram& P, const FPhase& ph) {
    const int n = ph.op_end - ph.op_begin;
    std::vector<int> into(size_t(n), -1);
    std::vector<bool> taken(size_t(n), false);
    for (int i = 0; i < n; ++i) {
        const FOp& x = P.ops[size_t(ph.op_begin + i)];
        if (x.kind != OP_X || (x.regok & 0xFFFFu) != 0xFFFFu || (x.tmask != 0) == (x.omask != 0)) continue;
        for (int j = i + 1; j < n; ++j) {
            const FOp& y = P.ops[size_t(ph.op_begin + j)];
            if (!((op_bits(y) >> x.j0) & 1u)) continue;
            const bool ok = plain_op(y) && !taken[size_t(j)] &&
                            ((y.kind == OP_U1 && y.j0 == x.j0) || (y.kind == OP_U2 && x.omask != 0));
            if (ok) {
                into[size_t(i)] = j;
                taken[size_t(j)] = true;
            }
            break;
        }
    }
    return into;
}

void gen_remap(Gen& g, const FPhase& from, const FPhase& to, const std::string& sfrom,
               const std::string& sto, const Split& sp) {
    std::ostringstream& o = g.o;
    const uint32_t es = g.dbl ? 16 : 8;
    const int sh = g.dbl ? 4 : 3;
    if (sp.kind == 0) {
        for (int rho = 0; rho < kNR; ++rho)
            o << "    " << lds_at(sfrom, sreg(from, rho), es) << " = " << g.X(rho) << ";\n";
        o << "    __syncthreads();\n";
        g.identity();
        for (int rho = 0; rho < kNR; ++rho)
            o << "    x[" << rho << "] = " << lds_at(sto, sreg(to, rho), es) << ";\n";
        o << "    __syncthreads();\n";
        return;
    }
    o << "    { const unsigned cf = " << drop_bit(sfrom, sp.h, sh) << ", ct = " << drop_bit(sto, sp.h, sh) << ";\n";
    if (sp.kind == 1) {
        const int old[kNR] = {g.perm[0], g.perm[1], g.perm[2],  g.perm[3],  g.perm[4],  g.perm[5],
                              g.perm[6], g.perm[7], g.perm[8],  g.perm[9],  g.perm[10], g.perm[11],
                              g.perm[12], g.perm[13], g.perm[14], g.perm[15]};
        // a wave bit: the branch is uniform per wave (a scalar branch)
        if (sp.j >= 6)
            o << "    const unsigned hb = __builtin_amdgcn_readfirstlane((t >> " << sp.j << ") & 1u);\n";
        else
            o << "    const unsigned hb = (t >> " << sp.j << ") & 1u;\n";
        for (int k = 0; k < 2; ++k) {
            o << "    if (hb == " << k << "u) {\n";
            for (int rho = 0; rho < kNR; ++rho)
                o << "    " << lds_at("cf", drop_bit(sreg(from, rho), sp.h), es) << " = x[" << old[rho] << "];\n";
            o << "    }\n    __syncthreads();\n    if (hb == " << k << "u) {\n";
            for (int rho = 0; rho < kNR; ++rho)
                o << "    x[" << rho << "] = " << lds_at("ct", drop_bit(sreg(to, rho), sp.h), es) << ";\n";
            o << "    }\n    __syncthreads();\n";
        }
        g.identity();
    } else {
        int np[kNR];
        for (int k = 0; k < 2; ++k) {
            std::vector<int> freed;
            for (int rho = 0; rho < kNR; ++rho)
                if (((rho >> sp.kf) & 1) == k) {
                    o << "    " << lds_at("cf", drop_bit(sreg(from, rho), sp.h), es) << " = " << g.X(rho) << ";\n";
                    freed.push_back(g.perm[rho]);
                }
            o << "    __syncthreads();\n";
            size_t f = 0;
            for (int rho = 0; rho < kNR; ++rho)
                if (((rho >> sp.kt) & 1) == k) {
                    np[rho] = freed[f++];
                    o << "    x[" << np[rho] << "] = " << lds_at("ct", drop_bit(sreg(to, rho), sp.h), es) << ";\n";
                }
            o << "    __syncthreads();\n";
        }
        for (int rho = 0; rho < kNR; ++rho) g.perm[rho] = np[rho];
    }
    o << "    }\n";
}

// The layout a pass stores from when its last phase cannot store directly:
// the last phase's register bits above tile bit 1, completed with the
// highest tile bits that are not its wave bits, so lanes 0 and 1 sit on tile
// bits 0 and 1 (HBM runs of 64 B and more) and the change to it splits.
FPhase store_layout(const FusedProgram& P, const FPass& ps, const FPhase& last) {
    const uint32_t m = uint32_t(ps.m);
    std::vector<uint32_t> Sv(m);
    for (uint32_t i = 0; i < m; ++i) Sv[i] = ps.S[i];
    std::set<uint32_t> R;
    for (uint32_t k = 0; k < P.reg_bits; ++k)
        if (last.regbits[k] >= 2) R.insert(last.regbits[k]);
    for (int b = int(m) - 1; R.size() < P.reg_bits && b >= 2; --b)
        if (uint32_t(b) != last.tbits[6] && uint32_t(b) != last.tbits[7]) R.insert(uint32_t(b));
    for (int b = int(m) - 1; R.size() < P.reg_bits && b >= 2; --b) R.insert(uint32_t(b));
    return qs::layout_phase(R, m, P.reg_bits, Sv, 2, &last);
}

// Kernel of one pass.  Each workgroup loads a tile (the pass's first
// layout), runs the phases with their layout changes and stores the tile;
// the launch gives every tile its own workgroup (a loop over tiles would let
// the compiler hoist every change's LDS addresses out of it, into spills).
// When every layout change of a complex128 pass splits, the kernel needs
// 32 KiB of LDS instead of 64 and runs 4 workgroups per CU (at most 128
// VGPRs: 4 waves per SIMD).
void gen_pass(Gen& g, const FusedProgram& P, int pi, uint32_t* lds_bytes) {
    std::ostringstream& o = g.o;
    const FPass& ps = P.passes[pi];
    const int pb = ps.phase_begin, pe = ps.phase_end;
    const FPhase& ld = P.phases[pb];
    g.tab.clear();
    // the layout changes, in order, and how each moves the tile
    const FPhase& lastp = P.phases[pe - 1];
    const bool direct_store = lastp.coalesced != 0;
    const FPhase st = direct_store ? lastp : store_layout(P, ps, lastp);
    std::vector<Split> sp;
    for (int ph = pb + 1; ph < pe; ++ph) sp.push_back(find_split(P.phases[ph - 1], P.phases[ph]));
    if (!direct_store) sp.push_back(find_split(lastp, st));
    bool half = g.dbl;
    for (const Split& x : sp)
        if (x.kind == 0) half = false;
    if (!half)
        for (Split& x : sp) x.kind = 0;
    for (const Split& x : sp) ++g.stats[x.kind == 0 ? 0 : x.kind == 2 ? 3 : x.j >= 6 ? 1 : 2];
    if (half) ++g.stats[4];
    const size_t es = g.dbl ? 16 : 8;
    *lds_bytes = uint32_t((half ? es << (ps.m - 1) : es << ps.m));
    o << "extern \"C\" __global__ __launch_bounds__(256) ";
    if (half) o << "__attribute__((amdgpu_waves_per_eu(4))) ";
    o << "void qs_pass_" << (g.dbl ? "c128_" : "c64_") << pi
      << "(T* __restrict__ a, unsigned long long ntiles, const R* __restrict__ mtab) {\n";
    // static LDS: its base address is a constant the compiler folds away
    o << "    __shared__ __attribute__((aligned(16))) unsigned char smem_raw[" << *lds_bytes << "];\n";
    o << "    const unsigned t = threadIdx.x;\n";
    o << "    const unsigned long long g_ld = " << thr_g(ld) << ";\n";
    // the LDS thread parts of the phases that take part in a layout change
    for (int ph = pb; ph < pe; ++ph)
        o << "    const unsigned s_" << ph - pb << " = " << thr_s(P.phases[ph], uint32_t(es)) << ";\n";
    if (!direct_store) o << "    const unsigned s_st = " << thr_s(st, uint32_t(es)) << ";\n";
    o << "    T x[16];\n";
    o << "    const unsigned long long tile = blockIdx.x;\n";
    o << "    if (tile >= ntiles) return;\n";
    o << "    {\n";
    o << "    const unsigned long long base = " << tile_base(ps, "tile") << ";\n";
    o << "    { const T* p = a + (base | g_ld);\n";
    for (int rho = 0; rho < kNR; ++rho) o << "    x[" << rho << "] = ldnt(p + " << u64(greg(ld, rho)) << ");\n";
    o << "    }\n";
    o << "    int z;\n    asm volatile(\"s_mov_b32 %0, 0\" : \"=s\"(z));\n";
    o << "    const R* __restrict__ cm = mtab + z;\n";
    g.identity();
    size_t si = 0;
    for (int ph = pb; ph < pe; ++ph) {
        if (ph != pb) {
            gen_remap(g, P.phases[ph - 1], P.phases[ph], "s_" + std::to_string(ph - 1 - pb),
                      "s_" + std::to_string(ph - pb), sp[si++]);
        }
        const FPhase& f = P.phases[ph];
        const std::vector<int> into = fold_targets(P, f);
        std::vector<int> from(into.size(), -1);
        for (size_t i = 0; i < into.size(); ++i)
            if (into[i] >= 0) from[size_t(into[i])] = int(i);
        for (int i = 0; i < f.op_end - f.op_begin; ++i) {
            if (into[size_t(i)] >= 0) continue;  // folded into a later op
            const FOp* fx = from[size_t(i)] >= 0 ? &P.ops[size_t(f.op_begin + from[size_t(i)])] : nullptr;
            gen_op(g, P, P.ops[size_t(f.op_begin + i)], fx);
        }
    }
    if (!direct_store) gen_remap(g, lastp, st, "s_" + std::to_string(pe - 1 - pb), "s_st", sp[si++]);
    o << "    T* q = a + (base | " << thr_g(st) << ");\n";
    for (int rho = 0; rho < kNR; ++rho) o << "    stnt(q + " << u64(greg(st, rho)) << ", " << g.X(rho) << ");\n";
    o << "    }\n}\n\n";
}

const char* kPreamble = R"(
__device__ __forceinline__ T mk(R x, R y) { T r; r.x = x; r.y = y; return r; }
// tile loads and stores with the non-temporal hint: every amplitude is
// read and written once per pass (about 1.5% faster than plain accesses)
typedef R V2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ T ldnt(const T* p) {
    const V2 v = __builtin_nontemporal_load(reinterpret_cast<const V2*>(p));
    return mk(v.x, v.y);
}
__device__ __forceinline__ void stnt(T* p, T x) {
    V2 v;
    v.x = x.x;
    v.y = x.y;
    __builtin_nontemporal_store(v, reinterpret_cast<V2*>(p));
}
__device__ __forceinline__ T mac2(T u0, T a, T u1, T b) {
    T r;
    r.x = u0.x * a.x - u0.y * a.y + u1.x * b.x - u1.y * b.y;
    r.y = u0.x * a.y + u0.y * a.x + u1.x * b.y + u1.y * b.x;
    return r;
}
// one chain per real part: 1 multiply and 7 FMA
__device__ __forceinline__ T mac4(T u0, T a, T u1, T b, T u2, T c, T u3, T d) {
    T r;
    r.x = u0.x * a.x - u0.y * a.y + u1.x * b.x - u1.y * b.y + u2.x * c.x - u2.y * c.y + u3.x * d.x - u3.y * d.y;
    r.y = u0.x * a.y + u0.y * a.x + u1.x * b.y + u1.y * b.x + u2.x * c.y + u2.y * c.x + u3.x * d.y + u3.y * d.x;
    return r;
}
__device__ __forceinline__ T cmul(T u, T a) { return mk(u.x * a.x - u.y * a.y, u.x * a.y + u.y * a.x); }
__device__ __forceinline__ T sel(bool c, T a, T b) { return mk(c ? a.x : b.x, c ? a.y : b.y); }
// the sign bit of v XOR bit 31 of m (a factor -1 where m has it)
__device__ __forceinline__ double flip(double v, unsigned m) {
    return __longlong_as_double(__double_as_longlong(v) ^ ((long long)(unsigned long long)m << 32));
}
__device__ __forceinline__ float flip(float v, unsigned m) {
    return __int_as_float(__float_as_int(v) ^ (int)m);
}
)";

}  // namespace

namespace qs {

// The program as HIP source: a preamble, then one kernel per pass (pass_src
// has each pass's kernel alone, for compiling groups of passes apart).
std::string native_source(const FusedProgram& P, qs_dtype dt, std::vector<double>* tab,
                          std::vector<uint64_t>* off, std::vector<std::string>* pass_src,
                          std::string* preamble, std::vector<uint32_t>* lds, uint64_t* stats) {
    Gen g;
    g.dbl = dt == QS_C128;
    std::string pre = g.dbl ? "typedef double2 T;\ntypedef double R;\n" : "typedef float2 T;\ntypedef float R;\n";
    pre += kPreamble;
    pre += "\n";
    if (preamble) *preamble = pre;
    if (tab) tab->clear();
    if (off) off->clear();
    if (pass_src) pass_src->clear();
    if (lds) lds->clear();
    std::string all = pre;
    for (size_t p = 0; p < P.passes.size(); ++p) {
        g.o.str("");
        uint32_t lb = 0;
        gen_pass(g, P, int(p), &lb);
        if (lds) lds->push_back(lb);
        const std::string k = g.o.str();
        all += k;
        if (pass_src) pass_src->push_back(k);
        if (off) off->push_back(tab ? tab->size() : 0);
        if (tab) tab->insert(tab->end(), g.tab.begin(), g.tab.end());
    }
    if (stats)
        for (int i = 0; i < 5; ++i) stats[i] = g.stats[i];
    return all;
}

}  // namespace qs

namespace {

// code objects of compiled sources, by the source text (a program compiled
// again -- the same circuit and options -- is not recompiled)
std::mutex g_code_mu;
std::map<std::string, std::vector<char>> g_code;

// compiles one source with hipRTC for gfx950; "" on success, else the log
std::string rtc_compile(const std::string& src, std::vector<char>& code) {
    {
        std::lock_guard<std::mutex> lk(g_code_mu);
        auto it = g_code.find(src);
        if (it != g_code.end()) {
            code = it->second;
            return "";
        }
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "qs_native.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        return "hiprtcCreateProgram failed";
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops"};
    if (hiprtcCompileProgram(prog, 6, opts) != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls, '\0');
        if (ls) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        return log.empty() ? std::string("compile failed") : log.substr(0, 2000);
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    std::lock_guard<std::mutex> lk(g_code_mu);
    g_code.emplace(src, code);
    return "";
}

}  // namespace

namespace qs {

// Generates and compiles the program's pass kernels, in groups of passes
// compiled in parallel threads; P.native_* hold the modules, one function
// per pass and the constant table afterwards.
int native_build(FusedProgram& P, qs_dtype dt) {
    constexpr const char* W = "qpp_amd::program";
    if (P.reg_bits != 4) return fail(QS_ERR_NOT_IMPLEMENTED, W, "native passes need reg_bits 4");
    for (const FPass& ps : P.passes)
        if (ps.m != 12) return fail(QS_ERR_NOT_IMPLEMENTED, W, "native passes need 2^12-amplitude tiles");
    if (int rc = ensure_device()) return rc;
    std::vector<double> tab;
    std::vector<uint64_t> off;
    std::vector<std::string> psrc;
    std::string pre;
    std::vector<uint32_t> lds;
    uint64_t stats[5];
    const std::string whole = native_source(native_fuse(P), dt, &tab, &off, &psrc, &pre, &lds, stats);
    // QPP_AMD_NATIVE_DUMP=<file>: write the generated source there (for
    // reading the kernels; appended, one program after another)
    if (const char* dump = std::getenv("QPP_AMD_NATIVE_DUMP")) {
        if (FILE* f = std::fopen(dump, "a")) {
            std::fputs(whole.c_str(), f);
            std::fclose(f);
        }
    }
    // groups of passes of about equal source size, one per thread
    const size_t np = P.passes.size();
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t ng = std::max<size_t>(1, std::min<size_t>({np, size_t(hw), 16}));
    std::vector<std::string> gsrc(ng, pre);
    std::vector<size_t> gsize(ng, 0), group_of(np, 0);
    std::vector<size_t> order(np);
    for (size_t p = 0; p < np; ++p) order[p] = p;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return psrc[a].size() > psrc[b].size(); });
    for (size_t p : order) {
        const size_t gi = size_t(std::min_element(gsize.begin(), gsize.end()) - gsize.begin());
        gsrc[gi] += psrc[p];
        gsize[gi] += psrc[p].size();
        group_of[p] = gi;
    }
    std::vector<std::vector<char>> code(ng);
    std::vector<std::string> err(ng);
    {
        std::vector<std::thread> th;
        for (size_t gi = 0; gi < ng; ++gi)
            th.emplace_back([&, gi] { err[gi] = rtc_compile(gsrc[gi], code[gi]); });
        for (auto& t : th) t.join();
    }
    for (size_t gi = 0; gi < ng; ++gi)
        if (!err[gi].empty())
            return fail_runtime(QS_ERR_DEVICE, "qpp_amd::program: native pass compile failed: " + err[gi]);
    std::vector<void*> mods(ng, nullptr);
    auto unload = [&] {
        for (void* m : mods)
            if (m) (void)hipModuleUnload(reinterpret_cast<hipModule_t>(m));
    };
    for (size_t gi = 0; gi < ng; ++gi) {
        hipModule_t mod = nullptr;
        const hipError_t e = hipModuleLoadData(&mod, code[gi].data());
        if (e != hipSuccess) {
            unload();
            return fail_hip(e, W);
        }
        mods[gi] = reinterpret_cast<void*>(mod);
    }
    std::vector<void*> fns(np);
    for (size_t p = 0; p < np; ++p) {
        const std::string name = std::string(dt == QS_C128 ? "qs_pass_c128_" : "qs_pass_c64_") + std::to_string(p);
        hipFunction_t f = nullptr;
        const hipError_t e = hipModuleGetFunction(&f, reinterpret_cast<hipModule_t>(mods[group_of[p]]), name.c_str());
        if (e != hipSuccess) {
            unload();
            return fail_hip(e, W);
        }
        fns[p] = reinterpret_cast<void*>(f);
    }
    // the constants, in the state's precision
    const size_t es = dt == QS_C128 ? 8 : 4;
    std::vector<unsigned char> hb(std::max<size_t>(1, tab.size()) * es);
    for (size_t i = 0; i < tab.size(); ++i) {
        if (es == 8) {
            std::memcpy(hb.data() + 8 * i, &tab[i], 8);
        } else {
            const float f = static_cast<float>(tab[i]);
            std::memcpy(hb.data() + 4 * i, &f, 4);
        }
    }
    void* dtab = nullptr;
    hipError_t e = hipMalloc(&dtab, hb.size());
    if (e == hipSuccess) e = hipMemcpy(dtab, hb.data(), hb.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (dtab) (void)hipFree(dtab);
        unload();
        return fail_hip(e, W);
    }
    native_free(P);
    P.native_modules = std::move(mods);
    P.native_fn = std::move(fns);
    P.native_dtype = dt;
    P.native_tab = dtab;
    P.native_off = std::move(off);
    P.native_lds = std::move(lds);
    for (int i = 0; i < 5; ++i) P.native_stats[i] = stats[i];
    return QS_OK;
}

void native_free(FusedProgram& P) {
    if (!P.native_modules.empty()) {
        if (P.last) (void)hipStreamSynchronize(P.last);
        if (P.stream) (void)hipStreamSynchronize(P.stream);
        for (void* m : P.native_modules) (void)hipModuleUnload(reinterpret_cast<hipModule_t>(m));
    }
    if (P.native_tab) (void)hipFree(P.native_tab);
    P.native_modules.clear();
    P.native_tab = nullptr;
    P.native_fn.clear();
    P.native_off.clear();
    P.native_lds.clear();
    for (uint64_t& c : P.native_stats) c = 0;
}

// One pass through its specialised kernel: a workgroup per tile (grid_cap is
// for k_tile's persistent launch; here a finished workgroup's CU slot takes
// the next tile's workgroup, which measured faster than persistent loops).
int native_launch(const FusedProgram& P, qs_state* st, int pass, uint32_t grid_cap) {
    (void)grid_cap;
    const FPass& ps = P.passes[pass];
    const size_t es = st->dtype == QS_C128 ? 16 : 8;
    const size_t shmem = 0;  // the kernels declare their LDS statically
    if (ps.ntiles > 0xffffffffull)
        return fail(QS_ERR_NOT_IMPLEMENTED, "qpp_amd::fused", "more than 2^32 tiles in one pass");
    const uint64_t grid = ps.ntiles;
    void* a = st->d;
    unsigned long long nt = ps.ntiles;
    void* tab = static_cast<unsigned char*>(P.native_tab) + P.native_off[pass] * (es / 2);
    void* args[] = {&a, &nt, &tab};
    QS_HIP(hipModuleLaunchKernel(reinterpret_cast<hipFunction_t>(P.native_fn[pass]),
                                 static_cast<unsigned>(grid), 1, 1, 256, 1, 1,
                                 static_cast<unsigned>(shmem), st->stream, args, nullptr),
           "qpp_amd::fused");
    return QS_OK;
}

}  // namespace qs

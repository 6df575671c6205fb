// qs_fused.hip -- tile-fused gate passes (the circuit fast path).
//
// A circuit is compiled on the host into passes.  A pass owns a set S of m
// index bits (the "tile bits": the low L bits of the element index, kept for
// coalescing, plus the target bits of the pass's gates).  A workgroup of 256
// threads holds one tile -- the 2^m elements that share all bits outside S --
// in registers (NR = 2^m / 256 per thread), applies every gate of the pass,
// and writes the tile back: the state streams through HBM once per pass.
// Workgroups are persistent and fetch the next tile into registers while
// they compute the current one.
//
// Inside a pass the gates are grouped into phases.  In a phase, r = log2(NR)
// tile bits are register bits (the NR elements of a thread differ only in
// them) and the other 8 tile bits index the thread.  Gates on register bits
// run in VALU without data exchange; changing the register bits is one round
// trip through LDS.  The host precomputes every uniform quantity of a phase
// and an op (global offsets, swizzled LDS offsets, the split of a control
// into register / thread / tile parts), so per element the kernel spends
// only the gate arithmetic, one XOR per LDS access and one add per global
// access.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstring>
#include <map>
#include <set>
#include <vector>

#include "qs_fused.hpp"
#include "qs_internal.hpp"
#include "qs_tile_ops.hpp"

namespace {

using qs::fail;

// One pass over all tiles of the state.
template <typename R, int NR>
__global__ __launch_bounds__(kThreads) void k_tile(cxt<R>* __restrict__ a,
                                                   const FPass* __restrict__ passes,
                                                   const FPhase* __restrict__ phases,
                                                   const FOp* __restrict__ ops,
                                                   const cxt<R>* __restrict__ mats,
                                                   int pass_id) {
    extern __shared__ unsigned char smem_raw[];
    cxt<R>* lds = reinterpret_cast<cxt<R>*>(smem_raw);
    const FPass* __restrict__ ps = passes + pass_id;
    const uint32_t t = threadIdx.x;
    constexpr int RB = RBof<NR>::v;
    const int pb = ps->phase_begin, pe = ps->phase_end;
    const uint64_t ntiles = ps->ntiles;
    const FPhase* ld = phases + pb;  // load layout (a phase the host marked coalesced)
    const uint64_t gt_ld = thr_g(ld, t);
    auto tile_base = [&](uint64_t tile) { return tile_base_of(ps, tile); };
    cxt<R> x[NR], y[NR], z[NR];
    uint64_t tile = blockIdx.x;
    if (tile < ntiles) {
        const cxt<R>* p = a + (tile_base(tile) | gt_ld);
#pragma unroll
        for (int rho = 0; rho < NR; ++rho) y[rho] = p[greg<RB>(ld, rho)];
    }
    // Wait for the first tile here, once.  Without this the loop head would
    // carry the waits for these loads, and since one vector-memory counter
    // covers loads and stores, every later iteration would stall there until
    // the previous tile's stores had completed.
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    for (; tile < ntiles; tile += gridDim.x) {
        const uint64_t base = tile_base(tile);
#pragma unroll
        for (int rho = 0; rho < NR; ++rho) x[rho] = y[rho];
        // software pipeline: fetch the next tile while computing this one
        const uint64_t next = tile + gridDim.x;
        if (next < ntiles) {
            const cxt<R>* p = a + (tile_base(next) | gt_ld);
#pragma unroll
            for (int rho = 0; rho < NR; ++rho) y[rho] = p[greg<RB>(ld, rho)];
        }
        // the current layout is carried as a phase index; readfirstlane makes
        // its uniformity explicit so its tables are read with scalar loads
        int ci = pb;
        for (int pi = pb; pi < pe; ++pi) {
            const FPhase* ph = phases + pi;
            if (pi != ci) {
                // remap through LDS: write with the current layout, read with
                // the new one (index = swizzled thread part ^ register part)
                const FPhase* cur = phases + __builtin_amdgcn_readfirstlane(ci);
                const uint32_t sw = thr_s(cur, t);
#pragma unroll
                for (int rho = 0; rho < NR; ++rho) lds[sw ^ sreg<RB>(cur, rho)] = x[rho];
                __syncthreads();
                const uint32_t sr = thr_s(ph, t);
#pragma unroll
                for (int rho = 0; rho < NR; ++rho) x[rho] = lds[sr ^ sreg<RB>(ph, rho)];
                __syncthreads();
                ci = pi;
            }
            // consecutive ops alternate between x and z (see run_op)
            const int ob = ph->op_begin, oe = ph->op_end;
            int o = ob;
            for (; o + 1 < oe; o += 2) {
                run_op<R, NR>(x, z, ops + o, ops[o].m, mats, t, base);
                run_op<R, NR>(z, x, ops + o + 1, ops[o + 1].m, mats, t, base);
            }
            if (o < oe) {
                run_op<R, NR>(x, z, ops + o, ops[o].m, mats, t, base);
                copy_regs<R, NR>(z, x);
            }
        }
        ci = __builtin_amdgcn_readfirstlane(ci);
        const FPhase* cur = phases + ci;
        if (!cur->coalesced) {
            const uint32_t sw = thr_s(cur, t);
#pragma unroll
            for (int rho = 0; rho < NR; ++rho) lds[sw ^ sreg<RB>(cur, rho)] = x[rho];
            __syncthreads();
            const uint32_t sr = thr_s(ld, t);
#pragma unroll
            for (int rho = 0; rho < NR; ++rho) x[rho] = lds[sr ^ sreg<RB>(ld, rho)];
            __syncthreads();
            cur = ld;
        }
        cxt<R>* p = a + (base | (cur == ld ? gt_ld : thr_g(cur, t)));
#pragma unroll
        for (int rho = 0; rho < NR; ++rho) p[greg<RB>(cur, rho)] = x[rho];
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host: compiler
// ---------------------------------------------------------------------------
namespace qs {

namespace {

using C = std::complex<double>;

// LDS swizzle: the low nibble of the tile index is XOR-ed with the two
// nibbles above it, so lanes that differ only in high tile bits still hit
// different banks.  It is linear over XOR, which lets the kernel combine the
// thread part and the register part of an index with one XOR.
uint32_t swz(uint32_t e) { return e ^ (((e >> 4) ^ (e >> 8)) & 15u); }

// 4x4 column-major with gate digits (hi, lo) swapped
std::vector<C> swap_digits_4x4(const std::vector<C>& m) {
    auto sw = [](int i) { return ((i & 1) << 1) | ((i >> 1) & 1); };
    std::vector<C> r(16);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r[sw(i) + 4 * sw(j)] = m[i + 4 * j];
    return r;
}

bool is_x(const std::vector<C>& m) {
    return m.size() == 4 && m[0] == C(0, 0) && m[3] == C(0, 0) && m[1] == C(1, 0) &&
           m[2] == C(1, 0);
}

bool is_diag(const std::vector<C>& m) {
    return m.size() == 4 && m[1] == C(0, 0) && m[2] == C(0, 0);
}

// Greedy capacity scheduling: walks `items` in order and takes every item
// whose bits fit (with the bits already taken) into `cap` bits and that does
// not touch a blocked qubit; items that do not fit block all their qubits so
// later items on them keep their order.  Returns the indices taken.
struct Item {
    std::vector<uint32_t> need;   // bits that must be in the set
    std::vector<uint32_t> touch;  // every bit the item depends on
};

std::vector<size_t> greedy(const std::vector<Item>& items, const std::vector<size_t>& todo,
                           std::set<uint32_t>& taken, size_t cap, size_t max_take) {
    std::set<uint32_t> blocked;
    std::vector<size_t> out;
    for (size_t id : todo) {
        const Item& it = items[id];
        bool blk = false;
        for (uint32_t b : it.touch)
            if (blocked.count(b)) blk = true;
        if (!blk && out.size() < max_take) {
            std::set<uint32_t> u = taken;
            for (uint32_t b : it.need) u.insert(b);
            if (u.size() <= cap) {
                taken = u;
                out.push_back(id);
                continue;
            }
        }
        for (uint32_t b : it.touch) blocked.insert(b);
    }
    return out;
}

// The gates a pass over the fixed tile-bit set S (a bit mask) takes: in
// order, every gate whose needed bits lie in S and that touches no blocked
// bit; a gate that cannot be taken blocks every bit it touches.
size_t take_count(const std::vector<uint64_t>& need, const std::vector<uint64_t>& touch,
                  const std::vector<size_t>& todo, uint64_t S, size_t max_take) {
    uint64_t blocked = 0;
    size_t n = 0;
    for (size_t id : todo) {
        if (!(touch[id] & blocked) && !(need[id] & ~S) && n < max_take) {
            ++n;
            continue;
        }
        blocked |= touch[id];
    }
    return n;
}

// Improves a pass's tile-bit set by hill climbing: swap one of its free bits
// (not among the `fixed` low bits) for a bit outside it while that lets the
// pass take more gates.
uint64_t improve_set(const std::vector<uint64_t>& need, const std::vector<uint64_t>& touch,
                     const std::vector<size_t>& todo, uint64_t S, uint64_t fixed, uint32_t nb,
                     size_t max_take) {
    size_t best = take_count(need, touch, todo, S, max_take);
    for (bool improved = true; improved && best < max_take;) {
        improved = false;
        for (uint32_t b = 0; b < nb && !improved; ++b) {
            if (!((S >> b) & 1) || ((fixed >> b) & 1)) continue;
            for (uint32_t c = 0; c < nb && !improved; ++c) {
                if ((S >> c) & 1) continue;
                const uint64_t S2 = (S & ~(uint64_t{1} << b)) | (uint64_t{1} << c);
                const size_t v = take_count(need, touch, todo, S2, max_take);
                if (v > best) {
                    best = v;
                    S = S2;
                    improved = true;
                }
            }
        }
    }
    return S;
}

// Layout of a phase: register tile bits R (ascending; register bit k is
// R[k]); the thread number enumerates the other tile bits (order below).
// The phase loads and stores HBM directly ("coalesced") when no
// register bit is below `lc`: lanes 0..lc-1 then sit on tile bits 0..lc-1,
// which are global bits 0..lc-1 (lc <= low_bits), so every 2^lc lanes access
// 16 * 2^lc contiguous bytes.
FPhase make_phase(const std::set<uint32_t>& Rset, uint32_t m, uint32_t RB,
                  const std::vector<uint32_t>& Sv, uint32_t lc, const FPhase* hint) {
    FPhase ph{};
    std::vector<uint32_t> Rv(Rset.begin(), Rset.end());
    ph.coalesced = 1;
    for (uint32_t k = 0; k < RB; ++k) {
        ph.regbits[k] = static_cast<uint8_t>(Rv[k]);
        if (Rv[k] < lc) ph.coalesced = 0;
    }
    const uint32_t NR = 1u << RB;
    for (uint32_t rho = 0; rho < NR; ++rho) {
        uint64_t g = 0;
        uint32_t e = 0;
        for (uint32_t k = 0; k < RB; ++k)
            if ((rho >> k) & 1) {
                g |= uint64_t{1} << Sv[Rv[k]];
                e |= 1u << Rv[k];
            }
        ph.g_reg[rho] = g;
        ph.s_reg[rho] = swz(e);
        for (uint32_t k = 0; k < RB; ++k)
            if (rho == (1u << k)) {
                ph.g_rb[k] = g;
                ph.s_rb[k] = swz(e);
            }
    }
    // Thread bits: lanes 0..3 take tile bits with residues 0, 1, 2, 3 mod 4
    // when there are such bits (lowest first).  A tile bit b moves the
    // swizzled LDS index in its low nibble bit b mod 4, so 16 lanes that
    // differ in four bits of distinct residues touch 16 distinct 16-byte
    // slots of a 256-byte LDS row (no bank conflicts); the lowest-first
    // choice keeps tile bits 0..1 on lanes 0..1 whenever they are free,
    // which direct HBM access needs.  The other tile bits follow in
    // ascending order, except that a tile bit which sits at thread bit j >= 4
    // of the neighbouring phase `hint` and is a thread bit here too keeps
    // position j: a layout change between the two phases can then move the
    // tile in two halves split on that bit (whole waves or lanes skipping a
    // half), through half the LDS.
    std::vector<uint32_t> rest;
    for (uint32_t b = 0; b < m; ++b)
        if (!Rset.count(b)) rest.push_back(b);
    std::vector<uint32_t> order;
    for (uint32_t res = 0; res < 4; ++res) {
        auto it = std::find_if(rest.begin(), rest.end(), [&](uint32_t b) { return b % 4 == res; });
        if (it == rest.end()) break;  // keep lanes contiguous in residue order
        order.push_back(*it);
        rest.erase(it);
    }
    const size_t nt = order.size() + rest.size();
    std::vector<int> fixed(nt, -1);
    if (hint)
        for (size_t j = std::max<size_t>(4, order.size()); j < nt && j < 8; ++j) {
            auto it = std::find(rest.begin(), rest.end(), uint32_t(hint->tbits[j]));
            if (it == rest.end()) continue;
            fixed[j] = int(*it);
            rest.erase(it);
        }
    for (size_t j = order.size(), r = 0; j < nt; ++j) {
        if (fixed[j] >= 0) {
            order.push_back(uint32_t(fixed[j]));
        } else {
            order.push_back(rest[r++]);
        }
    }
    for (uint32_t j = 0; j < order.size(); ++j) {
        ph.tbits[j] = static_cast<uint8_t>(order[j]);
        ph.g_thr[j] = uint64_t{1} << Sv[order[j]];
        ph.s_thr[j] = swz(1u << order[j]);
    }
    return ph;
}

// Lowers one gate into an op of the given phase layout.
FOp lower_op(const BitGate& g, const FPhase& ph, uint32_t RB, const std::vector<uint32_t>& Sv,
             std::vector<C>& mats) {
    FOp op{};
    // where each global bit of the tile lives in this layout
    std::map<uint32_t, int> reg_of, thr_of;
    for (uint32_t k = 0; k < RB; ++k) reg_of[Sv[ph.regbits[k]]] = static_cast<int>(k);
    for (uint32_t j = 0; j < 8; ++j) thr_of[Sv[ph.tbits[j]]] = static_cast<int>(j);
    const uint32_t NR = 1u << RB;
    // control split
    uint32_t regok = 0;
    for (uint32_t rho = 0; rho < NR; ++rho) {
        bool ok = true;
        for (auto& kv : reg_of) {
            const uint64_t bit = uint64_t{1} << kv.first;
            if (!(g.cmask & bit)) continue;
            const bool want = (g.cval & bit) != 0;
            if ((((rho >> kv.second) & 1) != 0) != want) ok = false;
        }
        if (ok) regok |= 1u << rho;
    }
    op.regok = regok;
    for (auto& kv : thr_of) {
        const uint64_t bit = uint64_t{1} << kv.first;
        if (!(g.cmask & bit)) continue;
        op.tmask |= 1u << kv.second;
        if (g.cval & bit) op.tval |= 1u << kv.second;
    }
    uint64_t in_tile = 0;
    for (uint32_t s : Sv) in_tile |= uint64_t{1} << s;
    op.omask = g.cmask & ~in_tile;
    op.oval = g.cval & op.omask;
    op.ctrl = (regok != (NR == 32 ? 0xFFFFFFFFu : ((1u << NR) - 1))) || op.tmask != 0;
    auto put = [&](int i, C v) {
        op.m[2 * i] = v.real();
        op.m[2 * i + 1] = v.imag();
    };
    if (g.bits.size() == 1 && is_diag(g.m)) {
        const uint32_t b = g.bits[0];
        put(0, g.m[0]);
        put(1, g.m[3]);
        if (reg_of.count(b)) {
            op.kind = OP_DIAG_R;
            op.j0 = reg_of[b];
        } else if (thr_of.count(b)) {
            op.kind = OP_DIAG_T;
            op.tbit = 1u << thr_of[b];
        } else {
            op.kind = OP_DIAG_O;
            op.dbit = uint64_t{1} << b;
        }
    } else if (g.bits.size() == 1) {
        op.j0 = reg_of.at(g.bits[0]);
        if (is_x(g.m)) {
            op.kind = OP_X;
        } else {
            op.kind = OP_U1;
            for (int i = 0; i < 4; ++i) put(i, g.m[i]);
        }
    } else {
        // bits[0] is the MSB of the gate index; kernel: LSB on j0 < j1
        const int r0 = reg_of.at(g.bits[0]), r1 = reg_of.at(g.bits[1]);
        op.kind = OP_U2;
        op.moff = static_cast<uint32_t>(mats.size());
        std::vector<C> mm = g.m;
        if (r1 < r0) {
            op.j0 = r1;
            op.j1 = r0;
        } else {
            op.j0 = r0;
            op.j1 = r1;
            mm = swap_digits_4x4(mm);
        }
        mats.insert(mats.end(), mm.begin(), mm.end());
    }
    return op;
}

// Replaces each run of >= 2 consecutive uncontrolled dense 1-bit ops on
// distinct register bits of a phase by one OP_U1B op (ops on distinct bits
// commute, so their order inside the run is free).  The gate -> op map
// no longer points at bundled gates.
void bundle_phase(FusedProgram& P, FPhase& ph, uint32_t RB, std::vector<C>& mats) {
    auto plain_u1 = [](const FOp& o) {
        return o.kind == OP_U1 && !o.ctrl && o.omask == 0;
    };
    std::vector<FOp> out;
    std::vector<int> remap(P.ops.size() - ph.op_begin, -1);
    int i = ph.op_begin;
    const int e = ph.op_end;
    while (i < e) {
        if (!plain_u1(P.ops[i])) {
            remap[i - ph.op_begin] = ph.op_begin + int(out.size());
            out.push_back(P.ops[i++]);
            continue;
        }
        int j = i;
        uint32_t mask = 0;
        while (j < e && plain_u1(P.ops[j]) && !((mask >> P.ops[j].j0) & 1u)) {
            mask |= 1u << P.ops[j].j0;
            ++j;
        }
        if (mask != (1u << RB) - 1) {  // bundles cover every register bit
            remap[i - ph.op_begin] = ph.op_begin + int(out.size());
            out.push_back(P.ops[i++]);
            continue;
        }
        FOp b{};
        b.kind = OP_U1B;
        b.tbit = mask;
        b.regok = P.ops[i].regok;
        b.moff = static_cast<uint32_t>(mats.size());
        for (uint32_t k = 0; k < 4; ++k) {
            C m[4] = {C(1, 0), C(0, 0), C(0, 0), C(1, 0)};
            for (int q = i; q < j; ++q)
                if (uint32_t(P.ops[q].j0) == k)
                    for (int t = 0; t < 4; ++t)
                        m[t] = C(P.ops[q].m[2 * t], P.ops[q].m[2 * t + 1]);
            mats.insert(mats.end(), m, m + 4);
        }
        out.push_back(b);
        i = j;
    }
    P.ops.resize(ph.op_begin);
    P.ops.insert(P.ops.end(), out.begin(), out.end());
    ph.op_end = static_cast<int>(P.ops.size());
    for (int& g : P.gate_op)
        if (g >= ph.op_begin) g = g < e ? remap[g - ph.op_begin] : -1;
}

}  // namespace

FPhase layout_phase(const std::set<uint32_t>& R, uint32_t m, uint32_t RB,
                    const std::vector<uint32_t>& Sv, uint32_t lc, const FPhase* hint) {
    return make_phase(R, m, RB, Sv, lc, hint);
}

int fused_compile(const std::vector<BitGate>& gates, uint32_t nb, const FusedOptions& opt,
                  FusedProgram& P) {
    P = FusedProgram{};
    P.nb = nb;
    const uint32_t RB = opt.reg_bits;  // 3 or 4
    const uint32_t m = 8 + RB;
    if (nb < m) return fail(QS_ERR_NOT_IMPLEMENTED, "qpp_amd::fused", "state below tile size");
    const uint32_t L = std::min<uint32_t>(opt.low_bits, m);
    // Any layout whose lanes keep tile bits 0..1 loads and stores HBM
    // directly (64-B runs): a wave reads the rest of each 128-B line through
    // its other registers in the same burst, and the halves meet in L2.
    // Measured on the 30-qubit depth-40 circuit, requiring longer runs
    // (up to 1 KiB) only adds layout changes: 823 ms per circuit with 64-B
    // runs against 881 ms with 1-KiB runs (low_bits 3, max_ops 48).
    const uint32_t lc = std::min<uint32_t>(L, 2);
    // scheduling items: the bits a gate needs in the tile (none for a
    // diagonal 1-bit gate) and every bit it depends on
    std::vector<Item> items(gates.size());
    for (size_t i = 0; i < gates.size(); ++i) {
        const BitGate& g = gates[i];
        Item& it = items[i];
        const bool diag = g.bits.size() == 1 && is_diag(g.m);
        if (!diag) it.need = g.bits;
        it.touch = g.bits;
        for (uint32_t b = 0; b < 64; ++b)
            if ((g.cmask >> b) & 1) it.touch.push_back(b);
    }
    std::vector<size_t> todo(gates.size());
    for (size_t i = 0; i < todo.size(); ++i) todo[i] = i;
    P.gate_op.assign(gates.size(), -1);
    std::vector<uint64_t> need_m(items.size()), touch_m(items.size());
    for (size_t i = 0; i < items.size(); ++i) {
        for (uint32_t b : items[i].need) need_m[i] |= uint64_t{1} << b;
        for (uint32_t b : items[i].touch) touch_m[i] |= uint64_t{1} << b;
    }
    const uint64_t low_mask = L >= 64 ? ~uint64_t{0} : (uint64_t{1} << L) - 1;
    std::vector<C> mats;
    while (!todo.empty()) {
        std::set<uint32_t> S;
        for (uint32_t b = 0; b < L; ++b) S.insert(b);
        std::vector<size_t> take = greedy(items, todo, S, m, opt.max_ops);
        if (take.empty()) return fail(QS_ERR_NOT_IMPLEMENTED, "qpp_amd::fused", "gate wider than a tile");
        // fill S up to m bits with the lowest unused bits
        for (uint32_t b = 0; S.size() < m && b < nb; ++b) S.insert(b);
        if (opt.improve_sets) {
            uint64_t Sm = 0;
            for (uint32_t b : S) Sm |= uint64_t{1} << b;
            const uint64_t S2 = improve_set(need_m, touch_m, todo, Sm, low_mask, nb, opt.max_ops);
            if (S2 != Sm) {
                S.clear();
                for (uint32_t b = 0; b < nb; ++b)
                    if ((S2 >> b) & 1) S.insert(b);
                std::set<uint32_t> full = S;
                take = greedy(items, todo, full, m, opt.max_ops);
            }
        }
        FPass ps{};
        ps.m = static_cast<int>(m);
        std::vector<uint32_t> Sv(S.begin(), S.end());
        for (uint32_t i = 0; i < m; ++i) ps.S[i] = static_cast<uint8_t>(Sv[i]);
        std::vector<uint32_t> rest;
        for (uint32_t b = 0; b < nb; ++b)
            if (!S.count(b)) rest.push_back(b);
        ps.nrest = static_cast<int>(rest.size());
        for (size_t i = 0; i < rest.size(); ++i) ps.rest[i] = static_cast<uint8_t>(rest[i]);
        ps.nruns = 0;
        for (size_t i = 0; i < rest.size();) {
            size_t j = i + 1;
            while (j < rest.size() && rest[j] == rest[j - 1] + 1) ++j;
            ps.runs[ps.nruns++] = uint32_t(i) | (rest[i] << 8) | (uint32_t(j - i) << 16);
            i = j;
        }
        ps.ntiles = uint64_t{1} << rest.size();
        auto tile_bit = [&](uint32_t g) {
            return static_cast<uint32_t>(std::find(Sv.begin(), Sv.end(), g) - Sv.begin());
        };
        // phases: greedy over the pass's gates with capacity RB register bits
        std::vector<Item> pit(take.size());
        for (size_t i = 0; i < take.size(); ++i) {
            const Item& it = items[take[i]];
            for (uint32_t b : it.need) pit[i].need.push_back(tile_bit(b));
            pit[i].touch = it.touch;  // dependencies on global bits
        }
        std::vector<size_t> ptodo(take.size());
        for (size_t i = 0; i < ptodo.size(); ++i) ptodo[i] = i;
        ps.phase_begin = static_cast<int>(P.phases.size());
        // When a phase's register bits avoid tile bits 0..5 its lanes sit on
        // tile bits 0..5 and HBM access coalesces: the first phase can load
        // and the last can store without an LDS round trip.  Otherwise a
        // load phase (top RB tile bits, no ops) comes first and the kernel
        // returns to it before storing.
        bool first = true;
        std::vector<uint64_t> pneed(pit.size()), ptouch(pit.size());
        for (size_t i = 0; i < pit.size(); ++i) {
            for (uint32_t b : pit[i].need) pneed[i] |= uint64_t{1} << b;
            for (uint32_t b : pit[i].touch) ptouch[i] |= uint64_t{1} << b;
        }
        while (!ptodo.empty()) {
            std::set<uint32_t> R;
            std::vector<size_t> ptake = greedy(pit, ptodo, R, RB, SIZE_MAX);
            // complete R with unused tile bits, preferring high ones
            for (int b = int(m) - 1; R.size() < RB && b >= 0; --b) R.insert(uint32_t(b));
            if (opt.improve_sets) {
                // hill climbing on the register bits, as for the passes' tile
                // bits: swap one for another tile bit while the phase then
                // takes more gates (tile bits below lc stay out, so a
                // coalesced phase stays coalesced)
                uint64_t Rm = 0;
                for (uint32_t b : R) Rm |= uint64_t{1} << b;
                // gates taken, then (on ties) gates on tile bits below lc
                // taken: a phase that does them early leaves a coalesced
                // layout for the pass's store more often
                const uint64_t lowm = (uint64_t{1} << lc) - 1;
                auto score = [&](uint64_t Rx) {
                    uint64_t blocked = 0;
                    size_t n = 0, low = 0;
                    for (size_t id : ptodo) {
                        if (!(ptouch[id] & blocked) && !(pneed[id] & ~Rx)) {
                            ++n;
                            if (pneed[id] & lowm) ++low;
                            continue;
                        }
                        blocked |= ptouch[id];
                    }
                    return n * 1024 + low;
                };
                size_t best = score(Rm);
                for (bool better = true; better;) {
                    better = false;
                    for (uint32_t a0 = 0; a0 < m && !better; ++a0) {
                        if (!((Rm >> a0) & 1)) continue;
                        // the first phase keeps tile bits below lc out (it
                        // loads HBM directly); later phases may take them in
                        for (uint32_t b0 = first ? lc : 0; b0 < m && !better; ++b0) {
                            if ((Rm >> b0) & 1) continue;
                            const uint64_t R2 = (Rm & ~(uint64_t{1} << a0)) | (uint64_t{1} << b0);
                            const size_t c = score(R2);
                            if (c > best) {
                                best = c;
                                Rm = R2;
                                better = true;
                            }
                        }
                    }
                }
                R.clear();
                for (uint32_t b = 0; b < m; ++b)
                    if ((Rm >> b) & 1) R.insert(b);
                std::set<uint32_t> Rg = R;
                ptake = greedy(pit, ptodo, Rg, RB, SIZE_MAX);
            }
            const FPhase* prev = first ? nullptr : &P.phases.back();
            FPhase ph = make_phase(R, m, RB, Sv, lc, prev);
            if (first && !ph.coalesced) {
                std::set<uint32_t> R0;
                for (uint32_t k = 0; k < RB; ++k) R0.insert(m - RB + k);
                FPhase ldp = make_phase(R0, m, RB, Sv, lc, &ph);
                ldp.op_begin = ldp.op_end = static_cast<int>(P.ops.size());
                P.phases.push_back(ldp);
            }
            ph.op_begin = static_cast<int>(P.ops.size());
            for (size_t pi : ptake) {
                P.gate_op[take[pi]] = static_cast<int>(P.ops.size());
                P.ops.push_back(lower_op(gates[take[pi]], ph, RB, Sv, mats));
            }
            ph.op_end = static_cast<int>(P.ops.size());
            if (opt.bundle) bundle_phase(P, ph, RB, mats);
            P.phases.push_back(ph);
            std::vector<size_t> rem;
            std::set<size_t> tk(ptake.begin(), ptake.end());
            for (size_t i : ptodo)
                if (!tk.count(i)) rem.push_back(i);
            ptodo.swap(rem);
            first = false;
        }
        ps.phase_end = static_cast<int>(P.phases.size());
        P.passes.push_back(ps);
        P.gates_per_pass.push_back(take.size());
        std::vector<size_t> rem;
        std::set<size_t> tk(take.begin(), take.end());
        for (size_t i : todo)
            if (!tk.count(i)) rem.push_back(i);
        todo.swap(rem);
    }
    P.mats = mats;
    P.reg_bits = RB;
    return QS_OK;
}

// upload program tables
int fused_upload(FusedProgram& P, qs_dtype dtype) {
    fused_free(P);
    const size_t pb = P.passes.size() * sizeof(FPass);
    const size_t hb = P.phases.size() * sizeof(FPhase);
    const size_t ob = std::max<size_t>(1, P.ops.size()) * sizeof(FOp);
    const size_t es = dtype == QS_C128 ? 16 : 8;
    const size_t mb = std::max<size_t>(1, P.mats.size()) * es;
    std::vector<unsigned char> blob(pb + hb + ob + mb + 64);
    size_t off = 0;
    auto put = [&](const void* src, size_t n) {
        std::memcpy(blob.data() + off, src, n);
        const size_t o = off;
        off += (n + 15) / 16 * 16;
        return o;
    };
    const size_t o_p = put(P.passes.data(), pb);
    const size_t o_h = put(P.phases.data(), hb);
    std::vector<FOp> opsv = P.ops;
    if (opsv.empty()) opsv.resize(1);
    for (FOp& o : opsv) seal_op(o);
    const size_t o_o = put(opsv.data(), ob);
    size_t o_m;
    if (dtype == QS_C128) {
        std::vector<double2> mm(std::max<size_t>(1, P.mats.size()));
        for (size_t i = 0; i < P.mats.size(); ++i) mm[i] = make_double2(P.mats[i].real(), P.mats[i].imag());
        o_m = put(mm.data(), mm.size() * 16);
    } else {
        std::vector<float2> mm(std::max<size_t>(1, P.mats.size()));
        for (size_t i = 0; i < P.mats.size(); ++i)
            mm[i] = make_float2(static_cast<float>(P.mats[i].real()), static_cast<float>(P.mats[i].imag()));
        o_m = put(mm.data(), mm.size() * 8);
    }
    // stream-ordered: no device-wide synchronisation (serving lanes overlap)
    P.stream = P.last = qs::current_stream();
    QS_HIP(hipMallocAsync(&P.dblob, blob.size(), P.stream), "qpp_amd::fused");
    QS_HIP(hipMemcpyAsync(P.dblob, blob.data(), blob.size(), hipMemcpyHostToDevice, P.stream),
           "qpp_amd::fused");
    QS_HIP(hipStreamSynchronize(P.stream), "qpp_amd::fused");  // `blob` is freed on return
    auto* b = static_cast<unsigned char*>(P.dblob);
    P.d_passes = reinterpret_cast<FPass*>(b + o_p);
    P.d_phases = reinterpret_cast<FPhase*>(b + o_h);
    P.d_ops = reinterpret_cast<FOp*>(b + o_o);
    P.d_mats = b + o_m;
    P.dtype = dtype;
    return QS_OK;
}

void fused_free(FusedProgram& P) {
    if (P.dblob) {
        // runs on another stream must finish before the blob returns to the pool
        if (P.last && P.last != P.stream) (void)hipStreamSynchronize(P.last);
        (void)hipFreeAsync(P.dblob, P.stream);
    }
    P.dblob = nullptr;
}

template <typename R, int NR>
int launch_pass(const FusedProgram& P, qs_state* st, int pass, uint32_t grid_cap) {
    const FPass& ps = P.passes[pass];
    const size_t shmem = sizeof(cxt<R>) << ps.m;
    uint64_t grid = ps.ntiles;
    if (grid > grid_cap) grid = grid_cap;
    hipLaunchKernelGGL((k_tile<R, NR>), dim3(static_cast<uint32_t>(grid)), dim3(kThreads), shmem,
                       st->stream, static_cast<cxt<R>*>(st->d), P.d_passes, P.d_phases,
                       P.d_ops, reinterpret_cast<const cxt<R>*>(P.d_mats), pass);
    return QS_OK;
}

int fused_run(const FusedProgram& P, qs_state* st, uint32_t grid_cap, hipEvent_t* ev) {
    if (!P.dblob) return fail(QS_ERR_INVALID_HANDLE, "qpp_amd::fused", "program not uploaded");
    if (st->dtype != P.dtype) return fail(QS_ERR_TYPE_MISMATCH, "qpp_amd::fused", "dtype");
    if (st->stream != P.stream) {
        // the upload finished on P.stream (synchronised); later runs on other
        // streams are remembered so the free can wait for them
        P.last = st->stream;
    }
    if (!st->qubits || st->nbits != P.nb)
        return fail(QS_ERR_DIMS_MISMATCH_CVECTOR, "qpp_amd::fused", "state/program");
    if (ev) QS_HIP(hipEventRecord(ev[0], st->stream), "qpp_amd::fused");
    const bool native = !P.native_fn.empty() && P.native_dtype == st->dtype;
    for (size_t p = 0; p < P.passes.size(); ++p) {
        int rc;
        if (native)
            rc = native_launch(P, st, int(p), grid_cap);
        else if (st->dtype == QS_C128)
            rc = P.reg_bits == 3 ? launch_pass<double, 8>(P, st, int(p), grid_cap)
                                 : launch_pass<double, 16>(P, st, int(p), grid_cap);
        else
            rc = P.reg_bits == 3 ? launch_pass<float, 8>(P, st, int(p), grid_cap)
                                 : launch_pass<float, 16>(P, st, int(p), grid_cap);
        if (rc) return rc;
        if (ev) QS_HIP(hipEventRecord(ev[p + 1], st->stream), "qpp_amd::fused");
    }
    QS_HIP(hipGetLastError(), "qpp_amd::fused");
    return QS_OK;
}

}  // namespace qs

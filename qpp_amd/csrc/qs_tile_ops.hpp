// qs_tile_ops.hpp -- device side of the tile-fused pass engine, shared by
// the single-state pass kernel (qs_fused.hip) and the batched trajectory
// kernel (qs_batch.hip): complex helpers, the ops on a thread's register
// array x[NR], and the phase index helpers.
#ifndef QPP_AMD_QS_TILE_OPS_HPP_
#define QPP_AMD_QS_TILE_OPS_HPP_

#include <hip/hip_runtime.h>

#include "qs_fused.hpp"

namespace {

template <typename R>
struct cx;
template <>
struct cx<double> {
    using T = double2;
};
template <>
struct cx<float> {
    using T = float2;
};
template <typename R>
using cxt = typename cx<R>::T;

template <typename R>
__device__ __forceinline__ cxt<R> mk(R x, R y) {
    cxt<R> r;
    r.x = x;
    r.y = y;
    return r;
}

// u0 * a + u1 * b
template <typename R>
__device__ __forceinline__ cxt<R> mac2(cxt<R> u0, cxt<R> a, cxt<R> u1, cxt<R> b) {
    cxt<R> r;
    r.x = u0.x * a.x - u0.y * a.y + u1.x * b.x - u1.y * b.y;
    r.y = u0.x * a.y + u0.y * a.x + u1.x * b.y + u1.y * b.x;
    return r;
}

template <typename R>
__device__ __forceinline__ cxt<R> cmul(cxt<R> u, cxt<R> a) {
    return mk<R>(u.x * a.x - u.y * a.y, u.x * a.y + u.y * a.x);
}

template <typename R>
__device__ __forceinline__ cxt<R> sel(bool c, cxt<R> a, cxt<R> b) {
    return mk<R>(c ? a.x : b.x, c ? a.y : b.y);
}

template <typename R>
__device__ __forceinline__ cxt<R> opm(const double* m, int i) {
    return mk<R>(static_cast<R>(m[2 * i]), static_cast<R>(m[2 * i + 1]));
}
template <typename R>
__device__ __forceinline__ cxt<R> opm(const FOp* op, int i) {
    return opm<R>(op->m, i);
}

constexpr int kThreads = 256;

// s_waitcnt immediate for vmcnt(0) with expcnt / lgkmcnt left at their
// maxima (gfx9 encoding: vmcnt[3:0], expcnt[6:4], lgkmcnt[11:8], vmcnt[15:14])
constexpr int kWaitVm0 = 0x0F70;

template <int NR>
struct RBof;
template <>
struct RBof<8> {
    static constexpr int v = 3;
};
template <>
struct RBof<16> {
    static constexpr int v = 4;
};

// ---------------------------------------------------------------------------
// ops on the register array: each op reads x[NR] and writes all of y[NR]
// (elements it leaves alone are copied).  The kernels alternate two arrays
// between consecutive ops, so an op's results already sit where the next op
// reads them; an op that updated one loop-carried array in place would make
// every case of the op switch move its results back into that array's
// registers (two full-array copies per op, as many VALU instructions as the
// gate arithmetic of a dense 1-bit op).
// ---------------------------------------------------------------------------
template <typename R, int NR, int J, bool CT>
__device__ __forceinline__ void do_u1(const cxt<R>* x, cxt<R>* y, uint32_t regok,
                                      const double* mu, bool tok) {
    const cxt<R> u00 = opm<R>(mu, 0), u10 = opm<R>(mu, 1);
    const cxt<R> u01 = opm<R>(mu, 2), u11 = opm<R>(mu, 3);
#pragma unroll
    for (int rho = 0; rho < NR; ++rho) {
        if ((rho >> J) & 1) continue;
        const int r1 = rho | (1 << J);
        const cxt<R> a = x[rho], b = x[r1];
        if (CT && !((regok >> rho) & 1u)) {
            y[rho] = a;
            y[r1] = b;
            continue;
        }
        const cxt<R> y0 = mac2<R>(u00, a, u01, b);
        const cxt<R> y1 = mac2<R>(u10, a, u11, b);
        y[rho] = CT ? sel<R>(tok, y0, a) : y0;
        y[r1] = CT ? sel<R>(tok, y1, b) : y1;
    }
}

// uncontrolled 2x2 on register bit J, matrix (column-major) at M
template <typename R, int NR, int J>
__device__ __forceinline__ void do_u1m(const cxt<R>* x, cxt<R>* y, const cxt<R>* __restrict__ M) {
    const cxt<R> u00 = M[0], u10 = M[1], u01 = M[2], u11 = M[3];
#pragma unroll
    for (int rho = 0; rho < NR; ++rho) {
        if ((rho >> J) & 1) continue;
        const int r1 = rho | (1 << J);
        const cxt<R> a = x[rho], b = x[r1];
        y[rho] = mac2<R>(u00, a, u01, b);
        y[r1] = mac2<R>(u10, a, u11, b);
    }
}

// a 2x2 on every register bit k, matrix at M + 4k (one straight run)
template <typename R, int NR>
__device__ __forceinline__ void do_u1b(const cxt<R>* x, cxt<R>* y, const cxt<R>* __restrict__ M) {
    cxt<R> s[NR], u[NR];
    do_u1m<R, NR, 0>(x, s, M);
    do_u1m<R, NR, 1>(s, u, M + 4);
    if constexpr (NR > 8) {
        do_u1m<R, NR, 2>(u, s, M + 8);
        do_u1m<R, NR, 3>(s, y, M + 12);
    } else {
        do_u1m<R, NR, 2>(u, y, M + 8);
    }
}

template <typename R, int NR>
__device__ __forceinline__ void copy_regs(const cxt<R>* x, cxt<R>* y) {
#pragma unroll
    for (int rho = 0; rho < NR; ++rho) y[rho] = x[rho];
}

template <typename R, int NR, int J, bool CT>
__device__ __forceinline__ void do_x(const cxt<R>* x, cxt<R>* y, uint32_t regok, bool tok) {
#pragma unroll
    for (int rho = 0; rho < NR; ++rho) {
        if ((rho >> J) & 1) continue;
        const int r1 = rho | (1 << J);
        const cxt<R> a = x[rho], b = x[r1];
        if (CT && !((regok >> rho) & 1u)) {
            y[rho] = a;
            y[r1] = b;
            continue;
        }
        y[rho] = CT ? sel<R>(tok, b, a) : b;
        y[r1] = CT ? sel<R>(tok, a, b) : a;
    }
}

template <typename R, int NR, int J, bool CT>
__device__ __forceinline__ void do_diag_r(const cxt<R>* x, cxt<R>* y, uint32_t regok,
                                          const double* mu, bool tok) {
    const cxt<R> d0 = opm<R>(mu, 0), d1 = opm<R>(mu, 1);
#pragma unroll
    for (int rho = 0; rho < NR; ++rho) {
        if (CT && !((regok >> rho) & 1u)) {
            y[rho] = x[rho];
            continue;
        }
        const cxt<R> v = cmul<R>(((rho >> J) & 1) ? d1 : d0, x[rho]);
        y[rho] = CT ? sel<R>(tok, v, x[rho]) : v;
    }
}

// diagonal factor d (per lane or uniform) on every element
template <typename R, int NR, bool CT>
__device__ __forceinline__ void do_scale(const cxt<R>* x, cxt<R>* y, uint32_t regok, cxt<R> d,
                                         bool tok) {
#pragma unroll
    for (int rho = 0; rho < NR; ++rho) {
        if (CT && !((regok >> rho) & 1u)) {
            y[rho] = x[rho];
            continue;
        }
        const cxt<R> v = cmul<R>(d, x[rho]);
        y[rho] = CT ? sel<R>(tok, v, x[rho]) : v;
    }
}

template <typename R, int NR, int J0, int J1, bool CT>
__device__ __forceinline__ void do_u2(const cxt<R>* x, cxt<R>* y, uint32_t regok,
                                      const cxt<R>* __restrict__ M, bool tok) {
    // column-major 4x4; gate index bit 0 <-> register bit J0, bit 1 <-> J1
    cxt<R> m[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = M[i];
#pragma unroll
    for (int rho = 0; rho < NR; ++rho) {
        if (((rho >> J0) & 1) || ((rho >> J1) & 1)) continue;
        const int i0 = rho, i1 = rho | (1 << J0), i2 = rho | (1 << J1),
                  i3 = rho | (1 << J0) | (1 << J1);
        const cxt<R> v[4] = {x[i0], x[i1], x[i2], x[i3]};
        if (CT && !((regok >> rho) & 1u)) {
            y[i0] = v[0];
            y[i1] = v[1];
            y[i2] = v[2];
            y[i3] = v[3];
            continue;
        }
        cxt<R> o[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            cxt<R> s = mac2<R>(m[r], v[0], m[r + 4], v[1]);
            const cxt<R> t2 = mac2<R>(m[r + 8], v[2], m[r + 12], v[3]);
            s.x += t2.x;
            s.y += t2.y;
            o[r] = s;
        }
        if (CT) {
            y[i0] = sel<R>(tok, o[0], v[0]);
            y[i1] = sel<R>(tok, o[1], v[1]);
            y[i2] = sel<R>(tok, o[2], v[2]);
            y[i3] = sel<R>(tok, o[3], v[3]);
        } else {
            y[i0] = o[0];
            y[i1] = o[1];
            y[i2] = o[2];
            y[i3] = o[3];
        }
    }
}

#define QS_J_CASES(F, ...)                                         \
    switch (j0) {                                                  \
        case 0: F<R, NR, 0, CT>(__VA_ARGS__); break;               \
        case 1: F<R, NR, 1, CT>(__VA_ARGS__); break;               \
        case 2: F<R, NR, 2, CT>(__VA_ARGS__); break;               \
        default:                                                   \
            if constexpr (NR > 8) F<R, NR, 3, CT>(__VA_ARGS__);    \
            else copy_regs<R, NR>(x, y);                           \
            break;                                                 \
    }

template <typename R, int NR, bool CT>
__device__ __forceinline__ void run_op_ct(const cxt<R>* x, cxt<R>* y, const OpHead& h,
                                          const FOp* op, const double* mu,
                                          const cxt<R>* mats, uint32_t t, uint64_t base,
                                          bool tok) {
    const uint32_t code = h.code, regok = h.regok;
    const int j0 = (code >> 4) & 3, j1 = (code >> 6) & 3;
    switch (code & 15u) {
        case OP_U1:
            QS_J_CASES(do_u1, x, y, regok, mu, tok)
            break;
        case OP_X:
            QS_J_CASES(do_x, x, y, regok, tok)
            break;
        case OP_DIAG_R:
            QS_J_CASES(do_diag_r, x, y, regok, mu, tok)
            break;
        case OP_DIAG_T: {
            const bool one = (t & h.tbit) != 0;
            const cxt<R> d = sel<R>(one, opm<R>(mu, 1), opm<R>(mu, 0));
            do_scale<R, NR, CT>(x, y, regok, d, tok);
            break;
        }
        case OP_DIAG_O: {
            const cxt<R> d = (base & op->dbit) ? opm<R>(mu, 1) : opm<R>(mu, 0);
            do_scale<R, NR, CT>(x, y, regok, d, tok);
            break;
        }
        case OP_U2: {
            const cxt<R>* M = mats + h.moff;
            switch (j0 * 4 + j1) {
                case 0 * 4 + 1: do_u2<R, NR, 0, 1, CT>(x, y, regok, M, tok); break;
                case 0 * 4 + 2: do_u2<R, NR, 0, 2, CT>(x, y, regok, M, tok); break;
                case 1 * 4 + 2: do_u2<R, NR, 1, 2, CT>(x, y, regok, M, tok); break;
                case 0 * 4 + 3:
                    if constexpr (NR > 8) do_u2<R, NR, 0, 3, CT>(x, y, regok, M, tok);
                    else copy_regs<R, NR>(x, y);
                    break;
                case 1 * 4 + 3:
                    if constexpr (NR > 8) do_u2<R, NR, 1, 3, CT>(x, y, regok, M, tok);
                    else copy_regs<R, NR>(x, y);
                    break;
                case 2 * 4 + 3:
                    if constexpr (NR > 8) do_u2<R, NR, 2, 3, CT>(x, y, regok, M, tok);
                    else copy_regs<R, NR>(x, y);
                    break;
                default: copy_regs<R, NR>(x, y); break;
            }
            break;
        }
        default:
            copy_regs<R, NR>(x, y);
            break;
    }
}

// y := op(x), the op's head already in registers.  `mu`: the 2x2 matrix (or
// the diagonal) the op applies: h.m, or a per-trajectory override of it.
// Bundles, the commonest op, are never controlled and go first, on a path
// that holds nothing but their arithmetic.
template <typename R, int NR>
__device__ __forceinline__ void run_op_h(const cxt<R>* x, cxt<R>* y, const OpHead& h,
                                         const FOp* op, const double* mu, const cxt<R>* mats,
                                         uint32_t t, uint64_t base) {
    const uint32_t code = h.code;
    if ((code & 15u) == OP_U1B) {
        do_u1b<R, NR>(x, y, mats + h.moff);
        return;
    }
    if ((code & kCodeOut) && (base & op->omask) != op->oval) {  // out-of-tile control
        copy_regs<R, NR>(x, y);
        return;
    }
    if (code & kCodeCtrl) {
        const bool tok = (t & h.tmask) == h.tval;
        run_op_ct<R, NR, true>(x, y, h, op, mu, mats, t, base, tok);
    } else {
        run_op_ct<R, NR, false>(x, y, h, op, mu, mats, t, base, true);
    }
}

__device__ __forceinline__ OpHead load_head(const FOp* op) {
    return *reinterpret_cast<const OpHead*>(op);
}

// y := op(x), loading the op's head here
template <typename R, int NR>
__device__ __forceinline__ void run_op(const cxt<R>* x, cxt<R>* y, const FOp* op,
                                       const double* mu, const cxt<R>* mats, uint32_t t,
                                       uint64_t base) {
    const OpHead h = load_head(op);
    run_op_h<R, NR>(x, y, h, op, mu, mats, t, base);
}

// Global index bits of tile number `tile` (the pass's out-of-tile bits as
// runs of consecutive bits; scalar shifts and masks, no table loads).
__device__ __forceinline__ uint64_t tile_base_of(const FPass* ps, uint64_t tile) {
    const int nr = ps->nruns;
    uint64_t base = 0;
#pragma unroll
    for (int i = 0; i < 24; ++i) {
        if (i < nr) {
            const uint32_t r = ps->runs[i];
            const uint32_t src = r & 255u, dst = (r >> 8) & 255u, len = r >> 16;
            base |= ((tile >> src) & ((uint64_t{1} << len) - 1)) << dst;
        }
    }
    return base;
}

// register part of a phase's global index / swizzled LDS index for register
// group rho (a compile-time constant in the unrolled loops)
template <int RB>
__device__ __forceinline__ uint64_t greg(const FPhase* ph, int rho) {
    uint64_t g = 0;
#pragma unroll
    for (int k = 0; k < RB; ++k)
        if ((rho >> k) & 1) g |= ph->g_rb[k];
    return g;
}
template <int RB>
__device__ __forceinline__ uint32_t sreg(const FPhase* ph, int rho) {
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < RB; ++k)
        if ((rho >> k) & 1) s ^= ph->s_rb[k];
    return s;
}

// thread part of a phase's global index / swizzled LDS index
__device__ __forceinline__ uint64_t thr_g(const FPhase* ph, uint32_t t) {
    uint64_t g = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        if ((t >> j) & 1u) g |= ph->g_thr[j];
    return g;
}

__device__ __forceinline__ uint32_t thr_s(const FPhase* ph, uint32_t t) {
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        if ((t >> j) & 1u) s ^= ph->s_thr[j];
    return s;
}

}  // namespace

#endif  // QPP_AMD_QS_TILE_OPS_HPP_

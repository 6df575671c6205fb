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
// ops on the register array x[NR]
// ---------------------------------------------------------------------------
template <typename R, int NR, int J, bool CT>
__device__ __forceinline__ void do_u1(cxt<R>* x, const FOp* op, const double* mu, bool tok) {
    const cxt<R> u00 = opm<R>(mu, 0), u10 = opm<R>(mu, 1);
    const cxt<R> u01 = opm<R>(mu, 2), u11 = opm<R>(mu, 3);
    const uint32_t regok = op->regok;
#pragma unroll
    for (int rho = 0; rho < NR; ++rho) {
        if ((rho >> J) & 1) continue;
        if (CT && !((regok >> rho) & 1u)) continue;
        const int r1 = rho | (1 << J);
        const cxt<R> a = x[rho], b = x[r1];
        const cxt<R> y0 = mac2<R>(u00, a, u01, b);
        const cxt<R> y1 = mac2<R>(u10, a, u11, b);
        x[rho] = CT ? sel<R>(tok, y0, a) : y0;
        x[r1] = CT ? sel<R>(tok, y1, b) : y1;
    }
}

// uncontrolled 2x2 on register bit J, matrix (column-major) at M
template <typename R, int NR, int J>
__device__ __forceinline__ void do_u1m(cxt<R>* x, const cxt<R>* __restrict__ M) {
    const cxt<R> u00 = M[0], u10 = M[1], u01 = M[2], u11 = M[3];
#pragma unroll
    for (int rho = 0; rho < NR; ++rho) {
        if ((rho >> J) & 1) continue;
        const int r1 = rho | (1 << J);
        const cxt<R> a = x[rho], b = x[r1];
        x[rho] = mac2<R>(u00, a, u01, b);
        x[r1] = mac2<R>(u10, a, u11, b);
    }
}

template <typename R, int NR, int J, bool CT>
__device__ __forceinline__ void do_x(cxt<R>* x, const FOp* op, bool tok) {
    const uint32_t regok = op->regok;
#pragma unroll
    for (int rho = 0; rho < NR; ++rho) {
        if ((rho >> J) & 1) continue;
        if (CT && !((regok >> rho) & 1u)) continue;
        const int r1 = rho | (1 << J);
        const cxt<R> a = x[rho], b = x[r1];
        x[rho] = CT ? sel<R>(tok, b, a) : b;
        x[r1] = CT ? sel<R>(tok, a, b) : a;
    }
}

template <typename R, int NR, int J, bool CT>
__device__ __forceinline__ void do_diag_r(cxt<R>* x, const FOp* op, bool tok) {
    const cxt<R> d0 = opm<R>(op, 0), d1 = opm<R>(op, 1);
    const uint32_t regok = op->regok;
#pragma unroll
    for (int rho = 0; rho < NR; ++rho) {
        if (CT && !((regok >> rho) & 1u)) continue;
        const cxt<R> y = cmul<R>(((rho >> J) & 1) ? d1 : d0, x[rho]);
        x[rho] = CT ? sel<R>(tok, y, x[rho]) : y;
    }
}

// diagonal factor d (per lane or uniform) on every element
template <typename R, int NR, bool CT>
__device__ __forceinline__ void do_scale(cxt<R>* x, const FOp* op, cxt<R> d, bool tok) {
    const uint32_t regok = op->regok;
#pragma unroll
    for (int rho = 0; rho < NR; ++rho) {
        if (CT && !((regok >> rho) & 1u)) continue;
        const cxt<R> y = cmul<R>(d, x[rho]);
        x[rho] = CT ? sel<R>(tok, y, x[rho]) : y;
    }
}

template <typename R, int NR, int J0, int J1, bool CT>
__device__ __forceinline__ void do_u2(cxt<R>* x, const FOp* op, const cxt<R>* __restrict__ mats,
                                      bool tok) {
    // column-major 4x4; gate index bit 0 <-> register bit J0, bit 1 <-> J1
    const cxt<R>* M = mats + op->moff;
    cxt<R> m[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = M[i];
    const uint32_t regok = op->regok;
#pragma unroll
    for (int rho = 0; rho < NR; ++rho) {
        if (((rho >> J0) & 1) || ((rho >> J1) & 1)) continue;
        if (CT && !((regok >> rho) & 1u)) continue;
        const int i0 = rho, i1 = rho | (1 << J0), i2 = rho | (1 << J1),
                  i3 = rho | (1 << J0) | (1 << J1);
        const cxt<R> v[4] = {x[i0], x[i1], x[i2], x[i3]};
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
            x[i0] = sel<R>(tok, o[0], v[0]);
            x[i1] = sel<R>(tok, o[1], v[1]);
            x[i2] = sel<R>(tok, o[2], v[2]);
            x[i3] = sel<R>(tok, o[3], v[3]);
        } else {
            x[i0] = o[0];
            x[i1] = o[1];
            x[i2] = o[2];
            x[i3] = o[3];
        }
    }
}

#define QS_J_CASES(F, ...)                                         \
    switch (op->j0) {                                              \
        case 0: F<R, NR, 0, CT>(__VA_ARGS__); break;               \
        case 1: F<R, NR, 1, CT>(__VA_ARGS__); break;               \
        case 2: F<R, NR, 2, CT>(__VA_ARGS__); break;               \
        case 3: if constexpr (NR > 8) F<R, NR, 3, CT>(__VA_ARGS__); break; \
        default: break;                                            \
    }

template <typename R, int NR, bool CT>
__device__ __forceinline__ void run_op_ct(cxt<R>* x, const FOp* op, const double* mu,
                                          const cxt<R>* mats, uint32_t t, uint64_t base,
                                          bool tok) {
    switch (op->kind) {
        case OP_U1:
            QS_J_CASES(do_u1, x, op, mu, tok)
            break;
        case OP_X:
            QS_J_CASES(do_x, x, op, tok)
            break;
        case OP_DIAG_R:
            QS_J_CASES(do_diag_r, x, op, tok)
            break;
        case OP_DIAG_T: {
            const bool one = (t & op->tbit) != 0;
            const cxt<R> d = sel<R>(one, opm<R>(op, 1), opm<R>(op, 0));
            do_scale<R, NR, CT>(x, op, d, tok);
            break;
        }
        case OP_DIAG_O: {
            const cxt<R> d = (base & op->dbit) ? opm<R>(op, 1) : opm<R>(op, 0);
            do_scale<R, NR, CT>(x, op, d, tok);
            break;
        }
        case OP_U1B: {
            const cxt<R>* M = mats + op->moff;
            const uint32_t mask = op->tbit;
            if (mask & 1u) do_u1m<R, NR, 0>(x, M);
            if (mask & 2u) do_u1m<R, NR, 1>(x, M + 4);
            if (mask & 4u) do_u1m<R, NR, 2>(x, M + 8);
            if constexpr (NR > 8)
                if (mask & 8u) do_u1m<R, NR, 3>(x, M + 12);
            break;
        }
        case OP_U2: {
            const int key = op->j0 * 8 + op->j1;
            switch (key) {
                case 0 * 8 + 1: do_u2<R, NR, 0, 1, CT>(x, op, mats, tok); break;
                case 0 * 8 + 2: do_u2<R, NR, 0, 2, CT>(x, op, mats, tok); break;
                case 1 * 8 + 2: do_u2<R, NR, 1, 2, CT>(x, op, mats, tok); break;
                case 0 * 8 + 3: if constexpr (NR > 8) do_u2<R, NR, 0, 3, CT>(x, op, mats, tok); break;
                case 1 * 8 + 3: if constexpr (NR > 8) do_u2<R, NR, 1, 3, CT>(x, op, mats, tok); break;
                case 2 * 8 + 3: if constexpr (NR > 8) do_u2<R, NR, 2, 3, CT>(x, op, mats, tok); break;
                default: break;
            }
            break;
        }
        default:
            break;
    }
}

// `mu`: the 2x2 matrix an OP_U1 applies (op->m, or a per-trajectory
// override of it)
template <typename R, int NR>
__device__ __forceinline__ void run_op(cxt<R>* x, const FOp* op, const double* mu,
                                       const cxt<R>* mats, uint32_t t, uint64_t base) {
    if ((base & op->omask) != op->oval) return;  // out-of-tile control, uniform
    if (op->ctrl) {
        const bool tok = (t & op->tmask) == op->tval;
        run_op_ct<R, NR, true>(x, op, mu, mats, t, base, tok);
    } else {
        run_op_ct<R, NR, false>(x, op, mu, mats, t, base, true);
    }
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

// qs_exact.hpp -- wave- and workgroup-level exact running sums (device side
// of qs_exact_core.hpp).
//
// A running sum is carried as the double s itself.  Within a binade the maps
// of qs_exact_core.hpp are scanned in parallel; the element whose add leaves
// the binade is added in floating point by the thread that holds it, and the
// scan restarts after it.  The results equal the sequential left-to-right
// sum bit for bit (round to nearest even, as on the host).
#ifndef QPP_AMD_QS_EXACT_HPP_
#define QPP_AMD_QS_EXACT_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>

#include "qs_exact_core.hpp"

namespace qs_exact {

__device__ __forceinline__ Tr shfl_up_tr(Tr t, int off) {
    return Tr{__shfl_up(static_cast<unsigned long long>(t.i0), off, 64),
              __shfl_up(static_cast<unsigned long long>(t.i1), off, 64)};
}

__device__ __forceinline__ Tr shfl_tr(Tr t, int lane) {
    return Tr{__shfl(static_cast<unsigned long long>(t.i0), lane, 64),
              __shfl(static_cast<unsigned long long>(t.i1), lane, 64)};
}

// inclusive scan of the maps over the 64 lanes (lane order = element order)
__device__ __forceinline__ Tr wave_incl(Tr t, int lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const Tr o = shfl_up_tr(t, off);
        if (lane >= off) t = compose(o, t);
    }
    return t;
}

// exclusive prefix of a map over the wave, from its inclusive scan
__device__ __forceinline__ Tr wave_excl(Tr inc, int lane) {
    const Tr o = shfl_up_tr(inc, 1);
    return lane == 0 ? identity() : o;
}

// Workgroup of NT threads (NT / 64 waves): exclusive scan of one map per
// thread, and the total.  sh: NT / 64 maps of LDS.
template <int NT>
__device__ inline void wg_scan(Tr t, Tr* sh, Tr& excl, Tr& total) {
    constexpr int W = NT / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const Tr inc = wave_incl(t, lane);
    const Tr ex = wave_excl(inc, lane);
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    Tr pre = identity(), tot = identity();
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const Tr v = sh[j];
        if (j < w) pre = compose(pre, v);
        tot = compose(tot, v);
    }
    __syncthreads();  // sh is free again
    excl = compose(pre, ex);
    total = tot;
}

// LDS a workgroup needs for wg_chunk
template <int NT>
struct ChunkLds {
    Tr tr[NT / 64];
    double s;
    int first;
};

// One chunk of NT * V elements, thread t holding the V consecutive elements
// t*V .. t*V+V-1 (zeros past the end are harmless: they add nothing).
// s: the exact sum before the chunk (uniform).  out[j]: the sum after the
// thread's element j.  Returns the sum after the chunk (uniform).
template <int NT, int V>
__device__ inline double wg_chunk(double s, const double (&x)[V], double (&out)[V],
                                  ChunkLds<NT>& L) {
    const int t = threadIdx.x;
    int start = 0;
    for (;;) {
        const int E = binade(s);
        const uint64_t m0 = mant(s);
        const bool act = t >= start;
        Tr mine = identity();
        if (act) {
#pragma unroll
            for (int j = 0; j < V; ++j) mine = compose(mine, elem(x[j], E));
        }
        Tr ex, tot;
        wg_scan<NT>(mine, L.tr, ex, tot);
        const uint64_t mend = apply(m0, tot);
        auto walk = [&]() {
            uint64_t m = apply(m0, ex);
#pragma unroll
            for (int j = 0; j < V; ++j) {
                m = apply(m, elem(x[j], E));
                out[j] = from_mant(m, E);
            }
        };
        if (mend < kTop) {
            if (act) walk();
            return from_mant(mend, E);
        }
        // the first thread whose run leaves the binade
        if (t == 0) L.first = NT;
        __syncthreads();
        const bool over = act && apply(m0, compose(ex, mine)) >= kTop;
        if (over) atomicMin(&L.first, t);
        __syncthreads();
        const int f = L.first;
        if (act && t < f) walk();
        if (t == f) {
            uint64_t m = apply(m0, ex);
            double v = from_mant(m, E);
            bool fp = false;
#pragma unroll
            for (int j = 0; j < V; ++j) {
                if (!fp) {
                    const uint64_t mn = apply(m, elem(x[j], E));
                    if (mn < kTop) {
                        m = mn;
                        v = from_mant(m, E);
                    } else {
                        v = __dadd_rn(v, x[j]);
                        fp = true;
                    }
                } else {
                    v = __dadd_rn(v, x[j]);
                }
                out[j] = v;
            }
            L.s = v;
        }
        __syncthreads();
        s = L.s;
        start = f + 1;
        __syncthreads();  // L is reused by the next round
        if (start >= NT) return s;
    }
}

// One workgroup, the whole sequence load(0 .. N-1): chunks of NT * V
// elements in order.  Returns the exact sequential sum (uniform).
template <int NT, int V, typename Load>
__device__ inline double wg_sum(uint64_t N, Load load, ChunkLds<NT>& L) {
    double s = 0.0;
    for (uint64_t b = 0; b < N; b += uint64_t(NT) * V) {
        double x[V], out[V];
        const uint64_t i0 = b + uint64_t(threadIdx.x) * V;
#pragma unroll
        for (int j = 0; j < V; ++j) x[j] = i0 + j < N ? load(i0 + j) : 0.0;
        s = wg_chunk<NT, V>(s, x, out, L);
    }
    return s;
}

// One workgroup: the first k with cp_k >= u, where cp is the sequence of
// running sums of load(0 .. N-1) and cp_{N-1} := 1 (lower_bound on the
// reference's table; u < 1).  Stops at the chunk that holds the answer.
template <int NT, int V, typename Load>
__device__ inline uint64_t wg_lower_bound(uint64_t N, Load load, double u, ChunkLds<NT>& L) {
    double s = 0.0;
    for (uint64_t b = 0; b < N; b += uint64_t(NT) * V) {
        double x[V], out[V];
        const uint64_t i0 = b + uint64_t(threadIdx.x) * V;
#pragma unroll
        for (int j = 0; j < V; ++j) x[j] = i0 + j < N ? load(i0 + j) : 0.0;
        s = wg_chunk<NT, V>(s, x, out, L);
        const bool last = b + uint64_t(NT) * V >= N;
        if (last || s >= u) {  // cp is non-decreasing: the answer is here
            if (threadIdx.x == 0) L.first = NT * V;
            __syncthreads();
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const uint64_t i = i0 + j;
                const double v = i == N - 1 ? 1.0 : out[j];
                if (i < N && v >= u) {
                    atomicMin(&L.first, int(threadIdx.x) * V + j);
                    break;
                }
            }
            __syncthreads();
            const int f = L.first;
            __syncthreads();
            if (f < NT * V) return b + uint64_t(f);
        }
    }
    return N - 1;
}

}  // namespace qs_exact

#endif  // QPP_AMD_QS_EXACT_HPP_

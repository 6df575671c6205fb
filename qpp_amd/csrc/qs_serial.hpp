// qs_serial.hpp -- one workgroup replays the sequential floating-point part
// of libstdc++'s discrete_distribution (random.tcc:2656-2712), which has no
// parallel form:
//   S    = p_0 + p_1 + ... + p_{N-1}            (left to right)
//   cp_k = cp_{k-1} + p_k / S,  cp_{N-1} := 1   (left to right)
//   draw(u) = first k with cp_k >= u
// Waves 1..3 (192 loader threads) stage the next tile of values (p_k, or the
// quotients p_k / S formed in parallel: the division is correctly rounded)
// in LDS while thread 0 consumes the current tile.  Each loader issues its
// PER global loads of a tile back to back, so a tile costs one memory
// latency; thread 0 reads 32 values per LDS wait, so only the add chain is
// serial.  cp is non-decreasing, so one comparison per batch finds the
// batches that contain a draw.
#ifndef QPP_AMD_QS_SERIAL_HPP_
#define QPP_AMD_QS_SERIAL_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qs_serial {

constexpr uint32_t kBatch = 32;
constexpr uint32_t kLoaders = 192;

// Double-buffered tile walk over tiles of T = 192 * PER values.  load(i) is
// called by the loader threads; consume(b, tile, n) by thread 0 for
// elements [b, b+n) in order, returning false to stop.  buf: 2*T doubles of
// LDS.  All 256 threads must call it.
template <uint32_t PER, typename Load, typename Consume>
__device__ void tiles(uint64_t N, double* buf, Load load, Consume consume) {
    constexpr uint32_t T = kLoaders * PER;
    __shared__ int s_stop;
    const uint32_t t = threadIdx.x;
    if (t == 0) s_stop = 0;
    const uint64_t ntile = (N + T - 1) / T;
    auto fill = [&](double* dst, uint64_t b) {
        const uint32_t lt = t - 64;
        // unconditional loads (index clamped into range, masked afterwards)
        // so that all PER of them are in flight at once
        double v[PER];
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j) {
            const uint64_t g = b + lt + j * kLoaders;
            v[j] = load(g < N ? g : N - 1);
        }
#pragma unroll
        for (uint32_t j = 0; j < PER; ++j)
            dst[lt + j * kLoaders] = b + lt + j * kLoaders < N ? v[j] : 0.0;
    };
    if (t >= 64) fill(buf, 0);
    __syncthreads();
    for (uint64_t k = 0; k < ntile; ++k) {
        const double* cur = buf + (k & 1) * T;
        if (t >= 64 && k + 1 < ntile) fill(buf + ((k + 1) & 1) * T, (k + 1) * T);
        if (t == 0) {
            const uint64_t b = k * T;
            const uint32_t n = N - b < T ? uint32_t(N - b) : T;
            if (!consume(b, cur, n)) s_stop = 1;
        }
        __syncthreads();
        if (s_stop) break;
    }
    __syncthreads();
}

// S = sequential sum of load(0..N-1); valid on thread 0.
template <uint32_t PER, typename Load>
__device__ double sum(uint64_t N, double* buf, Load load) {
    double s = 0;
    tiles<PER>(N, buf, load, [&](uint64_t, const double* c, uint32_t n) {
        uint32_t i = 0;
        for (; i + kBatch <= n; i += kBatch) {
            double v[kBatch];
#pragma unroll
            for (uint32_t j = 0; j < kBatch; ++j) v[j] = c[i + j];
#pragma unroll
            for (uint32_t j = 0; j < kBatch; ++j) s = __dadd_rn(s, v[j]);
        }
        for (; i < n; ++i) s = __dadd_rn(s, c[i]);
        return true;
    });
    return s;
}

// Walks cp over the quotients load(k) (= p_k / S) and reports, for each
// threshold in ascending order, the first k with cp_k >= threshold:
// next_u() gives the current threshold (> 1 when none is left), hit(k)
// records the answer for it and advances.  Runs on thread 0's values.
template <uint32_t PER, typename Load, typename NextU, typename Hit>
__device__ void scan(uint64_t N, double* buf, Load load, NextU next_u, Hit hit) {
    double cp = 0;
    double un = next_u();
    // elements [i, i+m) of the tile, one at a time (tail and hit batches)
    auto walk = [&](uint64_t b, const double* c, uint32_t i, uint32_t m) -> bool {
        for (uint32_t j = 0; j < m; ++j) {
            cp = __dadd_rn(cp, c[i + j]);
            const double cv = (b + i + j == N - 1) ? 1.0 : cp;  // cp_{N-1} := 1
            while (cv >= un) {
                hit(b + i + j);
                un = next_u();
                if (un > 1.0) return false;
            }
        }
        return true;
    };
    tiles<PER>(N, buf, load, [&](uint64_t b, const double* c, uint32_t n) {
        uint32_t i = 0;
        for (; i + kBatch <= n; i += kBatch) {
            // fast path: a straight chain of adds and one comparison (cp is
            // non-decreasing); a batch that holds a draw is walked again
            // from the same start, which repeats the same roundings
            double v[kBatch];
#pragma unroll
            for (uint32_t j = 0; j < kBatch; ++j) v[j] = c[i + j];
            double r = cp;
#pragma unroll
            for (uint32_t j = 0; j < kBatch; ++j) r = __dadd_rn(r, v[j]);
            const double top = (b + i + kBatch == N) ? 1.0 : r;
            if (top >= un) {
                if (!walk(b, c, i, kBatch)) return false;
            } else {
                cp = r;
            }
        }
        return walk(b, c, i, n - i);
    });
}

}  // namespace qs_serial

#endif  // QPP_AMD_QS_SERIAL_HPP_

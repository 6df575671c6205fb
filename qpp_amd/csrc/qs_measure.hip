// qs_measure.hip -- norms, probabilities, sampling and measure_seq.
//
// Sampling is bit-identical to qpp's, which builds libstdc++'s
// discrete_distribution over all basis probabilities (instruments.hpp:783-799,
// /usr/include/c++/11/bits/random.tcc:2656-2712):
//   s    = p_0 + p_1 + ... (sequential),   q_k = p_k / s,
//   cp_k = cp_{k-1} + q_k (sequential),    cp_{N-1} := 1,
//   index = first k with cp_k >= u.
// The GPU proceeds in two steps:
//  1. k_chunk_sums / k_scan_chunks / k_search: a parallel prefix E_k of p in
//     FP64 locates the candidate k with E_{k-1}/E < u <= E_k/E.  The
//     candidate is accepted when u is farther than B from both edges, where
//     B bounds |cp_k - E_k/E| over all k (bound_for()).
//  2. When any shot is within B of an edge, the reference's table itself is
//     built -- both sequential sums reproduced bit for bit by parallel scans
//     of binade maps (k_ex_*, qs_exact_core.hpp) -- and every shot is
//     searched in it.
// Probabilities are re*re + im*im rounded as std::norm does (no FMA).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "qs_internal.hpp"
#include "qs_exact.hpp"

namespace {

using qs::fail;

constexpr int kChunk = 4096;  // elements per chunk of the prefix

// std::norm(z) rounding: two products and one sum.  Contraction into an FMA
// is switched off here (the __dmul_rn / __dadd_rn helpers are plain
// operators that clang would otherwise fuse after inlining).
__device__ __forceinline__ double norm_exact(double x, double y) {
#pragma clang fp contract(off)
    const double xx = x * x;
    const double yy = y * y;
    return xx + yy;
}
__device__ __forceinline__ double norm_exact(float x, float y) {
#pragma clang fp contract(off)
    const double dx = x, dy = y;
    const double xx = dx * dx;
    const double yy = dy * dy;
    return xx + yy;
}

template <typename T>
__global__ __launch_bounds__(256) void k_probs(const T* __restrict__ a,
                                               double* __restrict__ p, uint64_t D,
                                               int rho) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < D;
         i += stride) {
        if (rho)
            p[i] = static_cast<double>(a[i + i * D].x);
        else
            p[i] = norm_exact(a[i].x, a[i].y);
    }
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// block sum of 256 threads
__device__ double block_sum(double v) {
    __shared__ double s[4];
    v = wave_sum(v);
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) s[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0) r = s[0] + s[1] + s[2] + s[3];
    __syncthreads();
    return r;
}

// per-chunk sums of p: one block per chunk
__global__ __launch_bounds__(256) void k_chunk_sums(const double* __restrict__ p,
                                                    uint64_t D,
                                                    double* __restrict__ sums,
                                                    uint64_t nchunks) {
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const uint64_t b = c * kChunk;
        double v = 0;
        for (uint64_t i = b + threadIdx.x; i < b + kChunk && i < D; i += 256) v += p[i];
        v = block_sum(v);
        if (threadIdx.x == 0) sums[c] = v;
    }
}

// exclusive prefix of the chunk sums, one block of 1024 threads; writes the
// total to prefix[nchunks]
__global__ __launch_bounds__(1024) void k_scan_chunks(const double* __restrict__ sums,
                                                      uint64_t nchunks,
                                                      double* __restrict__ prefix) {
    __shared__ double part[1024];
    const uint64_t per = (nchunks + 1023) / 1024;
    const uint64_t b = threadIdx.x * per;
    double v = 0;
    for (uint64_t i = b; i < b + per && i < nchunks; ++i) v += sums[i];
    part[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = 0;
        for (int t = 0; t < 1024; ++t) {
            const double x = part[t];
            part[t] = run;
            run += x;
        }
        prefix[nchunks] = run;
    }
    __syncthreads();
    double run = part[threadIdx.x];
    for (uint64_t i = b; i < b + per && i < nchunks; ++i) {
        prefix[i] = run;
        run += sums[i];
    }
}

// One wave per shot: binary search over chunk prefixes, then an in-chunk
// wave scan.  Writes the candidate index and whether it is certain.
__global__ __launch_bounds__(256) void k_search(const double* __restrict__ p,
                                                uint64_t D,
                                                const double* __restrict__ prefix,
                                                uint64_t nchunks,
                                                const double* __restrict__ u,
                                                uint64_t nshots, double B,
                                                uint64_t* __restrict__ idx,
                                                int* __restrict__ certain) {
    const uint64_t wave = (uint64_t(blockIdx.x) * 256 + threadIdx.x) / 64;
    const int lane = threadIdx.x & 63;
    if (wave >= nshots) return;
    const double total = prefix[nchunks];
    const double uu = u[wave];
    const double target = uu * total;
    // last chunk whose exclusive prefix is < target (chunk 0 if none)
    uint64_t lo = 0, hi = nchunks;  // invariant: prefix[lo] < target or lo == 0
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (prefix[mid] < target)
            lo = mid;
        else
            hi = mid;
    }
    // scan the chunk for the first element with E_k >= target
    double run = prefix[lo];
    uint64_t found = D;  // D = not found in this chunk
    double e_prev = 0, e_at = 0;
    const uint64_t b = lo * kChunk;
    for (uint64_t s = b; s < b + kChunk && s < D; s += 64) {
        const uint64_t i = s + lane;
        const double v = i < D ? p[i] : 0.0;
        // inclusive wave scan
        double inc = v;
        for (int m = 1; m < 64; m <<= 1) {
            const double o = __shfl_up(inc, m, 64);
            if (lane >= m) inc += o;
        }
        const double e = run + inc;
        const unsigned long long hit = __ballot(i < D && e >= target);
        if (hit) {
            const int l = __ffsll(hit) - 1;
            found = s + l;
            e_at = __shfl(e, l, 64);
            e_prev = __shfl(e - v, l, 64);
            break;
        }
        run += __shfl(inc, 63, 64);
    }
    if (lane == 0) {
        if (found == D) {
            // past the end of the chunk: the sequential table's last entry is
            // forced to 1, so the answer is certain only away from the edge
            idx[wave] = D - 1;
            certain[wave] = 0;
        } else {
            idx[wave] = found;
            const double lo_edge = e_prev / total, hi_edge = e_at / total;
            certain[wave] = (lo_edge < uu - B) && (hi_edge >= uu + B) &&
                            (found != D - 1 || hi_edge >= uu + B);
        }
    }
}

// p[idx[s]] for every shot
__global__ __launch_bounds__(256) void k_gather(const double* __restrict__ p,
                                                const uint64_t* __restrict__ idx, uint64_t n,
                                                double* __restrict__ out) {
    const uint64_t stride = uint64_t(gridDim.x) * 256;
    for (uint64_t s = uint64_t(blockIdx.x) * 256 + threadIdx.x; s < n; s += stride)
        out[s] = p[idx[s]];
}

// ---------------------------------------------------------------------------
// The reference's table, exactly: S = sequential sum of p, then
// cp_k = cp_{k-1} + p_k / S (sequential), cp_{D-1} := 1.  The sums run as
// parallel scans of binade maps (qs_exact.hpp) over blocks of kChunk
// elements, thread t of a block holding the kExV consecutive elements
// t*kExV ...:
//   k_ex_block_tr  one map per block, in the binade its approximate prefix
//                  (k_scan_chunks) falls in;
//   k_ex_chain     one wave walks the blocks in order carrying the exact sum:
//                  64 blocks per step when their maps apply, otherwise the
//                  block's own elements are scanned by the wave (wg_chunk);
//   k_ex_expand    every block rebuilds its elements' sums from its exact
//                  entry value (wg_chunk) and writes cp;
//   k_ex_search    lower_bound(cp, u) per shot.
// The approximate prefix only picks the binade a block's map is built in;
// a wrong pick sends the block to the element path and never changes a
// result.  Element k of the summed sequence is p_k (S == nullptr) or
// p_k / *S, the division correctly rounded as in __normalize.
// ---------------------------------------------------------------------------
constexpr int kExV = kChunk / 256;  // 16 elements per thread

__device__ __forceinline__ double ex_elem(const double* __restrict__ p, uint64_t i,
                                          const double* __restrict__ S, double sv) {
    return S ? __ddiv_rn(p[i], sv) : p[i];
}

// the block's kExV elements of thread t (zeros past D)
__device__ __forceinline__ void ex_load(const double* __restrict__ p, uint64_t D,
                                        const double* __restrict__ S, double sv,
                                        uint64_t b, double (&x)[kExV]) {
    const uint64_t i0 = b * kChunk + uint64_t(threadIdx.x) * kExV;
    if (i0 + kExV <= D) {
        const double2* q = reinterpret_cast<const double2*>(p + i0);
#pragma unroll
        for (int j = 0; j < kExV / 2; ++j) {
            const double2 v = q[j];
            x[2 * j] = v.x;
            x[2 * j + 1] = v.y;
        }
        if (S) {
#pragma unroll
            for (int j = 0; j < kExV; ++j) x[j] = __ddiv_rn(x[j], sv);
        }
    } else {
#pragma unroll
        for (int j = 0; j < kExV; ++j) x[j] = i0 + j < D ? ex_elem(p, i0 + j, S, sv) : 0.0;
    }
}

__global__ __launch_bounds__(256) void k_ex_block_tr(const double* __restrict__ p, uint64_t D,
                                                     const double* __restrict__ S,
                                                     const double* __restrict__ prefix,
                                                     uint64_t nb, qs_exact::Tr* __restrict__ T,
                                                     int* __restrict__ Eb) {
    using namespace qs_exact;
    __shared__ Tr sh[4];
    const double sv = S ? *S : 1.0;
    for (uint64_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const double approx = S ? prefix[b] / sv : prefix[b];
        const int E = binade(approx);
        double x[kExV];
        ex_load(p, D, S, sv, b, x);
        Tr mine = identity();
#pragma unroll
        for (int j = 0; j < kExV; ++j) mine = compose(mine, elem(x[j], E));
        Tr ex, tot;
        wg_scan<256>(mine, sh, ex, tot);
        if (threadIdx.x == 0) {
            T[b] = tot;
            Eb[b] = E;
        }
    }
}

// One wave.  s_in[b] = exact sum before block b; s_in[nb] = the total.
// The maps of the next 64 blocks are loaded while the current ones are
// scanned.  A block whose map does not apply is summed by the wave itself
// in two halves of 32 consecutive elements per lane, one scan each
// (wg_chunk with one wave).
__global__ __launch_bounds__(64) void k_ex_chain(const double* __restrict__ p, uint64_t D,
                                                 const double* __restrict__ S, uint64_t nb,
                                                 const qs_exact::Tr* __restrict__ T,
                                                 const int* __restrict__ Eb,
                                                 double* __restrict__ s_in) {
    using namespace qs_exact;
    __shared__ ChunkLds<64> L;
    const int lane = threadIdx.x;
    const double sv = S ? *S : 1.0;
    double s = 0.0;
    uint64_t b = 0;
    // maps of blocks pb + lane, loaded ahead
    uint64_t pb = 0;
    Tr tn = Tr{kTop, kTop};
    int en = 0;
    if (uint64_t(lane) < nb) {
        tn = T[lane];
        en = Eb[lane];
    }
    while (b < nb) {
        if (b != pb) {  // the group does not start where the prefetch assumed
            pb = b;
            if (pb + lane < nb) {
                tn = T[pb + lane];
                en = Eb[pb + lane];
            }
        }
        const Tr tc = tn;
        const int ec = en;
        pb = b + 64;
        if (pb + lane < nb) {
            tn = T[pb + lane];
            en = Eb[pb + lane];
        }
        const int E = binade(s);
        const uint64_t m0 = mant(s);
        const int cnt = nb - b < 64 ? int(nb - b) : 64;
        const uint64_t j = b + lane;
        const bool valid = lane < cnt && ec == E;
        const Tr inc = wave_incl(valid ? tc : Tr{kTop, kTop}, lane);
        const Tr ex = wave_excl(inc, lane);
        const uint64_t mi = apply(m0, inc);
        const unsigned long long bad = __ballot(lane < cnt && mi >= kTop);
        const int f = bad ? __ffsll(bad) - 1 : cnt;  // blocks b .. b+f-1 apply
        if (lane < f) s_in[j] = from_mant(apply(m0, ex), E);
        if (f > 0) s = from_mant(apply(m0, shfl_tr(inc, f - 1)), E);
        b += f;
        if (f < cnt) {
            if (lane == 0) s_in[b] = s;
            constexpr int V = kChunk / 128;  // two halves of 64 x 32 elements
#pragma unroll 1
            for (int h = 0; h < 2; ++h) {
                const uint64_t i0 = b * kChunk + uint64_t(h) * (kChunk / 2) + uint64_t(lane) * V;
                double x[V], out[V];
#pragma unroll
                for (int k = 0; k < V; ++k) x[k] = i0 + k < D ? ex_elem(p, i0 + k, S, sv) : 0.0;
                s = wg_chunk<64, V>(s, x, out, L);
            }
            b += 1;
        }
    }
    if (lane == 0) s_in[nb] = s;
}

__global__ __launch_bounds__(256) void k_ex_expand(const double* __restrict__ p, uint64_t D,
                                                   const double* __restrict__ S,
                                                   const double* __restrict__ s_in,
                                                   uint64_t nb, double* __restrict__ cp) {
    using namespace qs_exact;
    __shared__ ChunkLds<256> L;
    const double sv = S ? *S : 1.0;
    for (uint64_t b = blockIdx.x; b < nb; b += gridDim.x) {
        double x[kExV], out[kExV];
        ex_load(p, D, S, sv, b, x);
        (void)wg_chunk<256, kExV>(s_in[b], x, out, L);
        const uint64_t i0 = b * kChunk + uint64_t(threadIdx.x) * kExV;
#pragma unroll
        for (int j = 0; j < kExV; ++j) {
            const uint64_t i = i0 + j;
            if (i < D) cp[i] = i == D - 1 ? 1.0 : out[j];
        }
        __syncthreads();  // L is reused by the next block
    }
}

// first k with cp[k] >= u[s] (cp[D-1] == 1 > u)
__global__ __launch_bounds__(256) void k_ex_search(const double* __restrict__ cp, uint64_t D,
                                                   const double* __restrict__ u,
                                                   uint64_t nshots,
                                                   uint64_t* __restrict__ idx) {
    const uint64_t s = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (s >= nshots) return;
    const double us = u[s];
    uint64_t lo = 0, hi = D;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (cp[mid] < us)
            lo = mid + 1;
        else
            hi = mid;
    }
    idx[s] = lo < D ? lo : D - 1;
}

// ---------------------------------------------------------------------------
// reductions and scaling
// ---------------------------------------------------------------------------
// partial sums: ket -> sum |a|^2 (x), rho -> trace (x, y)
template <typename T>
__global__ __launch_bounds__(256) void k_norm_partial(const T* __restrict__ a,
                                                      uint64_t D, int rho,
                                                      double2* __restrict__ part) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    double sx = 0, sy = 0;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < D;
         i += stride) {
        if (rho) {
            const T v = a[i + i * D];
            sx += v.x;
            sy += v.y;
        } else {
            const T v = a[i];
            sx += double(v.x) * v.x + double(v.y) * v.y;
        }
    }
    sx = block_sum(sx);
    sy = block_sum(sy);
    if (threadIdx.x == 0) part[blockIdx.x] = make_double2(sx, sy);
}

__global__ __launch_bounds__(256) void k_norm_final(const double2* __restrict__ part,
                                                    int n, double2* __restrict__ out) {
    double sx = 0, sy = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        sx += part[i].x;
        sy += part[i].y;
    }
    sx = block_sum(sx);
    sy = block_sum(sy);
    if (threadIdx.x == 0) *out = make_double2(sx, sy);
}

// a *= 1 / c  (complex c, computed in FP64)
template <typename T>
__global__ __launch_bounds__(256) void k_div_scalar(T* __restrict__ a, uint64_t n,
                                                    double cr, double ci) {
    const double den = cr * cr + ci * ci;
    const double ir = cr / den, ii = -ci / den;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += stride) {
        const T v = a[i];
        T r;
        if (ci == 0.0) {
            r.x = static_cast<decltype(v.x)>(double(v.x) / cr);
            r.y = static_cast<decltype(v.x)>(double(v.y) / cr);
        } else {
            r.x = static_cast<decltype(v.x)>(double(v.x) * ir - double(v.y) * ii);
            r.y = static_cast<decltype(v.x)>(double(v.x) * ii + double(v.y) * ir);
        }
        a[i] = r;
    }
}

// zero every element whose digit (stride, dim) differs from m; for a
// density matrix both the row and the column digit must equal m
template <typename T>
__global__ __launch_bounds__(256) void k_project(T* __restrict__ a, uint64_t n,
                                                 uint64_t D, uint64_t stride,
                                                 uint64_t dim, uint64_t m, int rho) {
    const uint64_t gs = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += gs) {
        const uint64_t r = i % D, c = i / D;
        bool keep = (r / stride) % dim == m;
        if (rho) keep = keep && ((c / stride) % dim == m);
        if (!keep) {
            T z;
            z.x = 0;
            z.y = 0;
            a[i] = z;
        }
    }
}

// destructive contraction: out has subsystem `i` removed; element y of out
// reads the input with digit m inserted (row and column for rho)
template <typename T>
__global__ __launch_bounds__(256) void k_contract(const T* __restrict__ in,
                                                  T* __restrict__ out, uint64_t nout,
                                                  uint64_t Dout, uint64_t Din,
                                                  uint64_t stride, uint64_t dim,
                                                  uint64_t m, int rho) {
    const uint64_t gs = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t y = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; y < nout;
         y += gs) {
        const uint64_t r = y % Dout, c = y / Dout;
        // insert digit m at `stride` (stride of the removed subsystem in the
        // input equals the product of the dims after it, which is also the
        // block of lower digits in the output)
        const uint64_t ri = (r / stride) * stride * dim + m * stride + (r % stride);
        uint64_t x = ri;
        if (rho) {
            const uint64_t ci = (c / stride) * stride * dim + m * stride + (c % stride);
            x = ri + ci * Din;
        }
        out[y] = in[x];
    }
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
bool valid(const qs_state* st) { return st != nullptr && st->d != nullptr; }

uint32_t grid_for(uint64_t work, uint64_t per) {
    uint64_t b = (work + per - 1) / per;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;
    return static_cast<uint32_t>(b);
}

uint64_t row_stride(const std::vector<uint64_t>& dims, uint64_t q) {
    uint64_t s = 1;
    for (uint64_t i = dims.size(); i-- > q + 1;) s *= dims[i];
    return s;
}

bool check_subsys(const std::vector<uint64_t>& subsys,
                  const std::vector<uint64_t>& dims) {
    if (subsys.size() > dims.size()) return false;
    std::vector<uint64_t> s = subsys;
    std::sort(s.begin(), s.end());
    if (std::unique(s.begin(), s.end()) != s.end()) return false;
    for (uint64_t i : subsys)
        if (i + 1 > dims.size()) return false;
    return true;
}

// Scratch memory comes from the stream-ordered pool per call (hipMallocAsync
// on the state's stream, released the same way), so nothing outlives a call
// and no free synchronises the device.
constexpr int kNormBlocks = 1024;

int norm_complex(const qs_state* st, double* re, double* im) {
    const int nb = kNormBlocks;
    double2* part = nullptr;
    QS_HIP(hipMallocAsync(reinterpret_cast<void**>(&part), sizeof(double2) * (nb + 1),
                          st->stream),
           "qpp::norm()");
    const bool rho = st->kind == QS_RHO;
    if (st->dtype == QS_C128)
        hipLaunchKernelGGL(k_norm_partial<double2>, dim3(nb), dim3(256), 0, st->stream,
                           static_cast<const double2*>(st->d), rho ? st->D : st->nelem,
                           int(rho), part);
    else
        hipLaunchKernelGGL(k_norm_partial<float2>, dim3(nb), dim3(256), 0, st->stream,
                           static_cast<const float2*>(st->d), rho ? st->D : st->nelem,
                           int(rho), part);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(256), 0, st->stream, part, nb,
                       part + nb);
    double2 h{0, 0};
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(&h, part + nb, sizeof(double2), hipMemcpyDeviceToHost, st->stream);
    (void)hipFreeAsync(part, st->stream);
    const hipError_t e2 = hipStreamSynchronize(st->stream);
    if (e != hipSuccess) return qs::fail_hip(e, "qpp::norm()");
    if (e2 != hipSuccess) return qs::fail_hip(e2, "qpp::norm()");
    *re = h.x;
    *im = h.y;
    return QS_OK;
}

int div_scalar(qs_state* st, double cr, double ci) {
    if (st->dtype == QS_C128)
        hipLaunchKernelGGL(k_div_scalar<double2>, dim3(grid_for(st->nelem, 256)),
                           dim3(256), 0, st->stream, static_cast<double2*>(st->d),
                           st->nelem, cr, ci);
    else
        hipLaunchKernelGGL(k_div_scalar<float2>, dim3(grid_for(st->nelem, 256)),
                           dim3(256), 0, st->stream, static_cast<float2*>(st->d),
                           st->nelem, cr, ci);
    QS_HIP(hipGetLastError(), "qpp::normalize()");
    return QS_OK;
}

int probs_device(const qs_state* st, double* p) {
    const bool rho = st->kind == QS_RHO;
    if (st->dtype == QS_C128)
        hipLaunchKernelGGL(k_probs<double2>, dim3(grid_for(st->D, 256)), dim3(256), 0,
                           st->stream, static_cast<const double2*>(st->d), p, st->D,
                           int(rho));
    else
        hipLaunchKernelGGL(k_probs<float2>, dim3(grid_for(st->D, 256)), dim3(256), 0,
                           st->stream, static_cast<const float2*>(st->d), p, st->D,
                           int(rho));
    QS_HIP(hipGetLastError(), "qpp::sample()");
    return QS_OK;
}

// |cp_k (sequential) - E_k / E (parallel)| <= bound_for(N) for all k, with
// probabilities p >= 0 normalised to sum 1 by construction of cp:
//   sequential sum s: |s - S| <= (N-1) eps S             (recursive summation)
//   q_k = p_k / s rounded: relative eps each
//   sequential cp_k: absolute error <= k eps max|cp| <= k eps (1 + tiny)
//   parallel prefix E_k and total E: relative (log2 N + chunk) eps each
// so |cp_k - E_k/E| <= (2N + 2 log2 N + 2 kChunk + 8) eps (1 + o(1)); the
// bound adds a factor 2 of slack.  eps = 2^-53.
double bound_for(uint64_t N) {
    const double eps = std::ldexp(1.0, -53);
    const double lg = std::log2(double(N) + 1.0);
    return 2.0 * (2.0 * double(N) + 2.0 * lg + 2.0 * kChunk + 8.0) * eps;
}

// Device workspace of one sample_device call, carved from one allocation
// (256-byte aligned segments).
struct SampleWs {
    double* p;       // D probabilities
    double* cp;      // D: the exact cumulative table
    double* sums;    // nchunks block sums
    double* prefix;  // nchunks + 1: their exclusive prefix and total
    qs_exact::Tr* T; // nchunks block maps
    int* Eb;         // nchunks binades of the maps
    double* sp_in;   // nchunks + 1: exact sums of p before each block; [nchunks] = S
    double* sq_in;   // nchunks + 1: the same for p / S
    double* u;       // nshots uniforms (then the gathered probabilities)
    uint64_t* idx;   // nshots indices
    int* cert;       // nshots: decided by the approximate search
};

constexpr size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// byte sizes of the SampleWs segments, in member order
constexpr int kWsSegs = 11;
struct WsSizes {
    size_t seg[kWsSegs];
};
constexpr WsSizes ws_sizes(uint64_t D, uint64_t nshots) {
    const uint64_t nc = (D + kChunk - 1) / kChunk;
    return WsSizes{{8 * D, 8 * D, 8 * nc, 8 * (nc + 1), sizeof(qs_exact::Tr) * nc, 4 * nc,
                    8 * (nc + 1), 8 * (nc + 1), 8 * nshots, 8 * nshots, 4 * nshots}};
}
constexpr size_t ws_bytes(uint64_t D, uint64_t nshots) {
    const WsSizes z = ws_sizes(D, nshots);
    size_t off = 0;
    for (int i = 0; i < kWsSegs; ++i) off += align256(z.seg[i]);
    return off;
}

size_t sample_ws_layout(uint64_t D, uint64_t nshots, unsigned char* base, SampleWs* w) {
    const WsSizes z = ws_sizes(D, nshots);
    unsigned char* at[kWsSegs];
    size_t off = 0;
    for (int i = 0; i < kWsSegs; ++i) {
        at[i] = base ? base + off : nullptr;
        off += align256(z.seg[i]);
    }
    if (w) {
        w->p = reinterpret_cast<double*>(at[0]);
        w->cp = reinterpret_cast<double*>(at[1]);
        w->sums = reinterpret_cast<double*>(at[2]);
        w->prefix = reinterpret_cast<double*>(at[3]);
        w->T = reinterpret_cast<qs_exact::Tr*>(at[4]);
        w->Eb = reinterpret_cast<int*>(at[5]);
        w->sp_in = reinterpret_cast<double*>(at[6]);
        w->sq_in = reinterpret_cast<double*>(at[7]);
        w->u = reinterpret_cast<double*>(at[8]);
        w->idx = reinterpret_cast<uint64_t*>(at[9]);
        w->cert = reinterpret_cast<int*>(at[10]);
    }
    return off;
}

// Builds the exact table in w.cp (k_ex_* above) and searches every shot in
// it.  Needs w.p, w.prefix (k_chunk_sums + k_scan_chunks) and w.u.
int exact_table_search(const qs_state* st, const SampleWs& w, uint64_t nshots) {
    const uint64_t D = st->D;
    const uint64_t nb = (D + kChunk - 1) / kChunk;
    const uint32_t g = grid_for(nb, 1);
    hipStream_t sm = st->stream;
    hipLaunchKernelGGL(k_ex_block_tr, dim3(g), dim3(256), 0, sm, w.p, D, nullptr, w.prefix, nb,
                       w.T, w.Eb);
    hipLaunchKernelGGL(k_ex_chain, dim3(1), dim3(64), 0, sm, w.p, D, nullptr, nb, w.T, w.Eb,
                       w.sp_in);
    const double* S = w.sp_in + nb;
    hipLaunchKernelGGL(k_ex_block_tr, dim3(g), dim3(256), 0, sm, w.p, D, S, w.prefix, nb, w.T,
                       w.Eb);
    hipLaunchKernelGGL(k_ex_chain, dim3(1), dim3(64), 0, sm, w.p, D, S, nb, w.T, w.Eb, w.sq_in);
    hipLaunchKernelGGL(k_ex_expand, dim3(g), dim3(256), 0, sm, w.p, D, S, w.sq_in, nb, w.cp);
    hipLaunchKernelGGL(k_ex_search, dim3(grid_for(nshots, 256)), dim3(256), 0, sm, w.cp, D, w.u,
                       nshots, w.idx);
    QS_HIP(hipGetLastError(), "qpp::sample()");
    return QS_OK;
}

// Indices of nshots draws from host uniforms u_host, with p[index] of each
// in out_prob_host when given.  ws: a device workspace of at least
// sample_ws_layout(D, nshots) bytes, or nullptr to take one from the
// stream-ordered pool for the call.
int sample_device(const qs_state* st, const double* u_host, uint64_t nshots,
                  uint64_t* out_idx, double* out_prob_host, unsigned char* ws) {
    const uint64_t D = st->D;
    unsigned char* own = nullptr;
    if (!ws) {
        const size_t bytes = sample_ws_layout(D, nshots, nullptr, nullptr);
        QS_HIP(hipMallocAsync(reinterpret_cast<void**>(&own), bytes, st->stream), "qpp::sample()");
        ws = own;
    }
    const int rc = [&]() -> int {
        SampleWs w;
        sample_ws_layout(D, nshots, ws, &w);
        if (int rc = probs_device(st, w.p)) return rc;
        if (D < 2) {  // a one-element table returns 0 (random.tcc:2700)
            for (uint64_t s = 0; s < nshots; ++s) out_idx[s] = 0;
            if (out_prob_host) {
                QS_HIP(hipMemcpyAsync(out_prob_host, w.p, 8, hipMemcpyDeviceToHost, st->stream),
                       "qpp::sample()");
                QS_HIP(hipStreamSynchronize(st->stream), "qpp::sample()");
                for (uint64_t s = 1; s < nshots; ++s) out_prob_host[s] = out_prob_host[0];
            }
            return QS_OK;
        }
        const uint64_t nchunks = (D + kChunk - 1) / kChunk;
        QS_HIP(hipMemcpyAsync(w.u, u_host, nshots * 8, hipMemcpyHostToDevice, st->stream),
               "qpp::sample()");
        hipLaunchKernelGGL(k_chunk_sums, dim3(grid_for(nchunks, 1)), dim3(256), 0, st->stream,
                           w.p, D, w.sums, nchunks);
        hipLaunchKernelGGL(k_scan_chunks, dim3(1), dim3(1024), 0, st->stream, w.sums, nchunks,
                           w.prefix);
        const double B = bound_for(D);
        hipLaunchKernelGGL(k_search, dim3(grid_for(nshots * 64, 256)), dim3(256), 0, st->stream,
                           w.p, D, w.prefix, nchunks, w.u, nshots, B, w.idx, w.cert);
        QS_HIP(hipGetLastError(), "qpp::sample()");
        std::vector<int> hc(nshots);
        QS_HIP(hipMemcpyAsync(hc.data(), w.cert, nshots * 4, hipMemcpyDeviceToHost, st->stream),
               "qpp::sample()");
        QS_HIP(hipStreamSynchronize(st->stream), "qpp::sample()");
        const bool force = qs::force_exact_sampling();
        bool all = !force;
        for (uint64_t s = 0; s < nshots && all; ++s) all = hc[s] != 0;
        // shots the error bound cannot decide: every shot is searched in
        // the exact table (a decided shot gets the same index there)
        if (!all)
            if (int rc = exact_table_search(st, w, nshots)) return rc;
        QS_HIP(hipMemcpyAsync(out_idx, w.idx, nshots * 8, hipMemcpyDeviceToHost, st->stream),
               "qpp::sample()");
        // p[index] of every shot (instruments.hpp:808), gathered on the
        // device; the u slots are free once the searches have run
        if (out_prob_host) {
            hipLaunchKernelGGL(k_gather, dim3(grid_for(nshots, 256)), dim3(256), 0, st->stream,
                               w.p, w.idx, nshots, w.u);
            QS_HIP(hipGetLastError(), "qpp::sample()");
            QS_HIP(hipMemcpyAsync(out_prob_host, w.u, nshots * 8, hipMemcpyDeviceToHost,
                                  st->stream),
                   "qpp::sample()");
        }
        QS_HIP(hipStreamSynchronize(st->stream), "qpp::sample()");
        return QS_OK;
    }();
    if (own) {
        (void)hipFreeAsync(own, st->stream);
        if (rc != QS_OK) (void)hipStreamSynchronize(st->stream);
    }
    return rc;
}

int sample_device(const qs_state* st, const double* u_host, uint64_t nshots,
                  std::vector<uint64_t>& out_idx, double* out_prob_host = nullptr) {
    out_idx.assign(nshots, 0);
    return sample_device(st, u_host, nshots, out_idx.data(), out_prob_host, nullptr);
}

// ---------------------------------------------------------------------------
// Allocation-free sampling (qs_sample_noalloc): the workspace of one call on
// a state of at most kFixedMaxD basis states and kFixedMaxShots shots lives
// in a file-scope device array.  Every GPU the library runs on holds its own
// copy (about 16.4 MiB) until the process ends.  Its device address is
// looked up once per device with hipGetSymbolAddress (the array's name is a
// host-side stand-in, not a device address) and only that address is handed
// to kernels and copies.  Nothing in it needs clearing: every segment is
// written before it is read.  All calls on one GPU share the array, so calls
// that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr uint64_t kFixedMaxD = uint64_t{1} << 20;
constexpr uint64_t kFixedMaxShots = uint64_t{1} << 14;
constexpr size_t kFixedBytes = ws_bytes(kFixedMaxD, kFixedMaxShots);
static __device__ __align__(256) unsigned char scratch[kFixedBytes];

constexpr int kMaxDevices = 64;
std::once_flag g_scratch_once[kMaxDevices];
unsigned char* g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];

// the device address of `scratch` on device dev (set up on first use, with
// dev current)
int fixed_scratch(int dev, unsigned char** out) {
    constexpr const char* W = "qs_sample_noalloc";
    if (dev < 0 || dev >= kMaxDevices) return fail(QS_ERR_DEVICE, W, "device");
    std::call_once(g_scratch_once[dev], [dev] {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != dev) e = hipSetDevice(dev);
        void* ptr = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        g_scratch_err[dev] = e;
        g_scratch_ptr[dev] = e == hipSuccess ? static_cast<unsigned char*>(ptr) : nullptr;
    });
    if (g_scratch_err[dev] != hipSuccess) return qs::fail_hip(g_scratch_err[dev], W);
    *out = g_scratch_ptr[dev];
    return QS_OK;
}

// Every shot searched in the exact table (no host decision in between, so
// the launch sequence is fixed), in a caller-given workspace.  idx_host /
// prob_host may be null.
int sample_exact_only(const qs_state* st, const double* u_host, uint64_t nshots,
                      uint64_t* idx_host, double* prob_host, unsigned char* ws) {
    constexpr const char* W = "qpp::sample()";
    const uint64_t D = st->D;
    SampleWs w;
    sample_ws_layout(D, nshots, ws, &w);
    if (int rc = probs_device(st, w.p)) return rc;
    if (D < 2) {  // a one-element table returns 0 (random.tcc:2700)
        if (idx_host)
            for (uint64_t s = 0; s < nshots; ++s) idx_host[s] = 0;
        if (prob_host) {
            QS_HIP(hipMemcpyAsync(prob_host, w.p, 8, hipMemcpyDeviceToHost, st->stream), W);
            QS_HIP(hipStreamSynchronize(st->stream), W);
            for (uint64_t s = 1; s < nshots; ++s) prob_host[s] = prob_host[0];
        }
        return QS_OK;
    }
    const uint64_t nchunks = (D + kChunk - 1) / kChunk;
    QS_HIP(hipMemcpyAsync(w.u, u_host, nshots * 8, hipMemcpyHostToDevice, st->stream), W);
    hipLaunchKernelGGL(k_chunk_sums, dim3(grid_for(nchunks, 1)), dim3(256), 0, st->stream, w.p,
                       D, w.sums, nchunks);
    hipLaunchKernelGGL(k_scan_chunks, dim3(1), dim3(1024), 0, st->stream, w.sums, nchunks,
                       w.prefix);
    if (int rc = exact_table_search(st, w, nshots)) return rc;
    if (idx_host)
        QS_HIP(hipMemcpyAsync(idx_host, w.idx, nshots * 8, hipMemcpyDeviceToHost, st->stream), W);
    if (prob_host) {
        hipLaunchKernelGGL(k_gather, dim3(grid_for(nshots, 256)), dim3(256), 0, st->stream, w.p,
                           w.idx, nshots, w.u);
        QS_HIP(hipGetLastError(), W);
        QS_HIP(hipMemcpyAsync(prob_host, w.u, nshots * 8, hipMemcpyDeviceToHost, st->stream), W);
    }
    QS_HIP(hipStreamSynchronize(st->stream), W);
    return QS_OK;
}

int project(qs_state* st, uint64_t stride, uint64_t dim, uint64_t m) {
    const int rho = st->kind == QS_RHO;
    if (st->dtype == QS_C128)
        hipLaunchKernelGGL(k_project<double2>, dim3(grid_for(st->nelem, 256)), dim3(256),
                           0, st->stream, static_cast<double2*>(st->d), st->nelem,
                           st->D, stride, dim, m, rho);
    else
        hipLaunchKernelGGL(k_project<float2>, dim3(grid_for(st->nelem, 256)), dim3(256),
                           0, st->stream, static_cast<float2*>(st->d), st->nelem,
                           st->D, stride, dim, m, rho);
    QS_HIP(hipGetLastError(), "qpp::measure_seq()");
    return QS_OK;
}

// remove subsystem i, keeping digit m (ket: <m| contraction; rho: the (m, m)
// block, i.e. ptrace over i of the projected state)
int contract(qs_state* st, uint64_t i, uint64_t m) {
    std::vector<uint64_t> nd = st->dims;
    const uint64_t stride = row_stride(st->dims, i), dim = st->dims[i];
    nd.erase(nd.begin() + static_cast<std::ptrdiff_t>(i));
    if (nd.empty()) nd.push_back(1);
    uint64_t Dn = 1;
    for (uint64_t d : nd) Dn *= d;
    const int rho = st->kind == QS_RHO;
    const uint64_t nout = rho ? Dn * Dn : Dn;
    void* nbuf = nullptr;
    QS_HIP(hipMallocAsync(&nbuf, nout * st->elem_bytes(), st->stream), "qpp::measure_seq()");
    if (st->dtype == QS_C128)
        hipLaunchKernelGGL(k_contract<double2>, dim3(grid_for(nout, 256)), dim3(256), 0,
                           st->stream, static_cast<const double2*>(st->d),
                           static_cast<double2*>(nbuf), nout, Dn, st->D, stride, dim, m,
                           rho);
    else
        hipLaunchKernelGGL(k_contract<float2>, dim3(grid_for(nout, 256)), dim3(256), 0,
                           st->stream, static_cast<const float2*>(st->d),
                           static_cast<float2*>(nbuf), nout, Dn, st->D, stride, dim, m,
                           rho);
    QS_HIP(hipGetLastError(), "qpp::measure_seq()");
    (void)hipFreeAsync(st->d, st->stream);
    QS_HIP(hipStreamSynchronize(st->stream), "qpp::measure_seq()");
    st->d = nbuf;
    st->dims = nd;
    st->D = Dn;
    st->nelem = nout;
    st->qubits = std::all_of(nd.begin(), nd.end(), [](uint64_t d) { return d == 2; });
    st->nbits = st->qubits && !(nd.size() == 1 && nd[0] == 1)
                    ? static_cast<uint32_t>((rho ? 2 : 1) * nd.size())
                    : 0;
    if (nd.size() == 1 && nd[0] == 1) st->qubits = false;
    return QS_OK;
}

}  // namespace

extern "C" {

int qs_norm2(const qs_state* st, double* out) {
    if (!valid(st) || !out) return fail(QS_ERR_INVALID_HANDLE, "qpp::norm()");
    double re = 0, im = 0;
    if (int rc = norm_complex(st, &re, &im)) return rc;
    *out = re;
    return QS_OK;
}

// qpp::normalize, functions.hpp:304-333
int qs_normalize(qs_state* st) {
    constexpr const char* W = "qpp::normalize()";
    if (!valid(st)) return fail(QS_ERR_INVALID_HANDLE, W);
    double re = 0, im = 0;
    if (int rc = norm_complex(st, &re, &im)) return rc;
    if (st->kind == QS_KET) {
        const double nrm = std::sqrt(re);
        if (nrm == 0) return fail(QS_ERR_OVERFLOW, W);
        return div_scalar(st, nrm, 0.0);
    }
    if (std::hypot(re, im) == 0) return fail(QS_ERR_OVERFLOW, W);
    return div_scalar(st, re, im);
}

int qs_probs(const qs_state* st, double* host_out, uint64_t n_out) {
    if (!valid(st) || !host_out) return fail(QS_ERR_INVALID_HANDLE, "qpp::abssq()");
    if (n_out != st->D) return fail(QS_ERR_SIZE_MISMATCH, "qpp::abssq()", "n_out");
    void* p = nullptr;
    QS_HIP(hipMallocAsync(&p, st->D * 8, st->stream), "qpp::abssq()");
    int rc = probs_device(st, static_cast<double*>(p));
    if (rc == QS_OK) {
        const hipError_t e =
            hipMemcpyAsync(host_out, p, st->D * 8, hipMemcpyDeviceToHost, st->stream);
        if (e != hipSuccess) rc = qs::fail_hip(e, "qpp::abssq()");
    }
    (void)hipFreeAsync(p, st->stream);
    const hipError_t e = hipStreamSynchronize(st->stream);
    if (rc == QS_OK && e != hipSuccess) rc = qs::fail_hip(e, "qpp::abssq()");
    return rc;
}

// qpp::sample, instruments.hpp:746-929
int qs_sample(const qs_state* st, uint32_t ntarget, const uint64_t* target,
              uint64_t nshots, const double* u, uint64_t* out_index,
              uint64_t* out_digits, double* out_prob) {
    constexpr const char* W = "qpp::sample()";
    if (!valid(st)) return fail(QS_ERR_INVALID_HANDLE, W, "A");
    qs::clear_error();
    std::vector<uint64_t> tv(target, target + ntarget);
    if (!check_subsys(tv, st->dims)) return fail(QS_ERR_SUBSYS_MISMATCH_DIMS, W, "dims/target");
    if (nshots == 0) return fail(QS_ERR_OUT_OF_RANGE, W, "num_samples");
    if (!u) return fail(QS_ERR_ZERO_SIZE, W, "u");
    std::vector<uint64_t> idx;
    if (int rc = sample_device(st, u, nshots, idx, out_prob)) return rc;
    if (out_index) std::copy(idx.begin(), idx.end(), out_index);
    const uint64_t n = st->dims.size();
    if (out_digits)
        for (uint64_t s = 0; s < nshots; ++s) {
            uint64_t rem = idx[s];
            std::vector<uint64_t> mid(n);
            for (uint64_t k = n; k-- > 0;) {
                mid[k] = rem % st->dims[k];
                rem /= st->dims[k];
            }
            for (uint32_t t = 0; t < ntarget; ++t) out_digits[s * ntarget + t] = mid[tv[t]];
        }
    return QS_OK;
}

// qs_sample without any device allocation: the workspace is the fixed
// per-device array `scratch` (see its comment); every shot is searched in
// the exact table.  Host work uses no heap either: the target check runs in
// place, and when out_index is null the indices are staged in out_digits.
int qs_sample_noalloc(const qs_state* st, uint32_t ntarget, const uint64_t* target,
                      uint64_t nshots, const double* u, uint64_t* out_index,
                      uint64_t* out_digits, double* out_prob) {
    constexpr const char* W = "qpp::sample()";
    if (!valid(st)) return fail(QS_ERR_INVALID_HANDLE, W, "A");
    qs::clear_error();
    const uint64_t n = st->dims.size();
    bool ok = ntarget <= n && (ntarget == 0 || target != nullptr);
    for (uint32_t i = 0; ok && i < ntarget; ++i) {
        ok = target[i] < n;
        for (uint32_t j = 0; ok && j < i; ++j) ok = target[j] != target[i];
    }
    if (!ok) return fail(QS_ERR_SUBSYS_MISMATCH_DIMS, W, "dims/target");
    if (nshots == 0) return fail(QS_ERR_OUT_OF_RANGE, W, "num_samples");
    if (!u) return fail(QS_ERR_ZERO_SIZE, W, "u");
    if (st->D > kFixedMaxD || nshots > kFixedMaxShots || n > 64)
        return qs::fail_runtime(QS_ERR_OUT_OF_MEMORY,
                                "qs_sample_noalloc: the fixed scratch holds at most 2^20 basis "
                                "states and 16384 shots per call");
    unsigned char* ws = nullptr;
    if (int rc = fixed_scratch(st->device, &ws)) return rc;
    const bool need_digits = out_digits != nullptr && ntarget > 0;
    uint64_t* idx = out_index ? out_index : (need_digits ? out_digits : nullptr);
    if (int rc = sample_exact_only(st, u, nshots, idx, out_prob, ws)) return rc;
    if (need_digits) {
        // rows from the last shot down: row s starts at s * ntarget >= s,
        // so a staged index not yet read is never overwritten
        for (uint64_t s = nshots; s-- > 0;) {
            uint64_t rem = idx[s];
            uint64_t mid[64];  // n <= 64 (qpp's maxn), checked above
            for (uint64_t k = n; k-- > 0;) {
                mid[k] = rem % st->dims[k];
                rem /= st->dims[k];
            }
            for (uint32_t t = 0; t < ntarget; ++t) out_digits[s * ntarget + t] = mid[target[t]];
        }
    }
    return QS_OK;
}

// qpp::measure_seq(A, target, dims, destructive), instruments.hpp:985-1099
int qs_measure_seq(qs_state* st, uint32_t ntarget, const uint64_t* target,
                   int destructive, double u, uint64_t* results, double* probs) {
    constexpr const char* W = "qpp::measure_seq()";
    if (!valid(st)) return fail(QS_ERR_INVALID_HANDLE, W, "A");
    qs::clear_error();
    std::vector<uint64_t> tv(target, target + ntarget);
    if (!check_subsys(tv, st->dims)) return fail(QS_ERR_SUBSYS_MISMATCH_DIMS, W, "dims/target");
    // one sample of all targets
    std::vector<uint64_t> idx;
    if (int rc = sample_device(st, &u, 1, idx)) return rc;
    return qs::measure_seq_given(st, ntarget, target, destructive, idx[0], results, probs);
}

// qpp::reset, instruments.hpp:1148-1196: measure_seq(A, target, dims, false),
// then powm(Xd(dims[i]), dims[i] - result[i]) on target[i].  Xd(D) maps
// |j> to |j+1 mod D> (gates.hpp:360-373), so the correction is the exact
// permutation |j> -> |j + dims[i] - result[i] mod dims[i]>; the identity
// (result 0) is not launched.  qpp sizes the correction by dims[i], not
// dims[target[i]]; its apply() then throws where the two differ, which is
// checked here before the state changes.
int qs_reset(qs_state* st, uint32_t ntarget, const uint64_t* target, double u) {
    constexpr const char* W = "qpp::reset()";
    if (!valid(st)) return fail(QS_ERR_INVALID_HANDLE, W, "A");
    qs::clear_error();
    std::vector<uint64_t> tv(target, target + ntarget);
    if (!check_subsys(tv, st->dims)) return fail(QS_ERR_SUBSYS_MISMATCH_DIMS, W, "dims/target");
    for (uint32_t i = 0; i < ntarget; ++i)
        if (st->dims[i] != st->dims[tv[i]])
            return fail(QS_ERR_MATRIX_MISMATCH_SUBSYS, "qpp::apply()", "A/dims/target");
    std::vector<uint64_t> res(ntarget);
    std::vector<double> ps(ntarget);
    if (int rc = qs_measure_seq(st, ntarget, target, 0, u, res.data(), ps.data())) return rc;
    for (uint32_t i = 0; i < ntarget; ++i) {
        const uint64_t D = st->dims[i], k = (D - res[i]) % D;
        if (k == 0) continue;
        std::vector<double> M(2 * D * D, 0.0);  // column-major, interleaved complex
        for (uint64_t j = 0; j < D; ++j) M[2 * (((j + k) % D) + j * D)] = 1.0;
        if (int rc = qs_apply(st, M.data(), D, 1, &tv[i])) return rc;
    }
    return QS_OK;
}

}  // extern "C"

// measure_seq with the sampled basis index already known (the engine draws
// it together with its other shots from the same state)
int qs::measure_seq_given(qs_state* st, uint32_t ntarget, const uint64_t* target,
                          int destructive, uint64_t index, uint64_t* results,
                          double* probs) {
    constexpr const char* W = "qpp::measure_seq()";
    std::vector<uint64_t> tv(target, target + ntarget);
    if (!check_subsys(tv, st->dims)) return fail(QS_ERR_SUBSYS_MISMATCH_DIMS, W, "dims/target");
    const std::vector<uint64_t> idx{index};
    const uint64_t n0 = st->dims.size();
    std::vector<uint64_t> mid(n0), res(ntarget);
    {
        uint64_t rem = idx[0];
        for (uint64_t k = n0; k-- > 0;) {
            mid[k] = rem % st->dims[k];
            rem /= st->dims[k];
        }
        for (uint32_t t = 0; t < ntarget; ++t) res[t] = mid[tv[t]];
    }
    std::vector<double> ps(ntarget, 0.0);
    const bool rho = st->kind == QS_RHO;
    std::vector<uint32_t> order(ntarget);
    std::iota(order.begin(), order.end(), 0u);
    if (rho && n0 <= 10) {
        // instruments.hpp:1039-1054: project in the given order, then one
        // ptrace over all targets
        for (uint32_t cnt = 0; cnt < ntarget; ++cnt) {
            const uint64_t i = tv[cnt];
            if (int rc = project(st, row_stride(st->dims, i), st->dims[i], res[cnt])) return rc;
            double tr = 0, ti = 0;
            if (int rc = norm_complex(st, &tr, &ti)) return rc;
            ps[cnt] = tr;
            if (int rc = div_scalar(st, tr, 0.0)) return rc;
        }
        if (destructive) {
            // remove targets in descending order (each contraction keeps the
            // measured digit, equal to the ptrace of the projected state)
            std::vector<uint32_t> desc = order;
            std::sort(desc.begin(), desc.end(), [&](uint32_t a, uint32_t b) { return tv[a] > tv[b]; });
            for (uint32_t k : desc)
                if (int rc = contract(st, tv[k], res[k])) return rc;
        }
    } else {
        if (destructive)
            std::sort(order.begin(), order.end(),
                      [&](uint32_t a, uint32_t b) { return tv[a] > tv[b]; });
        for (uint32_t cnt = 0; cnt < ntarget; ++cnt) {
            const uint32_t ip = order[cnt];
            const uint64_t i = tv[ip];
            if (!rho && destructive) {
                if (int rc = contract(st, i, res[ip])) return rc;
                double nr = 0, ni = 0;
                if (int rc = norm_complex(st, &nr, &ni)) return rc;
                ps[ip] = nr;
                if (int rc = div_scalar(st, std::sqrt(nr), 0.0)) return rc;
            } else {
                if (int rc = project(st, row_stride(st->dims, i), st->dims[i], res[ip])) return rc;
                double nr = 0, ni = 0;
                if (int rc = norm_complex(st, &nr, &ni)) return rc;
                ps[ip] = nr;
                if (int rc = div_scalar(st, rho ? nr : std::sqrt(nr), 0.0)) return rc;
                if (destructive)
                    if (int rc = contract(st, i, res[ip])) return rc;
            }
        }
    }
    QS_HIP(hipStreamSynchronize(st->stream), W);
    for (uint32_t t = 0; t < ntarget; ++t) {
        if (results) results[t] = res[t];
        if (probs) probs[t] = ps[t];
    }
    return QS_OK;
}

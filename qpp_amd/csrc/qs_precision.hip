// qs_precision.hip -- dtype conversion and FP64 state comparison.
//
// The complex64 path stores amplitudes in FP32 but every reduction that
// decides a result runs in FP64: norms and probabilities (qs_measure.hip),
// and here the overlap <a|b>, the squared norms and the largest amplitude
// difference between two states of any dtype mix.  From these the host
// derives the infidelity 1 - |<a|b>|^2 / (|a|^2 |b|^2) and, for pure
// states, the trace distance sqrt(infidelity).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "qs_internal.hpp"

namespace {

using qs::fail;

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 2048;

template <typename T>
__device__ __forceinline__ double2 ld(const T* p, uint64_t i) {
    const T v = p[i];
    return make_double2(double(v.x), double(v.y));
}

__device__ __forceinline__ double wsum(double v) {
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
__device__ __forceinline__ double wmax(double v) {
    for (int m = 32; m > 0; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
    return v;
}

// per block: sum conj(a) b (re, im), sum |a|^2, sum |b|^2, max |a - b|
template <typename TA, typename TB>
__global__ __launch_bounds__(kBlock) void k_compare(const TA* __restrict__ a,
                                                    const TB* __restrict__ b,
                                                    uint64_t n, double* __restrict__ part) {
    double s[5] = {0, 0, 0, 0, 0};
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
        const double2 x = ld(a, i), y = ld(b, i);
        s[0] += x.x * y.x + x.y * y.y;
        s[1] += x.x * y.y - x.y * y.x;
        s[2] += x.x * x.x + x.y * x.y;
        s[3] += y.x * y.x + y.y * y.y;
        s[4] = fmax(s[4], hypot(x.x - y.x, x.y - y.y));
    }
    __shared__ double sh[5][kBlock / 64];
    const int w = threadIdx.x / 64, l = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double v = wsum(s[k]);
        if (l == 0) sh[k][w] = v;
    }
    {
        const double v = wmax(s[4]);
        if (l == 0) sh[4][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        const int k = threadIdx.x;
        double r = sh[k][0];
        for (int j = 1; j < kBlock / 64; ++j) r = k == 4 ? fmax(r, sh[k][j]) : r + sh[k][j];
        part[uint64_t(blockIdx.x) * 5 + k] = r;
    }
}

template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void k_convert(const TI* __restrict__ in,
                                                    TO* __restrict__ out, uint64_t n) {
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
        const TI v = in[i];
        TO r;
        r.x = static_cast<decltype(r.x)>(v.x);
        r.y = static_cast<decltype(r.y)>(v.y);
        out[i] = r;
    }
}

// Streaming copy of n 16-byte chunks, one per thread, consecutive threads on
// consecutive chunks, one launch covering the whole buffer (no grid-stride
// loop): the dispatcher then walks the buffer in order, which measured
// faster on MI355X than persistent grid-stride copies (6.2 TB/s against at
// most 5.5 TB/s with 1-8 chunks in flight per thread).  The practical HBM
// ceiling every other kernel is compared with.
constexpr uint64_t kCopyMaxGrid = uint64_t{1} << 30;  // workgroups per launch
__global__ __launch_bounds__(kBlock) void k_stream_copy(const uint4* __restrict__ in,
                                                        uint4* __restrict__ out, uint64_t n,
                                                        uint64_t base) {
    const uint64_t i = base + uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) out[i] = in[i];
}

unsigned grid_for(uint64_t n) {
    return static_cast<unsigned>(std::max<uint64_t>(
        1, std::min<uint64_t>(kMaxBlocks, (n + kBlock - 1) / kBlock)));
}

bool valid(const qs_state* s) { return s && s->d; }

}  // namespace

extern "C" {

// out[5] = { Re<a|b>, Im<a|b>, <a|a>, <b|b>, max_i |a_i - b_i| }, all FP64.
// For density matrices <a|b> is the Hilbert-Schmidt product tr(a^dagger b).
int qs_compare(const qs_state* a, const qs_state* b, double* out) {
    constexpr const char* W = "qpp_amd::compare()";
    if (!valid(a) || !valid(b) || !out) return fail(QS_ERR_INVALID_HANDLE, W);
    if (a->kind != b->kind) return fail(QS_ERR_TYPE_MISMATCH, W, "kind");
    if (a->dims != b->dims) return fail(QS_ERR_DIMS_NOT_EQUAL, W, "dims");
    if (int rc = qs::ensure_device()) return rc;
    const uint64_t n = a->nelem;
    const unsigned grid = grid_for(n);
    double* part = nullptr;
    hipStream_t s = a->stream;
    // b's pending work must be done before a's stream reads it
    if (b->stream != a->stream) QS_HIP(hipStreamSynchronize(b->stream), W);
    QS_HIP(hipMallocAsync(reinterpret_cast<void**>(&part), size_t(grid) * 5 * 8, s), W);
    const bool a64 = a->dtype == QS_C128, b64 = b->dtype == QS_C128;
    if (a64 && b64)
        hipLaunchKernelGGL((k_compare<double2, double2>), dim3(grid), dim3(kBlock), 0, s,
                           static_cast<const double2*>(a->d), static_cast<const double2*>(b->d),
                           n, part);
    else if (a64)
        hipLaunchKernelGGL((k_compare<double2, float2>), dim3(grid), dim3(kBlock), 0, s,
                           static_cast<const double2*>(a->d), static_cast<const float2*>(b->d),
                           n, part);
    else if (b64)
        hipLaunchKernelGGL((k_compare<float2, double2>), dim3(grid), dim3(kBlock), 0, s,
                           static_cast<const float2*>(a->d), static_cast<const double2*>(b->d),
                           n, part);
    else
        hipLaunchKernelGGL((k_compare<float2, float2>), dim3(grid), dim3(kBlock), 0, s,
                           static_cast<const float2*>(a->d), static_cast<const float2*>(b->d),
                           n, part);
    std::vector<double> h(size_t(grid) * 5);
    hipError_t e = hipMemcpyAsync(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipFreeAsync(part, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return qs::fail_hip(e, W);
    double r[5] = {0, 0, 0, 0, 0};
    for (unsigned g = 0; g < grid; ++g) {
        for (int k = 0; k < 4; ++k) r[k] += h[size_t(g) * 5 + k];
        r[4] = std::max(r[4], h[size_t(g) * 5 + 4]);
    }
    std::copy(r, r + 5, out);
    return QS_OK;
}

// dst := src for two states of the same shape and dtype, with the streaming
// copy kernel (every byte read once and written once: the reference point
// for memory-traffic counters and the practical HBM ceiling).
int qs_state_copy(qs_state* dst, const qs_state* src) {
    constexpr const char* W = "qpp_amd::copy()";
    if (!valid(dst) || !valid(src)) return fail(QS_ERR_INVALID_HANDLE, W);
    if (dst->dtype != src->dtype || dst->kind != src->kind || dst->dims != src->dims)
        return fail(QS_ERR_TYPE_MISMATCH, W, "dst/src");
    if (src->stream != dst->stream) QS_HIP(hipStreamSynchronize(src->stream), W);
    // 16-byte chunks; an odd number of complex64 elements leaves 8 bytes
    // over, copied by the runtime
    const uint64_t chunks = src->bytes() / 16;
    for (uint64_t b = 0; b < chunks; b += kCopyMaxGrid * kBlock) {
        const uint64_t left = chunks - b;
        const uint64_t g = std::min<uint64_t>(kCopyMaxGrid, (left + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_stream_copy, dim3(static_cast<unsigned>(g)), dim3(kBlock), 0,
                           dst->stream, static_cast<const uint4*>(src->d),
                           static_cast<uint4*>(dst->d), chunks, b);
    }
    QS_HIP(hipGetLastError(), W);
    if (src->bytes() % 16)
        QS_HIP(hipMemcpyAsync(static_cast<char*>(dst->d) + chunks * 16,
                              static_cast<const char*>(src->d) + chunks * 16, src->bytes() % 16,
                              hipMemcpyDeviceToDevice, dst->stream),
               W);
    return QS_OK;
}

// A copy of `st` stored as `dtype` (complex128 <-> complex64).
int qs_state_convert(const qs_state* st, qs_dtype dtype, qs_state** out) {
    constexpr const char* W = "qpp_amd::convert()";
    if (!valid(st) || !out) return fail(QS_ERR_INVALID_HANDLE, W);
    if (dtype != QS_C128 && dtype != QS_C64) return fail(QS_ERR_TYPE_MISMATCH, W, "dtype");
    qs_state* r = nullptr;
    {
        const qs::StreamScope on(st->stream);
        if (int rc = qs_state_create(dtype, st->kind, static_cast<uint32_t>(st->dims.size()),
                                     st->dims.data(), &r))
            return rc;
    }
    QS_HIP(hipStreamSynchronize(st->stream), W);
    const uint64_t n = st->nelem;
    const unsigned grid = grid_for(n);
    hipStream_t s = r->stream;
    if (st->dtype == dtype) {
        QS_HIP(hipMemcpyAsync(r->d, st->d, st->bytes(), hipMemcpyDeviceToDevice, s), W);
    } else if (dtype == QS_C64) {
        hipLaunchKernelGGL((k_convert<double2, float2>), dim3(grid), dim3(kBlock), 0, s,
                           static_cast<const double2*>(st->d), static_cast<float2*>(r->d), n);
    } else {
        hipLaunchKernelGGL((k_convert<float2, double2>), dim3(grid), dim3(kBlock), 0, s,
                           static_cast<const float2*>(st->d), static_cast<double2*>(r->d), n);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        qs_state_destroy(r);
        return qs::fail_hip(e, W);
    }
    *out = r;
    return QS_OK;
}

}  // extern "C"

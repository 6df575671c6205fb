// qs_kraus.hip -- Kraus channels beyond the fused superoperator pass.
//
// qs_apply_kraus (qs_api.hip) applies sum_K conj(K) (x) K in one pass when
// the targets span at most 32 basis states.  Wider channels, and qpp's
// full-system apply(A, Ks) with rectangular Kraus operators, are sums of
// K rho K^dagger formed operator by operator:
//   targeted (operations.hpp:476-560): rho_K = copy of rho with K applied
//     to the target digits of the rows and conj(K) to those of the columns
//     (the gate kernels), accumulated into the result;
//   full system (operations.hpp:422-464): T = K rho, then R += T K^dagger,
//     two complex GEMMs per operator (K is Dout x Din, R is Dout x Dout).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "qs_internal.hpp"

namespace {

using qs::fail;

constexpr int kT = 32;   // output tile edge
constexpr int kKc = 16;  // inner-dimension chunk

template <typename T>
struct Acc;
template <>
struct Acc<double2> {
    using R = double;
};
template <>
struct Acc<float2> {
    using R = float;
};

// C (+)= A op(B), column-major complex; op(B) = B (CONJ = false, B is K x N)
// or B^dagger (CONJ = true, B is N x K).  A is M x K.  Each workgroup makes
// one 32 x 32 tile of C, every thread 2 x 2 of its entries, the inner
// dimension staged through LDS 16 at a time; the sums run in FP64.
template <typename T, bool CONJ>
__global__ __launch_bounds__(256) void k_zgemm(uint64_t M, uint64_t N, uint64_t K,
                                               const T* __restrict__ A, uint64_t lda,
                                               const T* __restrict__ B, uint64_t ldb,
                                               T* __restrict__ C, uint64_t ldc, int accumulate) {
    __shared__ double2 As[kKc][kT];
    __shared__ double2 Bs[kKc][kT];
    const int t = threadIdx.x, tx = t % 16, ty = t / 16;
    const uint64_t m0 = uint64_t(blockIdx.x) * kT, n0 = uint64_t(blockIdx.y) * kT;
    double2 acc[2][2] = {{{0, 0}, {0, 0}}, {{0, 0}, {0, 0}}};
    for (uint64_t k0 = 0; k0 < K; k0 += kKc) {
        for (int e = t; e < kKc * kT; e += 256) {
            const int kk = e / kT, mm = e % kT;
            const uint64_t m = m0 + mm, k = k0 + kk;
            double2 a{0, 0};
            if (m < M && k < K) {
                const T v = A[m + k * lda];
                a = double2{double(v.x), double(v.y)};
            }
            As[kk][mm] = a;
            const uint64_t n = n0 + mm;  // the same loop fills the B tile
            double2 b{0, 0};
            if (n < N && k < K) {
                if (CONJ) {
                    const T v = B[n + k * ldb];
                    b = double2{double(v.x), -double(v.y)};
                } else {
                    const T v = B[k + n * ldb];
                    b = double2{double(v.x), double(v.y)};
                }
            }
            Bs[kk][mm] = b;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kKc; ++kk) {
            const double2 a0 = As[kk][tx], a1 = As[kk][tx + 16];
            const double2 b0 = Bs[kk][ty], b1 = Bs[kk][ty + 16];
            const double2 av[2] = {a0, a1}, bv[2] = {b0, b1};
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    acc[i][j].x += av[i].x * bv[j].x - av[i].y * bv[j].y;
                    acc[i][j].y += av[i].x * bv[j].y + av[i].y * bv[j].x;
                }
        }
        __syncthreads();
    }
    using R = typename Acc<T>::R;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const uint64_t m = m0 + tx + 16 * i, n = n0 + ty + 16 * j;
            if (m >= M || n >= N) continue;
            T& c = C[m + n * ldc];
            if (accumulate) {
                c.x = R(double(c.x) + acc[i][j].x);
                c.y = R(double(c.y) + acc[i][j].y);
            } else {
                c.x = R(acc[i][j].x);
                c.y = R(acc[i][j].y);
            }
        }
}

// acc += x, elementwise
template <typename T>
__global__ __launch_bounds__(256) void k_axpy(T* __restrict__ acc, const T* __restrict__ x,
                                              uint64_t n) {
    const uint64_t stride = uint64_t(gridDim.x) * 256;
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) {
        acc[i].x += x[i].x;
        acc[i].y += x[i].y;
    }
}

unsigned grid_of(uint64_t n) {
    return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(65536, (n + 255) / 256)));
}

template <typename T, bool CONJ>
void gemm(hipStream_t s, uint64_t M, uint64_t N, uint64_t K, const T* A, uint64_t lda,
          const T* B, uint64_t ldb, T* C, uint64_t ldc, bool accumulate) {
    const dim3 g(static_cast<unsigned>((M + kT - 1) / kT), static_cast<unsigned>((N + kT - 1) / kT));
    hipLaunchKernelGGL((k_zgemm<T, CONJ>), g, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc,
                       int(accumulate));
}

// the nK operators (complex128, column-major, back to back) as the state's
// element type, in device memory
template <typename T>
int upload_ops(hipStream_t s, const double* Ks, uint64_t count, T** out) {
    std::vector<T> h(count);
    for (uint64_t i = 0; i < count; ++i) {
        h[i].x = static_cast<decltype(h[i].x)>(Ks[2 * i]);
        h[i].y = static_cast<decltype(h[i].y)>(Ks[2 * i + 1]);
    }
    QS_HIP(hipMallocAsync(reinterpret_cast<void**>(out), count * sizeof(T), s), "qpp::apply()");
    QS_HIP(hipMemcpyAsync(*out, h.data(), count * sizeof(T), hipMemcpyHostToDevice, s),
           "qpp::apply()");
    QS_HIP(hipStreamSynchronize(s), "qpp::apply()");  // h goes out of scope
    return QS_OK;
}

template <typename T>
int kraus_full(const qs_state* rho, const double* Ks, uint32_t nK, uint64_t Dout, qs_state* out) {
    const uint64_t Din = rho->D;
    hipStream_t s = out->stream;
    T *dK = nullptr, *tmp = nullptr;
    if (int rc = upload_ops<T>(s, Ks, uint64_t(nK) * Dout * Din, &dK)) return rc;
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&tmp), Dout * Din * sizeof(T), s);
    if (e == hipSuccess) {
        const T* r = static_cast<const T*>(rho->d);
        T* R = static_cast<T*>(out->d);
        for (uint32_t k = 0; k < nK; ++k) {
            const T* Kk = dK + uint64_t(k) * Dout * Din;
            gemm<T, false>(s, Dout, Din, Din, Kk, Dout, r, Din, tmp, Dout, false);  // K rho
            gemm<T, true>(s, Dout, Dout, Din, tmp, Dout, Kk, Dout, R, Dout, k > 0);  // (K rho) K^+
        }
        e = hipGetLastError();
        (void)hipFreeAsync(tmp, s);
    }
    (void)hipFreeAsync(dK, s);
    const hipError_t e2 = hipStreamSynchronize(s);
    if (e != hipSuccess) return qs::fail_hip(e, "qpp::apply()");
    if (e2 != hipSuccess) return qs::fail_hip(e2, "qpp::apply()");
    return QS_OK;
}

}  // namespace

// sum_K K rho K^dagger on the targets, one operator at a time: each term is
// a copy of rho with K applied to the target digits (qs_apply: K on the rows,
// conj(K) on the columns), added into the result.
int qs::kraus_per_operator(qs_state* st, const double* Ks, uint32_t nK, uint64_t DK,
                           uint32_t ntarget, const uint64_t* target) {
    constexpr const char* W = "qpp::apply()";
    qs_state* acc = nullptr;
    for (uint32_t k = 0; k < nK; ++k) {
        qs_state* term = nullptr;
        if (int rc = qs_state_clone(st, &term)) {
            qs_state_destroy(acc);
            return rc;
        }
        int rc = qs_apply(term, Ks + 2 * uint64_t(k) * DK * DK, DK, ntarget, target);
        if (rc == QS_OK && acc) {
            if (st->dtype == QS_C128)
                hipLaunchKernelGGL(k_axpy<double2>, dim3(grid_of(st->nelem)), dim3(256), 0,
                                   st->stream, static_cast<double2*>(acc->d),
                                   static_cast<const double2*>(term->d), st->nelem);
            else
                hipLaunchKernelGGL(k_axpy<float2>, dim3(grid_of(st->nelem)), dim3(256), 0,
                                   st->stream, static_cast<float2*>(acc->d),
                                   static_cast<const float2*>(term->d), st->nelem);
            if (hipGetLastError() != hipSuccess) rc = fail(QS_ERR_DEVICE, W, "k_axpy");
            qs_state_destroy(term);
        } else if (rc == QS_OK) {
            acc = term;
        } else {
            qs_state_destroy(term);
        }
        if (rc != QS_OK) {
            qs_state_destroy(acc);
            return rc;
        }
    }
    // the result replaces rho's buffer
    std::swap(st->d, acc->d);
    qs_state_destroy(acc);
    QS_HIP(hipStreamSynchronize(st->stream), W);
    return QS_OK;
}

extern "C" {

// qpp::apply(A, Ks), operations.hpp:422-464
int qs_apply_kraus_full(const qs_state* rho, const double* Ks, uint32_t nK, uint64_t rowsK,
                        uint64_t colsK, qs_state** out) {
    constexpr const char* W = "qpp::apply()";
    if (!rho || !rho->d || !out) return fail(QS_ERR_INVALID_HANDLE, W, "A");
    qs::clear_error();
    *out = nullptr;
    if (rho->kind != QS_RHO) return fail(QS_ERR_MATRIX_NOT_SQUARE, W, "A");
    if (nK == 0 || !Ks || rowsK == 0) return fail(QS_ERR_ZERO_SIZE, W, "Ks");
    if (colsK != rho->D) return fail(QS_ERR_DIMS_MISMATCH_MATRIX, W, "Ks[0]");
    const uint64_t dims[1] = {rowsK};
    qs_state* r = nullptr;
    {
        const qs::StreamScope on(rho->stream);
        if (int rc = qs_state_create(rho->dtype, QS_RHO, 1, dims, &r)) return rc;
    }
    const int rc = rho->dtype == QS_C128 ? kraus_full<double2>(rho, Ks, nK, rowsK, r)
                                         : kraus_full<float2>(rho, Ks, nK, rowsK, r);
    if (rc != QS_OK) {
        qs_state_destroy(r);
        return rc;
    }
    *out = r;
    return QS_OK;
}

}  // extern "C"

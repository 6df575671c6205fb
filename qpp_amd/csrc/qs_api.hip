// qs_api.hip -- C ABI: state handles, argument validation and dispatch of
// the gate-application entry points (qpp_amd.h).
//
// Validation restates qpp's checks in qpp's order, so a bad call fails with
// the status (and the message) of the exception qpp would throw:
//   apply          operations.hpp:80-138
//   applyCTRL      operations.hpp:629-741
//   applyCTRL_fan  operations.hpp:972-1090
//   apply(A, Ks)   operations.hpp:480-539
// Dispatch: qubit states map every operation onto bit-addressed kernels over
// the element index (density matrices: row bits then column bits), qudit
// states onto the mixed-radix kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "qs_internal.hpp"

namespace qs {

namespace {
thread_local std::string g_last_error;
std::once_flag g_init_flag;
int g_init_status = QS_OK;
hipStream_t g_stream = nullptr;
thread_local hipStream_t t_stream = nullptr;  // per-thread override (serving lanes)
}  // namespace

const char* description(int code) {
    switch (code) {
        case QS_ERR_ZERO_SIZE: return "Object has zero size";
        case QS_ERR_MATRIX_NOT_SQUARE: return "Matrix is not square";
        case QS_ERR_MATRIX_NOT_CVECTOR: return "Matrix is not a column vector";
        case QS_ERR_MATRIX_NOT_SQUARE_NOR_CVECTOR:
            return "Matrix is not square nor column vector";
        case QS_ERR_MATRIX_MISMATCH_SUBSYS: return "Matrix mismatch subsystems";
        case QS_ERR_DIMS_INVALID: return "Invalid dimension(s)";
        case QS_ERR_DIMS_NOT_EQUAL: return "Dimensions not equal";
        case QS_ERR_DIMS_MISMATCH_MATRIX: return "Dimension(s) mismatch matrix size";
        case QS_ERR_DIMS_MISMATCH_CVECTOR:
            return "Dimension(s) mismatch column vector size";
        case QS_ERR_SUBSYS_MISMATCH_DIMS: return "Subsystems mismatch dimensions";
        case QS_ERR_PERM_INVALID: return "Invalid permutation";
        case QS_ERR_PERM_MISMATCH_DIMS: return "Permutation mismatch dimensions";
        case QS_ERR_OUT_OF_RANGE: return "Argument out of range";
        case QS_ERR_TYPE_MISMATCH: return "Type mismatch";
        case QS_ERR_SIZE_MISMATCH: return "Size mismatch";
        case QS_ERR_QUDIT_ALREADY_MEASURED:
            return "Qudit was already measured (destructively)";
        case QS_ERR_NOT_IMPLEMENTED: return "Not yet implemented";
        case QS_ERR_OVERFLOW: return "Division by zero";
        case QS_ERR_IO: return "I/O error";
        case QS_ERR_INVALID_CONDITIONAL: return "Invalid conditional";
        case QS_ERR_INVALID_HANDLE: return "Invalid handle";
        case QS_ERR_OUT_OF_MEMORY: return "Out of device memory";
        case QS_ERR_DEVICE: return "HIP runtime error";
        case QS_ERR_NO_DEVICE: return "No usable gfx950 device";
        default: return "UNKNOWN EXCEPTION";
    }
}

int fail(int code, const char* where, const char* context) {
    // qpp::exception::Exception::what(), classes/exception.hpp:98-113
    g_last_error = std::string("[WHERE: ") + where + "] " + description(code) + "!";
    if (context) g_last_error += std::string(" [CONTEXT: ") + context + "]";
    return code;
}

int fail_runtime(int code, const std::string& text) {
    g_last_error = text;  // a std::runtime_error's what(), verbatim
    return code;
}

int fail_hip(hipError_t e, const char* where) {
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
                         ? QS_ERR_OUT_OF_MEMORY
                         : QS_ERR_DEVICE;
    return fail(code, where, hipGetErrorString(e));
}

void clear_error() { g_last_error.clear(); }

int ensure_device() {
    std::call_once(g_init_flag, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            g_init_status = QS_ERR_NO_DEVICE;
            return;
        }
        if (hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) != hipSuccess) {
            g_init_status = QS_ERR_DEVICE;
            return;
        }
        // State buffers are stream-ordered allocations from the device's
        // default pool; keep freed blocks cached so a freed state's memory
        // serves the next request without a trip through the driver.
        int dev = 0;
        hipMemPool_t pool;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            uint64_t keep = ~uint64_t{0};
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
    });
    if (g_init_status != QS_OK)
        return fail(g_init_status, "qpp_amd::init",
                    g_init_status == QS_ERR_NO_DEVICE ? "hipGetDeviceCount" : "stream");
    return QS_OK;
}

hipStream_t default_stream() { return g_stream; }

hipStream_t current_stream() { return t_stream ? t_stream : g_stream; }

void set_thread_stream(hipStream_t s) { t_stream = s; }

bool force_exact_sampling() {
    const char* v = std::getenv("QPP_AMD_FORCE_EXACT");
    return v && v[0] == '1';
}

bool gatek_valu_forced() {
    const char* v = std::getenv("QPP_AMD_GATEK_VALU");
    return v && v[0] == '1';
}

StreamScope::StreamScope(hipStream_t s) : old(t_stream) { t_stream = s; }
StreamScope::~StreamScope() { t_stream = old; }

}  // namespace qs

namespace {

using qs::fail;

// ---------------------------------------------------------------------------
// small kernels: fill / basis
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_set_basis(T* a, uint64_t n, uint64_t index) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += stride) {
        T v;
        v.x = (i == index) ? 1 : 0;
        v.y = 0;
        a[i] = v;
    }
}

int set_basis(qs_state* st, uint64_t element) {
    const uint64_t n = st->nelem;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (st->dtype == QS_C128)
        hipLaunchKernelGGL(k_set_basis<double2>, dim3(blocks), dim3(256), 0,
                           st->stream, static_cast<double2*>(st->d), n, element);
    else
        hipLaunchKernelGGL(k_set_basis<float2>, dim3(blocks), dim3(256), 0,
                           st->stream, static_cast<float2*>(st->d), n, element);
    QS_HIP(hipGetLastError(), "qpp_amd::set_basis");
    return QS_OK;
}

bool valid(const qs_state* st) { return st != nullptr && st->d != nullptr; }

// ---------------------------------------------------------------------------
// validators (internal/util.hpp:158-287)
// ---------------------------------------------------------------------------
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

std::vector<uint64_t> to_vec(const uint64_t* p, uint32_t n) {
    return p ? std::vector<uint64_t>(p, p + n) : std::vector<uint64_t>();
}

// bit position of qubit q: rows (and kets) n-1-q, columns 2n-1-q
uint32_t row_bit(const qs_state* st, uint64_t q) {
    return static_cast<uint32_t>(st->dims.size() - 1 - q);
}
uint32_t col_bit(const qs_state* st, uint64_t q) {
    return static_cast<uint32_t>(2 * st->dims.size() - 1 - q);
}

// element strides of subsystem q (row index; the column stride is D times it)
uint64_t row_stride(const qs_state* st, uint64_t q) {
    uint64_t s = 1;
    for (uint64_t i = st->dims.size(); i-- > q + 1;) s *= st->dims[i];
    return s;
}

// Qudit gate descriptor for `targets` on the row (cols = false) or column
// index.  Spectators: every other digit of the element index.
qs::QuditGate qudit_desc(const qs_state* st, const std::vector<uint64_t>& tg,
                         bool cols) {
    qs::QuditGate g{};
    const uint64_t n = st->dims.size();
    const bool rho = st->kind == QS_RHO;
    const uint64_t scale = cols ? st->D : 1;
    g.ntg = static_cast<uint32_t>(tg.size());
    for (size_t i = 0; i < tg.size(); ++i) {
        g.tg_stride[i] = row_stride(st, tg[i]) * scale;
        g.tg_dim[i] = st->dims[tg[i]];
    }
    // spectators in order of decreasing stride (mixed radix, last fastest)
    auto add_sp = [&](uint64_t stride, uint64_t dim) {
        g.sp_stride[g.nsp] = stride;
        g.sp_dim[g.nsp] = dim;
        ++g.nsp;
    };
    if (rho)
        for (uint64_t q = 0; q < n; ++q) {
            if (cols && std::find(tg.begin(), tg.end(), q) != tg.end()) continue;
            add_sp(row_stride(st, q) * st->D, st->dims[q]);
        }
    for (uint64_t q = 0; q < n; ++q) {
        if (!cols && std::find(tg.begin(), tg.end(), q) != tg.end()) continue;
        add_sp(row_stride(st, q), st->dims[q]);
    }
    return g;
}

// U (and conj(U) on the columns of a density matrix) on `target` under the
// controls (ctrl, shift); ctrl empty = uncontrolled.
int gate_dispatch(qs_state* st, const double* A, uint64_t dimA,
                  const std::vector<uint64_t>& target,
                  const std::vector<uint64_t>& ctrl,
                  const std::vector<uint64_t>& shift) {
    const bool rho = st->kind == QS_RHO;
    if (st->qubits) {
        const uint32_t k = static_cast<uint32_t>(target.size());
        for (int pass = 0; pass < (rho ? 2 : 1); ++pass) {
            const bool cols = pass == 1;
            std::vector<uint32_t> bits(k);
            for (uint32_t i = 0; i < k; ++i)
                bits[i] = cols ? col_bit(st, target[i]) : row_bit(st, target[i]);
            uint64_t cmask = 0, cval = 0;
            for (size_t c = 0; c < ctrl.size(); ++c) {
                const uint32_t b = cols ? col_bit(st, ctrl[c]) : row_bit(st, ctrl[c]);
                cmask |= uint64_t{1} << b;
                // shift 0: control on |1>, shift 1: on |0> (apply_ctrl.hpp:123-132)
                if (shift[c] == 0) cval |= uint64_t{1} << b;
            }
            if (int rc = qs::launch_gatek(st, bits.data(), k, A, cmask, cval, cols))
                return rc;
        }
        return QS_OK;
    }
    // qudits: A^r table for controlled gates (operations.hpp:777-782)
    std::vector<double> mats;
    uint32_t nm = 1;
    uint64_t dc = 2;
    if (!ctrl.empty()) {
        dc = st->dims[ctrl[0]];
        nm = static_cast<uint32_t>(dc);
        const uint64_t n2 = dimA * dimA;
        mats.assign(2 * n2 * dc, 0.0);
        // A^0 = I, A^r = A^(r-1) A (complex, column-major)
        for (uint64_t i = 0; i < dimA; ++i) mats[2 * (i + i * dimA)] = 1.0;
        for (uint64_t r = 1; r < dc; ++r) {
            const double* P = mats.data() + 2 * (r - 1) * n2;
            double* Q = mats.data() + 2 * r * n2;
            for (uint64_t i = 0; i < dimA; ++i)
                for (uint64_t j = 0; j < dimA; ++j) {
                    double re = 0, im = 0;
                    for (uint64_t l = 0; l < dimA; ++l) {
                        const double ar = P[2 * (i + l * dimA)], ai = P[2 * (i + l * dimA) + 1];
                        const double br = A[2 * (l + j * dimA)], bi = A[2 * (l + j * dimA) + 1];
                        re += ar * br - ai * bi;
                        im += ar * bi + ai * br;
                    }
                    Q[2 * (i + j * dimA)] = re;
                    Q[2 * (i + j * dimA) + 1] = im;
                }
        }
    } else {
        mats.assign(A, A + 2 * dimA * dimA);
    }
    for (int pass = 0; pass < (rho ? 2 : 1); ++pass) {
        const bool cols = pass == 1;
        qs::QuditGate g = qudit_desc(st, target, cols);
        if (!ctrl.empty()) {
            g.nctrl = static_cast<uint32_t>(ctrl.size());
            g.c_dim = dc;
            for (size_t c = 0; c < ctrl.size(); ++c) {
                g.c_stride[c] = row_stride(st, ctrl[c]) * (cols ? st->D : 1);
                g.c_shift[c] = shift[c];
            }
        }
        if (int rc = qs::launch_qudit_gate(st, g, mats.data(), nm, dimA, cols))
            return rc;
    }
    return QS_OK;
}

}  // namespace

extern "C" {

const char* qs_last_error(void) { return qs::g_last_error.c_str(); }

int qs_device_info(char* name, size_t name_len, int* num_cu, uint64_t* hbm_bytes,
                   uint64_t* hbm_free) {
    if (int rc = qs::ensure_device()) return rc;
    int dev = 0;
    QS_HIP(hipGetDevice(&dev), "qs_device_info");
    hipDeviceProp_t p;
    QS_HIP(hipGetDeviceProperties(&p, dev), "qs_device_info");
    if (name && name_len) {
        std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (num_cu) *num_cu = p.multiProcessorCount;
    size_t fr = 0, tot = 0;
    QS_HIP(hipMemGetInfo(&fr, &tot), "qs_device_info");
    if (hbm_bytes) *hbm_bytes = tot;
    if (hbm_free) *hbm_free = fr;
    return QS_OK;
}

int qs_synchronize(void) {
    if (int rc = qs::ensure_device()) return rc;
    QS_HIP(hipDeviceSynchronize(), "qs_synchronize");
    return QS_OK;
}

static hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;

int qs_profile_begin(void) {
    if (int rc = qs::ensure_device()) return rc;
    if (!g_ev0) {
        QS_HIP(hipEventCreate(&g_ev0), "qs_profile_begin");
        QS_HIP(hipEventCreate(&g_ev1), "qs_profile_begin");
    }
    QS_HIP(hipEventRecord(g_ev0, qs::default_stream()), "qs_profile_begin");
    return QS_OK;
}

int qs_profile_end(double* ms) {
    if (!g_ev0) return fail(QS_ERR_INVALID_HANDLE, "qs_profile_end", "no begin");
    QS_HIP(hipEventRecord(g_ev1, qs::default_stream()), "qs_profile_end");
    QS_HIP(hipEventSynchronize(g_ev1), "qs_profile_end");
    float f = 0;
    QS_HIP(hipEventElapsedTime(&f, g_ev0, g_ev1), "qs_profile_end");
    if (ms) *ms = f;
    return QS_OK;
}

int qs_state_create(qs_dtype dtype, qs_kind kind, uint32_t ndims,
                    const uint64_t* dims, qs_state** out) {
    if (!out) return fail(QS_ERR_INVALID_HANDLE, "qs_state_create", "out");
    *out = nullptr;
    if (dtype != QS_C128 && dtype != QS_C64)
        return fail(QS_ERR_TYPE_MISMATCH, "qs_state_create", "dtype");
    if (kind != QS_KET && kind != QS_RHO)
        return fail(QS_ERR_MATRIX_NOT_SQUARE_NOR_CVECTOR, "qs_state_create", "kind");
    const auto dv = to_vec(dims, ndims);
    // check_dims, internal/util.hpp:158-166
    if (dv.empty() || std::find(dv.begin(), dv.end(), uint64_t{0}) != dv.end())
        return fail(QS_ERR_DIMS_INVALID, "qs_state_create", "dims");
    if (int rc = qs::ensure_device()) return rc;
    auto* st = new qs_state;
    st->dtype = dtype;
    st->kind = kind;
    st->dims = dv;
    st->D = 1;
    for (uint64_t d : dv) {
        if (st->D > (uint64_t{1} << 40) / d) {
            delete st;
            return fail(QS_ERR_OUT_OF_RANGE, "qs_state_create", "dims");
        }
        st->D *= d;
    }
    st->nelem = kind == QS_KET ? st->D : st->D * st->D;
    st->qubits = std::all_of(dv.begin(), dv.end(), [](uint64_t d) { return d == 2; });
    st->nbits = st->qubits ? static_cast<uint32_t>((kind == QS_KET ? 1 : 2) * dv.size()) : 0;
    st->stream = qs::current_stream();
    hipError_t e = hipMallocAsync(&st->d, st->bytes(), st->stream);
    if (e != hipSuccess) {
        delete st;
        return qs::fail_hip(e, "qs_state_create");
    }
    if (int rc = set_basis(st, 0)) {
        (void)hipFreeAsync(st->d, st->stream);
        delete st;
        return rc;
    }
    *out = st;
    return QS_OK;
}

// The buffer goes back to the pool in stream order after the state's last
// queued operation; the host does not wait.
int qs_state_destroy(qs_state* st) {
    if (!st) return QS_OK;
    if (st->d) (void)hipFreeAsync(st->d, st->stream);
    delete st;
    return QS_OK;
}

int qs_state_clone(const qs_state* st, qs_state** out) {
    if (!valid(st) || !out) return fail(QS_ERR_INVALID_HANDLE, "qs_state_clone");
    const qs::StreamScope on(st->stream);  // the copy lives on the source's stream
    if (int rc = qs_state_create(st->dtype, st->kind,
                                 static_cast<uint32_t>(st->dims.size()),
                                 st->dims.data(), out))
        return rc;
    QS_HIP(hipMemcpyAsync((*out)->d, st->d, st->bytes(), hipMemcpyDeviceToDevice,
                          st->stream),
           "qs_state_clone");
    return QS_OK;
}

int qs_state_info(const qs_state* st, qs_dtype* dtype, qs_kind* kind,
                  uint32_t* ndims, uint64_t* dims_out, uint64_t* num_elems) {
    if (!valid(st)) return fail(QS_ERR_INVALID_HANDLE, "qs_state_info");
    if (dtype) *dtype = st->dtype;
    if (kind) *kind = st->kind;
    if (ndims) *ndims = static_cast<uint32_t>(st->dims.size());
    if (dims_out) std::copy(st->dims.begin(), st->dims.end(), dims_out);
    if (num_elems) *num_elems = st->nelem;
    return QS_OK;
}

int qs_state_upload(qs_state* st, const void* host, uint64_t nbytes) {
    if (!valid(st)) return fail(QS_ERR_INVALID_HANDLE, "qs_state_upload");
    if (nbytes != st->bytes())
        return fail(st->kind == QS_KET ? QS_ERR_DIMS_MISMATCH_CVECTOR
                                       : QS_ERR_DIMS_MISMATCH_MATRIX,
                    "qs_state_upload", "dims/state");
    QS_HIP(hipMemcpyAsync(st->d, host, nbytes, hipMemcpyHostToDevice, st->stream),
           "qs_state_upload");
    QS_HIP(hipStreamSynchronize(st->stream), "qs_state_upload");
    return QS_OK;
}

int qs_state_download(const qs_state* st, void* host, uint64_t nbytes) {
    if (!valid(st)) return fail(QS_ERR_INVALID_HANDLE, "qs_state_download");
    if (nbytes != st->bytes())
        return fail(QS_ERR_SIZE_MISMATCH, "qs_state_download", "nbytes");
    QS_HIP(hipMemcpyAsync(host, st->d, nbytes, hipMemcpyDeviceToHost, st->stream),
           "qs_state_download");
    QS_HIP(hipStreamSynchronize(st->stream), "qs_state_download");
    return QS_OK;
}

int qs_state_set_basis(qs_state* st, uint64_t index) {
    if (!valid(st)) return fail(QS_ERR_INVALID_HANDLE, "qs_state_set_basis");
    if (index >= st->D) return fail(QS_ERR_OUT_OF_RANGE, "qs_state_set_basis", "index");
    return set_basis(st, st->kind == QS_KET ? index : index + index * st->D);
}

int qs_state_device_ptr(const qs_state* st, void** dptr) {
    if (!valid(st) || !dptr) return fail(QS_ERR_INVALID_HANDLE, "qs_state_device_ptr");
    *dptr = st->d;
    return QS_OK;
}

// qpp::apply(state, A, target, dims), operations.hpp:72-372
int qs_apply(qs_state* st, const double* A, uint64_t dimA, uint32_t ntarget,
             const uint64_t* target) {
    constexpr const char* W = "qpp::apply()";
    if (!valid(st)) return fail(QS_ERR_INVALID_HANDLE, W, "state");
    qs::clear_error();
    if (dimA == 0 || !A) return fail(QS_ERR_ZERO_SIZE, W, "A");
    if (ntarget == 0) return fail(QS_ERR_ZERO_SIZE, W, "target");
    const auto tv = to_vec(target, ntarget);
    if (!check_subsys(tv, st->dims))
        return fail(QS_ERR_SUBSYS_MISMATCH_DIMS, W, "dims/target");
    uint64_t DA = 1;
    for (uint64_t t : tv) DA *= st->dims[t];
    if (DA != dimA) return fail(QS_ERR_MATRIX_MISMATCH_SUBSYS, W, "A/dims/target");
    if (st->D == 1) return QS_OK;
    return gate_dispatch(st, A, dimA, tv, {}, {});
}

// qpp::applyCTRL, operations.hpp:616-887
int qs_apply_ctrl(qs_state* st, const double* A, uint64_t dimA, uint32_t nctrl,
                  const uint64_t* ctrl, uint32_t ntarget, const uint64_t* target,
                  const uint64_t* shift) {
    constexpr const char* W = "qpp::applyCTRL()";
    if (!valid(st)) return fail(QS_ERR_INVALID_HANDLE, W, "state");
    qs::clear_error();
    if (dimA == 0 || !A) return fail(QS_ERR_ZERO_SIZE, W, "A");
    if (nctrl == 0) return fail(QS_ERR_ZERO_SIZE, W, "ctrl");
    if (ntarget == 0) return fail(QS_ERR_ZERO_SIZE, W, "target");
    const auto cv = to_vec(ctrl, nctrl);
    const auto tv = to_vec(target, ntarget);
    if (!check_subsys(cv, st->dims))
        return fail(QS_ERR_SUBSYS_MISMATCH_DIMS, W, "ctrl/dims");
    const uint64_t d = st->dims[cv[0]];
    for (uint64_t c : cv)
        if (st->dims[c] != d) return fail(QS_ERR_DIMS_NOT_EQUAL, W, "ctrl");
    if (!check_subsys(tv, st->dims))
        return fail(QS_ERR_SUBSYS_MISMATCH_DIMS, W, "dims/target");
    for (uint64_t c : cv)
        for (uint64_t t : tv)
            if (c == t) return fail(QS_ERR_OUT_OF_RANGE, W, "ctrl/target");
    uint64_t DA = 1;
    for (uint64_t t : tv) DA *= st->dims[t];
    if (DA != dimA) return fail(QS_ERR_MATRIX_MISMATCH_SUBSYS, W, "A/target");
    std::vector<uint64_t> sv(nctrl, 0);
    if (shift) {
        sv = to_vec(shift, nctrl);
        for (uint64_t s : sv)
            if (s >= d) return fail(QS_ERR_OUT_OF_RANGE, W, "shift");
    }
    if (st->D == 1) return QS_OK;
    return gate_dispatch(st, A, dimA, tv, cv, sv);
}

// qpp::applyCTRL_fan, operations.hpp:961-1234: A (d x d) on every target in
// the given order, each time under the same controls.
int qs_apply_ctrl_fan(qs_state* st, const double* A, uint64_t dimA, uint32_t nctrl,
                      const uint64_t* ctrl, uint32_t ntarget, const uint64_t* target,
                      const uint64_t* shift) {
    constexpr const char* W = "qpp::applyCTRL_fan()";
    if (!valid(st)) return fail(QS_ERR_INVALID_HANDLE, W, "state");
    qs::clear_error();
    if (dimA == 0 || !A) return fail(QS_ERR_ZERO_SIZE, W, "A");
    if (nctrl == 0) return fail(QS_ERR_ZERO_SIZE, W, "ctrl");
    if (ntarget == 0) return fail(QS_ERR_ZERO_SIZE, W, "target");
    const auto cv = to_vec(ctrl, nctrl);
    const auto tv = to_vec(target, ntarget);
    if (!check_subsys(cv, st->dims))
        return fail(QS_ERR_SUBSYS_MISMATCH_DIMS, W, "ctrl/dims");
    const uint64_t d = st->dims[cv[0]];
    for (uint64_t c : cv)
        if (st->dims[c] != d) return fail(QS_ERR_DIMS_NOT_EQUAL, W, "ctrl");
    if (!check_subsys(tv, st->dims))
        return fail(QS_ERR_SUBSYS_MISMATCH_DIMS, W, "dims/target");
    for (uint64_t t : tv)
        if (st->dims[t] != d) return fail(QS_ERR_DIMS_NOT_EQUAL, W, "target");
    for (uint64_t c : cv)
        for (uint64_t t : tv)
            if (c == t) return fail(QS_ERR_OUT_OF_RANGE, W, "ctrl/target");
    if (dimA != d) return fail(QS_ERR_MATRIX_MISMATCH_SUBSYS, W, "A/target");
    std::vector<uint64_t> sv(nctrl, 0);
    if (shift) {
        sv = to_vec(shift, nctrl);
        for (uint64_t s : sv)
            if (s >= d) return fail(QS_ERR_OUT_OF_RANGE, W, "shift");
    }
    if (st->D == 1) return QS_OK;
    for (uint64_t t : tv)
        if (int rc = gate_dispatch(st, A, dimA, {t}, cv, sv)) return rc;
    return QS_OK;
}

// qpp::apply(rho, Ks, target, dims), operations.hpp:476-560, fused into one
// pass: the superoperator S = sum_K conj(K) (x) K acts on the (column digits,
// row digits) of the targets of vec(rho).  Targets spanning more than 32
// basis states take qs::kraus_per_operator (qs_kraus.hip).
int qs_apply_kraus(qs_state* st, const double* Ks, uint32_t nK, uint64_t dimK,
                   uint32_t ntarget, const uint64_t* target) {
    constexpr const char* W = "qpp::apply()";
    if (!valid(st)) return fail(QS_ERR_INVALID_HANDLE, W, "A");
    qs::clear_error();
    if (ntarget == 0) return fail(QS_ERR_ZERO_SIZE, W, "target");
    // a ket is not a square matrix (operations.hpp:493-496)
    if (st->kind != QS_RHO) return fail(QS_ERR_MATRIX_NOT_SQUARE, W, "A");
    const auto tv = to_vec(target, ntarget);
    if (!check_subsys(tv, st->dims))
        return fail(QS_ERR_SUBSYS_MISMATCH_DIMS, W, "dims/target");
    if (nK == 0 || !Ks) return fail(QS_ERR_ZERO_SIZE, W, "Ks");
    uint64_t DK = 1;
    for (uint64_t t : tv) DK *= st->dims[t];
    if (dimK != DK) return fail(QS_ERR_MATRIX_MISMATCH_SUBSYS, W, "dims/Ks[0]/target");
    // above 32 target basis states the superoperator (DK^2 x DK^2) is too
    // large for one pass: one term K rho K^dagger per operator instead
    if (DK > 32) return qs::kraus_per_operator(st, Ks, nK, DK, ntarget, target);
    const uint64_t DS = DK * DK;
    // S[(a1, b1), (a2, b2)] = sum_K conj(K[a1, a2]) K[b1, b2], column-major
    std::vector<double> S(2 * DS * DS, 0.0);
    for (uint32_t k = 0; k < nK; ++k) {
        const double* K = Ks + 2 * k * DK * DK;
        for (uint64_t a1 = 0; a1 < DK; ++a1)
            for (uint64_t b1 = 0; b1 < DK; ++b1)
                for (uint64_t a2 = 0; a2 < DK; ++a2)
                    for (uint64_t b2 = 0; b2 < DK; ++b2) {
                        const double xr = K[2 * (a1 + a2 * DK)], xi = -K[2 * (a1 + a2 * DK) + 1];
                        const double yr = K[2 * (b1 + b2 * DK)], yi = K[2 * (b1 + b2 * DK) + 1];
                        const uint64_t r = a1 * DK + b1, c = a2 * DK + b2;
                        S[2 * (r + c * DS)] += xr * yr - xi * yi;
                        S[2 * (r + c * DS) + 1] += xr * yi + xi * yr;
                    }
    }
    if (st->D == 1) return QS_OK;
    if (st->qubits) {
        const uint32_t k = static_cast<uint32_t>(tv.size());
        std::vector<uint32_t> bits(2 * k);
        for (uint32_t i = 0; i < k; ++i) {
            bits[i] = col_bit(st, tv[i]);
            bits[k + i] = row_bit(st, tv[i]);
        }
        return qs::launch_gatek(st, bits.data(), 2 * k, S.data(), 0, 0, false);
    }
    // qudits: gate digits = column digits then row digits of the targets
    qs::QuditGate g{};
    const uint64_t n = st->dims.size();
    for (uint64_t t : tv) {
        g.tg_stride[g.ntg] = row_stride(st, t) * st->D;
        g.tg_dim[g.ntg++] = st->dims[t];
    }
    for (uint64_t t : tv) {
        g.tg_stride[g.ntg] = row_stride(st, t);
        g.tg_dim[g.ntg++] = st->dims[t];
    }
    for (int cols = 1; cols >= 0; --cols)
        for (uint64_t q = 0; q < n; ++q) {
            if (std::find(tv.begin(), tv.end(), q) != tv.end()) continue;
            g.sp_stride[g.nsp] = row_stride(st, q) * (cols ? st->D : 1);
            g.sp_dim[g.nsp++] = st->dims[q];
        }
    return qs::launch_qudit_gate(st, g, S.data(), 1, DS, false);
}

}  // extern "C"

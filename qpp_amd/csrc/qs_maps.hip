// qs_maps.hip -- ptrace, ptranspose and syspermute (out of place).
//
//   k_bitperm    out[y] = in[G(y)] for a permutation G of the index bits:
//                syspermute of qubit kets / density matrices and ptranspose
//                of qubit density matrices.  A workgroup moves one tile whose
//                bit set holds the low output bits and the output bits fed by
//                the low input bits, so both the gather from HBM and the
//                scatter to HBM run over contiguous 2^L-element segments; the
//                transposition happens in LDS.  Each tile bit's share of
//                the indices is a kernel argument; the tile number's comes
//                from a table k_bitperm_table builds once per call.
//   k_digitperm  the same map for qudit states (mixed-radix digits).
//   k_ket_pt     ptranspose of a ket: out(i, j) = psi(i') conj(psi(j')).
//   k_ptrace     partial trace; `LANES` lanes cooperate on one output entry
//                and reduce with wavefront shuffles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "qs_internal.hpp"

namespace {

using qs::fail;

template <typename T>
struct Real;
template <>
struct Real<double2> {
    using R = double;
};
template <>
struct Real<float2> {
    using R = float;
};

// ---------------------------------------------------------------------------
// bit-permutation gather through LDS
// ---------------------------------------------------------------------------
struct BitPerm {
    uint32_t nb;           // index bits
    uint32_t tbits;        // |T|
    uint32_t nrest;        // bits outside T
    uint8_t src[64];       // input bit of output bit b
    uint8_t rest[64];      // output bits outside T, ascending
};

__device__ __forceinline__ uint64_t gather_bits(uint64_t y, const uint8_t* src,
                                                uint32_t nb) {
    uint64_t x = 0;
    for (uint32_t b = 0; b < nb; ++b) x |= ((y >> b) & 1) << src[b];
    return x;
}

// Largest tile: 2^12 elements, i.e. 16 per thread of a 256-thread workgroup.
constexpr uint32_t kBpMaxTileBits = 12;
constexpr int kBpPer = 1 << (kBpMaxTileBits - 8);

// Every index is XOR-linear (here: OR of disjoint bits) in the element's
// position e inside the tile and in the tile number: G and the LDS packing
// act bit by bit.  A table built once per call (it stays in L2) holds the
// contributions of each thread number (bits 0..7 of e), of each register
// index i (bits 8.. of e, e = t + 256 i) and of each tile-number byte value
// to the input index, the output index and the LDS slots of the load and
// store orders: a thread reads its entry once, and each element ORs in the
// (uniform) entry of its register index.
struct BpBits {
    uint64_t in_bit[kBpMaxTileBits];   // input index bit of e bit k (load order)
    uint64_t out_bit[kBpMaxTileBits];  // output index bit of e bit k (store order)
    uint32_t ld_bit[kBpMaxTileBits];   // LDS slot bit of e bit k, load order
    uint32_t st_bit[kBpMaxTileBits];   // LDS slot bit of e bit k, store order
};
struct BpPart {
    uint64_t in, out;
    uint32_t ld, st;
};
constexpr int kBpTileBytes = 8;  // up to 64 tile-number bits, 8 at a time
struct BpTable {
    uint64_t xt[kBpTileBytes][256];  // input bits of tile-number byte c = v
    uint64_t yt[kBpTileBytes][256];  // output bits of tile-number byte c = v
    BpPart thr[256];                 // thread number t
    BpPart reg[1 << (kBpMaxTileBits - 8)];  // register index i
};

__global__ __launch_bounds__(256) void k_bitperm_table(BitPerm bp, BpBits bb, BpTable* __restrict__ tab) {
    const uint32_t c = blockIdx.x, v = threadIdx.x;  // tile-number byte c, value v
    uint64_t x = 0, y = 0;
    for (uint32_t k = 8 * c; k < 8 * c + 8 && k < bp.nrest; ++k) {
        const uint64_t bit = (v >> (k - 8 * c)) & 1;
        y |= bit << bp.rest[k];
        x |= bit << bp.src[bp.rest[k]];
    }
    tab->xt[c][v] = x;
    tab->yt[c][v] = y;
    // e bits of v: thread number v (block 0) or register index v (block 1)
    if (c < 2 && (c == 0 || v < (1u << (kBpMaxTileBits - 8)))) {
        BpPart r{0, 0, 0, 0};
        for (uint32_t k = 0; k < 8; ++k) {
            const uint32_t kb = 8 * c + k;
            if (kb >= kBpMaxTileBits || !((v >> k) & 1)) continue;
            r.in |= bb.in_bit[kb];
            r.out |= bb.out_bit[kb];
            r.ld |= bb.ld_bit[kb];
            r.st |= bb.st_bit[kb];
        }
        (c == 0 ? tab->thr[v] : tab->reg[v]) = r;
    }
}

// One tile per workgroup over a grid of all tiles (the dispatcher then
// walks the state in order, which streams HBM better than a persistent
// grid): every load of the tile is issued, the elements go to their LDS
// slots, and after one barrier every store.  FULL: the tile has exactly
// 256 PER elements (no per-element test, so the compiler keeps every load
// ahead of the LDS writes instead of sinking each next to its write).
template <typename T, int PER, bool FULL>
__global__ __launch_bounds__(256) void k_bitperm(const T* __restrict__ in,
                                                 T* __restrict__ out,
                                                 const BpTable* __restrict__ tab,
                                                 uint64_t tile0, uint32_t tbits, uint32_t nrest) {
    extern __shared__ unsigned char smem_raw[];
    T* s = reinterpret_cast<T*>(smem_raw);
    const uint32_t tsz = 1u << tbits;
    const uint32_t t = threadIdx.x;
    const uint64_t tile = tile0 + blockIdx.x;
    const BpPart th = tab->thr[t];
    uint64_t xb = th.in, yb = th.out;
    for (uint32_t c = 0; 8 * c < nrest; ++c) {
        const uint32_t v = uint32_t(tile >> (8 * c)) & 255u;
        xb |= tab->xt[c][v];
        yb |= tab->yt[c][v];
    }
    // PER = tsz / 256 elements per thread when tsz >= 256; a smaller tile
    // (PER 1, not FULL) leaves threads without elements (bits of e from
    // tbits up contribute 0, so their loads stay inside the state; they
    // store nothing)
    T v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) v[i] = in[xb | tab->reg[i].in];  // all loads in flight
#pragma unroll
    for (int i = 0; i < PER; ++i)
        if (FULL || t + 256u * i < tsz) s[th.ld | tab->reg[i].ld] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i)
        if (FULL || t + 256u * i < tsz) out[yb | tab->reg[i].out] = s[th.st | tab->reg[i].st];
}

// ---------------------------------------------------------------------------
// mixed-radix digit permutation: output digit k (dims odim[k], stride ostr)
// takes the input digit at input stride istr[k]
// ---------------------------------------------------------------------------
struct DigitPerm {
    uint32_t nd;
    uint64_t odim[80];
    uint64_t istr[80];
};

template <typename T>
__global__ __launch_bounds__(256) void k_digitperm(const T* __restrict__ in,
                                                   T* __restrict__ out,
                                                   uint64_t n, DigitPerm dp) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t y = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; y < n;
         y += stride) {
        uint64_t rem = y, x = 0;
        for (int k = int(dp.nd) - 1; k >= 0; --k) {
            x += (rem % dp.odim[k]) * dp.istr[k];
            rem /= dp.odim[k];
        }
        out[y] = in[x];
    }
}

// ---------------------------------------------------------------------------
// ket partial transpose: out(i, j) = psi(r) conj(psi(c)) with r, c the row
// and column indices after swapping the target digits (mixed radix)
// ---------------------------------------------------------------------------
struct KetPT {
    uint32_t n;
    uint64_t dim[40];
    uint64_t stride[40];
    uint8_t swap[40];
    uint64_t D;
};

template <typename T>
__global__ __launch_bounds__(256) void k_ket_pt(const T* __restrict__ psi,
                                                T* __restrict__ out, KetPT p) {
    const uint64_t total = p.D * p.D;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t y = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; y < total;
         y += stride) {
        const uint64_t i = y % p.D, j = y / p.D;
        uint64_t ri = i, rj = j, r = 0, c = 0;
        for (int k = int(p.n) - 1; k >= 0; --k) {
            const uint64_t di = ri % p.dim[k], dj = rj % p.dim[k];
            ri /= p.dim[k];
            rj /= p.dim[k];
            r += (p.swap[k] ? dj : di) * p.stride[k];
            c += (p.swap[k] ? di : dj) * p.stride[k];
        }
        const T a = psi[r], b = psi[c];
        T v;
        v.x = a.x * b.x + a.y * b.y;
        v.y = a.y * b.x - a.x * b.y;
        out[y] = v;
    }
}

// ---------------------------------------------------------------------------
// partial trace
// ---------------------------------------------------------------------------
struct PT {
    uint32_t nkeep, ntr;
    uint64_t kdim[40], kstr[40];  // kept subsystems (ascending), row strides
    uint64_t tdim[40], tstr[40];  // traced subsystems (target order)
    uint64_t Dbar, Dsub, D;
    uint32_t rho;
};

__device__ __forceinline__ uint64_t mixed(uint64_t v, uint32_t n,
                                          const uint64_t* dim,
                                          const uint64_t* str) {
    uint64_t off = 0;
    for (int k = int(n) - 1; k >= 0; --k) {
        off += (v % dim[k]) * str[k];
        v /= dim[k];
    }
    return off;
}

template <typename T, int LANES>
__global__ __launch_bounds__(256) void k_ptrace(const T* __restrict__ a,
                                                T* __restrict__ out, PT p) {
    using R = typename Real<T>::R;
    const uint64_t groups = uint64_t(gridDim.x) * (256 / LANES);
    const uint32_t lane = threadIdx.x % LANES;
    const uint64_t total = p.Dbar * p.Dbar;
    for (uint64_t o = uint64_t(blockIdx.x) * (256 / LANES) + threadIdx.x / LANES;
         o < total + 0 * groups; o += groups) {
        const uint64_t i = o % p.Dbar, j = o / p.Dbar;
        const uint64_t rb = mixed(i, p.nkeep, p.kdim, p.kstr);
        const uint64_t cb = mixed(j, p.nkeep, p.kdim, p.kstr);
        R sr = 0, si = 0;
        for (uint64_t t = lane; t < p.Dsub; t += LANES) {
            const uint64_t off = mixed(t, p.ntr, p.tdim, p.tstr);
            if (p.rho) {
                const T v = a[(rb + off) + (cb + off) * p.D];
                sr += v.x;
                si += v.y;
            } else {
                const T x = a[rb + off], y = a[cb + off];
                // x * conj(y)
                sr += x.x * y.x + x.y * y.y;
                si += x.y * y.x - x.x * y.y;
            }
        }
        for (int m = LANES / 2; m > 0; m >>= 1) {
            sr += __shfl_xor(sr, m, LANES);
            si += __shfl_xor(si, m, LANES);
        }
        if (lane == 0) {
            T v;
            v.x = sr;
            v.y = si;
            out[o] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
bool valid(const qs_state* st) { return st != nullptr && st->d != nullptr; }

uint64_t row_stride(const qs_state* st, uint64_t q) {
    uint64_t s = 1;
    for (uint64_t i = st->dims.size(); i-- > q + 1;) s *= st->dims[i];
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

uint32_t grid_for(uint64_t work, uint64_t per) {
    uint64_t b = (work + per - 1) / per;
    if (b < 1) b = 1;
    if (b > (uint64_t{1} << 30)) b = uint64_t{1} << 30;
    return static_cast<uint32_t>(b);
}

int make_out(const qs_state* st, qs_kind kind, const std::vector<uint64_t>& dims,
             qs_state** out) {
    // created on `st`'s stream: kernels reading `st` write it in stream order
    const qs::StreamScope on(st->stream);
    return qs_state_create(st->dtype, kind, static_cast<uint32_t>(dims.size()), dims.data(),
                           out);
}

// out[y] = in[G(y)], src[b] = input bit of output bit b
template <typename T>
int launch_bitperm(const qs_state* in, qs_state* out, const std::vector<uint32_t>& src) {
    const uint32_t nb = static_cast<uint32_t>(src.size());
    BitPerm bp{};
    bp.nb = nb;
    for (uint32_t b = 0; b < nb; ++b) bp.src[b] = static_cast<uint8_t>(src[b]);
    // T = low L output bits + output bits whose source is a low input bit.
    // L starts at one 512-byte run (32 double2 / 64 float2) and grows while
    // the tile stays within 16 KiB of LDS: when the permutation leaves the
    // low bits in place, a 2^L tile of 32 elements would keep only 32 of
    // the 256 threads busy.
    const uint32_t lmin = std::min<uint32_t>(nb, sizeof(T) == 16 ? 5 : 6);
    const uint32_t tmax = sizeof(T) == 16 ? 10 : 11;
    auto tsize = [&](uint32_t l) {
        uint32_t c = l;
        for (uint32_t b = l; b < nb; ++b) c += src[b] < l;
        return c;
    };
    uint32_t L = lmin;
    while (L < nb && tsize(L + 1) <= tmax) ++L;
    std::vector<uint32_t> low_out, low_in_fed(L);
    for (uint32_t b = 0; b < L; ++b) low_out.push_back(b);
    std::vector<bool> inT(nb, false);
    for (uint32_t b = 0; b < L; ++b) inT[b] = true;
    for (uint32_t b = 0; b < nb; ++b)
        if (src[b] < L) {
            low_in_fed[src[b]] = b;
            inT[b] = true;
        }
    std::vector<uint32_t> tset;
    for (uint32_t b = 0; b < nb; ++b)
        if (inT[b]) tset.push_back(b);
    bp.tbits = static_cast<uint32_t>(tset.size());
    // load order: low e bits feed the output bits sourced by input bits 0..L-1
    std::vector<uint32_t> load(low_in_fed.begin(), low_in_fed.end());
    for (uint32_t b : tset)
        if (std::find(load.begin(), load.end(), b) == load.end()) load.push_back(b);
    // store order: low e bits are output bits 0..L-1
    std::vector<uint32_t> store(low_out.begin(), low_out.end());
    for (uint32_t b : tset)
        if (std::find(store.begin(), store.end(), b) == store.end()) store.push_back(b);
    std::vector<uint32_t> rest;
    for (uint32_t b = 0; b < nb; ++b)
        if (!inT[b]) rest.push_back(b);
    bp.nrest = static_cast<uint32_t>(rest.size());
    for (size_t k = 0; k < rest.size(); ++k) bp.rest[k] = static_cast<uint8_t>(rest[k]);
    // the kernel keeps at most 2^kBpMaxTileBits / 256 elements per thread
    if (bp.tbits > kBpMaxTileBits) return fail(QS_ERR_NOT_IMPLEMENTED, "qpp_amd::bitperm", "tile");
    // each tile bit's contribution to the four index parts
    BpBits bb{};
    auto slot_of = [&](uint32_t b) {
        return uint32_t(std::find(tset.begin(), tset.end(), b) - tset.begin());
    };
    for (uint32_t k = 0; k < bp.tbits; ++k) {
        bb.in_bit[k] = uint64_t{1} << src[load[k]];
        bb.out_bit[k] = uint64_t{1} << store[k];
        bb.ld_bit[k] = 1u << slot_of(load[k]);
        bb.st_bit[k] = 1u << slot_of(store[k]);
    }
    const uint64_t ntiles = uint64_t{1} << bp.nrest;
    const size_t shmem = sizeof(T) << bp.tbits;
    constexpr const char* W = "qpp_amd::bitperm";
    BpTable* tab = nullptr;
    QS_HIP(hipMallocAsync(reinterpret_cast<void**>(&tab), sizeof(BpTable), in->stream), W);
    hipLaunchKernelGGL(k_bitperm_table, dim3(kBpTileBytes), dim3(256), 0, in->stream, bp, bb, tab);  // >= 2 blocks
    // one tile per workgroup, at most 2^30 workgroups per launch
    for (uint64_t t0 = 0; t0 < ntiles; t0 += uint64_t{1} << 30) {
        const uint64_t g = std::min<uint64_t>(ntiles - t0, uint64_t{1} << 30);
        const dim3 gd(static_cast<uint32_t>(g));
        const T* src = static_cast<const T*>(in->d);
        T* dst = static_cast<T*>(out->d);
        // 256 threads x PER elements: a full tile for 8 to 12 tile bits
        switch (bp.tbits) {
            case 12:
                hipLaunchKernelGGL((k_bitperm<T, 16, true>), gd, dim3(256), shmem, in->stream, src,
                                   dst, tab, t0, bp.tbits, bp.nrest);
                break;
            case 11:
                hipLaunchKernelGGL((k_bitperm<T, 8, true>), gd, dim3(256), shmem, in->stream, src,
                                   dst, tab, t0, bp.tbits, bp.nrest);
                break;
            case 10:
                hipLaunchKernelGGL((k_bitperm<T, 4, true>), gd, dim3(256), shmem, in->stream, src,
                                   dst, tab, t0, bp.tbits, bp.nrest);
                break;
            case 9:
                hipLaunchKernelGGL((k_bitperm<T, 2, true>), gd, dim3(256), shmem, in->stream, src,
                                   dst, tab, t0, bp.tbits, bp.nrest);
                break;
            case 8:
                hipLaunchKernelGGL((k_bitperm<T, 1, true>), gd, dim3(256), shmem, in->stream, src,
                                   dst, tab, t0, bp.tbits, bp.nrest);
                break;
            default:
                hipLaunchKernelGGL((k_bitperm<T, 1, false>), gd, dim3(256), shmem, in->stream, src,
                                   dst, tab, t0, bp.tbits, bp.nrest);
                break;
        }
    }
    const hipError_t e = hipGetLastError();
    (void)hipFreeAsync(tab, in->stream);
    if (e != hipSuccess) return qs::fail_hip(e, W);
    return QS_OK;
}

int bitperm(const qs_state* in, qs_state* out, const std::vector<uint32_t>& src) {
    return in->dtype == QS_C128 ? launch_bitperm<double2>(in, out, src)
                                : launch_bitperm<float2>(in, out, src);
}

template <typename T>
int launch_digitperm(const qs_state* in, qs_state* out, const DigitPerm& dp) {
    hipLaunchKernelGGL(k_digitperm<T>, dim3(grid_for(out->nelem, 256)), dim3(256), 0,
                       in->stream, static_cast<const T*>(in->d),
                       static_cast<T*>(out->d), out->nelem, dp);
    QS_HIP(hipGetLastError(), "qpp_amd::digitperm");
    return QS_OK;
}

int digitperm(const qs_state* in, qs_state* out, const DigitPerm& dp) {
    return in->dtype == QS_C128 ? launch_digitperm<double2>(in, out, dp)
                                : launch_digitperm<float2>(in, out, dp);
}

template <typename T>
int launch_ket_pt(const qs_state* in, qs_state* out, const KetPT& p) {
    hipLaunchKernelGGL(k_ket_pt<T>, dim3(grid_for(out->nelem, 256)), dim3(256), 0,
                       in->stream, static_cast<const T*>(in->d),
                       static_cast<T*>(out->d), p);
    QS_HIP(hipGetLastError(), "qpp_amd::ptranspose");
    return QS_OK;
}

template <typename T>
int launch_ptrace(const qs_state* in, qs_state* out, const PT& p) {
    const uint64_t total = p.Dbar * p.Dbar;
    if (p.Dsub >= 64) {
        hipLaunchKernelGGL((k_ptrace<T, 64>), dim3(grid_for(total, 4)), dim3(256), 0,
                           in->stream, static_cast<const T*>(in->d),
                           static_cast<T*>(out->d), p);
    } else if (p.Dsub >= 8) {
        hipLaunchKernelGGL((k_ptrace<T, 8>), dim3(grid_for(total, 32)), dim3(256), 0,
                           in->stream, static_cast<const T*>(in->d),
                           static_cast<T*>(out->d), p);
    } else {
        hipLaunchKernelGGL((k_ptrace<T, 1>), dim3(grid_for(total, 256)), dim3(256), 0,
                           in->stream, static_cast<const T*>(in->d),
                           static_cast<T*>(out->d), p);
    }
    QS_HIP(hipGetLastError(), "qpp::ptrace()");
    return QS_OK;
}

}  // namespace

extern "C" {

// qpp::ptrace(A, target, dims), operations.hpp:1908-2114; qubit kernels
// ptrace.hpp:61-257 (same sums: traced digits in target order, kept digits
// ascending).
int qs_ptrace(const qs_state* st, uint32_t ntarget, const uint64_t* target,
              qs_state** out) {
    constexpr const char* W = "qpp::ptrace()";
    if (!valid(st) || !out) return fail(QS_ERR_INVALID_HANDLE, W, "A");
    qs::clear_error();
    std::vector<uint64_t> tv(target, target + ntarget);
    if (!check_subsys(tv, st->dims)) return fail(QS_ERR_SUBSYS_MISMATCH_DIMS, W, "dims/target");
    const uint64_t n = st->dims.size();
    std::vector<uint64_t> keep;
    for (uint64_t q = 0; q < n; ++q)
        if (std::find(tv.begin(), tv.end(), q) == tv.end()) keep.push_back(q);
    PT p{};
    p.rho = st->kind == QS_RHO;
    p.D = st->D;
    p.Dbar = 1;
    p.Dsub = 1;
    p.nkeep = static_cast<uint32_t>(keep.size());
    p.ntr = static_cast<uint32_t>(tv.size());
    if (p.nkeep > 40 || p.ntr > 40) return fail(QS_ERR_NOT_IMPLEMENTED, W, "subsystems");
    for (size_t k = 0; k < keep.size(); ++k) {
        p.kdim[k] = st->dims[keep[k]];
        p.kstr[k] = row_stride(st, keep[k]);
        p.Dbar *= p.kdim[k];
    }
    for (size_t k = 0; k < tv.size(); ++k) {
        p.tdim[k] = st->dims[tv[k]];
        p.tstr[k] = row_stride(st, tv[k]);
        p.Dsub *= p.tdim[k];
    }
    std::vector<uint64_t> odims;
    for (uint64_t q : keep) odims.push_back(st->dims[q]);
    if (odims.empty()) odims.push_back(1);  // full trace: a 1 x 1 matrix
    if (int rc = make_out(st, QS_RHO, odims, out)) return rc;
    return st->dtype == QS_C128 ? launch_ptrace<double2>(st, *out, p)
                                : launch_ptrace<float2>(st, *out, p);
}

// qpp::ptranspose(A, target, dims), operations.hpp:2167-2335 (documented);
// ptranspose_{psi,rho}_kq, ptranspose.hpp:59-219 (as-built: for
// 0 < |target| < n the qubit kernels transpose the subsystems
// complement({n-1-t : t in target})).
int qs_ptranspose(const qs_state* st, uint32_t ntarget, const uint64_t* target,
                  qs_semantics sem, qs_state** out) {
    constexpr const char* W = "qpp::ptranspose()";
    if (!valid(st) || !out) return fail(QS_ERR_INVALID_HANDLE, W, "A");
    qs::clear_error();
    std::vector<uint64_t> tv(target, target + ntarget);
    if (!check_subsys(tv, st->dims)) return fail(QS_ERR_SUBSYS_MISMATCH_DIMS, W, "dims/target");
    const uint64_t n = st->dims.size();
    // subsystems whose row and column digits are swapped
    std::vector<bool> sw(n, false);
    if (sem == QS_SEM_AS_BUILT && st->qubits && !tv.empty() && tv.size() != n) {
        std::vector<bool> inR(n, false);
        for (uint64_t t : tv) inR[n - 1 - t] = true;
        for (uint64_t q = 0; q < n; ++q) sw[q] = !inR[q];
    } else {
        for (uint64_t t : tv) sw[t] = true;
    }
    if (int rc = make_out(st, QS_RHO, st->dims, out)) return rc;
    if (st->kind == QS_KET) {
        KetPT p{};
        p.n = static_cast<uint32_t>(n);
        p.D = st->D;
        if (n > 40) return fail(QS_ERR_NOT_IMPLEMENTED, W, "subsystems");
        for (uint64_t q = 0; q < n; ++q) {
            p.dim[q] = st->dims[q];
            p.stride[q] = row_stride(st, q);
            p.swap[q] = sw[q];
        }
        return st->dtype == QS_C128 ? launch_ket_pt<double2>(st, *out, p)
                                    : launch_ket_pt<float2>(st, *out, p);
    }
    if (st->qubits) {
        std::vector<uint32_t> src(2 * n);
        for (uint64_t q = 0; q < n; ++q) {
            const uint32_t r = static_cast<uint32_t>(n - 1 - q), c = static_cast<uint32_t>(2 * n - 1 - q);
            src[r] = sw[q] ? c : r;
            src[c] = sw[q] ? r : c;
        }
        return bitperm(st, *out, src);
    }
    DigitPerm dp{};
    dp.nd = static_cast<uint32_t>(2 * n);
    if (dp.nd > 80) return fail(QS_ERR_NOT_IMPLEMENTED, W, "subsystems");
    // output digits: column digits (most significant) then row digits
    for (uint64_t q = 0; q < n; ++q) {
        const uint64_t rs = row_stride(st, q), cs = rs * st->D;
        dp.odim[q] = st->dims[q];
        dp.istr[q] = sw[q] ? rs : cs;
        dp.odim[n + q] = st->dims[q];
        dp.istr[n + q] = sw[q] ? cs : rs;
    }
    return digitperm(st, *out, dp);
}

// qpp::syspermute(A, perm, dims), operations.hpp:2386-2514 (documented:
// output subsystem k is input subsystem perm[k]); syspermute_{psi,rho}_kq,
// syspermute.hpp:58-180 (as-built: output LSB-bit k is input LSB-bit
// inv_perm[k]).
int qs_syspermute(const qs_state* st, const uint64_t* perm, qs_semantics sem,
                  qs_state** out) {
    constexpr const char* W = "qpp::syspermute()";
    if (!valid(st) || !out) return fail(QS_ERR_INVALID_HANDLE, W, "A");
    qs::clear_error();
    const uint64_t n = st->dims.size();
    if (!perm) return fail(QS_ERR_PERM_INVALID, W, "perm");
    std::vector<uint64_t> pv(perm, perm + n);
    {
        std::vector<uint64_t> s = pv;
        std::sort(s.begin(), s.end());
        for (uint64_t i = 0; i < n; ++i)
            if (s[i] != i) return fail(QS_ERR_PERM_INVALID, W, "perm");
    }
    const bool rho = st->kind == QS_RHO;
    if (st->qubits) {
        // output bit b <- input bit src[b], for the row bits and column bits
        std::vector<uint32_t> srcn(n);
        if (sem == QS_SEM_AS_BUILT) {
            std::vector<uint64_t> inv(n);
            for (uint64_t k = 0; k < n; ++k) inv[pv[k]] = k;
            for (uint64_t k = 0; k < n; ++k) srcn[k] = static_cast<uint32_t>(inv[k]);
        } else {
            for (uint64_t k = 0; k < n; ++k)
                srcn[n - 1 - k] = static_cast<uint32_t>(n - 1 - pv[k]);
        }
        if (int rc = make_out(st, st->kind, st->dims, out)) return rc;
        std::vector<uint32_t> src(rho ? 2 * n : n);
        for (uint64_t b = 0; b < n; ++b) {
            src[b] = srcn[b];
            if (rho) src[n + b] = static_cast<uint32_t>(n + srcn[b]);
        }
        return bitperm(st, *out, src);
    }
    std::vector<uint64_t> odims(n);
    for (uint64_t k = 0; k < n; ++k) odims[k] = st->dims[pv[k]];
    if (int rc = make_out(st, st->kind, odims, out)) return rc;
    DigitPerm dp{};
    dp.nd = static_cast<uint32_t>(rho ? 2 * n : n);
    if (dp.nd > 80) return fail(QS_ERR_NOT_IMPLEMENTED, W, "subsystems");
    for (uint64_t k = 0; k < n; ++k) {
        dp.odim[k + (rho ? n : 0)] = odims[k];
        dp.istr[k + (rho ? n : 0)] = row_stride(st, pv[k]);
        if (rho) {
            dp.odim[k] = odims[k];
            dp.istr[k] = row_stride(st, pv[k]) * st->D;
        }
    }
    return digitperm(st, *out, dp);
}

}  // extern "C"

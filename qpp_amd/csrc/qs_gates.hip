// qs_gates.hip -- gate-application kernels for gfx950 (CDNA4, wave64).
//
// Kernels operate on a bit-addressed array of complex elements (see
// qs_internal.hpp): kets, and density matrices as vec(rho) over 2n bits.
//
//   k_gate1_hi  1-bit gate, target bit >= 6: each work item is one amplitude
//               pair (i0, i0 | 2^bit); control bits and the target bit are
//               squeezed out of the work index, so only pairs whose controls
//               are met are read or written (2*s*2^(nb-c) bytes) and lanes of
//               a wave touch consecutive addresses in both halves.
//   k_gate1_lo  1-bit gate, target bit < 6: the partner of an element lies in
//               the same 64-element aligned chunk, i.e. in another lane of
//               the same wave; lanes load consecutive elements (fully
//               coalesced) and exchange partners with a lane shuffle.
//   k_gatek     k-bit dense gate (k <= 10) through LDS: a workgroup stages a
//               batch of amplitude groups, each thread then forms outputs
//               from the staged group and the gate matrix.
//   k_gatek_mfma  4- and 5-bit dense gates on complex64 states with FP32
//               MFMA: Y = U X over the state's amplitude groups as a real
//               GEMM (2^(k+1) x 2^(k+1) times 16 groups per wave step).
//   k_qudit     mixed-radix gate for qudit states (multi-index strides).
//
// Arithmetic: complex128 in double, complex64 in float; gate matrices are
// converted to the state's precision on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "qs_internal.hpp"

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
__device__ __forceinline__ cxt<R> cmul(cxt<R> a, cxt<R> b) {
    cxt<R> r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

template <typename R>
__device__ __forceinline__ cxt<R> cfma(cxt<R> a, cxt<R> b, cxt<R> c) {
    // a * b + c
    cxt<R> r;
    r.x = c.x + a.x * b.x - a.y * b.y;
    r.y = c.y + a.x * b.y + a.y * b.x;
    return r;
}

// 2x2 gate, row-major entries u00 u01 u10 u11 in the state's precision
template <typename R>
struct M2 {
    cxt<R> u00, u01, u10, u11;
};

// Insert zero bits at the sorted positions pos[0..np-1] (ascending) into x.
// The positions (bytes of a kernel argument) are read four to a 32-bit
// word: a uniform word load is a scalar load, where a byte load would be a
// vector load whose wait also waits for every other load and store the
// wave has in flight (the prefetched amplitudes among them).
__device__ __forceinline__ uint64_t insert_zeros(uint64_t x, const uint8_t* pos,
                                                 int np) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(pos);
    for (int i = 0; i < np; ++i) {
        const uint32_t p = (w[i >> 2] >> (8 * (i & 3))) & 0xFFu;
        const uint64_t lo = x & ((uint64_t{1} << p) - 1);
        x = ((x >> p) << (p + 1)) | lo;
    }
    return x;
}

struct Fixed {
    uint8_t pos[64];  // sorted ascending bit positions to squeeze out
    int np;
};

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// 1-bit gate, target bit >= 6
// ---------------------------------------------------------------------------
template <typename R, int PPT>
__global__ __launch_bounds__(kBlock) void k_gate1_hi(cxt<R>* __restrict__ a,
                                                      uint64_t npairs,
                                                      uint64_t step, M2<R> u,
                                                      Fixed fx, uint64_t cval) {
    const uint64_t nthreads = uint64_t(gridDim.x) * kBlock;
    uint64_t p = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    // PPT independent pairs per thread, strided by the grid for coalescing
    cxt<R> x0[PPT], x1[PPT];
    uint64_t i0[PPT];
    for (; p < npairs; p += nthreads * PPT) {
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const uint64_t q = p + j * nthreads;
            i0[j] = q < npairs ? (insert_zeros(q, fx.pos, fx.np) | cval) : 0;
            if (q < npairs) {
                x0[j] = a[i0[j]];
                x1[j] = a[i0[j] + step];
            }
        }
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const uint64_t q = p + j * nthreads;
            if (q < npairs) {
                const cxt<R> y0 = cfma<R>(u.u00, x0[j], cmul<R>(u.u01, x1[j]));
                const cxt<R> y1 = cfma<R>(u.u10, x0[j], cmul<R>(u.u11, x1[j]));
                a[i0[j]] = y0;
                a[i0[j] + step] = y1;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// 1-bit gate, target bit < 6 (partner in the same wave)
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ cxt<R> shfl_xor_c(cxt<R> v, int m) {
    cxt<R> r;
    r.x = __shfl_xor(v.x, m, 64);
    r.y = __shfl_xor(v.y, m, 64);
    return r;
}

// Work index w covers all elements with the high control bits squeezed out
// (fx holds only control bits >= 6, so the low 6 bits of w are the lane's
// element offset inside an aligned 64-element chunk).  Low control bits are
// tested per element.
template <typename R, int EPT>
__global__ __launch_bounds__(kBlock) void k_gate1_lo(cxt<R>* __restrict__ a,
                                                      uint64_t nwork, int bit,
                                                      M2<R> u, Fixed fx,
                                                      uint64_t cval,
                                                      uint64_t lmask,
                                                      uint64_t lval) {
    const uint64_t nthreads = uint64_t(gridDim.x) * kBlock;
    const uint64_t tbit = uint64_t{1} << bit;
    // nwork is a multiple of 64 and every wave's indices are a full aligned
    // chunk, so all 64 lanes take part in each shuffle.
    for (uint64_t w0 = uint64_t(blockIdx.x) * kBlock + threadIdx.x; w0 < nwork;
         w0 += nthreads * EPT) {
        cxt<R> x[EPT];
        uint64_t e[EPT];
        bool ok[EPT];
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const uint64_t w = w0 + j * nthreads;
            ok[j] = w < nwork;
            e[j] = ok[j] ? (insert_zeros(w, fx.pos, fx.np) | cval) : 0;
            x[j].x = 0;
            x[j].y = 0;
            if (ok[j]) x[j] = a[e[j]];
        }
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            // every lane joins the shuffle; a lane's partner (lane ^ tbit) is
            // in range whenever the lane is (nwork is a power of two > tbit)
            const cxt<R> y = shfl_xor_c<R>(x[j], int(tbit));
            if (ok[j]) {
                const bool one = (e[j] & tbit) != 0;
                const bool act = (e[j] & lmask) == lval;
                cxt<R> r;
                if (!one)
                    r = cfma<R>(u.u00, x[j], cmul<R>(u.u01, y));
                else
                    r = cfma<R>(u.u10, y, cmul<R>(u.u11, x[j]));
                if (act) a[e[j]] = r;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k-bit dense gate through LDS
// ---------------------------------------------------------------------------
// Groups: the 2^k elements sharing all non-target bits.  The group index g
// enumerates spectator bits (targets and high controls squeezed out, control
// value OR-ed in); `inner[c]` is the offset of gate digit c (bits[0] = MSB).
// A block stages G = kBlock*EPT / 2^k groups.  Thread t loads element
// (c = t / G, g = t % G) so consecutive lanes read consecutive groups.
template <typename R, int EPT>
__global__ __launch_bounds__(kBlock) void k_gatek(
    cxt<R>* __restrict__ a, uint64_t ngroups, int k,
    const uint64_t* __restrict__ inner, const cxt<R>* __restrict__ m,
    Fixed fx, uint64_t cval, uint64_t lmask, uint64_t lval) {
    extern __shared__ unsigned char smem_raw[];
    cxt<R>* s = reinterpret_cast<cxt<R>*>(smem_raw);
    const int dim = 1 << k;
    const int G = (kBlock * EPT) >> k;  // groups per block (>= 1)
    for (uint64_t gb = uint64_t(blockIdx.x) * G; gb < ngroups;
         gb += uint64_t(gridDim.x) * G) {
        // stage
        for (int t = threadIdx.x; t < G * dim; t += kBlock) {
            const int c = t / G, gl = t % G;
            const uint64_t g = gb + gl;
            if (g < ngroups) {
                const uint64_t base = insert_zeros(g, fx.pos, fx.np) | cval;
                s[gl * dim + c] = a[base + inner[c]];
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < G * dim; t += kBlock) {
            const int r = t / G, gl = t % G;
            const uint64_t g = gb + gl;
            if (g < ngroups) {
                const uint64_t base = insert_zeros(g, fx.pos, fx.np) | cval;
                if ((base & lmask) == lval) {
                    cxt<R> acc;
                    acc.x = 0;
                    acc.y = 0;
                    const cxt<R>* in = s + gl * dim;
                    // m is column-major: m[r + c*dim]
                    for (int c = 0; c < dim; ++c) acc = cfma<R>(m[r + c * dim], in[c], acc);
                    a[base + inner[r]] = acc;
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// k-bit dense gate with MFMA (complex64, k = 4, 5)
// ---------------------------------------------------------------------------
// The groups (columns) are as in k_gatek.  Over the real and imaginary parts
// the gate is one real GEMM, [Yr; Yi] = W [Xr; Xi] with W = [[Ur, -Ui],
// [Ui, Ur]] (2 DIM x 2 DIM, DIM = 2^K), and every wave step multiplies W by
// 16 groups with v_mfma_f32_16x16x4f32 (D[16x16] += A[16x4] B[4x16]; lane l
// holds A[l%16][l/16], B[l/16][l%16] and D[4(l/16)+v][l%16], v = 0..3):
//   B: lane (g = l/16, j = l%16) holds real row 4s+g of group j for slice
//      s; it loads the complex amplitudes 4q+g of its group, whose real
//      parts are slices q and whose imaginary parts are slices DIM/4+q;
//   A: W in 16-row blocks mb and 4-column slices s, each lane's entries
//      packed by the host in that order and read from LDS;
//   D: row block mb < NB/2 holds real parts and mb + NB/2 the imaginary
//      parts of the same complex rows 16 mb + 4 g + v, so each lane writes
//      whole complex amplitudes.
// Arithmetic per amplitude: 4 DIM FP32 FMA, as many as the complex product
// takes on VALU, at twice the VALU FP32 rate (packed FP32 aside).
typedef float floatx4 __attribute__((ext_vector_type(4)));
constexpr uint32_t kMfmaGrid = 2048;

template <int K>
__global__ __launch_bounds__(kBlock) void k_gatek_mfma(float2* __restrict__ a, uint64_t ncols,
                                                       const uint64_t* __restrict__ inner,
                                                       const float* __restrict__ wpack, Fixed fx,
                                                       uint64_t cval) {
    constexpr int DIM = 1 << K;
    constexpr int NX = DIM / 4;       // complex inputs per lane
    constexpr int NS = 2 * NX;        // slices of 4 real columns of W
    constexpr int NB = 2 * DIM / 16;  // blocks of 16 real rows of W
    constexpr int NO = NB / 2;        // output row blocks (complex)
    const int lane = threadIdx.x & 63;
    const int j = lane & 15, g = lane >> 4;
    // the packed matrix in LDS, shared by the workgroup's waves (registers
    // go to the amplitudes in flight instead)
    __shared__ float wl[NB * NS * 64];
    for (int i = threadIdx.x; i < NB * NS * 64; i += kBlock) wl[i] = wpack[i];
    __syncthreads();
    uint64_t in_off[NX], out_off[NO][4];
#pragma unroll
    for (int q = 0; q < NX; ++q) in_off[q] = inner[4 * q + g];
#pragma unroll
    for (int mh = 0; mh < NO; ++mh)
#pragma unroll
        for (int v = 0; v < 4; ++v) out_off[mh][v] = inner[16 * mh + 4 * g + v];
    constexpr int kWaves = kBlock / 64;
    const uint64_t nwaves = uint64_t(gridDim.x) * kWaves;
    // a wave's group `grp` covers columns 16 grp .. 16 grp + 15.  The
    // groups after the current one are loaded before its products, into
    // register buffers taken in turn (the loop unrolled, so no buffer is
    // copied and no load has to land before its own group's turn)
    struct Buf {
        float xr[NX], xi[NX];
        uint64_t base;
        bool ok;
    };
    // a lane past the last column loads column 0's amplitudes (in bounds,
    // no branch around the loads) and stores nothing
    auto fetch = [&](uint64_t grp, Buf& b) {
        const uint64_t col = grp * 16 + j;
        b.ok = col < ncols;
        b.base = insert_zeros(b.ok ? col : 0, fx.pos, fx.np) | cval;
#pragma unroll
        for (int q = 0; q < NX; ++q) {
            const float2 v = a[b.base + in_off[q]];
            b.xr[q] = v.x;
            b.xi[q] = v.y;
        }
    };
    auto process = [&](const Buf& b) {
        floatx4 acc[NB];
#pragma unroll
        for (int mb = 0; mb < NB; ++mb) acc[mb] = floatx4{0.f, 0.f, 0.f, 0.f};
        // slice s's matrix entries are read while slice s-1's products run.
        // The matrix is read again at every group: an offset the compiler
        // cannot see through (zero, from an instruction it must keep) stops
        // it from hoisting all 2 DIM^2 / 16 entries into registers, and the
        // scheduling barrier from issuing every slice's reads at once.
        int z;
        asm volatile("s_mov_b32 %0, 0" : "=s"(z));
        const float* wv = wl + z;
        float wc[NB], wn[NB];
#pragma unroll
        for (int mb = 0; mb < NB; ++mb) wc[mb] = wv[mb * NS * 64 + lane];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (s + 1 < NS) {
#pragma unroll
                for (int mb = 0; mb < NB; ++mb) wn[mb] = wv[(mb * NS + s + 1) * 64 + lane];
            }
            const float bx = s < NX ? b.xr[s] : b.xi[s - NX];
#pragma unroll
            for (int mb = 0; mb < NB; ++mb)
                acc[mb] = __builtin_amdgcn_mfma_f32_16x16x4f32(wc[mb], bx, acc[mb], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int mb = 0; mb < NB; ++mb) wc[mb] = wn[mb];
        }
        if (b.ok) {
#pragma unroll
            for (int mh = 0; mh < NO; ++mh)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    a[b.base + out_off[mh][v]] = make_float2(acc[mh][v], acc[mh + NO][v]);
        }
    };
    uint64_t grp = uint64_t(blockIdx.x) * kWaves + threadIdx.x / 64;
    if (grp * 16 >= ncols) return;
    // every fetch is issued, past the wave's last group on its current
    // group again (in bounds, never used): with no load under a branch the
    // waits before each group's products count exactly the loads and
    // stores issued after its own loads
    if constexpr (K == 4) {
        // 4 amplitudes per lane and group, 8 waves per SIMD: the next group
        // in flight, loaded under the lanes' bounds and copied into place
        float yr[NX], yi[NX];
        Buf cur;
        cur.ok = false;
        bool ok = (grp * 16 + j) < ncols;
        uint64_t base = ok ? (insert_zeros(grp * 16 + j, fx.pos, fx.np) | cval) : 0;
#pragma unroll
        for (int q = 0; q < NX; ++q) {
            float2 v = make_float2(0.f, 0.f);
            if (ok) v = a[base + in_off[q]];
            cur.xr[q] = v.x;
            cur.xi[q] = v.y;
        }
        for (;;) {
            const uint64_t next = grp + nwaves;
            const bool more = next * 16 < ncols;  // uniform over the wave
            uint64_t nbase = 0;
            bool nok = false;
            if (more) {
                const uint64_t col = next * 16 + j;
                nok = col < ncols;
                nbase = nok ? (insert_zeros(col, fx.pos, fx.np) | cval) : 0;
#pragma unroll
                for (int q = 0; q < NX; ++q) {
                    float2 v = make_float2(0.f, 0.f);
                    if (nok) v = a[nbase + in_off[q]];
                    yr[q] = v.x;
                    yi[q] = v.y;
                }
            }
            cur.base = base;
            cur.ok = ok;
            process(cur);
            if (!more) break;
            grp = next;
            base = nbase;
            ok = nok;
#pragma unroll
            for (int q = 0; q < NX; ++q) {
                cur.xr[q] = yr[q];
                cur.xi[q] = yi[q];
            }
        }
    } else {
        // 8 amplitudes per lane and group, 4 waves per SIMD: the next two
        // groups in flight (three buffers in turn)
        Buf b0, b1, b2;
        fetch(grp, b0);
        bool has1 = (grp + nwaves) * 16 < ncols;
        fetch(has1 ? grp + nwaves : grp, b1);
        // `cur` holds group grp, the next buffer group grp + nwaves (when
        // has1), and `far` receives group grp + 2 nwaves; false once grp was
        // the last
        auto step = [&](const Buf& cur, Buf& far) {
            const bool has2 = (grp + 2 * nwaves) * 16 < ncols;
            fetch(has2 ? grp + 2 * nwaves : grp, far);
            process(cur);
            if (!has1) return false;
            grp += nwaves;
            has1 = has2;
            return true;
        };
        for (;;) {
            if (!step(b0, b2)) break;
            if (!step(b1, b0)) break;
            if (!step(b2, b1)) break;
        }
    }
}

// ---------------------------------------------------------------------------
// qudit (mixed-radix) gate
// ---------------------------------------------------------------------------
struct QG {
    uint32_t ntg, nsp, nctrl;
    uint64_t tg_stride[16], tg_dim[16];
    uint64_t sp_stride[40], sp_dim[40];
    uint64_t c_stride[40], c_shift[40];
    uint64_t c_dim;
    uint64_t DA;
};

// One thread per spectator index: gather DA elements (<= 64) from the
// target-digit offsets, apply A^r, scatter.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_qudit(cxt<R>* __restrict__ a,
                                                   uint64_t nspec, QG q,
                                                   const cxt<R>* __restrict__ mats,
                                                   const uint64_t* __restrict__ off) {
    const uint64_t nthreads = uint64_t(gridDim.x) * kBlock;
    for (uint64_t s = uint64_t(blockIdx.x) * kBlock + threadIdx.x; s < nspec;
         s += nthreads) {
        uint64_t base = 0, rem = s;
        for (int j = int(q.nsp) - 1; j >= 0; --j) {
            base += (rem % q.sp_dim[j]) * q.sp_stride[j];
            rem /= q.sp_dim[j];
        }
        uint64_t r = 1;  // exponent of A (1 when uncontrolled)
        if (q.nctrl) {
            r = ((base / q.c_stride[0]) % q.c_dim + q.c_shift[0]) % q.c_dim;
            for (uint32_t c = 1; c < q.nctrl; ++c)
                if (((base / q.c_stride[c]) % q.c_dim + q.c_shift[c]) % q.c_dim != r)
                    r = 0;
        }
        if (r == 0) continue;  // A^0 = identity
        const cxt<R>* m = mats + (q.nctrl ? r : 0) * q.DA * q.DA;
        cxt<R> in[64];
        for (uint64_t c = 0; c < q.DA; ++c) in[c] = a[base + off[c]];
        for (uint64_t rr = 0; rr < q.DA; ++rr) {
            cxt<R> acc;
            acc.x = 0;
            acc.y = 0;
            for (uint64_t c = 0; c < q.DA; ++c) acc = cfma<R>(m[rr + c * q.DA], in[c], acc);
            a[base + off[rr]] = acc;
        }
    }
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
uint32_t grid_for(uint64_t work, uint64_t per_block) {
    uint64_t b = (work + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > (uint64_t{1} << 30)) b = uint64_t{1} << 30;
    return static_cast<uint32_t>(b);
}

template <typename R>
cxt<R> to_c(const double* m, uint64_t idx, bool conj) {
    cxt<R> v;
    v.x = static_cast<R>(m[2 * idx]);
    v.y = static_cast<R>(conj ? -m[2 * idx + 1] : m[2 * idx + 1]);
    return v;
}

template <typename R>
M2<R> make_m2(const double* m, bool conj) {
    // column-major input: m[r + 2c]
    M2<R> u;
    u.u00 = to_c<R>(m, 0, conj);
    u.u10 = to_c<R>(m, 1, conj);
    u.u01 = to_c<R>(m, 2, conj);
    u.u11 = to_c<R>(m, 3, conj);
    return u;
}

Fixed make_fixed(std::vector<uint32_t> pos) {
    std::sort(pos.begin(), pos.end());
    Fixed f;
    f.np = static_cast<int>(pos.size());
    for (size_t i = 0; i < pos.size(); ++i) f.pos[i] = static_cast<uint8_t>(pos[i]);
    return f;
}

std::vector<uint32_t> bits_of(uint64_t mask) {
    std::vector<uint32_t> v;
    for (uint32_t b = 0; b < 64; ++b)
        if ((mask >> b) & 1) v.push_back(b);
    return v;
}

template <typename R>
int gate1_impl(qs_state* st, uint32_t bit, const double* m, uint64_t cmask,
               uint64_t cval, bool conj) {
    const M2<R> u = make_m2<R>(m, conj);
    auto* a = static_cast<cxt<R>*>(st->d);
    const uint32_t nb = st->nbits;
    if (bit >= 6) {
        std::vector<uint32_t> fixed = bits_of(cmask);
        fixed.push_back(bit);
        const Fixed fx = make_fixed(fixed);
        const uint64_t npairs = uint64_t{1} << (nb - fx.np);
        constexpr int PPT = 4;
        const uint32_t grid = grid_for(npairs, uint64_t(kBlock) * PPT);
        hipLaunchKernelGGL((k_gate1_hi<R, PPT>), dim3(grid), dim3(kBlock), 0,
                           st->stream, a, npairs, uint64_t{1} << bit, u, fx,
                           cval & ~(uint64_t{1} << bit));
    } else {
        // controls above bit 5 are squeezed out, the others tested in-lane
        const uint64_t hmask = cmask & ~uint64_t{63};
        const uint64_t lmask = cmask & uint64_t{63};
        const Fixed fx = make_fixed(bits_of(hmask));
        const uint64_t nwork = uint64_t{1} << (nb - fx.np);
        constexpr int EPT = 4;
        const uint32_t grid = grid_for(nwork, uint64_t(kBlock) * EPT);
        hipLaunchKernelGGL((k_gate1_lo<R, EPT>), dim3(grid), dim3(kBlock), 0,
                           st->stream, a, nwork, int(bit), u, fx, cval & hmask,
                           lmask, cval & lmask);
    }
    return QS_OK;
}

// Device scratch for gate matrices / index tables of one call.  Kept per
// state so stream ordering on the state's stream protects reuse.
struct Scratch {
    void* d = nullptr;
    size_t bytes = 0;
};

int upload(qs_state* st, const void* host, size_t bytes, void** dptr) {
    // one allocation per call, freed stream-ordered after the kernel
    QS_HIP(hipMallocAsync(dptr, bytes, st->stream), "qpp_amd::upload");
    QS_HIP(hipMemcpyAsync(*dptr, host, bytes, hipMemcpyHostToDevice, st->stream),
           "qpp_amd::upload");
    // pageable source: make sure the copy has consumed `host` before return
    QS_HIP(hipStreamSynchronize(st->stream), "qpp_amd::upload");
    return QS_OK;
}

int gatek_mfma(qs_state* st, uint32_t k, const std::vector<uint64_t>& inner, const double* m,
               bool conj, const Fixed& fx, uint64_t cval, uint64_t ncols);

template <typename R>
int gatek_impl(qs_state* st, const uint32_t* bits, uint32_t k, const double* m,
               uint64_t cmask, uint64_t cval, bool conj) {
    const uint32_t dim = 1u << k;
    std::vector<uint64_t> inner(dim);
    for (uint32_t c = 0; c < dim; ++c) {
        uint64_t v = 0;
        for (uint32_t j = 0; j < k; ++j)
            if ((c >> j) & 1) v |= uint64_t{1} << bits[k - 1 - j];
        inner[c] = v;
    }
    std::vector<cxt<R>> mm(uint64_t(dim) * dim);
    for (uint64_t i = 0; i < mm.size(); ++i) mm[i] = to_c<R>(m, i, conj);
    // squeeze targets and control bits that are not targets
    std::vector<uint32_t> fixed(bits, bits + k);
    for (uint32_t b : bits_of(cmask)) fixed.push_back(b);
    const Fixed fx = make_fixed(fixed);
    const uint64_t ngroups = uint64_t{1} << (st->nbits - fx.np);
    if constexpr (std::is_same<R, float>::value)
        if ((k == 4 || k == 5) && !qs::gatek_valu_forced())
            return gatek_mfma(st, k, inner, m, conj, fx, cval, ngroups);
    const size_t inner_bytes = inner.size() * sizeof(uint64_t);
    const size_t mat_bytes = mm.size() * sizeof(cxt<R>);
    std::vector<unsigned char> blob(inner_bytes + mat_bytes);
    std::memcpy(blob.data(), inner.data(), inner_bytes);
    std::memcpy(blob.data() + inner_bytes, mm.data(), mat_bytes);
    void* dblob = nullptr;
    if (int rc = upload(st, blob.data(), blob.size(), &dblob)) return rc;
    constexpr int EPT = 4;
    // groups per block must be >= 1: for dim > kBlock*EPT use larger EPT
    int G = (kBlock * EPT) >> k;
    if (G < 1) return qs::fail(QS_ERR_NOT_IMPLEMENTED, "qpp::apply()",
                               "gate wider than 10 qubits");
    const uint32_t grid = grid_for(ngroups, uint64_t(G));
    const size_t shmem = size_t(G) * dim * sizeof(cxt<R>);
    hipLaunchKernelGGL((k_gatek<R, EPT>), dim3(grid), dim3(kBlock), shmem,
                       st->stream, static_cast<cxt<R>*>(st->d), ngroups, int(k),
                       static_cast<const uint64_t*>(dblob),
                       reinterpret_cast<const cxt<R>*>(
                           static_cast<unsigned char*>(dblob) + inner_bytes),
                       fx, cval, uint64_t{0}, uint64_t{0});
    QS_HIP(hipFreeAsync(dblob, st->stream), "qpp_amd::apply");
    return QS_OK;
}

// W = [[Ur, -Ui], [Ui, Ur]] (m column-major complex, conj: conj(U)) packed
// in k_gatek_mfma's A-operand order: entry (mb, s, lane) = W[16 mb + lane %
// 16][4 s + lane / 16].
int gatek_mfma(qs_state* st, uint32_t k, const std::vector<uint64_t>& inner, const double* m,
               bool conj, const Fixed& fx, uint64_t cval, uint64_t ncols) {
    const int dim = 1 << k, nr = 2 * dim, ns = nr / 4, nb = nr / 16;
    auto W = [&](int r, int c) -> float {
        const int rr = r % dim, cc = c % dim;
        const double re = m[2 * (rr + cc * dim)];
        const double im = conj ? -m[2 * (rr + cc * dim) + 1] : m[2 * (rr + cc * dim) + 1];
        if ((r < dim) == (c < dim)) return static_cast<float>(re);
        return static_cast<float>(r < dim ? -im : im);
    };
    std::vector<float> wp(size_t(nb) * ns * 64);
    for (int mb = 0; mb < nb; ++mb)
        for (int s = 0; s < ns; ++s)
            for (int l = 0; l < 64; ++l) wp[(size_t(mb) * ns + s) * 64 + l] = W(16 * mb + l % 16, 4 * s + l / 16);
    const size_t inner_bytes = inner.size() * sizeof(uint64_t);
    std::vector<unsigned char> blob(inner_bytes + wp.size() * sizeof(float));
    std::memcpy(blob.data(), inner.data(), inner_bytes);
    std::memcpy(blob.data() + inner_bytes, wp.data(), wp.size() * sizeof(float));
    // persistent waves (each walks many groups, two ahead in flight): no
    // more workgroups than are resident at once, so none starts late
    const uint64_t groups16 = (ncols + 15) / 16;
    int per_cu = 0, dev = 0, ncu = 0;
    QS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k == 4 ? k_gatek_mfma<4> : k_gatek_mfma<5>,
                                                        kBlock, 0),
           "qpp::apply()");
    QS_HIP(hipGetDevice(&dev), "qpp::apply()");
    QS_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "qpp::apply()");
    const uint32_t resident = static_cast<uint32_t>(std::max(1, per_cu) * std::max(1, ncu));
    const uint32_t grid = std::min<uint32_t>(grid_for(groups16, kBlock / 64), std::min(kMfmaGrid, resident));
    void* dblob = nullptr;
    if (int rc = upload(st, blob.data(), blob.size(), &dblob)) return rc;
    auto* inn = static_cast<const uint64_t*>(dblob);
    auto* wpk = reinterpret_cast<const float*>(static_cast<unsigned char*>(dblob) + inner_bytes);
    if (k == 4)
        hipLaunchKernelGGL(k_gatek_mfma<4>, dim3(grid), dim3(kBlock), 0, st->stream,
                           static_cast<float2*>(st->d), ncols, inn, wpk, fx, cval);
    else
        hipLaunchKernelGGL(k_gatek_mfma<5>, dim3(grid), dim3(kBlock), 0, st->stream,
                           static_cast<float2*>(st->d), ncols, inn, wpk, fx, cval);
    QS_HIP(hipGetLastError(), "qpp::apply()");
    QS_HIP(hipFreeAsync(dblob, st->stream), "qpp_amd::apply");
    return QS_OK;
}

template <typename R>
int qudit_impl(qs_state* st, const qs::QuditGate& g, const double* mats,
               uint32_t nmats, uint64_t DA, bool conj) {
    if (DA > 64)
        return qs::fail(QS_ERR_NOT_IMPLEMENTED, "qpp::apply()",
                        "qudit gate dimension above 64");
    QG q{};
    q.ntg = g.ntg;
    q.nsp = g.nsp;
    q.nctrl = g.nctrl;
    for (uint32_t i = 0; i < g.ntg; ++i) q.tg_stride[i] = g.tg_stride[i], q.tg_dim[i] = g.tg_dim[i];
    if (g.nsp > 40 || g.nctrl > 40)
        return qs::fail(QS_ERR_NOT_IMPLEMENTED, "qpp::apply()", "too many subsystems");
    uint64_t nspec = 1;
    for (uint32_t i = 0; i < g.nsp; ++i) {
        q.sp_stride[i] = g.sp_stride[i];
        q.sp_dim[i] = g.sp_dim[i];
        nspec *= g.sp_dim[i];
    }
    for (uint32_t i = 0; i < g.nctrl; ++i) q.c_stride[i] = g.c_stride[i], q.c_shift[i] = g.c_shift[i];
    q.c_dim = g.c_dim;
    q.DA = DA;
    // offsets of the gate digits (target[0] most significant)
    std::vector<uint64_t> off(DA);
    for (uint64_t mi = 0; mi < DA; ++mi) {
        uint64_t rem = mi, v = 0;
        for (int j = int(g.ntg) - 1; j >= 0; --j) {
            v += (rem % g.tg_dim[j]) * g.tg_stride[j];
            rem /= g.tg_dim[j];
        }
        off[mi] = v;
    }
    std::vector<cxt<R>> mm(uint64_t(nmats) * DA * DA);
    for (uint64_t i = 0; i < mm.size(); ++i) mm[i] = to_c<R>(mats, i, conj);
    const size_t off_bytes = off.size() * sizeof(uint64_t);
    const size_t mat_bytes = mm.size() * sizeof(cxt<R>);
    std::vector<unsigned char> blob(off_bytes + mat_bytes);
    std::memcpy(blob.data(), off.data(), off_bytes);
    std::memcpy(blob.data() + off_bytes, mm.data(), mat_bytes);
    void* dblob = nullptr;
    if (int rc = upload(st, blob.data(), blob.size(), &dblob)) return rc;
    const uint32_t grid = grid_for(nspec, kBlock);
    hipLaunchKernelGGL((k_qudit<R>), dim3(grid), dim3(kBlock), 0, st->stream,
                       static_cast<cxt<R>*>(st->d), nspec, q,
                       reinterpret_cast<const cxt<R>*>(
                           static_cast<unsigned char*>(dblob) + off_bytes),
                       static_cast<const uint64_t*>(dblob));
    QS_HIP(hipFreeAsync(dblob, st->stream), "qpp_amd::apply");
    return QS_OK;
}

}  // namespace

namespace qs {

int launch_gate1(qs_state* st, uint32_t bit, const double* m, uint64_t cmask,
                 uint64_t cval, bool conj) {
    int rc = st->dtype == QS_C128
                 ? gate1_impl<double>(st, bit, m, cmask, cval, conj)
                 : gate1_impl<float>(st, bit, m, cmask, cval, conj);
    if (rc) return rc;
    QS_HIP(hipGetLastError(), "qpp::apply()");
    return QS_OK;
}

int launch_gatek(qs_state* st, const uint32_t* bits, uint32_t k,
                 const double* m, uint64_t cmask, uint64_t cval, bool conj) {
    if (k == 1) return launch_gate1(st, bits[0], m, cmask, cval, conj);
    int rc = st->dtype == QS_C128
                 ? gatek_impl<double>(st, bits, k, m, cmask, cval, conj)
                 : gatek_impl<float>(st, bits, k, m, cmask, cval, conj);
    if (rc) return rc;
    QS_HIP(hipGetLastError(), "qpp::apply()");
    return QS_OK;
}

int launch_qudit_gate(qs_state* st, const QuditGate& g, const double* mats,
                      uint32_t nmats, uint64_t DA, bool conj) {
    int rc = st->dtype == QS_C128 ? qudit_impl<double>(st, g, mats, nmats, DA, conj)
                                  : qudit_impl<float>(st, g, mats, nmats, DA, conj);
    if (rc) return rc;
    QS_HIP(hipGetLastError(), "qpp::apply()");
    return QS_OK;
}

}  // namespace qs

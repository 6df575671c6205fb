// qs_batch.hip -- batched noisy trajectories in one persistent launch.
//
// Trajectory t of a batch is what
//   QNoisyEngine<QubitDepolarizingNoise(p)>(qc_t, seed_t).execute(1)
// computes (classes/qnoisy_engine.hpp:104-140; classes/noise.hpp:131-160,
// 388-408): before every step, one std::generate_canonical<double,53> draw
// from std::mt19937(seed_t) per live qubit picks I/X/Y/Z with
// discrete_distribution{1-p, p/3, p/3, p/3}; the circuit ends in a
// measurement of every qubit, whose outcome is drawn with the next uniform
// (instruments.hpp:985-1099).
//
// Host side (per trajectory, in parallel threads): the draws are replayed
// with std::mt19937 itself and the Pauli errors are carried as a Pauli frame
// F = i^k (x) X^x Z^z.  A frame factor on a qubit is absorbed into the next
// dense 1-qubit gate on it (U -> U X^x Z^z, exact: a Pauli permutes and
// negates entries), passed through Clifford gates (CNOT, CZ, X, ...) by
// conjugation, and what is left at the end flips measured bits (X part).
// All trajectories of a circuit therefore share one compiled tile program
// and differ only in a short, sorted list of 2x2 matrix overrides.
//
// Device side: one kernel, one workgroup per trajectory at a time, taken
// from a work queue of trajectories sorted by cost, largest first.  The
// workgroup runs every pass of its circuit's program over its own state
// slot (tiles streamed through registers/LDS as in k_tile; the first pass
// synthesises |0..0> instead of loading), then samples the measurement:
// FP64 |a|^2, a block prefix with a rigorous error bound, and -- only when
// the bound cannot decide -- libstdc++'s sequential sums (random.tcc:
// 2656-2712) reproduced bit for bit by parallel binade scans (qs_exact.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <thread>
#include <vector>

#include "qs_fused.hpp"
#include "qs_internal.hpp"
#include "qs_exact.hpp"
#include "qs_tile_ops.hpp"

namespace {

using qs::fail;

struct BClass {
    int pass_begin, pass_end;
    uint32_t nb;
    uint32_t pad;
};

struct TrajIn {
    uint32_t cls;
    uint32_t ovr_begin, ovr_end;
    uint32_t pad;
    uint64_t xmask;
    double u;
};

struct TrajOut {
    uint64_t index;
    double prob;
    uint32_t exact;  // 1 when the exact sequential replay decided the shot
    uint32_t pad;
};

struct Ovr {
    uint32_t op;
    uint32_t pad;
    double m[8];
};

constexpr double kUr = 1.1102230246251565e-16;  // 2^-53

__device__ __forceinline__ double pnorm(double2 v) {
#pragma clang fp contract(off)
    const double xx = v.x * v.x;
    const double yy = v.y * v.y;
    return xx + yy;
}
__device__ __forceinline__ double pnorm(float2 v) {
#pragma clang fp contract(off)
    const double x = v.x, y = v.y;
    const double xx = x * x;
    const double yy = y * y;
    return xx + yy;
}

// One tile of one pass of a trajectory: x holds the tile in the load layout
// of the pass's first phase (phases pb..pe); the phases' ops run with the
// trajectory's matrix overrides ovr[ov..ovr_end) (sorted by op, so each one
// is met in order) and the tile is stored back into the state a.  Returns
// the override index after the pass's ops.
template <typename R, int NR>
__device__ __forceinline__ uint32_t traj_tile(cxt<R>* x, cxt<R>* z, cxt<R>* __restrict__ a, cxt<R>* lds,
                                              const FPhase* __restrict__ phases, int pb, int pe,
                                              const FOp* __restrict__ ops, const cxt<R>* __restrict__ mats,
                                              const Ovr* __restrict__ ovr, uint32_t ov, uint32_t ovr_end,
                                              uint64_t base, uint64_t gt_ld, uint32_t t) {
    constexpr int RB = RBof<NR>::v;
    const FPhase* ld = phases + pb;
    uint32_t oc = ov;
    uint32_t next_op = oc < ovr_end ? ovr[oc].op : 0xFFFFFFFFu;
    int ci = pb;
    for (int pi = pb; pi < pe; ++pi) {
        const FPhase* ph = phases + pi;
        if (pi != ci) {
            const FPhase* cur = phases + __builtin_amdgcn_readfirstlane(ci);
            const uint32_t sw = thr_s(cur, t);
#pragma unroll
            for (int rho = 0; rho < NR; ++rho) lds[sw ^ sreg<RB>(cur, rho)] = x[rho];
            __syncthreads();
            const uint32_t sr = thr_s(ph, t);
#pragma unroll
            for (int rho = 0; rho < NR; ++rho) x[rho] = lds[sr ^ sreg<RB>(ph, rho)];
            __syncthreads();
            ci = pi;
        }
        // the matrix of op o: its own, or this trajectory's override
        auto mu_of = [&](int o) -> const double* {
            if (uint32_t(o) != next_op) return ops[o].m;
            const double* mu = ovr[oc].m;
            ++oc;
            next_op = oc < ovr_end ? ovr[oc].op : 0xFFFFFFFFu;
            return mu;
        };
        // consecutive ops alternate between x and z (see run_op)
        const int ob = ph->op_begin, oe = ph->op_end;
        int o = ob;
        for (; o + 1 < oe; o += 2) {
            run_op<R, NR>(x, z, ops + o, mu_of(o), mats, t, base);
            run_op<R, NR>(z, x, ops + o + 1, mu_of(o + 1), mats, t, base);
        }
        if (o < oe) {
            run_op<R, NR>(x, z, ops + o, mu_of(o), mats, t, base);
            copy_regs<R, NR>(z, x);
        }
    }
    const FPhase* cur = phases + __builtin_amdgcn_readfirstlane(ci);
    if (!cur->coalesced) {
        const uint32_t sw = thr_s(cur, t);
#pragma unroll
        for (int rho = 0; rho < NR; ++rho) lds[sw ^ sreg<RB>(cur, rho)] = x[rho];
        __syncthreads();
        const uint32_t sr = thr_s(ld, t);
#pragma unroll
        for (int rho = 0; rho < NR; ++rho) x[rho] = lds[sr ^ sreg<RB>(ld, rho)];
        __syncthreads();
        cur = ld;
    }
    cxt<R>* q = a + (base | (cur == ld ? gt_ld : thr_g(cur, t)));
#pragma unroll
    for (int rho = 0; rho < NR; ++rho) q[greg<RB>(cur, rho)] = x[rho];
    return oc;
}

// One pass of the wide trajectories of one circuit (states too large for a
// workgroup each): workgroup (x, y) runs tile x of the state st[w0 + y]
// through pass p; wov[w * stride + pl] is trajectory w's first override
// of the pass (its pl-th), wend[w] the end of its overrides.  The first
// pass of a circuit (zero_in) synthesises |0..0> instead of loading.
template <typename R, int NR>
__global__ __launch_bounds__(kThreads) void k_traj_pass(cxt<R>* const* __restrict__ st,
                                                        const uint32_t* __restrict__ wov,
                                                        const uint32_t* __restrict__ wend, uint32_t w0,
                                                        uint32_t stride, uint32_t pl,
                                                        const FPass* __restrict__ passes, int p,
                                                        const FPhase* __restrict__ phases,
                                                        const FOp* __restrict__ ops,
                                                        const cxt<R>* __restrict__ mats,
                                                        const Ovr* __restrict__ ovr, int zero_in) {
    extern __shared__ unsigned char smem_raw[];
    cxt<R>* lds = reinterpret_cast<cxt<R>*>(smem_raw);
    constexpr int RB = RBof<NR>::v;
    const uint32_t t = threadIdx.x;
    const uint32_t w = w0 + blockIdx.y;
    cxt<R>* __restrict__ a = st[w];
    const FPass* ps = passes + p;
    const int pb = ps->phase_begin, pe = ps->phase_end;
    const FPhase* ld = phases + pb;
    const uint64_t gt_ld = thr_g(ld, t);
    const uint64_t base = tile_base_of(ps, uint64_t(blockIdx.x));
    const uint64_t off = base | gt_ld;
    cxt<R> x[NR], z[NR];
    if (zero_in) {
#pragma unroll
        for (int rho = 0; rho < NR; ++rho) x[rho] = mk<R>((off | greg<RB>(ld, rho)) == 0 ? R(1) : R(0), R(0));
    } else {
        const cxt<R>* q = a + off;
#pragma unroll
        for (int rho = 0; rho < NR; ++rho) x[rho] = q[greg<RB>(ld, rho)];
    }
    (void)traj_tile<R, NR>(x, z, a, lds, phases, pb, pe, ops, mats, ovr, wov[uint64_t(w) * stride + pl],
                           wend[w], base, gt_ld, t);
}

template <typename R, int NR>
__global__ __launch_bounds__(kThreads) void k_traj(cxt<R>* __restrict__ ws, uint64_t slot,
                                                   const BClass* __restrict__ cls,
                                                   const FPass* __restrict__ passes,
                                                   const FPhase* __restrict__ phases,
                                                   const FOp* __restrict__ ops,
                                                   const cxt<R>* __restrict__ mats,
                                                   const TrajIn* __restrict__ tin,
                                                   const Ovr* __restrict__ ovr,
                                                   const uint32_t* __restrict__ order,
                                                   uint32_t ntraj, TrajOut* __restrict__ out,
                                                   uint32_t* __restrict__ next, int force_exact,
                                                   cxt<R>* const* __restrict__ dst,
                                                   cxt<R>* const* __restrict__ src) {
    extern __shared__ unsigned char smem_raw[];
    cxt<R>* lds = reinterpret_cast<cxt<R>*>(smem_raw);
    __shared__ double pre[kThreads + 1];
    __shared__ uint32_t s_k;
    __shared__ uint64_t s_idx;
    __shared__ int s_certain;
    __shared__ qs_exact::ChunkLds<kThreads> exl;
    const uint32_t t = threadIdx.x;
    // work queue: trajectories in `order` (largest first), claimed one at a
    // time; every workgroup leaves once the queue is exhausted
    for (;;) {
        if (t == 0) s_k = atomicAdd(next, 1u);
        __syncthreads();
        const uint32_t k = s_k;
        if (k >= ntraj) break;
        const uint32_t tr = order[k];
        // the trajectory's state: the workgroup's slot, or (dst set) the
        // caller's state of that trajectory, which keeps the final state, or
        // (src[tr] set) a state whose passes k_traj_pass has already run,
        // which is only measured here
        cxt<R>* const sp = src ? src[tr] : nullptr;
        cxt<R>* __restrict__ a = sp ? sp : dst ? dst[tr] : ws + uint64_t(blockIdx.x) * slot;
        const TrajIn ti = tin[tr];
        const BClass c = cls[ti.cls];
        uint32_t ov = ti.ovr_begin;
        // fields the lambdas below use, as scalars: capturing `ti` itself by
        // reference would put the record in scratch memory
        const uint32_t ovr_end = ti.ovr_end;
        const double u_traj = ti.u;
        for (int p = sp ? c.pass_end : c.pass_begin; p < c.pass_end; ++p) {
            const FPass* ps = passes + p;
            const bool zero_in = p == c.pass_begin;  // first pass: |0..0>
            constexpr int RB = RBof<NR>::v;
            const int pb = ps->phase_begin, pe = ps->phase_end;
            const uint64_t ntiles = ps->ntiles;
            const FPhase* ld = phases + pb;
            const uint64_t gt_ld = thr_g(ld, t);
            auto tile_base = [&](uint64_t tile) { return tile_base_of(ps, tile); };
            auto fetch = [&](uint64_t tile, cxt<R>* y) {
                const uint64_t off = tile_base(tile) | gt_ld;
                if (zero_in) {
#pragma unroll
                    for (int rho = 0; rho < NR; ++rho)
                        y[rho] = mk<R>((off | greg<RB>(ld, rho)) == 0 ? R(1) : R(0), R(0));
                } else {
                    const cxt<R>* q = a + off;
#pragma unroll
                    for (int rho = 0; rho < NR; ++rho) y[rho] = q[greg<RB>(ld, rho)];
                }
            };
            cxt<R> x[NR], y[NR], z[NR];
            fetch(0, y);
            uint32_t oc = ov;
            for (uint64_t tile = 0; tile < ntiles; ++tile) {
                const uint64_t base = tile_base(tile);
#pragma unroll
                for (int rho = 0; rho < NR; ++rho) x[rho] = y[rho];
                if (tile + 1 < ntiles) fetch(tile + 1, y);
                oc = traj_tile<R, NR>(x, z, a, lds, phases, pb, pe, ops, mats, ovr, ov, ovr_end, base, gt_ld, t);
            }
            ov = oc;
            __syncthreads();  // the pass's stores are visible to the next pass
        }
        if (dst) {
            __syncthreads();  // every thread has read s_k before it is reused
            continue;
        }

        // ---- measurement of every qubit: index of the outcome -------------
        const uint64_t N = uint64_t{1} << c.nb;
        const uint64_t blk = N / kThreads;  // nb >= 11 -> blk >= 8
        const uint64_t xm = ti.xmask;
        {
            double s = 0;
            const uint64_t i0 = uint64_t(t) * blk;
            for (uint64_t i = i0; i < i0 + blk; ++i) s += pnorm(a[i ^ xm]);
            pre[t + 1] = s;
        }
        __syncthreads();
        if (t == 0) {
            pre[0] = 0;
            for (int j = 1; j <= kThreads; ++j) pre[j] = pre[j - 1] + pre[j];
            const double total = pre[kThreads];
            const double u = u_traj;
            const double target = u * total;
            int b = 0;
            while (b < kThreads - 1 && pre[b + 1] < target) ++b;
            double cc = pre[b], cprev = cc;
            uint64_t idx = N;
            for (uint64_t i = uint64_t(b) * blk; i < uint64_t(b + 1) * blk; ++i) {
                cprev = cc;
                cc += pnorm(a[i ^ xm]);
                if (cc >= target) {
                    idx = i;
                    break;
                }
            }
            // distance of the exact libstdc++ table from ours, relative to
            // the total: sequential sum and partial sums of N terms on its
            // side, block sums + prefix + block scan on ours
            const double bound =
                2.0 * (2.0 * double(N) + 2.0 * double(blk) + 2.0 * kThreads + 8.0) * kUr * total;
            const bool certain = !force_exact && idx < N &&
                                 (idx == 0 || target - cprev > bound) &&
                                 (idx == N - 1 || cc - target > bound);
            s_idx = idx;
            s_certain = certain ? 1 : 0;
        }
        __syncthreads();
        if (!s_certain) {
            // libstdc++ discrete_distribution: sequential sum, normalise,
            // sequential partial sums, last := 1, lower_bound
            // (random.tcc:2656-2712), computed exactly by the workgroup's
            // parallel binade scans (qs_exact.hpp)
            const double S = qs_exact::wg_sum<kThreads, 8>(
                N, [&](uint64_t i) { return pnorm(a[i ^ xm]); }, exl);
            const uint64_t found = qs_exact::wg_lower_bound<kThreads, 8>(
                N, [&](uint64_t i) { return __ddiv_rn(pnorm(a[i ^ xm]), S); }, u_traj, exl);
            if (t == 0) s_idx = found;
        }
        if (t == 0) {
            const uint64_t idx = s_idx;
            TrajOut o;
            o.index = idx;
            o.prob = pnorm(a[idx ^ xm]) / pre[kThreads];
            o.exact = s_certain ? 0 : 1;
            o.pad = 0;
            out[tr] = o;
        }
        __syncthreads();  // the slot and `pre` are reused by the next trajectory
    }
}

__global__ void k_zero_u32(uint32_t* p) { *p = 0; }

// ---------------------------------------------------------------------------
// host: Pauli frames
// ---------------------------------------------------------------------------
using C = std::complex<double>;

// one qubit's frame factor X^x Z^z as a 2x2 column-major matrix
std::array<C, 4> pauli_xz(int x, int z) {
    // X^x Z^z: Z = diag(1,-1), X = swap
    const C d0(1, 0), d1(z ? -1 : 1, 0);
    if (!x) return {d0, C(0, 0), C(0, 0), d1};  // (0,0)=d0, (1,1)=d1
    return {C(0, 0), d0, d1, C(0, 0)};            // X * diag: (1,0)=d0, (0,1)=d1
}

std::vector<C> matmul(const std::vector<C>& A, const std::vector<C>& B, size_t d) {
    std::vector<C> r(d * d, C(0, 0));
    for (size_t j = 0; j < d; ++j)
        for (size_t k = 0; k < d; ++k)
            for (size_t i = 0; i < d; ++i) r[i + j * d] += A[i + k * d] * B[k + j * d];
    return r;
}

std::vector<C> kron(const std::vector<C>& A, size_t da, const std::vector<C>& B, size_t db) {
    const size_t d = da * db;
    std::vector<C> r(d * d);
    for (size_t ia = 0; ia < da; ++ia)
        for (size_t ja = 0; ja < da; ++ja)
            for (size_t ib = 0; ib < db; ++ib)
                for (size_t jb = 0; jb < db; ++jb)
                    r[(ia * db + ib) + (ja * db + jb) * d] = A[ia + ja * da] * B[ib + jb * db];
    return r;
}

// Pauli on |Q| qubits from a pattern: bit 2j = x of Q[j], bit 2j+1 = z
std::vector<C> pauli_of(uint32_t pat, size_t nq) {
    std::vector<C> r{C(1, 0)};
    size_t d = 1;
    for (size_t j = 0; j < nq; ++j) {
        const auto p = pauli_xz((pat >> (2 * j)) & 1, (pat >> (2 * j + 1)) & 1);
        r = kron(r, d, std::vector<C>(p.begin(), p.end()), 2);
        d *= 2;
    }
    return r;
}

// How a gate conjugates Paulis on its qubits: G P G^dagger = i^k P'.
struct Conj {
    std::vector<uint32_t> out;  // pattern -> pattern
    std::vector<int> phase;     // pattern -> k (mod 4)
};

bool conj_table(const std::vector<C>& G, size_t nq, Conj& cj) {
    const size_t d = size_t{1} << nq, np = size_t{1} << (2 * nq);
    std::vector<C> Gd(d * d);
    for (size_t i = 0; i < d; ++i)
        for (size_t j = 0; j < d; ++j) Gd[i + j * d] = std::conj(G[j + i * d]);
    cj.out.assign(np, 0);
    cj.phase.assign(np, 0);
    for (uint32_t pat = 0; pat < np; ++pat) {
        const auto M = matmul(matmul(G, pauli_of(pat, nq), d), Gd, d);
        bool found = false;
        for (uint32_t q = 0; q < np && !found; ++q) {
            const auto P = pauli_of(q, nq);
            C tr(0, 0);
            for (size_t i = 0; i < d * d; ++i) tr += std::conj(P[i]) * M[i];
            tr /= double(d);
            if (std::abs(std::abs(tr) - 1.0) > 1e-9) continue;
            // M must equal tr * P entirely
            double err = 0;
            for (size_t i = 0; i < d * d; ++i) err = std::max(err, std::abs(M[i] - tr * P[i]));
            if (err > 1e-9) continue;
            int k = -1;
            const C ph[4] = {C(1, 0), C(0, 1), C(-1, 0), C(0, -1)};
            for (int j = 0; j < 4; ++j)
                if (std::abs(tr - ph[j]) < 1e-9) k = j;
            if (k < 0) return false;
            cj.out[pat] = q;
            cj.phase[pat] = k;
            found = true;
        }
        if (!found) return false;
    }
    return true;
}

struct Frame {
    std::vector<uint8_t> x, z;  // per qubit
    int k = 0;                   // global phase i^k
    void reset(uint32_t n) {
        x.assign(n, 0);
        z.assign(n, 0);
        k = 0;
    }
    // N_q F for a Pauli error N in {1: X, 2: Y, 3: Z}
    void error(uint32_t q, int e) {
        const int xq = x[q];
        if (e == 1) {
            x[q] ^= 1;
        } else if (e == 3) {
            k += 2 * xq;
            z[q] ^= 1;
        } else {
            k += 1 + 2 * xq;
            x[q] ^= 1;
            z[q] ^= 1;
        }
        k &= 3;
    }
};

// What a lowered gate does to a frame.
struct GateRole {
    uint32_t step = 0;          // circuit step it comes from
    int op = -1;                // op index in the class program
    bool absorb = false;        // dense 1-qubit op: U -> U P
    uint32_t q = 0;             // absorb: its qubit
    std::vector<C> U;           // absorb: its matrix
    std::vector<uint32_t> qs;   // Clifford: its qubits, qs[0] most significant
    Conj cj;
};

struct StepInfo {
    bool measure = false;
    std::vector<uint32_t> gates;  // lowered gate indices
};

struct Cls {
    uint32_t n = 0, nb = 0;
    FusedProgram P;
    std::vector<GateRole> roles;
    std::vector<StepInfo> steps;
    std::vector<uint32_t> meas;  // measured qubits in target order
};

struct Plan {
    std::vector<Ovr> ovr;  // sorted by op
    uint64_t xmask = 0, zmask = 0;
    int phase = 0;
    double u = 0;
    std::vector<uint32_t> ovr_step;  // circuit step of each override
};

std::array<double, 4> noise_table(double p) {
    // discrete_distribution{1-p, p/3, p/3, p/3}: sum, normalise, partial
    // sums, last := 1 (random.tcc:2656-2690)
    const double w[4] = {1 - p, p / 3, p / 3, p / 3};
    double sum = 0;
    for (double v : w) sum += v;
    std::array<double, 4> cp{};
    double acc = 0;
    for (int i = 0; i < 4; ++i) {
        acc += w[i] / sum;
        cp[i] = acc;
    }
    cp[3] = 1.0;
    return cp;
}

void plan_trajectory(const Cls& c, uint64_t seed, double noise_p, Plan& out) {
    std::mt19937 gen(static_cast<std::mt19937::result_type>(seed));
    const auto cp = noise_table(noise_p < 0 ? 0 : noise_p);
    Frame F;
    F.reset(c.n);
    out.ovr.clear();
    out.ovr_step.clear();
    std::vector<std::pair<uint32_t, size_t>> tmp;  // (op, index into out.ovr)
    for (size_t s = 0; s < c.steps.size(); ++s) {
        if (noise_p >= 0)
            for (uint32_t q = 0; q < c.n; ++q) {
                const double u = std::generate_canonical<double, 53>(gen);
                const int e = int(std::lower_bound(cp.begin(), cp.end(), u) - cp.begin());
                if (e) F.error(q, e);
            }
        const StepInfo& si = c.steps[s];
        if (si.measure) {
            out.u = std::generate_canonical<double, 53>(gen);
            break;
        }
        for (uint32_t gi : si.gates) {
            const GateRole& r = c.roles[gi];
            if (r.absorb) {
                const uint32_t q = r.q;
                if (!F.x[q] && !F.z[q]) continue;
                const auto P = pauli_xz(F.x[q], F.z[q]);
                const auto M = matmul(r.U, std::vector<C>(P.begin(), P.end()), 2);
                Ovr o{};
                o.op = static_cast<uint32_t>(r.op);
                for (int i = 0; i < 4; ++i) {
                    o.m[2 * i] = M[i].real();
                    o.m[2 * i + 1] = M[i].imag();
                }
                out.ovr.push_back(o);
                out.ovr_step.push_back(r.step);
                F.x[q] = F.z[q] = 0;
            } else {
                uint32_t pat = 0;
                for (size_t j = 0; j < r.qs.size(); ++j)
                    pat |= (uint32_t(F.x[r.qs[j]]) << (2 * j)) | (uint32_t(F.z[r.qs[j]]) << (2 * j + 1));
                if (!pat) continue;
                const uint32_t np = r.cj.out[pat];
                F.k = (F.k + r.cj.phase[pat]) & 3;
                for (size_t j = 0; j < r.qs.size(); ++j) {
                    F.x[r.qs[j]] = (np >> (2 * j)) & 1;
                    F.z[r.qs[j]] = (np >> (2 * j + 1)) & 1;
                }
            }
        }
    }
    // sort overrides by op (the kernel walks ops in program order)
    std::vector<size_t> idx(out.ovr.size());
    std::iota(idx.begin(), idx.end(), size_t{0});
    std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return out.ovr[a].op < out.ovr[b].op; });
    std::vector<Ovr> so;
    std::vector<uint32_t> ss;
    for (size_t i : idx) {
        so.push_back(out.ovr[i]);
        ss.push_back(out.ovr_step[i]);
    }
    out.ovr.swap(so);
    out.ovr_step.swap(ss);
    out.xmask = out.zmask = 0;
    for (uint32_t q = 0; q < c.n; ++q) {
        if (F.x[q]) out.xmask |= uint64_t{1} << (c.n - 1 - q);
        if (F.z[q]) out.zmask |= uint64_t{1} << (c.n - 1 - q);
    }
    out.phase = F.k;
}

template <typename T>
int dev_put(void*& dptr, size_t& cap, const std::vector<T>& v, hipStream_t s) {
    const size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
    if (bytes > cap) {
        if (dptr) (void)hipFreeAsync(dptr, s);
        dptr = nullptr;
        cap = 0;
        QS_HIP(hipMallocAsync(&dptr, bytes, s), "qpp_amd::batch");
        cap = bytes;
    }
    if (!v.empty())
        QS_HIP(hipMemcpyAsync(dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s),
               "qpp_amd::batch");
    return QS_OK;
}

}  // namespace

// Device memory is stream-ordered on the stream of the first run (the
// calling thread's stream: a serving lane's own), so creating and
// destroying batches never synchronises other streams.
struct qs_batch {
    qs_dtype dtype = QS_C64;
    hipStream_t stream = nullptr;
    uint32_t reg_bits = 4;
    std::vector<Cls> cls;
    // device tables (all classes concatenated)
    void* d_tab = nullptr;
    BClass* d_cls = nullptr;
    FPass* d_passes = nullptr;
    FPhase* d_phases = nullptr;
    FOp* d_ops = nullptr;
    void* d_mats = nullptr;
    // per-run buffers
    void* d_tin = nullptr;
    size_t c_tin = 0;
    void* d_ovr = nullptr;
    size_t c_ovr = 0;
    void* d_order = nullptr;
    size_t c_order = 0;
    void* d_out = nullptr;
    size_t c_out = 0;
    uint32_t* d_next = nullptr;  // work-queue counter
    void* d_ws = nullptr;
    size_t c_ws = 0;
    void* d_dst = nullptr;  // per-trajectory destination states (run_states)
    size_t c_dst = 0;
    // wide trajectories (run pass by pass over all their tiles): their
    // states, per-trajectory state pointers (null for the others), and
    // per wide trajectory its states, first override of every pass and
    // end of overrides
    void* d_wide = nullptr;
    size_t c_wide = 0;
    void* d_src = nullptr;
    size_t c_src = 0;
    void* d_wst = nullptr;
    size_t c_wst = 0;
    void* d_wov = nullptr;
    size_t c_wov = 0;
    void* d_wend = nullptr;
    size_t c_wend = 0;
    // class -> offsets of its passes and ops in the concatenated tables
    std::vector<int> pass_base, op_base;
    // last run
    uint64_t last_exact = 0;
    double last_host_ms = 0, last_device_ms = 0;
    uint32_t last_grid = 0;
};

namespace {

int upload_tables(qs_batch* b, hipStream_t s) {
    std::vector<BClass> bc;
    std::vector<FPass> ps;
    std::vector<FPhase> ph;
    std::vector<FOp> op;
    std::vector<C> mats;
    b->op_base.clear();
    b->pass_base.clear();
    for (const Cls& c : b->cls) {
        const int pbase = int(ps.size()), hbase = int(ph.size()), obase = int(op.size());
        const uint32_t mbase = uint32_t(mats.size());
        b->op_base.push_back(obase);
        b->pass_base.push_back(pbase);
        bc.push_back(BClass{pbase, pbase + int(c.P.passes.size()), c.nb, 0});
        for (FPass p : c.P.passes) {
            p.phase_begin += hbase;
            p.phase_end += hbase;
            ps.push_back(p);
        }
        for (FPhase h : c.P.phases) {
            h.op_begin += obase;
            h.op_end += obase;
            ph.push_back(h);
        }
        for (FOp o : c.P.ops) {
            if (o.kind == OP_U2 || o.kind == OP_U1B) o.moff += mbase;
            seal_op(o);
            op.push_back(o);
        }
        mats.insert(mats.end(), c.P.mats.begin(), c.P.mats.end());
    }
    const size_t es = b->dtype == QS_C128 ? 16 : 8;
    auto al = [](size_t n) { return (n + 255) / 256 * 256; };
    const size_t o1 = 0, o2 = o1 + al(bc.size() * sizeof(BClass)),
                 o3 = o2 + al(std::max<size_t>(1, ps.size()) * sizeof(FPass)),
                 o4 = o3 + al(std::max<size_t>(1, ph.size()) * sizeof(FPhase)),
                 o5 = o4 + al(std::max<size_t>(1, op.size()) * sizeof(FOp)),
                 tot = o5 + al(std::max<size_t>(1, mats.size()) * es);
    std::vector<unsigned char> blob(tot, 0);
    std::memcpy(blob.data() + o1, bc.data(), bc.size() * sizeof(BClass));
    std::memcpy(blob.data() + o2, ps.data(), ps.size() * sizeof(FPass));
    std::memcpy(blob.data() + o3, ph.data(), ph.size() * sizeof(FPhase));
    std::memcpy(blob.data() + o4, op.data(), op.size() * sizeof(FOp));
    for (size_t i = 0; i < mats.size(); ++i) {
        if (es == 16) {
            const double2 v = make_double2(mats[i].real(), mats[i].imag());
            std::memcpy(blob.data() + o5 + i * 16, &v, 16);
        } else {
            const float2 v = make_float2(float(mats[i].real()), float(mats[i].imag()));
            std::memcpy(blob.data() + o5 + i * 8, &v, 8);
        }
    }
    QS_HIP(hipMallocAsync(&b->d_tab, tot, s), "qpp_amd::batch");
    QS_HIP(hipMemcpyAsync(b->d_tab, blob.data(), tot, hipMemcpyHostToDevice, s), "qpp_amd::batch");
    QS_HIP(hipStreamSynchronize(s), "qpp_amd::batch");  // blob is freed on return
    auto* base = static_cast<unsigned char*>(b->d_tab);
    b->d_cls = reinterpret_cast<BClass*>(base + o1);
    b->d_passes = reinterpret_cast<FPass*>(base + o2);
    b->d_phases = reinterpret_cast<FPhase*>(base + o3);
    b->d_ops = reinterpret_cast<FOp*>(base + o4);
    b->d_mats = base + o5;
    return QS_OK;
}

int compile_class(const qs_circuit* qc, const FusedOptions& opt, Cls& c) {
    constexpr const char* W = "qpp_amd::batch";
    if (!qc || qc->d != 2) return fail(QS_ERR_NOT_IMPLEMENTED, W, "qubit circuits only");
    c.n = qc->nq;
    c.nb = qc->nq;
    if (c.nb < 8 + opt.reg_bits || c.nb > 40)
        return fail(QS_ERR_NOT_IMPLEMENTED, W, "qubit count outside the tile range");
    std::vector<BitGate> gates;
    bool measured = false;
    for (uint32_t s = 0; s < qc->nsteps; ++s) {
        const qs_step& st = qc->steps[s];
        StepInfo si;
        if (st.type == QS_STEP_MEASURE || st.type == QS_STEP_MEASURE_MANY) {
            if (measured || st.type != QS_STEP_MEASURE_MANY || st.ntarget != qc->nq)
                return fail(QS_ERR_NOT_IMPLEMENTED, W, "the circuit must end in one measurement of every qubit");
            si.measure = true;
            c.meas.assign(qc->pool + st.target_off, qc->pool + st.target_off + st.ntarget);
            measured = true;
        } else {
            if (measured) return fail(QS_ERR_NOT_IMPLEMENTED, W, "gates after the measurement");
            std::vector<BitGate> g;
            if (int rc = qs::lower_steps(qc, s, s + 1, false, g)) return rc;
            for (auto& x : g) {
                si.gates.push_back(uint32_t(gates.size()));
                gates.push_back(x);
                GateRole r;
                r.step = s;
                c.roles.push_back(r);
            }
        }
        c.steps.push_back(si);
    }
    if (!measured) return fail(QS_ERR_NOT_IMPLEMENTED, W, "the circuit must end in one measurement of every qubit");
    std::vector<uint8_t> seen(c.n, 0);
    for (uint32_t q : c.meas) {
        if (q >= c.n || seen[q]) return fail(QS_ERR_SUBSYS_MISMATCH_DIMS, W, "measurement targets");
        seen[q] = 1;
    }
    if (int rc = qs::fused_compile(gates, c.nb, opt, c.P)) return rc;
    for (size_t i = 0; i < gates.size(); ++i) {
        GateRole& r = c.roles[i];
        const BitGate& g = gates[i];
        r.op = c.P.gate_op[i];
        if (g.bits.size() == 1 && g.cmask == 0 && c.P.ops[r.op].kind == OP_U1) {
            r.absorb = true;
            r.q = c.n - 1 - g.bits[0];
            r.U = g.m;
            continue;
        }
        // qubits of the gate: controls then targets, and its full matrix
        std::vector<uint32_t> cb, tb = g.bits;
        for (uint32_t bt = 0; bt < 64; ++bt)
            if ((g.cmask >> bt) & 1) cb.push_back(bt);
        if (cb.size() + tb.size() > 2)
            return fail(QS_ERR_NOT_IMPLEMENTED, W, "gates on more than 2 qubits");
        std::vector<uint32_t> bits = cb;
        bits.insert(bits.end(), tb.begin(), tb.end());
        const size_t nq = bits.size(), d = size_t{1} << nq, dt = size_t{1} << tb.size();
        std::vector<C> G(d * d, C(0, 0));
        for (size_t col = 0; col < d; ++col) {
            // digit j of col <-> bits[j], j = 0 most significant
            bool on = true;
            for (size_t j = 0; j < cb.size(); ++j) {
                const int v = (col >> (nq - 1 - j)) & 1;
                const int want = (g.cval >> cb[j]) & 1;
                if (v != want) on = false;
            }
            const size_t tcol = col & (dt - 1);
            for (size_t trow = 0; trow < dt; ++trow) {
                const size_t row = (col & ~(dt - 1)) | trow;
                G[row + col * d] = on ? g.m[trow + tcol * dt] : (trow == tcol ? C(1, 0) : C(0, 0));
            }
        }
        if (!conj_table(G, nq, r.cj))
            return fail(QS_ERR_NOT_IMPLEMENTED, W, "non-Clifford gate that is not a dense 1-qubit gate");
        for (uint32_t bt : bits) r.qs.push_back(c.n - 1 - bt);
    }
    return QS_OK;
}

}  // namespace

extern "C" {

int qs_batch_create(const qs_circuit* const* circuits, uint32_t ncircuits, qs_dtype dtype,
                    const qs_fuse_opts* opts, qs_batch** out) {
    constexpr const char* W = "qpp_amd::batch";
    if (!out || !circuits || ncircuits == 0) return fail(QS_ERR_INVALID_HANDLE, W);
    *out = nullptr;
    if (dtype != QS_C128 && dtype != QS_C64) return fail(QS_ERR_TYPE_MISMATCH, W, "dtype");
    FusedOptions opt;
    opt.bundle = false;  // per-trajectory overrides address single 1-bit ops
    opt.reg_bits = 4;
    opt.low_bits = 5;
    opt.max_ops = 48;
    if (opts) {
        if (opts->reg_bits) opt.reg_bits = opts->reg_bits;
        if (opts->low_bits) opt.low_bits = opts->low_bits;
        if (opts->max_ops) opt.max_ops = opts->max_ops;
    }
    if (opt.reg_bits != 3 && opt.reg_bits != 4) return fail(QS_ERR_OUT_OF_RANGE, W, "reg_bits");
    auto* b = new qs_batch;
    b->dtype = dtype;
    b->reg_bits = opt.reg_bits;
    b->cls.resize(ncircuits);
    for (uint32_t i = 0; i < ncircuits; ++i)
        if (int rc = compile_class(circuits[i], opt, b->cls[i])) {
            delete b;
            return rc;
        }
    *out = b;
    return QS_OK;
}

int qs_batch_destroy(qs_batch* b) {
    if (!b) return QS_OK;
    for (void* p : {b->d_tab, b->d_tin, b->d_ovr, b->d_order, b->d_out, b->d_ws, b->d_dst, b->d_wide,
                    b->d_src, b->d_wst, b->d_wov, b->d_wend, static_cast<void*>(b->d_next)})
        if (p) (void)hipFreeAsync(p, b->stream);
    delete b;
    return QS_OK;
}

int qs_batch_info(const qs_batch* b, uint32_t circuit, uint64_t* num_passes,
                  uint64_t* num_ops, uint64_t* last_exact, double* last_host_ms,
                  double* last_device_ms) {
    if (!b || circuit >= b->cls.size()) return fail(QS_ERR_INVALID_HANDLE, "qpp_amd::batch");
    if (num_passes) *num_passes = b->cls[circuit].P.passes.size();
    if (num_ops) *num_ops = b->cls[circuit].P.ops.size();
    if (last_exact) *last_exact = b->last_exact;
    if (last_host_ms) *last_host_ms = b->last_host_ms;
    if (last_device_ms) *last_device_ms = b->last_device_ms;
    return QS_OK;
}

// Host-only: the frame plan of one trajectory (for inspection and tests).
int qs_batch_plan(const qs_batch* b, uint32_t circuit, uint64_t seed, double noise_p,
                  uint64_t cap, uint64_t* n_ovr, uint32_t* ovr_step, double* ovr_mats,
                  uint64_t* xmask, uint64_t* zmask, uint32_t* phase, double* u) {
    if (!b || circuit >= b->cls.size()) return fail(QS_ERR_INVALID_HANDLE, "qpp_amd::batch");
    Plan pl;
    plan_trajectory(b->cls[circuit], seed, noise_p, pl);
    if (n_ovr) *n_ovr = pl.ovr.size();
    for (size_t i = 0; i < pl.ovr.size() && i < cap; ++i) {
        if (ovr_step) ovr_step[i] = pl.ovr_step[i];
        if (ovr_mats) std::memcpy(ovr_mats + 8 * i, pl.ovr[i].m, 8 * sizeof(double));
    }
    if (xmask) *xmask = pl.xmask;
    if (zmask) *zmask = pl.zmask;
    if (phase) *phase = uint32_t(pl.phase);
    if (u) *u = pl.u;
    return QS_OK;
}

}  // extern "C"

namespace {

// Trajectories of at least this many qubits run pass by pass over all their
// tiles (k_traj_pass, one workgroup per tile of each such state) before
// k_traj measures them: one workgroup streaming a state of 2^20 amplitudes
// and more tile after tile is bound by its own latency.  The environment
// variable QPP_AMD_TRAJ_WIDE_BITS sets another threshold (64: none).
uint32_t wide_bits() {
    const char* v = std::getenv("QPP_AMD_TRAJ_WIDE_BITS");
    if (!v || !*v) return 18;
    return static_cast<uint32_t>(std::strtoul(v, nullptr, 10));
}

// One launch of k_traj over ntraj trajectories (arguments validated by the
// caller), after the passes of the wide ones.  dst: nullptr (each
// trajectory ends in its measurement), or one state per trajectory that
// receives its final state instead (the wide ones' passes then run on it
// and k_traj leaves them alone).
int run_impl(qs_batch* b, uint64_t ntraj, const uint32_t* circuit_of, const uint64_t* seeds,
             double noise_p, const std::vector<void*>* dst, uint64_t* out_index, double* out_prob) {
    constexpr const char* W = "qpp_amd::batch";
    if (int rc = qs::ensure_device()) return rc;
    if (!b->stream) b->stream = qs::current_stream();
    hipStream_t s = b->stream;
    if (!b->d_tab)
        if (int rc = upload_tables(b, s)) return rc;

    // ---- host: per-trajectory frame plans, in parallel --------------------
    const auto h0 = std::chrono::steady_clock::now();
    std::vector<Plan> plans(ntraj);
    {
        unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        nth = static_cast<unsigned>(std::min<uint64_t>(nth, ntraj));
        std::vector<std::thread> th;
        for (unsigned w = 0; w < nth; ++w)
            th.emplace_back([&, w]() {
                for (uint64_t i = w; i < ntraj; i += nth)
                    plan_trajectory(b->cls[circuit_of[i]], seeds[i], noise_p, plans[i]);
            });
        for (auto& x : th) x.join();
    }
    std::vector<TrajIn> tin(ntraj);
    std::vector<Ovr> ovr;
    for (uint64_t i = 0; i < ntraj; ++i) {
        TrajIn& ti = tin[i];
        ti.cls = circuit_of[i];
        ti.ovr_begin = uint32_t(ovr.size());
        const int obase = b->op_base[ti.cls];
        for (Ovr o : plans[i].ovr) {
            o.op += uint32_t(obase);
            ovr.push_back(o);
        }
        ti.ovr_end = uint32_t(ovr.size());
        ti.xmask = plans[i].xmask;
        ti.u = plans[i].u;
    }
    // queue order: largest work first (greedy longest-processing-time)
    std::vector<uint32_t> order(ntraj);
    std::iota(order.begin(), order.end(), 0u);
    // wide trajectories, grouped by circuit: their passes run before k_traj
    const uint32_t wbits = wide_bits();
    std::vector<uint32_t> wide;
    std::vector<char> is_wide(ntraj, 0);
    for (uint64_t i = 0; i < ntraj; ++i) {
        const Cls& c = b->cls[circuit_of[i]];
        if (c.nb >= wbits && !c.P.passes.empty()) {
            wide.push_back(uint32_t(i));
            is_wide[i] = 1;
        }
    }
    std::stable_sort(wide.begin(), wide.end(),
                     [&](uint32_t x, uint32_t y) { return circuit_of[x] < circuit_of[y]; });
    auto cost = [&](uint32_t i) {
        const Cls& c = b->cls[circuit_of[i]];
        return double(is_wide[i] ? 1 : c.P.passes.size() + 1) * std::ldexp(1.0, int(c.nb));
    };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return cost(x) > cost(y); });
    const auto h1 = std::chrono::steady_clock::now();
    b->last_host_ms = std::chrono::duration<double, std::milli>(h1 - h0).count();

    // ---- device ----------------------------------------------------------
    uint32_t maxnb = 0;
    for (uint64_t i = 0; i < ntraj; ++i)
        if (!is_wide[i]) maxnb = std::max(maxnb, b->cls[circuit_of[i]].nb);
    const size_t es = b->dtype == QS_C128 ? 16 : 8;
    const size_t shmem = es << (8 + b->reg_bits);
    int per_cu = 0;
    hipError_t e;
    if (b->dtype == QS_C128)
        e = b->reg_bits == 3
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_traj<double, 8>, kThreads, shmem)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_traj<double, 16>, kThreads, shmem);
    else
        e = b->reg_bits == 3
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_traj<float, 8>, kThreads, shmem)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_traj<float, 16>, kThreads, shmem);
    if (e != hipSuccess) return qs::fail_hip(e, W);
    int dev = 0, ncu = 0;
    QS_HIP(hipGetDevice(&dev), W);
    QS_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), W);
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>(ntraj, uint64_t(std::max(1, per_cu)) * ncu));
    const uint64_t slot = uint64_t{1} << maxnb;
    const size_t ws_bytes = dst ? 0 : size_t(grid) * slot * es;
    if (ws_bytes > b->c_ws) {
        if (b->d_ws) (void)hipFreeAsync(b->d_ws, s);
        b->d_ws = nullptr;
        b->c_ws = 0;
        QS_HIP(hipMallocAsync(&b->d_ws, ws_bytes, s), W);
        b->c_ws = ws_bytes;
    }
    if (dst)
        if (int rc = dev_put(b->d_dst, b->c_dst, *dst, s)) return rc;
    // wide trajectories: a state each, and every pass's first override
    uint32_t stride = 1;
    for (uint32_t i : wide) stride = std::max<uint32_t>(stride, uint32_t(b->cls[circuit_of[i]].P.passes.size()));
    std::vector<void*> src;
    std::vector<void*> wst(wide.size());
    std::vector<uint32_t> wov(wide.size() * size_t(stride), 0), wend(wide.size());
    if (!wide.empty()) {
        size_t wbytes = 0;
        if (!dst)
            for (uint32_t i : wide) wbytes += es << b->cls[circuit_of[i]].nb;
        if (wbytes > b->c_wide) {
            if (b->d_wide) (void)hipFreeAsync(b->d_wide, s);
            b->d_wide = nullptr;
            b->c_wide = 0;
            QS_HIP(hipMallocAsync(&b->d_wide, wbytes, s), W);
            b->c_wide = wbytes;
        }
        src.assign(ntraj, nullptr);
        size_t off = 0;
        for (size_t w = 0; w < wide.size(); ++w) {
            const uint32_t i = wide[w];
            const uint32_t ci = circuit_of[i];
            const Cls& c = b->cls[ci];
            if (dst) {
                wst[w] = (*dst)[i];
            } else {
                wst[w] = static_cast<unsigned char*>(b->d_wide) + off;
                off += es << c.nb;
            }
            src[i] = wst[w];
            const TrajIn& ti = tin[i];
            wend[w] = ti.ovr_end;
            // the overrides are sorted by op and the passes' ops are
            // consecutive: pass pl starts at the first override at or after
            // its first op
            uint32_t k = ti.ovr_begin;
            for (size_t pl = 0; pl < c.P.passes.size(); ++pl) {
                const FPass& ps = c.P.passes[pl];
                if (ps.phase_begin >= ps.phase_end) return fail(QS_ERR_NOT_IMPLEMENTED, W, "empty pass");
                if (ps.ntiles == 0 || ps.ntiles > 0xFFFFFFFFull) return fail(QS_ERR_OUT_OF_RANGE, W, "tiles per pass");
                const uint32_t first = uint32_t(b->op_base[ci] + c.P.phases[size_t(ps.phase_begin)].op_begin);
                while (k < ti.ovr_end && ovr[k].op < first) ++k;
                wov[w * stride + pl] = k;
            }
        }
        if (int rc = dev_put(b->d_src, b->c_src, src, s)) return rc;
        if (int rc = dev_put(b->d_wst, b->c_wst, wst, s)) return rc;
        if (int rc = dev_put(b->d_wov, b->c_wov, wov, s)) return rc;
        if (int rc = dev_put(b->d_wend, b->c_wend, wend, s)) return rc;
    }
    if (int rc = dev_put(b->d_tin, b->c_tin, tin, s)) return rc;
    if (int rc = dev_put(b->d_ovr, b->c_ovr, ovr, s)) return rc;
    if (int rc = dev_put(b->d_order, b->c_order, order, s)) return rc;
    const size_t ob = size_t(ntraj) * sizeof(TrajOut);
    if (ob > b->c_out) {
        if (b->d_out) (void)hipFreeAsync(b->d_out, s);
        b->d_out = nullptr;
        b->c_out = 0;
        QS_HIP(hipMallocAsync(&b->d_out, ob, s), W);
        b->c_out = ob;
    }
    if (!b->d_next) QS_HIP(hipMallocAsync(reinterpret_cast<void**>(&b->d_next), 256, s), W);
    hipEvent_t e0, e1;
    QS_HIP(hipEventCreate(&e0), W);
    QS_HIP(hipEventCreate(&e1), W);
    QS_HIP(hipEventRecord(e0, s), W);
    const dim3 gd(static_cast<uint32_t>(grid)), bd(kThreads);
    auto* tinp = static_cast<const TrajIn*>(b->d_tin);
    auto* ovrp = static_cast<const Ovr*>(b->d_ovr);
    auto* ordp = static_cast<const uint32_t*>(b->d_order);
    auto* outp = static_cast<TrajOut*>(b->d_out);
    const uint32_t nt = static_cast<uint32_t>(ntraj);
    const int force = qs::force_exact_sampling() ? 1 : 0;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, b->d_next);
    // the wide trajectories' passes, circuit by circuit, pass by pass: grid
    // (tiles of the pass, trajectories of the circuit, at most 65535 a launch)
    for (size_t w0 = 0; w0 < wide.size();) {
        const uint32_t ci = circuit_of[wide[w0]];
        size_t w1 = w0;
        while (w1 < wide.size() && circuit_of[wide[w1]] == ci) ++w1;
        const Cls& c = b->cls[ci];
        for (size_t pl = 0; pl < c.P.passes.size(); ++pl) {
            const uint64_t ntiles = c.P.passes[pl].ntiles;
            const int p = b->pass_base[ci] + int(pl);
            const int zin = pl == 0 ? 1 : 0;
            for (size_t y0 = w0; y0 < w1; y0 += 65535) {
                const dim3 g(static_cast<uint32_t>(ntiles), static_cast<uint32_t>(std::min<size_t>(65535, w1 - y0)));
                const uint32_t wy = static_cast<uint32_t>(y0), pls = static_cast<uint32_t>(pl);
                auto* wovp = static_cast<const uint32_t*>(b->d_wov);
                auto* wendp = static_cast<const uint32_t*>(b->d_wend);
                if (b->dtype == QS_C128) {
                    auto* stp = static_cast<double2* const*>(b->d_wst);
                    auto* mp = static_cast<const double2*>(b->d_mats);
                    if (b->reg_bits == 3)
                        hipLaunchKernelGGL((k_traj_pass<double, 8>), g, dim3(kThreads), shmem, s, stp, wovp, wendp,
                                           wy, stride, pls, b->d_passes, p, b->d_phases, b->d_ops, mp, ovrp, zin);
                    else
                        hipLaunchKernelGGL((k_traj_pass<double, 16>), g, dim3(kThreads), shmem, s, stp, wovp, wendp,
                                           wy, stride, pls, b->d_passes, p, b->d_phases, b->d_ops, mp, ovrp, zin);
                } else {
                    auto* stp = static_cast<float2* const*>(b->d_wst);
                    auto* mp = static_cast<const float2*>(b->d_mats);
                    if (b->reg_bits == 3)
                        hipLaunchKernelGGL((k_traj_pass<float, 8>), g, dim3(kThreads), shmem, s, stp, wovp, wendp,
                                           wy, stride, pls, b->d_passes, p, b->d_phases, b->d_ops, mp, ovrp, zin);
                    else
                        hipLaunchKernelGGL((k_traj_pass<float, 16>), g, dim3(kThreads), shmem, s, stp, wovp, wendp,
                                           wy, stride, pls, b->d_passes, p, b->d_phases, b->d_ops, mp, ovrp, zin);
                }
            }
        }
        w0 = w1;
    }
    if (b->dtype == QS_C128) {
        if (b->reg_bits == 3)
            hipLaunchKernelGGL((k_traj<double, 8>), gd, bd, shmem, s, static_cast<double2*>(b->d_ws), slot,
                               b->d_cls, b->d_passes, b->d_phases, b->d_ops,
                               static_cast<const double2*>(b->d_mats), tinp, ovrp, ordp, nt, outp, b->d_next, force,
                               static_cast<double2* const*>(dst ? b->d_dst : nullptr),
                               static_cast<double2* const*>(wide.empty() ? nullptr : b->d_src));
        else
            hipLaunchKernelGGL((k_traj<double, 16>), gd, bd, shmem, s, static_cast<double2*>(b->d_ws), slot,
                               b->d_cls, b->d_passes, b->d_phases, b->d_ops,
                               static_cast<const double2*>(b->d_mats), tinp, ovrp, ordp, nt, outp, b->d_next, force,
                               static_cast<double2* const*>(dst ? b->d_dst : nullptr),
                               static_cast<double2* const*>(wide.empty() ? nullptr : b->d_src));
    } else {
        if (b->reg_bits == 3)
            hipLaunchKernelGGL((k_traj<float, 8>), gd, bd, shmem, s, static_cast<float2*>(b->d_ws), slot,
                               b->d_cls, b->d_passes, b->d_phases, b->d_ops,
                               static_cast<const float2*>(b->d_mats), tinp, ovrp, ordp, nt, outp, b->d_next, force,
                               static_cast<float2* const*>(dst ? b->d_dst : nullptr),
                               static_cast<float2* const*>(wide.empty() ? nullptr : b->d_src));
        else
            hipLaunchKernelGGL((k_traj<float, 16>), gd, bd, shmem, s, static_cast<float2*>(b->d_ws), slot,
                               b->d_cls, b->d_passes, b->d_phases, b->d_ops,
                               static_cast<const float2*>(b->d_mats), tinp, ovrp, ordp, nt, outp, b->d_next, force,
                               static_cast<float2* const*>(dst ? b->d_dst : nullptr),
                               static_cast<float2* const*>(wide.empty() ? nullptr : b->d_src));
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    std::vector<TrajOut> res(ntraj);
    if (e == hipSuccess) e = hipMemcpyAsync(res.data(), outp, ob, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return qs::fail_hip(e, W);
    b->last_device_ms = ms;
    b->last_grid = static_cast<uint32_t>(grid);
    uint64_t nex = 0;
    for (uint64_t i = 0; i < ntraj; ++i) {
        if (out_index) out_index[i] = res[i].index;
        if (out_prob) out_prob[i] = res[i].prob;
        nex += res[i].exact;
    }
    b->last_exact = nex;
    return QS_OK;
}

}  // namespace

extern "C" {

int qs_batch_run(qs_batch* b, uint64_t ntraj, const uint32_t* circuit_of, const uint64_t* seeds,
                 double noise_p, uint64_t* out_index, double* out_prob) {
    constexpr const char* W = "qpp_amd::batch";
    if (!b) return fail(QS_ERR_INVALID_HANDLE, W);
    if (ntraj == 0) return QS_OK;
    if (!circuit_of || !seeds) return fail(QS_ERR_INVALID_HANDLE, W, "inputs");
    if (ntraj > 0xFFFFFFFFull) return fail(QS_ERR_OUT_OF_RANGE, W, "ntraj");
    if (!(noise_p <= 1.0)) return fail(QS_ERR_OUT_OF_RANGE, W, "noise_p");
    for (uint64_t i = 0; i < ntraj; ++i)
        if (circuit_of[i] >= b->cls.size()) return fail(QS_ERR_OUT_OF_RANGE, W, "circuit_of");
    return run_impl(b, ntraj, circuit_of, seeds, noise_p, nullptr, out_index, out_prob);
}

// The gate steps of circuit circuit_of[i] applied to |0..0> into out[i] (a
// ket of that circuit's qubits in the batch's dtype), all in one launch;
// the measurement is not taken.
int qs_batch_run_states(qs_batch* b, uint64_t nstates, const uint32_t* circuit_of,
                        qs_state* const* out) {
    constexpr const char* W = "qpp_amd::batch";
    if (!b) return fail(QS_ERR_INVALID_HANDLE, W);
    if (nstates == 0) return QS_OK;
    if (!circuit_of || !out) return fail(QS_ERR_INVALID_HANDLE, W, "inputs");
    if (nstates > 0xFFFFFFFFull) return fail(QS_ERR_OUT_OF_RANGE, W, "nstates");
    std::vector<void*> dst(nstates);
    for (uint64_t i = 0; i < nstates; ++i) {
        if (circuit_of[i] >= b->cls.size()) return fail(QS_ERR_OUT_OF_RANGE, W, "circuit_of");
        const qs_state* st = out[i];
        if (!st || !st->d) return fail(QS_ERR_INVALID_HANDLE, W, "out");
        if (st->dtype != b->dtype) return fail(QS_ERR_TYPE_MISMATCH, W, "out");
        if (st->kind != QS_KET || st->nelem != (uint64_t{1} << b->cls[circuit_of[i]].nb))
            return fail(QS_ERR_DIMS_MISMATCH_CVECTOR, W, "out");
        for (uint64_t j = 0; j < i; ++j)
            if (out[j]->d == st->d) return fail(QS_ERR_OUT_OF_RANGE, W, "out (the same state twice)");
        dst[i] = st->d;
    }
    // the states' own pending work (their initialisation) comes first
    if (int rc = qs::ensure_device()) return rc;
    if (!b->stream) b->stream = qs::current_stream();
    for (uint64_t i = 0; i < nstates; ++i)
        if (out[i]->stream != b->stream) QS_HIP(hipStreamSynchronize(out[i]->stream), W);
    std::vector<uint64_t> seeds(nstates, 0);
    return run_impl(b, nstates, circuit_of, seeds.data(), -1.0, &dst, nullptr, nullptr);
}

}  // extern "C"

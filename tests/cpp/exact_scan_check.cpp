// Host check of the exact parallel table (qpp_amd/csrc/qs_exact_core.hpp):
// runs the same block decomposition as the device kernels -- block maps in
// the binade of an approximate prefix, the 64-block chain with its
// element-by-element fallback, the per-block expansion with in-chunk binade
// crossings, and the single-workgroup walk of the trajectory kernel -- with
// the header's integer arithmetic, and compares every result bit for bit
// with libstdc++'s own sequential std::accumulate / std::partial_sum
// (random.tcc:2656-2712: S, q = p / S, cp, cp[N-1] := 1).
//
// usage: exact_scan_check <mode> <log2 N> <seed> [perturb]
// prints one line "ok S=... slow=... restarts=..." or "FAIL ..." and exits
// non-zero on a mismatch.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../qpp_amd/csrc/qs_exact_core.hpp"

using namespace qs_exact;

namespace {

constexpr int kChunk = 4096;  // device block size (256 threads x 16)
constexpr int kNT = 256;

long g_restarts = 0;  // binade crossings resolved inside a run
long g_slow = 0;      // blocks the chain summed element by element

// wg_chunk: NT threads x V consecutive elements each (qs_exact.hpp)
template <int NT, int V>
double wg_chunk(double s, const double* x, double* out) {
    int start = 0;
    for (;;) {
        const int E = binade(s);
        const uint64_t m0 = mant(s);
        std::vector<Tr> mine(NT, identity()), ex(NT);
        for (int t = start; t < NT; ++t)
            for (int j = 0; j < V; ++j) mine[t] = compose(mine[t], elem(x[t * V + j], E));
        Tr acc = identity();
        for (int t = 0; t < NT; ++t) {
            ex[t] = acc;
            acc = compose(acc, mine[t]);
        }
        auto walk = [&](int t) {
            uint64_t m = apply(m0, ex[t]);
            for (int j = 0; j < V; ++j) {
                m = apply(m, elem(x[t * V + j], E));
                out[t * V + j] = from_mant(m, E);
            }
        };
        const uint64_t mend = apply(m0, acc);
        if (mend < kTop) {
            for (int t = start; t < NT; ++t) walk(t);
            return from_mant(mend, E);
        }
        ++g_restarts;
        int f = NT;
        for (int t = start; t < NT; ++t)
            if (apply(m0, compose(ex[t], mine[t])) >= kTop) {
                f = t;
                break;
            }
        for (int t = start; t < f; ++t) walk(t);
        uint64_t m = apply(m0, ex[f]);
        double v = from_mant(m, E);
        bool fp = false;
        for (int j = 0; j < V; ++j) {
            const double xj = x[f * V + j];
            if (!fp) {
                const uint64_t mn = apply(m, elem(xj, E));
                if (mn < kTop) {
                    m = mn;
                    v = from_mant(m, E);
                } else {
                    v = v + xj;
                    fp = true;
                }
            } else {
                v = v + xj;
            }
            out[f * V + j] = v;
        }
        s = v;
        start = f + 1;
        if (start >= NT) return s;
    }
}

// element i of the summed sequence: p_i, or p_i / S
struct Seq {
    const std::vector<double>& p;
    double S;
    bool div;
    double operator()(uint64_t i) const { return div ? p[i] / S : p[i]; }
};

// k_ex_block_tr + k_ex_chain: exact sum before each block, total in [nb]
std::vector<double> chain(const Seq& x, uint64_t N, const std::vector<double>& approx) {
    const uint64_t nb = (N + kChunk - 1) / kChunk;
    std::vector<Tr> T(nb);
    std::vector<int> Eb(nb);
    for (uint64_t b = 0; b < nb; ++b) {
        Eb[b] = binade(approx[b]);
        Tr t = identity();
        for (uint64_t i = b * kChunk; i < (b + 1) * kChunk && i < N; ++i)
            t = compose(t, elem(x(i), Eb[b]));
        T[b] = t;
    }
    std::vector<double> s_in(nb + 1);
    double s = 0.0;
    uint64_t b = 0;
    while (b < nb) {
        const int E = binade(s);
        const uint64_t m0 = mant(s);
        const int cnt = nb - b < 64 ? int(nb - b) : 64;
        Tr acc = identity();
        int f = cnt;
        std::vector<Tr> inc(cnt), exv(cnt);
        for (int l = 0; l < cnt; ++l) {
            exv[l] = acc;
            acc = compose(acc, Eb[b + l] == E ? T[b + l] : Tr{kTop, kTop});
            inc[l] = acc;
            if (apply(m0, acc) >= kTop) {
                f = l;
                break;
            }
        }
        for (int l = 0; l < f; ++l) s_in[b + l] = from_mant(apply(m0, exv[l]), E);
        if (f > 0) s = from_mant(apply(m0, inc[f - 1]), E);
        b += f;
        if (f < cnt) {
            ++g_slow;
            s_in[b] = s;
            // one wave, two halves of 64 lanes x 32 consecutive elements
            double xs[kChunk / 2], val[kChunk / 2];
            for (int h = 0; h < 2; ++h) {
                const uint64_t e0 = b * kChunk + uint64_t(h) * (kChunk / 2);
                for (uint64_t k = 0; k < uint64_t(kChunk / 2); ++k)
                    xs[k] = e0 + k < N ? x(e0 + k) : 0.0;
                s = wg_chunk<64, kChunk / 128>(s, xs, val);
            }
            b += 1;
        }
    }
    s_in[nb] = s;
    return s_in;
}

std::vector<double> make_input(const std::string& mode, uint64_t N, uint64_t seed) {
    std::mt19937_64 g(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::uniform_real_distribution<double> ud(0.0, 1.0);
    std::vector<double> p(N);
    // |a|^2 as std::norm rounds it (no fused multiply-add)
    auto norm2 = [](double re, double im) {
        volatile double a = re * re;
        volatile double b = im * im;
        return a + b;
    };
    if (mode == "random") {
        for (auto& v : p) v = norm2(nd(g), nd(g)) / double(N);
    } else if (mode == "zeros") {  // long zero runs, including a zero prefix
        for (uint64_t i = 0; i < N; ++i) {
            const uint64_t blk = (i / 5000) % 4;
            p[i] = (i < N / 8 || blk == 1 || blk == 3) ? 0.0 : norm2(nd(g), nd(g)) / double(N);
        }
    } else if (mode == "dominant") {  // one weight ~1, the rest ~1e-20
        for (auto& v : p) v = norm2(nd(g), nd(g)) * 1e-20;
        p[N / 3] = 0.75;
    } else if (mode == "ties") {  // grid values that make half-ulp ties
        // s climbs through [1, 2), [2, 4) ...; adds of 2^-53, 3 * 2^-53 and
        // 2^-52 land on ties and on exact grid points of every binade
        p[0] = 1.0;
        for (uint64_t i = 1; i < N; ++i) {
            const double r = ud(g);
            p[i] = r < 0.5 ? std::ldexp(1.0, -53) : r < 0.8 ? std::ldexp(3.0, -53)
                                                 : r < 0.95 ? std::ldexp(1.0, -52)
                                                            : std::ldexp(1.0, -30);
        }
    } else if (mode == "tiny") {  // a subnormal start, then log-uniform scales
        p[0] = 4.9e-324;
        p[1] = 1e-310;
        for (uint64_t i = 2; i < N; ++i) p[i] = std::pow(10.0, -300.0 * ud(g));
    } else if (mode == "powers") {  // exact powers of two
        for (auto& v : p) v = std::ldexp(1.0, -int(20 + g() % 40));
    } else if (mode == "halves") {  // amplitudes of 2^-k: p on coarse grids
        for (uint64_t i = 0; i < N; ++i) {
            const double a = std::ldexp(1.0, -int(g() % 12)) / std::sqrt(double(N));
            p[i] = norm2(a, (i & 1) ? a : 0.0);
        }
    } else {
        std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
        std::exit(2);
    }
    return p;
}

std::string bits(double v) {
    char b[64];
    uint64_t u;
    std::memcpy(&u, &v, 8);
    std::snprintf(b, sizeof b, "%.17g (0x%016" PRIx64 ")", v, u);
    return b;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s mode log2N seed [perturb]\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    const uint64_t N = uint64_t{1} << std::atoi(argv[2]);
    const uint64_t seed = std::strtoull(argv[3], nullptr, 10);
    const bool perturb = argc > 4 && std::atoi(argv[4]) != 0;
    const std::vector<double> p = make_input(mode, N, seed);

    // the reference: libstdc++'s own steps
    const double S_ref = std::accumulate(p.begin(), p.end(), 0.0);
    std::vector<double> q(N), cp_ref;
    for (uint64_t i = 0; i < N; ++i) q[i] = p[i] / S_ref;
    cp_ref.reserve(N);
    std::partial_sum(q.begin(), q.end(), std::back_inserter(cp_ref));
    cp_ref[N - 1] = 1.0;

    // approximate block prefix (any order): picks the binade of each block's
    // map; "perturb" skews it so that many picks are wrong
    const uint64_t nb = (N + kChunk - 1) / kChunk;
    std::vector<double> approx(nb + 1, 0.0);
    for (uint64_t b = 0; b < nb; ++b) {
        double bs = 0;
        for (uint64_t i = b * kChunk; i < (b + 1) * kChunk && i < N; ++i) bs += p[i];
        approx[b + 1] = approx[b] + bs;
    }
    if (perturb)
        for (uint64_t b = 0; b < nb; b += 3) approx[b] *= (b % 2) ? 1.7 : 0.6;

    // pass 1: S
    const std::vector<double> sp = chain(Seq{p, 1.0, false}, N, approx);
    const double S = sp[nb];
    if (std::memcmp(&S, &S_ref, 8) != 0) {
        std::printf("FAIL S %s != %s\n", bits(S).c_str(), bits(S_ref).c_str());
        return 1;
    }
    // pass 2: q = p / S, block entry values, then each block expanded
    std::vector<double> aq(nb + 1);
    for (uint64_t b = 0; b <= nb; ++b) aq[b] = approx[b] / S;
    const Seq xq{p, S, true};
    const std::vector<double> sq = chain(xq, N, aq);
    std::vector<double> x(kChunk), out(kChunk);
    for (uint64_t b = 0; b < nb; ++b) {
        for (uint64_t j = 0; j < uint64_t(kChunk); ++j) {
            const uint64_t i = b * kChunk + j;
            x[j] = i < N ? xq(i) : 0.0;
        }
        wg_chunk<kNT, 16>(sq[b], x.data(), out.data());
        for (uint64_t j = 0; j < uint64_t(kChunk); ++j) {
            const uint64_t i = b * kChunk + j;
            if (i >= N) break;
            const double v = i == N - 1 ? 1.0 : out[j];
            if (std::memcmp(&v, &cp_ref[i], 8) != 0) {
                std::printf("FAIL cp[%" PRIu64 "] %s != %s\n", i, bits(v).c_str(),
                            bits(cp_ref[i]).c_str());
                return 1;
            }
        }
    }
    // the single-workgroup walk (trajectory kernel: 256 threads x 8)
    {
        double s = 0;
        std::vector<double> xv(kNT * 8), ov(kNT * 8);
        for (uint64_t b = 0; b < N; b += kNT * 8) {
            for (uint64_t j = 0; j < uint64_t(kNT * 8); ++j)
                xv[j] = b + j < N ? p[b + j] : 0.0;
            s = wg_chunk<kNT, 8>(s, xv.data(), ov.data());
        }
        if (std::memcmp(&s, &S_ref, 8) != 0) {
            std::printf("FAIL wg_sum %s != %s\n", bits(s).c_str(), bits(S_ref).c_str());
            return 1;
        }
    }
    std::printf("ok mode=%s N=2^%s S=%s slow=%ld restarts=%ld\n", mode.c_str(), argv[2],
                bits(S).c_str(), g_slow, g_restarts);
    return 0;
}

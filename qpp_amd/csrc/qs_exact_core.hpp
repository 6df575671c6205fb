// qs_exact_core.hpp -- the integer arithmetic that reproduces a sequential
// floating-point sum of non-negative doubles in any grouping.
//
// libstdc++'s discrete_distribution (random.tcc:2656-2712) forms
//   S    = ((p_0 + p_1) + p_2) + ...           (std::accumulate)
//   cp_k = cp_{k-1} + p_k / S                  (std::partial_sum)
// with every add rounded to nearest, ties to even.  All terms are >= 0, so a
// running sum s never decreases.  While s stays inside one binade
// [2^E, 2^(E+1)) -- or [0, 2^-1021) for E = -1022, which also holds zero and
// the subnormals -- every double it can take is a multiple of u = 2^(E-52):
// s = m u with an integer m < 2^53, and
//     fl(m u + x) = (m + inc) u,   inc = x / u rounded to an integer, a tie
//                                        going to the even m + inc,
// as long as m + inc < 2^53 (the sum stays on the grid of u).  inc depends on
// m only through its parity, so an element is a map
//     T : parity of m  ->  increment,
// and two such maps compose into one of the same form (compose() below).
// Composition is associative, so a run of elements collapses into one map and
// a parallel scan of maps gives every running sum exactly.  Increments are
// >= 0, so if a run's total keeps m below 2^53 every partial sum inside it
// does too.  The one add that carries m to 2^53 or beyond leaves the binade:
// it is done in floating point (exact by definition: it is the reference's
// own operation) and the scan resumes in the new binade.  A sum of N terms
// crosses at most one binade per doubling of s, a few dozen per table.
//
// Everything here is integer arithmetic on the bit patterns, identical on the
// host and the device; the header is also compiled by the host test of the
// scan (tests/cpp/exact_scan_check.cpp).
#ifndef QPP_AMD_QS_EXACT_CORE_HPP_
#define QPP_AMD_QS_EXACT_CORE_HPP_

#include <cstdint>

#if defined(__HIP__)
#define QS_HD __host__ __device__ inline
#else
#define QS_HD inline
#endif

namespace qs_exact {

constexpr uint64_t kTop = uint64_t{1} << 53;       // m must stay below this
constexpr uint64_t kHidden = uint64_t{1} << 52;
constexpr uint64_t kFrac = kHidden - 1;

// increments for an even / odd m; kTop means "leaves the binade"
struct Tr {
    uint64_t i0, i1;
};

QS_HD uint64_t bits_of(double v) { return __builtin_bit_cast(uint64_t, v); }
QS_HD double from_bits(uint64_t b) { return __builtin_bit_cast(double, b); }

// E with s in [2^E, 2^(E+1)); -1022 for s < 2^-1021 (zero, subnormals and
// the first normal binade share the grid 2^-1074)
QS_HD int binade(double s) {
    const int ef = int((bits_of(s) >> 52) & 0x7FF);
    return (ef < 1 ? 1 : ef) - 1023;
}

// m with s = m * 2^(binade(s) - 52)
QS_HD uint64_t mant(double s) {
    const uint64_t b = bits_of(s);
    return (b & kFrac) | (((b >> 52) & 0x7FF) ? kHidden : 0);
}

// m * 2^(E - 52) for m < 2^53 (m >= 2^52 unless E == -1022)
QS_HD double from_mant(uint64_t m, int E) {
    return from_bits(m < kHidden ? m : ((uint64_t(E + 1023) << 52) | (m & kFrac)));
}

QS_HD uint64_t sat(uint64_t v) { return v < kTop ? v : kTop; }

// the map of one element x >= 0 in binade E: x / u = mx * 2^d exactly
QS_HD Tr elem(double x, int E) {
    const uint64_t b = bits_of(x);
    const int ef = int((b >> 52) & 0x7FF);
    const uint64_t mx = (b & kFrac) | (ef ? kHidden : 0);
    const int d = (ef < 1 ? 1 : ef) - 1023 - E;
    if (d > 0) return Tr{kTop, kTop};  // x >= 2^(E+1): leaves the binade
    if (d == 0) return Tr{mx, mx};     // x on the grid of u
    const int sh = -d;
    if (sh >= 54) return Tr{0, 0};     // x < u / 2
    const uint64_t K = mx >> sh;
    const uint64_t r = mx & ((uint64_t{1} << sh) - 1);
    const uint64_t half = uint64_t{1} << (sh - 1);
    if (r < half) return Tr{K, K};
    if (r > half) return Tr{K + 1, K + 1};
    // a tie: m + inc must be even
    return Tr{K + (K & 1), K + 1 - (K & 1)};
}

QS_HD Tr identity() { return Tr{0, 0}; }

// a, then b
QS_HD Tr compose(Tr a, Tr b) {
    return Tr{sat(a.i0 + ((a.i0 & 1) ? b.i1 : b.i0)),
              sat(a.i1 + ((a.i1 & 1) ? b.i0 : b.i1))};
}

// m after the run of map t (>= kTop: the run leaves the binade)
QS_HD uint64_t apply(uint64_t m, Tr t) { return m + ((m & 1) ? t.i1 : t.i0); }

}  // namespace qs_exact

#endif  // QPP_AMD_QS_EXACT_CORE_HPP_

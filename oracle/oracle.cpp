// oracle.cpp -- CPU restatement of qpp 7.0.3's gate-application hot path.
//
// TEST INFRASTRUCTURE ONLY: the checker for the GPU engine (see oracle.h).
// Every function cites the qpp file:line it restates; paths are relative to
// the qpp source tree.  qpp is header-only on top of Eigen (absent here), so
// this file uses std::complex<double> / std::vector with Eigen's column-major
// storage and keeps qpp's loop structure and OpenMP placement for the hot
// kernels, which also makes it the CPU baseline ("qpp 7.0.3 qubit kernels,
// restated").
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <numeric>
#include <optional>
#include <random>
#include <set>
#include <stdexcept>
#include <vector>

#include "../include/qpp_amd.h"

namespace {

using cplx = std::complex<double>;
using idx = std::uint64_t;
using cvec = std::vector<cplx>;

// ---------------------------------------------------------------------------
// small dense column-major matrix
// ---------------------------------------------------------------------------
struct Mat {
    idx r = 0, c = 0;
    cvec a;
    Mat() = default;
    Mat(idx rows, idx cols) : r(rows), c(cols), a(rows * cols, cplx(0, 0)) {}
    cplx& operator()(idx i, idx j) { return a[i + j * r]; }
    const cplx& operator()(idx i, idx j) const { return a[i + j * r]; }
};

Mat from_raw(const double* p, idx dim) {
    Mat m(dim, dim);
    std::memcpy(static_cast<void*>(m.a.data()), p, sizeof(cplx) * dim * dim);
    return m;
}

Mat matmul(const Mat& A, const Mat& B) {
    Mat C(A.r, B.c);
    for (idx j = 0; j < B.c; ++j)
        for (idx k = 0; k < A.c; ++k) {
            const cplx b = B(k, j);
            for (idx i = 0; i < A.r; ++i) C(i, j) += A(i, k) * b;
        }
    return C;
}

Mat adjoint(const Mat& A) {
    Mat R(A.c, A.r);
    for (idx i = 0; i < A.r; ++i)
        for (idx j = 0; j < A.c; ++j) R(j, i) = std::conj(A(i, j));
    return R;
}

Mat identity(idx d) {
    Mat I(d, d);
    for (idx i = 0; i < d; ++i) I(i, i) = 1;
    return I;
}

// powm(A, k) (functions.hpp powm): repeated squaring is not needed for the
// small exponents (< d) used by applyCTRL; plain products suffice.
Mat powm(const Mat& A, idx k) {
    Mat R = identity(A.r);
    for (idx i = 0; i < k; ++i) R = matmul(R, A);
    return R;
}

// ---------------------------------------------------------------------------
// index helpers: internal/util.hpp:58-116 (big-endian multi-index maps),
// functions.hpp:1953 complement, util.hpp:498-507 all_qubits
// ---------------------------------------------------------------------------
void n2multiidx(idx n, idx numdims, const idx* dims, idx* result) {
    for (idx i = numdims; i-- > 0;) {
        result[i] = n % dims[i];
        n /= dims[i];
    }
}

idx multiidx2n(const idx* midx, idx numdims, const idx* dims) {
    idx part = 1, result = 0;
    for (idx i = numdims; i-- > 0;) {
        result += midx[i] * part;
        part *= dims[i];
    }
    return result;
}

std::vector<idx> complement(const std::vector<idx>& subsys, idx n) {
    std::vector<idx> all(n), result;
    std::iota(all.begin(), all.end(), idx{0});
    std::vector<idx> s = subsys;
    std::sort(s.begin(), s.end());
    std::set_difference(all.begin(), all.end(), s.begin(), s.end(),
                        std::back_inserter(result));
    return result;
}

bool all_qubits(const std::vector<idx>& dims) {
    return std::all_of(dims.begin(), dims.end(), [](idx d) { return d == 2; });
}

idx prod(const std::vector<idx>& dims) {
    idx p = 1;
    for (idx d : dims) p *= d;
    return p;
}

// ---------------------------------------------------------------------------
// Qubit kernels (QPP_QUBIT_OPTIMIZATIONS), internal/kernels/qubit/
// ---------------------------------------------------------------------------

// apply_psi_1q, apply.hpp:57-126.  Deep copy, then pairs (k0, k0 + 2^j),
// j = n-1-i, OpenMP over the outer L loop only.
cvec apply_psi_1q(const cvec& state, const Mat& A, idx i, idx n) {
    const idx D = idx{1} << n;
    cvec result = state;
    const idx j = n - 1 - i;
    const idx step = idx{1} << j;
    const idx jump = idx{1} << (j + 1);
    const cplx a00 = A(0, 0), a01 = A(0, 1), a10 = A(1, 0), a11 = A(1, 1);
#pragma omp parallel for
    for (idx L = 0; L < D; L += jump) {
        for (idx R = 0; R < step; ++R) {
            const idx k0 = L + R, k1 = k0 + step;
            const cplx p0 = state[k0], p1 = state[k1];
            result[k0] = (a00 * p0) + (a01 * p1);
            result[k1] = (a10 * p0) + (a11 * p1);
        }
    }
    return result;
}

// inner index table of apply_psi_kq (apply.hpp:524-537): bit j of r selects
// target[k-1-j], i.e. target[0] is the most significant gate digit.
std::vector<idx> inner_table(const std::vector<idx>& target, idx n) {
    const idx k = target.size(), dim = idx{1} << k;
    std::vector<idx> inner(dim);
    for (idx r = 0; r < dim; ++r) {
        idx v = 0;
        for (idx j = 0; j < k; ++j)
            if ((r >> j) & 1) v += idx{1} << (n - 1 - target[k - 1 - j]);
        inner[r] = v;
    }
    return inner;
}

// outer (spectator) table of apply_psi_kq (apply.hpp:539-561): bit j of m
// selects the j-th spectator qubit in ascending qubit order.
std::vector<idx> outer_table(const std::vector<idx>& target, idx n) {
    std::vector<bool> acted(n, false);
    for (idx q : target) acted[q] = true;
    std::vector<idx> spect;
    for (idx i = 0; i < n; ++i)
        if (!acted[i]) spect.push_back(i);
    const idx outer_dim = idx{1} << (n - target.size());
    std::vector<idx> outer(outer_dim);
    for (idx m = 0; m < outer_dim; ++m) {
        idx v = 0;
        for (idx j = 0; j < spect.size(); ++j)
            if ((m >> j) & 1) v += idx{1} << (n - 1 - spect[j]);
        outer[m] = v;
    }
    return outer;
}

// apply_psi_kq, apply.hpp:474-594 (apply_psi_2q / 3q, apply.hpp:139-462,
// compute the same map with unrolled loops).
cvec apply_psi_kq(const cvec& state, const Mat& A,
                  const std::vector<idx>& target, idx n) {
    const idx dim = idx{1} << target.size();
    cvec result = state;
    const auto inner = inner_table(target, n);
    const auto outer = outer_table(target, n);
    const idx outer_dim = outer.size();
#pragma omp parallel for
    for (idx m = 0; m < outer_dim; ++m) {
        const idx base = outer[m];
        cvec in(dim);
        for (idx c = 0; c < dim; ++c) in[c] = state[base + inner[c]];
        for (idx r = 0; r < dim; ++r) {
            cplx s = 0;
            for (idx c = 0; c < dim; ++c) s += A(r, c) * in[c];
            result[base + inner[r]] = s;
        }
    }
    return result;
}

// apply_rho_kq, apply.hpp:1002-1167 (1q/2q/3q variants, apply.hpp:606-990,
// compute the same block map U * B * U^dagger).
cvec apply_rho_kq(const cvec& rho, const Mat& A, const std::vector<idx>& target,
                  idx n) {
    const idx D = idx{1} << n, dim = idx{1} << target.size();
    cvec result(D * D);
    const auto inner = inner_table(target, n);
    const auto outer = outer_table(target, n);
    const Mat Ad = adjoint(A);
    const idx nb = outer.size();
#pragma omp parallel for
    for (idx r = 0; r < nb; ++r) {
        Mat B(dim, dim);
        for (idx c = 0; c < nb; ++c) {
            for (idx a = 0; a < dim; ++a)
                for (idx b = 0; b < dim; ++b)
                    B(a, b) = rho[(outer[r] + inner[a]) +
                                  (outer[c] + inner[b]) * D];
            const Mat R = matmul(matmul(A, B), Ad);
            for (idx a = 0; a < dim; ++a)
                for (idx b = 0; b < dim; ++b)
                    result[(outer[r] + inner[a]) + (outer[c] + inner[b]) * D] =
                        R(a, b);
        }
    }
    return result;
}

// Control masks of apply_ctrl_psi_1q, apply_ctrl.hpp:110-131: shift 0 means
// "control on |1>", shift 1 means "control on |0>".
struct CtrlMask {
    idx ones = 0, zeros = 0;
    bool met(idx k) const { return (k & ones) == ones && (k & zeros) == 0; }
};

CtrlMask ctrl_mask(const std::vector<idx>& ctrl, const std::vector<idx>& shift,
                   idx n) {
    CtrlMask m;
    for (idx c = 0; c < ctrl.size(); ++c) {
        const idx bit = idx{1} << (n - 1 - ctrl[c]);
        if (shift[c] == 0)
            m.ones |= bit;
        else
            m.zeros |= bit;
    }
    return m;
}

// apply_ctrl_psi_1q, apply_ctrl.hpp:63-177: the 1q kernel with a control test
// on k0 (control bits are never the target bit).
cvec apply_ctrl_psi_1q(const cvec& state, const Mat& A,
                       const std::vector<idx>& ctrl, idx i,
                       const std::vector<idx>& shift, idx n) {
    const idx D = idx{1} << n;
    cvec result = state;
    const idx j = n - 1 - i;
    const idx step = idx{1} << j, jump = idx{1} << (j + 1);
    const CtrlMask cm = ctrl_mask(ctrl, shift, n);
    const cplx a00 = A(0, 0), a01 = A(0, 1), a10 = A(1, 0), a11 = A(1, 1);
#pragma omp parallel for
    for (idx L = 0; L < D; L += jump) {
        for (idx R = 0; R < step; ++R) {
            const idx k0 = L + R, k1 = k0 + step;
            if (cm.met(k0)) {
                const cplx p0 = state[k0], p1 = state[k1];
                result[k0] = (a00 * p0) + (a01 * p1);
                result[k1] = (a10 * p0) + (a11 * p1);
            }
        }
    }
    return result;
}

// apply_ctrl_psi_kq, apply_ctrl.hpp:329-509 (2q variant :196-311 computes the
// same map): blocks whose spectator bits satisfy the controls get A.
cvec apply_ctrl_psi_kq(const cvec& state, const Mat& A,
                       const std::vector<idx>& ctrl,
                       const std::vector<idx>& target,
                       const std::vector<idx>& shift, idx n) {
    const idx dim = idx{1} << target.size();
    cvec result = state;
    const auto inner = inner_table(target, n);
    const auto outer = outer_table(target, n);
    const CtrlMask cm = ctrl_mask(ctrl, shift, n);
    const idx nb = outer.size();
#pragma omp parallel for
    for (idx m = 0; m < nb; ++m) {
        const idx base = outer[m];
        if (!cm.met(base)) continue;
        cvec in(dim);
        for (idx c = 0; c < dim; ++c) in[c] = state[base + inner[c]];
        for (idx r = 0; r < dim; ++r) {
            cplx s = 0;
            for (idx c = 0; c < dim; ++c) s += A(r, c) * in[c];
            result[base + inner[r]] = s;
        }
    }
    return result;
}

// apply_ctrl_rho_kq, apply_ctrl.hpp:894-1124 (1q :527-672, 2q :691-876):
// both controls met -> U B U^dagger, row only -> U B, column only ->
// B U^dagger, neither -> unchanged.
cvec apply_ctrl_rho_kq(const cvec& rho, const Mat& A,
                       const std::vector<idx>& ctrl,
                       const std::vector<idx>& target,
                       const std::vector<idx>& shift, idx n) {
    const idx D = idx{1} << n, dim = idx{1} << target.size();
    cvec result = rho;
    const auto inner = inner_table(target, n);
    const auto outer = outer_table(target, n);
    const CtrlMask cm = ctrl_mask(ctrl, shift, n);
    const Mat Ad = adjoint(A);
    const idx nb = outer.size();
#pragma omp parallel for
    for (idx r = 0; r < nb; ++r) {
        const bool rc = cm.met(outer[r]);
        Mat B(dim, dim);
        for (idx c = 0; c < nb; ++c) {
            const bool cc = cm.met(outer[c]);
            if (!rc && !cc) continue;
            for (idx a = 0; a < dim; ++a)
                for (idx b = 0; b < dim; ++b)
                    B(a, b) = rho[(outer[r] + inner[a]) +
                                  (outer[c] + inner[b]) * D];
            Mat R;
            if (rc && cc)
                R = matmul(matmul(A, B), Ad);
            else if (rc)
                R = matmul(A, B);
            else
                R = matmul(B, Ad);
            for (idx a = 0; a < dim; ++a)
                for (idx b = 0; b < dim; ++b)
                    result[(outer[r] + inner[a]) + (outer[c] + inner[b]) * D] =
                        R(a, b);
        }
    }
    return result;
}

// apply_ctrl_fan_psi, apply_ctrl_fan.hpp:62-158: for each target in the given
// order, in place, a 2x2 A on pairs (i0, i0 | tbit) whose controls are met.
cvec apply_ctrl_fan_psi(const cvec& state, const Mat& A,
                        const std::vector<idx>& ctrl,
                        const std::vector<idx>& target,
                        const std::vector<idx>& shift, idx n) {
    const idx D = idx{1} << n;
    cvec result = state;
    if (D == 1) return result;
    const CtrlMask cm = ctrl_mask(ctrl, shift, n);
    const cplx a00 = A(0, 0), a01 = A(0, 1), a10 = A(1, 0), a11 = A(1, 1);
    for (idx t : target) {
        const idx tbit = idx{1} << (n - 1 - t);
#pragma omp parallel for
        for (idx i0 = 0; i0 < D; ++i0) {
            if ((i0 & tbit) || !cm.met(i0)) continue;
            const idx i1 = i0 | tbit;
            const cplx p0 = result[i0], p1 = result[i1];
            result[i0] = a00 * p0 + a01 * p1;
            result[i1] = a10 * p0 + a11 * p1;
        }
    }
    return result;
}

// apply_ctrl_fan_rho, apply_ctrl_fan.hpp:174-303: per target, a row pass with
// A then a column pass with conj(A), each under the controls of its index.
cvec apply_ctrl_fan_rho(const cvec& rho, const Mat& A,
                        const std::vector<idx>& ctrl,
                        const std::vector<idx>& target,
                        const std::vector<idx>& shift, idx n) {
    const idx D = idx{1} << n;
    cvec result = rho;
    if (D == 1) return result;
    const CtrlMask cm = ctrl_mask(ctrl, shift, n);
    const cplx a00 = A(0, 0), a01 = A(0, 1), a10 = A(1, 0), a11 = A(1, 1);
    for (idx t : target) {
        const idx tbit = idx{1} << (n - 1 - t);
        // rows: rho <- A rho on the row index
#pragma omp parallel for
        for (idx col = 0; col < D; ++col)
            for (idx i0 = 0; i0 < D; ++i0) {
                if ((i0 & tbit) || !cm.met(i0)) continue;
                const idx i1 = i0 | tbit;
                const cplx p0 = result[i0 + col * D], p1 = result[i1 + col * D];
                result[i0 + col * D] = a00 * p0 + a01 * p1;
                result[i1 + col * D] = a10 * p0 + a11 * p1;
            }
        // columns: rho <- rho A^dagger, i.e. conj(A) on the column index
#pragma omp parallel for
        for (idx row = 0; row < D; ++row)
            for (idx j0 = 0; j0 < D; ++j0) {
                if ((j0 & tbit) || !cm.met(j0)) continue;
                const idx j1 = j0 | tbit;
                const cplx p0 = result[row + j0 * D], p1 = result[row + j1 * D];
                result[row + j0 * D] = std::conj(a00) * p0 + std::conj(a01) * p1;
                result[row + j1 * D] = std::conj(a10) * p0 + std::conj(a11) * p1;
            }
    }
    return result;
}

// ---------------------------------------------------------------------------
// Generic qudit path (operations.hpp), used when some dimension is not 2
// ---------------------------------------------------------------------------

// Gathers the amplitudes of one "gate block": the complement digits are
// fixed by r, the target digits run over m (operations.hpp:211-247).
struct BlockMap {
    std::vector<idx> dims, target, bar, dA, dbar;
    idx n = 0, DA = 1, Dbar = 1;
    BlockMap(const std::vector<idx>& d, const std::vector<idx>& t)
        : dims(d), target(t), n(d.size()) {
        bar = complement(t, n);
        for (idx q : t) dA.push_back(d[q]), DA *= d[q];
        for (idx q : bar) dbar.push_back(d[q]), Dbar *= d[q];
    }
    idx index(idx m, idx r) const {
        std::vector<idx> mid(n), ma(target.size()), mb(bar.size());
        if (!bar.empty()) n2multiidx(r, bar.size(), dbar.data(), mb.data());
        n2multiidx(m, target.size(), dA.data(), ma.data());
        for (idx k = 0; k < bar.size(); ++k) mid[bar[k]] = mb[k];
        for (idx k = 0; k < target.size(); ++k) mid[target[k]] = ma[k];
        return multiidx2n(mid.data(), n, dims.data());
    }
};

// apply generic ket / density matrix, operations.hpp:182-371.
cvec apply_generic(const cvec& state, bool is_rho, const Mat& A,
                   const std::vector<idx>& target,
                   const std::vector<idx>& dims) {
    const idx D = prod(dims);
    if (D == 1) return state;
    BlockMap bm(dims, target);
    cvec result = state;
    if (!is_rho) {
#pragma omp parallel for
        for (idx r = 0; r < bm.Dbar; ++r) {
            for (idx m = 0; m < bm.DA; ++m) {
                cplx coeff = 0;
                for (idx k = 0; k < bm.DA; ++k)
                    coeff += A(m, k) * state[bm.index(k, r)];
                result[bm.index(m, r)] = coeff;
            }
        }
        return result;
    }
    const Mat Ad = adjoint(A);
#pragma omp parallel for
    for (idx r1 = 0; r1 < bm.Dbar; ++r1)
        for (idx r2 = 0; r2 < bm.Dbar; ++r2)
            for (idx m1 = 0; m1 < bm.DA; ++m1)
                for (idx m2 = 0; m2 < bm.DA; ++m2) {
                    cplx coeff = 0;
                    for (idx n1 = 0; n1 < bm.DA; ++n1) {
                        const idx row = bm.index(n1, r1);
                        for (idx n2 = 0; n2 < bm.DA; ++n2)
                            coeff += A(m1, n1) * state[row + bm.index(n2, r2) * D] *
                                     Ad(n2, m2);
                    }
                    result[bm.index(m1, r1) + bm.index(m2, r2) * D] = coeff;
                }
    return result;
}

// Generic controlled gate, operations.hpp:777-886, written as the operator it
// builds: on every basis component whose (control digit + shift) mod d is the
// same value r for all controls, A^r acts on the target digits; other
// components are untouched.  `shift_user` is the caller's shift (the
// inversion at operations.hpp:727-735 and the projection at :796-821 cancel).
cvec ctrl_generic(const cvec& state, bool is_rho, const Mat& A,
                  const std::vector<idx>& ctrl, const std::vector<idx>& target,
                  const std::vector<idx>& dims,
                  const std::vector<idx>& shift_user, bool fan) {
    const idx D = prod(dims);
    if (D == 1) return state;
    const idx d = dims[ctrl[0]];
    std::vector<Mat> Ak;
    for (idx k = 0; k < d; ++k) Ak.push_back(powm(A, k));
    const idx n = dims.size();
    // exponent r(i) for basis index i, or d if the controls disagree
    auto rexp = [&](idx i) -> idx {
        std::vector<idx> mid(n);
        n2multiidx(i, n, dims.data(), mid.data());
        idx r = (mid[ctrl[0]] + shift_user[0]) % d;
        for (idx c = 1; c < ctrl.size(); ++c)
            if ((mid[ctrl[c]] + shift_user[c]) % d != r) return d;
        return r;
    };
    // one pass of "A^r on `tg`" applied to the row index of a D x ncols array
    auto pass = [&](cvec& v, idx ncols, const std::vector<idx>& tg, bool conj,
                    bool on_cols) {
        BlockMap bm(dims, tg);
        cvec out = v;
        for (idx col = 0; col < ncols; ++col)
            for (idx r = 0; r < bm.Dbar; ++r) {
                const idx e = rexp(bm.index(0, r));
                if (e == d || e == 0) continue;
                const Mat& M = Ak[e];
                for (idx m = 0; m < bm.DA; ++m) {
                    cplx s = 0;
                    for (idx k = 0; k < bm.DA; ++k) {
                        const idx src = on_cols ? col + bm.index(k, r) * D
                                                : bm.index(k, r) + col * D;
                        s += (conj ? std::conj(M(m, k)) : M(m, k)) * v[src];
                    }
                    const idx dst = on_cols ? col + bm.index(m, r) * D
                                            : bm.index(m, r) + col * D;
                    out[dst] = s;
                }
            }
        v.swap(out);
    };
    cvec result = state;
    const idx ncols = is_rho ? D : 1;
    std::vector<std::vector<idx>> groups;
    if (fan)
        for (idx t : target) groups.push_back({t});
    else
        groups.push_back(target);
    for (const auto& g : groups) {
        pass(result, ncols, g, false, false);
        if (is_rho) pass(result, D, g, true, true);
    }
    return result;
}

// ---------------------------------------------------------------------------
// ptrace / ptranspose / syspermute
// ---------------------------------------------------------------------------

// ptrace_psi_kq / ptrace_rho_kq, ptrace.hpp:61-257, and the generic path
// operations.hpp:1970-2114 (same sums; the sum over `a` runs in ascending
// order with target[0] the most significant summation digit).
cvec ptrace_any(const cvec& A, bool is_rho, const std::vector<idx>& dims,
                const std::vector<idx>& target) {
    const idx n = dims.size(), D = prod(dims);
    if (target.size() == n) {
        cplx s = 0;
        if (is_rho)
            for (idx i = 0; i < D; ++i) s += A[i + i * D];
        else
            for (idx i = 0; i < D; ++i) s += std::conj(A[i]) * A[i];
        return {s};
    }
    if (target.empty()) {
        if (is_rho) return A;
        cvec r(D * D);
        for (idx j = 0; j < D; ++j)
            for (idx i = 0; i < D; ++i) r[i + j * D] = A[i] * std::conj(A[j]);
        return r;
    }
    const std::vector<idx> bar = complement(target, n);
    std::vector<idx> dsub, dbar;
    idx Dsub = 1, Dbar = 1;
    for (idx q : target) dsub.push_back(dims[q]), Dsub *= dims[q];
    for (idx q : bar) dbar.push_back(dims[q]), Dbar *= dims[q];
    // strides of each subsystem in the full index
    std::vector<idx> stride(n);
    {
        idx s = 1;
        for (idx i = n; i-- > 0;) stride[i] = s, s *= dims[i];
    }
    auto expand = [&](idx c, const std::vector<idx>& subs,
                      const std::vector<idx>& sd) {
        std::vector<idx> mid(subs.size());
        n2multiidx(c, subs.size(), sd.data(), mid.data());
        idx v = 0;
        for (idx k = 0; k < subs.size(); ++k) v += mid[k] * stride[subs[k]];
        return v;
    };
    std::vector<idx> off(Dsub);
    for (idx a = 0; a < Dsub; ++a) off[a] = expand(a, target, dsub);
    cvec result(Dbar * Dbar);
    for (idx j = 0; j < Dbar; ++j) {
        const idx cb = expand(j, bar, dbar);
#pragma omp parallel for
        for (idx i = 0; i < Dbar; ++i) {
            const idx rb = expand(i, bar, dbar);
            cplx s = 0;
            for (idx a = 0; a < Dsub; ++a) {
                if (is_rho)
                    s += A[(rb + off[a]) + (cb + off[a]) * D];
                else
                    s += A[rb + off[a]] * std::conj(A[cb + off[a]]);
            }
            result[i + j * Dbar] = s;
        }
    }
    return result;
}

// Generic ptranspose, operations.hpp:2200-2331: result(i, j) takes the input
// at the row/column multi-indices with the target digits swapped.  For a ket
// the input is |psi><psi|.
cvec ptranspose_generic(const cvec& A, bool is_rho,
                        const std::vector<idx>& dims,
                        const std::vector<idx>& target) {
    const idx n = dims.size(), D = prod(dims);
    auto elem = [&](idx r, idx c) -> cplx {
        return is_rho ? A[r + c * D] : A[r] * std::conj(A[c]);
    };
    cvec result(D * D);
    if (target.size() == n) {
        for (idx j = 0; j < D; ++j)
            for (idx i = 0; i < D; ++i) result[i + j * D] = elem(j, i);
        return result;
    }
    if (target.empty()) {
        for (idx j = 0; j < D; ++j)
            for (idx i = 0; i < D; ++i) result[i + j * D] = elem(i, j);
        return result;
    }
#pragma omp parallel for
    for (idx j = 0; j < D; ++j) {
        std::vector<idx> mc(n), mr(n), mct(n);
        n2multiidx(j, n, dims.data(), mc.data());
        for (idx i = 0; i < D; ++i) {
            n2multiidx(i, n, dims.data(), mr.data());
            mct = mc;
            for (idx t : target) std::swap(mct[t], mr[t]);
            result[i + j * D] = elem(multiidx2n(mr.data(), n, dims.data()),
                                     multiidx2n(mct.data(), n, dims.data()));
        }
    }
    return result;
}

// Qubit-kernel ptranspose, ptranspose.hpp:59-219: mask bit k (LSB numbering)
// per target k; result(i, j) = A(j_T, i_T) with i_T = (i & ~M) | (j & M).
// Trivial cases (empty / full target) are handled first as the kernels do.
cvec ptranspose_qubit(const cvec& A, bool is_rho, idx n,
                      const std::vector<idx>& target) {
    const idx D = idx{1} << n;
    auto elem = [&](idx r, idx c) -> cplx {
        return is_rho ? A[r + c * D] : A[r] * std::conj(A[c]);
    };
    cvec result(D * D);
    if (target.size() == n) {
        for (idx j = 0; j < D; ++j)
            for (idx i = 0; i < D; ++i) result[i + j * D] = elem(j, i);
        return result;
    }
    if (target.empty()) {
        for (idx j = 0; j < D; ++j)
            for (idx i = 0; i < D; ++i) result[i + j * D] = elem(i, j);
        return result;
    }
    idx M = 0;
    for (idx k : target) M |= idx{1} << k;
#pragma omp parallel for
    for (idx i = 0; i < D; ++i)
        for (idx j = 0; j < D; ++j) {
            const idx iT = (i & ~M) | (j & M);
            const idx jT = (j & ~M) | (i & M);
            // ket: conj(psi(i_T)) * psi(j_T) == elem(j_T, i_T)
            result[i + j * D] = elem(jT, iT);
        }
    return result;
}

// Generic syspermute, operations.hpp:2448-2508: the element at input
// multi-index midx goes to the output multi-index whose digit k is
// midx[perm[k]]; a density matrix permutes row and column digits alike.
cvec syspermute_generic(const cvec& A, bool is_rho,
                        const std::vector<idx>& dims,
                        const std::vector<idx>& perm) {
    const idx n = dims.size(), D = prod(dims);
    const idx mult = is_rho ? 2 : 1, total = n * mult;
    std::vector<idx> Cdims(total), Cperm(total);
    for (idx i = 0; i < n; ++i) {
        Cdims[i] = dims[i];
        Cperm[i] = perm[i];
        if (is_rho) {
            Cdims[i + n] = dims[i];
            Cperm[i + n] = perm[i] + n;
        }
    }
    const idx N = is_rho ? D * D : D;
    cvec result(N);
#pragma omp parallel for
    for (idx i = 0; i < N; ++i) {
        std::vector<idx> midx(total), mt(total), pd(total);
        n2multiidx(i, total, Cdims.data(), midx.data());
        for (idx k = 0; k < total; ++k) {
            pd[k] = Cdims[Cperm[k]];
            mt[k] = midx[Cperm[k]];
        }
        result[multiidx2n(mt.data(), total, pd.data())] = A[i];
    }
    return result;
}

// Qubit-kernel syspermute, syspermute.hpp:58-180: with inv_perm[perm[k]] = k,
// new LSB-bit k <- old LSB-bit inv_perm[k]; rows and columns alike.
cvec syspermute_qubit(const cvec& A, bool is_rho, idx n,
                      const std::vector<idx>& perm) {
    const idx D = idx{1} << n;
    std::vector<idx> inv(n);
    for (idx k = 0; k < n; ++k) inv[perm[k]] = k;
    auto pb = [&](idx old) {
        idx v = 0;
        for (idx k = 0; k < n; ++k) v |= ((old >> inv[k]) & 1) << k;
        return v;
    };
    if (!is_rho) {
        cvec result(D);
        for (idx i = 0; i < D; ++i) result[pb(i)] = A[i];
        return result;
    }
    cvec result(D * D);
#pragma omp parallel for
    for (idx i = 0; i < D * D; ++i) {
        const idx ri = i / D, ci = i % D;
        result[pb(ri) * D + pb(ci)] = A[i];
    }
    return result;
}

// ---------------------------------------------------------------------------
// sampling: libstdc++ discrete_distribution (random.tcc:2656-2712)
// ---------------------------------------------------------------------------
std::vector<double> discrete_cp(const std::vector<double>& p) {
    // param_type::_M_initialize: sum with std::accumulate (sequential),
    // __normalize (element / sum), std::partial_sum, last := 1.0
    if (p.size() < 2) return {};
    double sum = std::accumulate(p.begin(), p.end(), 0.0);
    std::vector<double> q(p.size());
    for (std::size_t i = 0; i < p.size(); ++i) q[i] = p[i] / sum;
    std::vector<double> cp(p.size());
    std::partial_sum(q.begin(), q.end(), cp.begin());
    cp.back() = 1.0;
    return cp;
}

idx discrete_draw(const std::vector<double>& cp, double u) {
    // operator(): an empty table (fewer than 2 weights) returns 0 without
    // drawing; otherwise lower_bound(cp, u)
    if (cp.empty()) return 0;
    return std::lower_bound(cp.begin(), cp.end(), u) - cp.begin();
}

std::vector<double> probs_of(const cvec& A, bool is_rho, idx D) {
    std::vector<double> p(D);
    if (is_rho)
        for (idx i = 0; i < D; ++i) p[i] = A[i + i * D].real();
    else
        for (idx i = 0; i < D; ++i) p[i] = std::norm(A[i]);
    return p;
}

// sample(A, target, dims), instruments.hpp:746-809: returns the sampled full
// index, the target digits and pbs[index].
struct SampleResult {
    idx index;
    std::vector<idx> digits;
    double prob;
    bool drew;
};

SampleResult sample_one(const cvec& A, bool is_rho,
                        const std::vector<idx>& dims,
                        const std::vector<idx>& target, double u) {
    const idx D = prod(dims);
    const auto p = probs_of(A, is_rho, D);
    const auto cp = discrete_cp(p);
    SampleResult s;
    s.drew = !cp.empty();
    s.index = discrete_draw(cp, u);
    std::vector<idx> mid(dims.size());
    n2multiidx(s.index, dims.size(), dims.data(), mid.data());
    for (idx t : target) s.digits.push_back(mid[t]);
    s.prob = p[s.index];
    return s;
}

// ---------------------------------------------------------------------------
// measure_seq, instruments.hpp:985-1099 (with ip, :61-183, for the
// destructive ket contraction)
// ---------------------------------------------------------------------------

// <m| on subsystem i of a ket: keeps the components whose digit i equals m
// and removes subsystem i (ip(mket({m}), psi, {i}, dims)).
cvec contract_ket(const cvec& psi, const std::vector<idx>& dims, idx i, idx m) {
    std::vector<idx> nd = dims;
    nd.erase(nd.begin() + static_cast<std::ptrdiff_t>(i));
    const idx Dn = prod(nd), n = dims.size();
    cvec out(Dn);
    std::vector<idx> mid(n), mo(n - 1);
    for (idx k = 0; k < Dn; ++k) {
        if (n > 1) n2multiidx(k, n - 1, nd.data(), mo.data());
        for (idx a = 0, b = 0; a < n; ++a) mid[a] = (a == i) ? m : mo[b++];
        out[k] = psi[multiidx2n(mid.data(), n, dims.data())];
    }
    return out;
}

Mat projector(idx m, idx d) {
    Mat P(d, d);
    P(m, m) = 1;
    return P;
}

double sqnorm(const cvec& v) {
    double s = 0;
    for (const auto& z : v) s += std::norm(z);
    return s;
}

cplx trace(const cvec& rho, idx D) {
    cplx s = 0;
    for (idx i = 0; i < D; ++i) s += rho[i + i * D];
    return s;
}

cvec apply_any(const cvec& st, bool is_rho, const Mat& A,
               const std::vector<idx>& target, const std::vector<idx>& dims);

struct MeasureResult {
    std::vector<idx> results;
    std::vector<double> probs;
    cvec state;
    std::vector<idx> dims;
    bool drew;
};

MeasureResult measure_seq_impl(cvec A, bool is_rho, std::vector<idx> dims,
                               std::vector<idx> target, bool destructive,
                               double u) {
    MeasureResult mr;
    const SampleResult s = sample_one(A, is_rho, dims, target, u);
    mr.drew = s.drew;
    const idx tsz = target.size();
    std::vector<double> ps(tsz);
    std::vector<idx> order(tsz);
    std::iota(order.begin(), order.end(), idx{0});
    if (destructive)
        std::sort(order.begin(), order.end(),
                  [&](idx a, idx b) { return target[a] > target[b]; });
    if (is_rho && dims.size() <= 10) {
        // instruments.hpp:1039-1054: small density matrices, targets in the
        // given order, then one ptrace
        for (idx cnt = 0; cnt < tsz; ++cnt) {
            const idx i = target[cnt];
            A = apply_any(A, true, projector(s.digits[cnt], dims[i]), {i}, dims);
            const double p = trace(A, prod(dims)).real();
            ps[cnt] = p;
            for (auto& z : A) z /= p;
        }
        if (destructive) {
            A = ptrace_any(A, true, dims, target);
            std::vector<idx> nd;
            for (idx q : complement(target, dims.size())) nd.push_back(dims[q]);
            dims = nd;
        }
        mr.results = s.digits;
        mr.probs = ps;
        mr.state = A;
        mr.dims = dims;
        return mr;
    }
    // instruments.hpp:1058-1096: kets and larger density matrices
    for (idx cnt = 0; cnt < tsz; ++cnt) {
        const idx ip = destructive ? order[cnt] : cnt;
        const idx i = target[ip];
        const idx res = s.digits[ip];
        if (!is_rho && destructive) {
            A = contract_ket(A, dims, i, res);
            const double p = sqnorm(A);
            ps[ip] = p;
            const double sq = std::sqrt(p);
            for (auto& z : A) z /= sq;
        } else {
            A = apply_any(A, is_rho, projector(res, dims[i]), {i}, dims);
            const double p = is_rho ? trace(A, prod(dims)).real() : sqnorm(A);
            ps[ip] = p;
            const double div = is_rho ? p : std::sqrt(p);
            for (auto& z : A) z /= div;
            if (destructive) A = ptrace_any(A, true, dims, {i});
        }
        if (destructive) dims.erase(dims.begin() + static_cast<std::ptrdiff_t>(i));
    }
    mr.results = s.digits;
    mr.probs = ps;
    mr.state = A;
    mr.dims = dims;
    return mr;
}

cvec apply_any(const cvec& st, bool is_rho, const Mat& A,
               const std::vector<idx>& target, const std::vector<idx>& dims) {
    if (all_qubits(dims)) {
        const idx n = dims.size();
        if (!is_rho)
            return target.size() == 1 ? apply_psi_1q(st, A, target[0], n)
                                      : apply_psi_kq(st, A, target, n);
        return apply_rho_kq(st, A, target, n);
    }
    return apply_generic(st, is_rho, A, target, dims);
}

cvec apply_ctrl_any(const cvec& st, bool is_rho, const Mat& A,
                    const std::vector<idx>& ctrl,
                    const std::vector<idx>& target,
                    const std::vector<idx>& dims,
                    const std::vector<idx>& shift_user) {
    if (all_qubits(dims)) {
        // shift inversion for d = 2 is the identity (operations.hpp:727-735)
        const idx n = dims.size();
        if (!is_rho)
            return target.size() == 1
                       ? apply_ctrl_psi_1q(st, A, ctrl, target[0], shift_user, n)
                       : apply_ctrl_psi_kq(st, A, ctrl, target, shift_user, n);
        return apply_ctrl_rho_kq(st, A, ctrl, target, shift_user, n);
    }
    return ctrl_generic(st, is_rho, A, ctrl, target, dims, shift_user, false);
}

cvec apply_ctrl_fan_any(const cvec& st, bool is_rho, const Mat& A,
                        const std::vector<idx>& ctrl,
                        const std::vector<idx>& target,
                        const std::vector<idx>& dims,
                        const std::vector<idx>& shift_user) {
    if (all_qubits(dims)) {
        const idx n = dims.size();
        return is_rho ? apply_ctrl_fan_rho(st, A, ctrl, target, shift_user, n)
                      : apply_ctrl_fan_psi(st, A, ctrl, target, shift_user, n);
    }
    return ctrl_generic(st, is_rho, A, ctrl, target, dims, shift_user, true);
}

cvec load(const double* p, idx n) {
    cvec v(n);
    std::memcpy(static_cast<void*>(v.data()), p, sizeof(cplx) * n);
    return v;
}

void store(const cvec& v, double* p) {
    std::memcpy(p, static_cast<const void*>(v.data()), sizeof(cplx) * v.size());
}

std::vector<idx> vec_of(const uint64_t* p, uint32_t n) {
    return std::vector<idx>(p, p + n);
}

// Xd(D)^k (gates.hpp:360-373): |j> -> |j+k mod D>.  qpp builds Xd(D > 2)
// numerically as Fd^-1 Zd Fd; this is its exact value (the two differ by
// rounding, ~1e-16 per entry).
Mat shift_pow(idx D, idx k) {
    Mat M(D, D);
    for (idx j = 0; j < D; ++j) M((j + k) % D, j) = 1;
    return M;
}

// reset, instruments.hpp:1148-1196: measure_seq(A, target, dims, false),
// then powm(Xd(dims[i]), dims[i] - result[i]) on target[i].  The
// correction's dimension is dims[i], the i-th subsystem's (as in qpp);
// callers pass dims where that equals dims[target[i]].
cvec reset_impl(const cvec& A, bool is_rho, const std::vector<idx>& dims,
                const std::vector<idx>& target, double u) {
    const MeasureResult mr = measure_seq_impl(A, is_rho, dims, target, false, u);
    cvec r = mr.state;
    for (idx i = 0; i < target.size(); ++i)
        r = apply_any(r, is_rho, shift_pow(dims[i], dims[i] - mr.results[i]),
                      {target[i]}, dims);
    return r;
}

// ---------------------------------------------------------------------------
// Engines
// ---------------------------------------------------------------------------
struct Step {
    uint32_t type, gate;
    std::vector<idx> ctrl, target;
    std::optional<std::vector<idx>> shift;
    idx c_reg;
    bool destructive;
    std::vector<idx> ps;  // POST_SELECT values; COND_IF dit values or truth table
    bool table = false;   // COND_IF / WHILE: ps is a truth table over the ctrl dits
    idx jump = 0;         // COND_IF / WHILE: its ELSE, else its END; COND_ELSE:
                          // its END; the END of a WHILE: the WHILE
    bool loop_end = false;
};

bool is_cctrl(const Step& s) {
    return s.type == QS_STEP_CCTRL || s.type == QS_STEP_CCTRL_FAN;
}

bool is_runtime(const Step& s) {
    return s.type == QS_STEP_COND_IF || s.type == QS_STEP_COND_ELSE ||
           s.type == QS_STEP_COND_END || s.type == QS_STEP_COND_WHILE;
}

// is_measurement, classes/qcircuit.hpp:7008-7025: the projective
// measurements and post-selections (reset and discard are not measurements)
bool is_measurement(const Step& s) {
    return s.type == QS_STEP_MEASURE || s.type == QS_STEP_MEASURE_MANY ||
           s.type == QS_STEP_POST_SELECT;
}

bool is_reset_or_discard(const Step& s) {
    return s.type == QS_STEP_RESET || s.type == QS_STEP_DISCARD;
}

// extract_ctrl_target_ps_vals_c_reg, classes/qcircuit.hpp:7371-7449: a
// measurement gives its targets and c_reg..c_reg+k-1, a gate step its
// controls and targets, and a reset or discard step nothing at all (it
// is neither a measurement nor a gate step there).
std::vector<idx> c_regs_of(const Step& s) {
    std::vector<idx> c;
    if (is_measurement(s)) {
        c.resize(s.target.size());
        std::iota(c.begin(), c.end(), s.c_reg);
    } else if (is_cctrl(s)) {
        c = s.ctrl;  // a cCTRL's control dits (qcircuit.hpp:7420-7427)
    }
    return c;
}

std::vector<idx> targets_of(const Step& s) {
    return is_reset_or_discard(s) ? std::vector<idx>{} : s.target;
}

// can_swap, classes/qcircuit.hpp:7465-7492: controls, targets and classical
// registers pairwise disjoint.
bool can_swap(const Step& a, const Step& b) {
    auto overlap = [](const std::vector<idx>& x, const std::vector<idx>& y) {
        return std::find_first_of(x.begin(), x.end(), y.begin(), y.end()) !=
               x.end();
    };
    const std::vector<idx> none;
    if (overlap(is_cctrl(a) ? none : a.ctrl, is_cctrl(b) ? none : b.ctrl)) return false;
    if (overlap(targets_of(a), targets_of(b))) return false;
    if (overlap(c_regs_of(a), c_regs_of(b))) return false;
    return true;
}

// canonical_form, classes/qcircuit.hpp:7556-7612 (no cCTRL / runtime steps
// exist in the program format): bubble every measurement after the first
// one towards the end while it can swap with a non-measurement.
std::vector<Step> canonical_form(const std::vector<Step>& steps) {
    // a circuit with runtime steps is not reordered (:7563-7567)
    if (std::any_of(steps.begin(), steps.end(), is_runtime)) return steps;
    auto first = std::find_if(steps.begin(), steps.end(), is_measurement);
    std::vector<Step> before(steps.begin(), first);
    std::vector<Step> rest(first, steps.end());
    // cCTRL steps to the end first (:7580-7595), then the measurements
    for (std::ptrdiff_t ri = static_cast<std::ptrdiff_t>(rest.size()) - 1;
         ri >= 0; --ri) {
        if (!is_cctrl(rest[ri])) continue;
        for (std::size_t after = ri + 1; after < rest.size(); ++after) {
            if (can_swap(rest[after - 1], rest[after]) && !is_cctrl(rest[after]))
                std::swap(rest[after - 1], rest[after]);
            else
                break;
        }
    }
    for (std::ptrdiff_t ri = static_cast<std::ptrdiff_t>(rest.size()) - 1;
         ri >= 0; --ri) {
        if (!is_measurement(rest[ri])) continue;
        for (std::size_t after = ri + 1; after < rest.size(); ++after) {
            Step& cur = rest[after - 1];
            Step& nxt = rest[after];
            if (can_swap(cur, nxt) && !is_measurement(nxt))
                std::swap(cur, nxt);
            else
                break;
        }
    }
    before.insert(before.end(), rest.begin(), rest.end());
    return before;
}

// an engine error, as the qs_status code of the matching qpp exception
struct EngineError {
    int code;
};

struct EngineState {
    cvec qstate;
    std::vector<idx> dims;  // dims of the non-measured qudits
    std::vector<double> probs;
    std::vector<idx> dits;
    std::vector<idx> subsys;  // relative position or SIZE_MAX if measured
    // internal/classes/qengine_state.hpp:66-72
    bool ps_ok = true;
    bool ensure_ps = false;
    idx max_ps = std::numeric_limits<idx>::max();
};

struct Engine {
    std::vector<Step> steps;
    std::vector<Mat> mats;
    idx nq, nc, d;
    bool is_rho;
    double noise_p;
    std::mt19937 gen;
    EngineState st;
    std::map<std::vector<idx>, idx> stats;
    // noise_results_ (qnoisy_engine.hpp:62, 104-121): per step, the Kraus
    // index drawn for each live qubit at each visit
    std::vector<std::vector<idx>> noise_results;
    idx cur_ip = 0;  // the step being executed
    idx loops = 0;   // while-loop repetitions in the current run
    // where qpp runs for ever: give up after this many body runs of a while
    // loop / failures of an unlimited post-selection (0: never)
    idx loop_give_up = idx{1} << 24, ps_give_up = idx{1} << 16;

    double canonical() { return std::generate_canonical<double, 53>(gen); }

    bool was_measured(idx i) const {
        return st.subsys[i] == std::numeric_limits<idx>::max();
    }

    // QEngineState::reset, internal/classes/qengine_state.hpp:115-156
    void reset_state(std::optional<cvec> q) {
        const idx D = static_cast<idx>(std::pow(double(d), double(nq)) + 0.5);
        if (q) {
            st.qstate = *q;
        } else {
            st.qstate.assign(is_rho ? D * D : D, cplx(0, 0));
            st.qstate[0] = 1;
        }
        st.dims.assign(nq, d);
        st.probs.assign(nc, 0.0);
        st.dits.assign(nc, 0);
        st.subsys.resize(nq);
        std::iota(st.subsys.begin(), st.subsys.end(), idx{0});
        st.ps_ok = true;
        st.ensure_ps = false;
        st.max_ps = std::numeric_limits<idx>::max();
    }

    // get_relative_pos_, qengine.hpp:514-526: QuditAlreadyMeasured for a
    // destructively measured qudit
    std::vector<idx> rel(const std::vector<idx>& v) const {
        std::vector<idx> r = v;
        for (auto& x : r) {
            if (was_measured(x)) throw EngineError{QS_ERR_QUDIT_ALREADY_MEASURED};
            x = st.subsys[x];
        }
        return r;
    }

    // set_measured_destructively_, qengine.hpp:489-503
    void set_measured(idx i) {
        st.subsys[i] = std::numeric_limits<idx>::max();
        for (idx m = i; m < nq; ++m)
            if (!was_measured(m)) --st.subsys[m];
    }

    // visit_gate_step_, qengine.hpp:533-579 (gate and controlled gates)
    void gate_step(const Step& s) {
        const auto trel = rel(s.target);
        const Mat& U = mats[s.gate];
        switch (s.type) {
            case QS_STEP_GATE:
                st.qstate = apply_any(st.qstate, is_rho, U, trel, st.dims);
                break;
            case QS_STEP_FAN:
                for (idx t : trel)
                    st.qstate = apply_any(st.qstate, is_rho, U, {t}, st.dims);
                break;
            case QS_STEP_CTRL:
            case QS_STEP_CTRL_FAN: {
                const auto crel = rel(s.ctrl);
                const std::vector<idx> sh =
                    s.shift ? *s.shift : std::vector<idx>(s.ctrl.size(), 0);
                st.qstate = s.type == QS_STEP_CTRL
                                ? apply_ctrl_any(st.qstate, is_rho, U, crel,
                                                 trel, st.dims, sh)
                                : apply_ctrl_fan_any(st.qstate, is_rho, U, crel,
                                                     trel, st.dims, sh);
                break;
            }
            default:
                break;
        }
    }

    // visit_measurement_step_, qengine.hpp:639-734 (MEASURE, MEASURE_MANY,
    // POST_SELECT, POST_SELECT_MANY and their _ND forms).  A failed
    // post-selection is measured again from the same state (a new draw)
    // while ensure is set and fewer than max_ps attempts were made; out of
    // attempts, the step returns leaving state and dits untouched; without
    // ensure the failed outcome is kept and ps_ok cleared.
    void measure_step(const Step& s) {
        const auto trel = rel(s.target);
        idx attempts = 0;
        for (;;) {
            const double u = draw_for_sample();
            MeasureResult mr = measure_seq_impl(st.qstate, is_rho, st.dims,
                                                trel, s.destructive, u);
            if (s.type == QS_STEP_POST_SELECT && mr.results != s.ps) {
                ++attempts;
                if (!st.ensure_ps) {
                    st.ps_ok = false;
                } else if (attempts < st.max_ps) {
                    // unlimited retries of a result that never comes: give
                    // up (qpp loops for ever) after 2^16 attempts
                    if (st.max_ps == std::numeric_limits<idx>::max() && ps_give_up &&
                        attempts >= ps_give_up)
                        throw EngineError{QS_ERR_OUT_OF_RANGE};
                    continue;
                } else {
                    st.ps_ok = false;
                    return;
                }
            }
            st.qstate = mr.state;
            st.dims = mr.dims;
            for (idx i = 0; i < mr.results.size(); ++i) {
                st.dits[s.c_reg + i] = mr.results[i];
                st.probs[s.c_reg + i] = mr.probs[i];
            }
            if (s.destructive)
                for (idx t : s.target) set_measured(t);
            return;
        }
    }

    // RESET / RESET_MANY, qengine.hpp:804-808: qpp::reset on the relative
    // positions with the engine's uniform dimension
    void reset_step(const Step& s) {
        const double u = draw_for_sample();
        st.qstate = reset_impl(st.qstate, is_rho, st.dims, rel(s.target), u);
    }

    // DISCARD / DISCARD_MANY, qengine.hpp:810-821: destructive measure_seq,
    // results dropped
    void discard_step(const Step& s) {
        const double u = draw_for_sample();
        MeasureResult mr = measure_seq_impl(st.qstate, is_rho, st.dims,
                                            rel(s.target), true, u);
        st.qstate = mr.state;
        st.dims = mr.dims;
        for (idx t : s.target) set_measured(t);
    }

    // A sample over >= 2 basis states consumes one generate_canonical draw;
    // a 1-element table consumes none (random.tcc:2700-2705).  The draw is
    // taken here unless the state has a single basis element.
    double draw_for_sample() {
        const idx D = prod(st.dims);
        return D >= 2 ? canonical() : 0.0;
    }

    // NoiseBase::compute_state_ for QubitDepolarizingNoise,
    // noise.hpp:131-160, 388-408: draw k from {1-p, p/3, p/3, p/3}, apply
    // Ks[k] on the target, normalize.
    void noise_on(idx target) {
        const double p = noise_p;
        const std::vector<double> w{1 - p, p / 3, p / 3, p / 3};
        const auto cp = discrete_cp(w);
        const idx k = discrete_draw(cp, canonical());
        noise_results[cur_ip].push_back(k);
        static const cplx I(0, 1);
        Mat K(2, 2);
        switch (k) {
            case 0: K(0, 0) = 1; K(1, 1) = 1; break;
            case 1: K(0, 1) = 1; K(1, 0) = 1; break;
            case 2: K(0, 1) = -I; K(1, 0) = I; break;
            default: K(0, 0) = 1; K(1, 1) = -1; break;
        }
        st.qstate = apply_any(st.qstate, is_rho, K, {target}, st.dims);
        // normalize(), functions.hpp:304-333
        if (!is_rho) {
            const double nrm = std::sqrt(sqnorm(st.qstate));
            for (auto& z : st.qstate) z /= nrm;
        } else {
            const cplx tr = trace(st.qstate, prod(st.dims));
            for (auto& z : st.qstate) z /= tr;
        }
    }

    // QNoisyEngineT::execute(it), qnoisy_engine.hpp:104-121, then
    // QEngineT::execute(it)
    void execute_step(const Step& s) {
        if (noise_p >= 0) {
            std::vector<idx> live;
            for (idx i = 0; i < nq; ++i)
                if (!was_measured(i)) live.push_back(i);
            for (idx i : rel(live)) noise_on(i);
        }
        if (is_measurement(s))
            measure_step(s);
        else if (s.type == QS_STEP_RESET)
            reset_step(s);
        else if (s.type == QS_STEP_DISCARD)
            discard_step(s);
        else if (is_cctrl(s))
            cctrl_step(s);
        else if (s.type != QS_STEP_NOP && !is_runtime(s))
            gate_step(s);
    }

    // classically controlled gate, qengine.hpp:581-630: the control dits
    // (each plus its shift mod d) must agree on k; then powm(U, k) on the
    // targets (cCTRL_FAN: on each target)
    void cctrl_step(const Step& s) {
        const auto trel = rel(s.target);
        auto val = [&](idx m) {
            return s.shift ? (st.dits[s.ctrl[m]] + (*s.shift)[m]) % d : st.dits[s.ctrl[m]];
        };
        const idx k = val(0);
        for (idx m = 1; m < s.ctrl.size(); ++m)
            if (val(m) != k) return;
        const Mat U = powm(mats[s.gate], k);
        if (s.type == QS_STEP_CCTRL_FAN)
            for (idx t : trel) st.qstate = apply_any(st.qstate, is_rho, U, {t}, st.dims);
        else
            st.qstate = apply_any(st.qstate, is_rho, U, trel, st.dims);
    }

    // a truth table (qpp's predicate lambda evaluated on every value of
    // the dits it reads, qengine.hpp:845-858) or "these dits equal these
    // values"
    bool cond_true(const Step& c) const {
        if (c.table) {
            idx i = 0;
            for (idx q : c.ctrl) {
                if (st.dits[q] >= d) throw EngineError{QS_ERR_OUT_OF_RANGE};
                i = i * d + st.dits[q];
            }
            return c.ps[i] != 0;
        }
        for (idx j = 0; j < c.ctrl.size(); ++j)
            if (st.dits[c.ctrl[j]] != c.ps[j]) return false;
        return true;
    }

    // visit_runtime_step_, qengine.hpp:826-880 (IF / WHILE / ELSE / ENDIF /
    // ENDWHILE): the position the loop continues after
    idx runtime_next(const std::vector<Step>& steps_, idx i) {
        const Step& s = steps_[i];
        if (s.type == QS_STEP_COND_IF || s.type == QS_STEP_COND_WHILE)
            return cond_true(s) ? i : s.jump;
        if (s.type == QS_STEP_COND_ELSE) return s.jump;
        if (s.type == QS_STEP_COND_END && s.loop_end && cond_true(steps_[s.jump])) {
            // a loop that never ends (qpp runs it for ever) stops once its
            // body has run loop_give_up times in one run
            if (++loops >= loop_give_up && loop_give_up) throw EngineError{QS_ERR_OUT_OF_RANGE};
            return s.jump;  // back to the loop's body
        }
        return i;
    }

    // execute_steps_once_, qengine.hpp:240-294 (no runtime steps): stops
    // after the step whose post-selection failed; statistics only for a
    // run that measured and got to the end
    void execute_steps_once(const std::vector<Step>& steps_, idx pos,
                            bool ensure) {
        const bool saved_ensure = st.ensure_ps;
        if (ensure) st.ensure_ps = true;
        bool measured = false;
        loops = 0;
        for (idx i = pos; i < steps_.size(); ++i) {
            if (is_measurement(steps_[i])) measured = true;
            cur_ip = i;
            execute_step(steps_[i]);
            i = runtime_next(steps_, i);
            if (!st.ps_ok) {
                if (ensure) st.ensure_ps = saved_ensure;
                return;
            }
        }
        if (measured) ++stats[st.dits];
        if (ensure) st.ensure_ps = saved_ensure;
    }

    // execute_steps_with_postselection_retries_, qengine.hpp:316-345: every
    // rep starts from the saved state; the engine ends in the first rep
    // whose post-selection held, else in the last rep's state
    void execute_with_retries(const std::vector<Step>& steps_, idx pos,
                              idx reps) {
        EngineState copy = st;
        EngineState ok;
        bool found = false;
        for (idx rep = 0; rep < reps; ++rep) {
            execute_steps_once(steps_, pos, st.ensure_ps);
            if (st.ps_ok && !found) {
                found = true;
                ok = st;
            }
            if (rep + 1 < reps) st = copy;
        }
        if (found) st = ok;
    }

    // execute_measurement_steps_with_sampling_and_postselection_retries_,
    // qengine.hpp:374-481, with build_sampling_maps_ (:150-213).  Every
    // step from pos on is a projective measurement or post-selection.  Each
    // attempt samples the saved pre-measurement state (one draw); a sample
    // whose post-selected qudits differ from their values fails the rep (or,
    // with ensure, is drawn again up to max_ps attempts).  The first
    // successful rep runs the steps once (ensure on, no attempt limit) and
    // fixes the engine's final state; later successes only add statistics.
    void execute_sampling(const std::vector<Step>& steps_, idx pos, idx reps) {
        std::map<idx, idx> c_q;       // classical dit <- qudit
        std::map<idx, idx> q_ps_val;  // post-selected qudit -> value
        bool measured = false;
        for (idx i = pos; i < steps_.size(); ++i) {
            if (!is_measurement(steps_[i])) continue;
            measured = true;
            const auto cr = c_regs_of(steps_[i]);
            for (idx q = 0; q < steps_[i].target.size(); ++q) {
                const idx t = steps_[i].target[q];
                c_q[cr[q]] = t;
                if (steps_[i].type != QS_STEP_POST_SELECT) continue;
                auto it = q_ps_val.find(t);
                if (it == q_ps_val.end()) {
                    q_ps_val[t] = steps_[i].ps[q];
                } else if (it->second != steps_[i].ps[q]) {
                    st.ps_ok = false;  // incompatible post-selections
                    return;
                }
            }
        }
        if (!measured) return;
        std::set<idx> sset;
        for (auto& kv : c_q) sset.insert(kv.second);
        const std::vector<idx> sample_from(sset.begin(), sset.end());
        EngineState copy = st;
        EngineState ok = st;
        bool found = false;
        for (idx rep = 0; rep < reps; ++rep) {
            idx attempts = 0;
            SampleResult s;
            bool failed;
            do {
                ++attempts;
                // sample(engine_state_copy.qstate_, sample_from, d)
                const idx D = prod(copy.dims);
                const double u = D >= 2 ? canonical() : 0.0;
                s = sample_one(copy.qstate, is_rho, copy.dims, sample_from, u);
                failed = false;
                for (idx q = 0; q < sample_from.size(); ++q) {
                    auto it = q_ps_val.find(sample_from[q]);
                    if (it != q_ps_val.end() && s.digits[q] != it->second) {
                        failed = true;
                        break;
                    }
                }
                if (failed) st.ps_ok = false;
                // unlimited retries of a sample that never succeeds: give up
                // (qpp loops for ever) after 2^16 attempts
                if (failed && st.ensure_ps && st.max_ps == std::numeric_limits<idx>::max() &&
                    ps_give_up && attempts >= ps_give_up)
                    throw EngineError{QS_ERR_OUT_OF_RANGE};
            } while (failed && st.ensure_ps && attempts < st.max_ps);
            if (failed) continue;
            st.ps_ok = true;
            if (!found) {
                found = true;
                st.max_ps = std::numeric_limits<idx>::max();
                execute_steps_once(steps_, pos, true);
                st.max_ps = copy.max_ps;
                ok = st;
            } else {
                // the sampled digits (ascending qudit order) fill the
                // classical dits in ascending c_reg order
                std::vector<idx> res = st.dits;
                idx i = 0;
                for (auto& kv : c_q) res[kv.first] = s.digits[i++];
                ++stats[res];
            }
        }
        if (found) st = ok;
    }

    // QEngineT::execute(reps), qengine.hpp:1341-1405, and
    // QNoisyEngineT::execute(reps), qnoisy_engine.hpp:129-140
    void execute(idx reps) {
        if (noise_p >= 0) {
            execute_with_retries(steps, 0, reps);
            return;
        }
        // reset(qstate, false): keep state and statistics; post_select_ok
        // back to true, ensure and max_ps kept
        const bool ensure = st.ensure_ps;
        const idx max_ps = st.max_ps;
        reset_state(st.qstate);
        st.ensure_ps = ensure;
        st.max_ps = max_ps;
        const std::vector<Step> s = reps > 1 ? canonical_form(steps) : steps;
        if (s.empty()) return;
        // compute_optimize_up_to_pos_ (:87-101) and can_sample_from_
        // (:122-143): the first measurement / reset / discard step, and
        // whether only projective measurements follow it
        const idx pos = std::find_if(s.begin(), s.end(), [](const Step& x) {
                            return is_measurement(x) || is_reset_or_discard(x) ||
                                   is_runtime(x);
                        }) - s.begin();
        bool can_sample = true;
        for (idx i = pos; i < s.size(); ++i)
            if (!is_measurement(s[i])) can_sample = false;
        for (idx i = 0; i < pos; ++i) {
            cur_ip = i;
            execute_step(s[i]);
        }
        if (reps > 1 && can_sample)
            execute_sampling(s, pos, reps);
        else
            execute_with_retries(s, pos, reps);
    }
};

// the noise results of the last orc_engine_run(_ps) call
std::vector<std::vector<idx>> g_last_noise;

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int orc_apply(const double* state, int is_rho, uint32_t ndims,
              const uint64_t* dims, const double* A, uint32_t ntarget,
              const uint64_t* target, double* out) {
    const auto dv = vec_of(dims, ndims);
    const auto tv = vec_of(target, ntarget);
    idx DA = 1;
    for (idx t : tv) DA *= dv[t];
    const idx D = prod(dv);
    const cvec r = apply_any(load(state, is_rho ? D * D : D), is_rho,
                             from_raw(A, DA), tv, dv);
    store(r, out);
    return 0;
}

int orc_apply_ctrl(const double* state, int is_rho, uint32_t ndims,
                   const uint64_t* dims, const double* A, uint32_t nctrl,
                   const uint64_t* ctrl, uint32_t ntarget,
                   const uint64_t* target, const uint64_t* shift,
                   double* out) {
    const auto dv = vec_of(dims, ndims);
    const auto tv = vec_of(target, ntarget);
    const auto cv = vec_of(ctrl, nctrl);
    const std::vector<idx> sh =
        shift ? vec_of(shift, nctrl) : std::vector<idx>(nctrl, 0);
    idx DA = 1;
    for (idx t : tv) DA *= dv[t];
    const idx D = prod(dv);
    const cvec r = apply_ctrl_any(load(state, is_rho ? D * D : D), is_rho,
                                  from_raw(A, DA), cv, tv, dv, sh);
    store(r, out);
    return 0;
}

int orc_apply_ctrl_fan(const double* state, int is_rho, uint32_t ndims,
                       const uint64_t* dims, const double* A, uint32_t nctrl,
                       const uint64_t* ctrl, uint32_t ntarget,
                       const uint64_t* target, const uint64_t* shift,
                       double* out) {
    const auto dv = vec_of(dims, ndims);
    const auto tv = vec_of(target, ntarget);
    const auto cv = vec_of(ctrl, nctrl);
    const std::vector<idx> sh =
        shift ? vec_of(shift, nctrl) : std::vector<idx>(nctrl, 0);
    const idx D = prod(dv);
    const cvec r = apply_ctrl_fan_any(load(state, is_rho ? D * D : D), is_rho,
                                      from_raw(A, dv[tv[0]]), cv, tv, dv, sh);
    store(r, out);
    return 0;
}

int orc_apply_kraus(const double* rho, uint32_t ndims, const uint64_t* dims,
                    uint32_t nK, const double* Ks, uint32_t ntarget,
                    const uint64_t* target, double* out) {
    const auto dv = vec_of(dims, ndims);
    const auto tv = vec_of(target, ntarget);
    idx DK = 1;
    for (idx t : tv) DK *= dv[t];
    const idx D = prod(dv);
    const cvec in = load(rho, D * D);
    // operations.hpp:541-557: result = sum_K apply(rho, K, target, dims)
    cvec result(D * D, cplx(0, 0));
    for (uint32_t k = 0; k < nK; ++k) {
        const cvec r = apply_any(in, true, from_raw(Ks + 2 * k * DK * DK, DK),
                                 tv, dv);
        for (idx i = 0; i < D * D; ++i) result[i] += r[i];
    }
    store(result, out);
    return 0;
}

// apply(A, Ks), operations.hpp:422-464: result (Dout x Dout) =
// sum_K K * A * adjoint(K), K of shape Dout x D (rectangular allowed); the
// products in the order K * A first, then times adjoint(K)
int orc_apply_kraus_full(const double* rho, uint64_t D, uint32_t nK, const double* Ks,
                         uint64_t Dout, double* out) {
    const cvec A = load(rho, D * D);
    cvec result(Dout * Dout, cplx(0, 0));
    cvec T(Dout * D);
    for (uint32_t k = 0; k < nK; ++k) {
        const cvec K = load(Ks + 2 * uint64_t(k) * Dout * D, Dout * D);  // column-major
        for (idx j = 0; j < D; ++j)
            for (idx i = 0; i < Dout; ++i) {
                cplx acc(0, 0);
                for (idx a = 0; a < D; ++a) acc += K[i + a * Dout] * A[a + j * D];
                T[i + j * Dout] = acc;
            }
        for (idx j = 0; j < Dout; ++j)
            for (idx i = 0; i < Dout; ++i) {
                cplx acc(0, 0);
                for (idx b = 0; b < D; ++b) acc += T[i + b * Dout] * std::conj(K[j + b * Dout]);
                result[i + j * Dout] += acc;
            }
    }
    store(result, out);
    return 0;
}

int orc_ptrace(const double* A, int is_rho, uint32_t ndims,
               const uint64_t* dims, uint32_t ntarget, const uint64_t* target,
               double* out) {
    const auto dv = vec_of(dims, ndims);
    const idx D = prod(dv);
    store(ptrace_any(load(A, is_rho ? D * D : D), is_rho, dv,
                     vec_of(target, ntarget)),
          out);
    return 0;
}

int orc_ptranspose(const double* A, int is_rho, uint32_t ndims,
                   const uint64_t* dims, uint32_t ntarget,
                   const uint64_t* target, int sem, double* out) {
    const auto dv = vec_of(dims, ndims);
    const idx D = prod(dv);
    const cvec in = load(A, is_rho ? D * D : D);
    const auto tv = vec_of(target, ntarget);
    if (sem == QS_SEM_AS_BUILT && all_qubits(dv))
        store(ptranspose_qubit(in, is_rho, ndims, tv), out);
    else
        store(ptranspose_generic(in, is_rho, dv, tv), out);
    return 0;
}

int orc_syspermute(const double* A, int is_rho, uint32_t ndims,
                   const uint64_t* dims, const uint64_t* perm, int sem,
                   double* out) {
    const auto dv = vec_of(dims, ndims);
    const idx D = prod(dv);
    const cvec in = load(A, is_rho ? D * D : D);
    const auto pv = vec_of(perm, ndims);
    if (sem == QS_SEM_AS_BUILT && all_qubits(dv))
        store(syspermute_qubit(in, is_rho, ndims, pv), out);
    else
        store(syspermute_generic(in, is_rho, dv, pv), out);
    return 0;
}

int orc_probs(const double* A, int is_rho, uint64_t D, double* out) {
    const auto p = probs_of(load(A, is_rho ? D * D : D), is_rho, D);
    std::memcpy(out, p.data(), sizeof(double) * D);
    return 0;
}

uint64_t orc_discrete_draw(const double* p, uint64_t N, double u) {
    return discrete_draw(discrete_cp(std::vector<double>(p, p + N)), u);
}

void orc_discrete_table(const double* p, uint64_t N, double* cp) {
    const auto t = discrete_cp(std::vector<double>(p, p + N));
    for (uint64_t i = 0; i < N; ++i) cp[i] = t.empty() ? 1.0 : t[i];
}

int orc_sample(const double* A, int is_rho, uint32_t ndims,
               const uint64_t* dims, uint32_t ntarget, const uint64_t* target,
               double u, uint64_t* out_index, uint64_t* out_digits,
               double* out_prob) {
    const auto dv = vec_of(dims, ndims);
    const idx D = prod(dv);
    const SampleResult s = sample_one(load(A, is_rho ? D * D : D), is_rho, dv,
                                      vec_of(target, ntarget), u);
    if (out_index) *out_index = s.index;
    if (out_digits)
        for (uint32_t i = 0; i < ntarget; ++i) out_digits[i] = s.digits[i];
    if (out_prob) *out_prob = s.prob;
    return 0;
}

int orc_measure_seq(const double* A, int is_rho, uint32_t ndims,
                    const uint64_t* dims, uint32_t ntarget,
                    const uint64_t* target, int destructive, double u,
                    uint64_t* results, double* probs, double* out_state,
                    uint64_t* out_len) {
    const auto dv = vec_of(dims, ndims);
    const idx D = prod(dv);
    const MeasureResult mr =
        measure_seq_impl(load(A, is_rho ? D * D : D), is_rho, dv,
                         vec_of(target, ntarget), destructive != 0, u);
    for (uint32_t i = 0; i < ntarget; ++i) {
        results[i] = mr.results[i];
        probs[i] = mr.probs[i];
    }
    store(mr.state, out_state);
    *out_len = mr.state.size();
    return 0;
}

int orc_reset(const double* A, int is_rho, uint32_t ndims, const uint64_t* dims,
              uint32_t ntarget, const uint64_t* target, double u,
              double* out) {
    const auto dv = vec_of(dims, ndims);
    const idx D = prod(dv);
    store(reset_impl(load(A, is_rho ? D * D : D), is_rho, dv,
                     vec_of(target, ntarget), u),
          out);
    return 0;
}

void* orc_rng_new(uint64_t seed) {
    return new std::mt19937(static_cast<std::mt19937::result_type>(seed));
}

double orc_rng_canonical(void* rng) {
    return std::generate_canonical<double, 53>(*static_cast<std::mt19937*>(rng));
}

void orc_rng_free(void* rng) { delete static_cast<std::mt19937*>(rng); }

int orc_time_gates(uint32_t n, uint32_t ngates, const uint32_t* kind,
                   const uint32_t* target, const uint32_t* ctrl,
                   const double* mats, double* seconds) {
    cvec psi(idx{1} << n, cplx(0, 0));
    psi[0] = 1;
    for (uint32_t g = 0; g < ngates; ++g) {
        const Mat U = from_raw(mats + 8 * g, 2);
        const auto t0 = std::chrono::steady_clock::now();
        if (kind[g] == 0)
            psi = apply_psi_1q(psi, U, target[g], n);
        else
            psi = apply_ctrl_psi_1q(psi, U, {ctrl[g]}, target[g], {0}, n);
        const auto t1 = std::chrono::steady_clock::now();
        seconds[g] = std::chrono::duration<double>(t1 - t0).count();
    }
    int threads = 1;
#pragma omp parallel
    {
#pragma omp single
        threads = omp_get_num_threads();
    }
    return threads;
}

int orc_engine_run(const qs_circuit* qc, int is_rho, double noise_p,
                   uint64_t seed, uint64_t reps, uint64_t max_outcomes,
                   uint64_t* n_outcomes, uint64_t* stats_dits,
                   uint64_t* stats_counts, uint64_t* dits, double* probs,
                   double* out_state, uint64_t* out_len) {
    return orc_engine_run_ps(qc, is_rho, noise_p, seed, reps, 0,
                             UINT64_MAX, max_outcomes, n_outcomes, stats_dits,
                             stats_counts, dits, probs, out_state, out_len,
                             nullptr);
}

// give-up counts for the following engine runs (0: never give up)
static idx g_loop_give_up = idx{1} << 24, g_ps_give_up = idx{1} << 16;
int orc_set_engine_limits(uint64_t max_loop_reps, uint64_t max_ps_failures) {
    g_loop_give_up = max_loop_reps;
    g_ps_give_up = max_ps_failures;
    return 0;
}

int orc_last_noise_results(uint64_t step, uint64_t* out, uint64_t* count) {
    if (step >= g_last_noise.size()) return QS_ERR_OUT_OF_RANGE;
    *count = g_last_noise[step].size();
    if (out) std::copy(g_last_noise[step].begin(), g_last_noise[step].end(), out);
    return 0;
}

int orc_engine_run_ps(const qs_circuit* qc, int is_rho, double noise_p,
                      uint64_t seed, uint64_t reps, int ensure_ps,
                      uint64_t max_ps, uint64_t max_outcomes,
                      uint64_t* n_outcomes, uint64_t* stats_dits,
                      uint64_t* stats_counts, uint64_t* dits, double* probs,
                      double* out_state, uint64_t* out_len, int* ps_ok) {
    Engine e;
    e.nq = qc->nq;
    e.nc = qc->nc;
    e.d = qc->d;
    e.is_rho = is_rho != 0;
    e.noise_p = noise_p;
    e.loop_give_up = g_loop_give_up;
    e.ps_give_up = g_ps_give_up;
    e.gen.seed(static_cast<std::mt19937::result_type>(seed));
    for (uint32_t m = 0; m < qc->nmats; ++m)
        e.mats.push_back(from_raw(qc->mats + 2 * qc->mat_off[m], qc->mat_dim[m]));
    for (uint32_t i = 0; i < qc->nsteps; ++i) {
        const qs_step& s = qc->steps[i];
        Step t;
        t.type = s.type;
        t.gate = s.gate;
        t.ctrl.assign(qc->pool + s.ctrl_off, qc->pool + s.ctrl_off + s.nctrl);
        t.target.assign(qc->pool + s.target_off,
                        qc->pool + s.target_off + s.ntarget);
        if (s.shift_off != UINT32_MAX)
            t.shift = std::vector<idx>(qc->pool + s.shift_off,
                                       qc->pool + s.shift_off + s.nctrl);
        t.c_reg = s.c_reg;
        t.destructive = s.destructive != 0;
        if (s.type == QS_STEP_POST_SELECT)
            t.ps.assign(qc->pool + s.ps_off, qc->pool + s.ps_off + s.ntarget);
        if (s.type == QS_STEP_COND_IF_TABLE || s.type == QS_STEP_COND_WHILE_TABLE) {
            t.type = s.type == QS_STEP_COND_IF_TABLE ? QS_STEP_COND_IF : QS_STEP_COND_WHILE;
            t.table = true;
            idx len = 1;
            for (uint32_t j = 0; j < s.nctrl; ++j) len *= qc->d;
            t.ps.assign(qc->pool + s.ps_off, qc->pool + s.ps_off + len);
        } else if (s.type == QS_STEP_COND_IF || s.type == QS_STEP_COND_WHILE) {
            t.ps.assign(qc->pool + s.ps_off, qc->pool + s.ps_off + s.nctrl);
        }
        e.steps.push_back(t);
    }
    // match IF / ELSE / END (QCircuit::validate_conditionals)
    std::vector<idx> open;
    int in_while = 0;
    for (idx i = 0; i < e.steps.size(); ++i) {
        const uint32_t ty = e.steps[i].type;
        // QCircuit rejects destructive measurements and discards in a while
        if (in_while && ((is_measurement(e.steps[i]) && e.steps[i].destructive) ||
                         ty == QS_STEP_DISCARD))
            return QS_ERR_INVALID_CONDITIONAL;
        if (ty == QS_STEP_COND_IF) {
            open.push_back(i);
        } else if (ty == QS_STEP_COND_WHILE) {
            open.push_back(i);
            ++in_while;
        } else if (ty == QS_STEP_COND_ELSE) {
            if (open.empty() || e.steps[open.back()].type != QS_STEP_COND_IF)
                return QS_ERR_INVALID_CONDITIONAL;
            e.steps[open.back()].jump = i;
            open.back() = i;
        } else if (ty == QS_STEP_COND_END) {
            if (open.empty()) return QS_ERR_INVALID_CONDITIONAL;
            Step& top = e.steps[open.back()];
            top.jump = i;
            if (top.type == QS_STEP_COND_WHILE) {
                e.steps[i].jump = open.back();
                e.steps[i].loop_end = true;
                --in_while;
            }
            open.pop_back();
        }
    }
    if (!open.empty()) return QS_ERR_INVALID_CONDITIONAL;
    e.noise_results.assign(e.steps.size(), {});
    e.reset_state(std::nullopt);
    e.st.ensure_ps = ensure_ps != 0;
    e.st.max_ps = max_ps;
    try {
        e.execute(reps);
    } catch (const EngineError& err) {
        return err.code;
    }
    g_last_noise = e.noise_results;
    if (ps_ok) *ps_ok = e.st.ps_ok ? 1 : 0;
    idx k = 0;
    for (auto& kv : e.stats) {
        if (k < max_outcomes) {
            for (idx c = 0; c < e.nc; ++c) stats_dits[k * e.nc + c] = kv.first[c];
            stats_counts[k] = kv.second;
        }
        ++k;
    }
    *n_outcomes = k;
    for (idx c = 0; c < e.nc; ++c) {
        dits[c] = e.st.dits[c];
        probs[c] = e.st.probs[c];
    }
    if (out_state) store(e.st.qstate, out_state);
    if (out_len) *out_len = e.st.qstate.size();
    return 0;
}

}  // extern "C"

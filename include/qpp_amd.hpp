// qpp_amd.hpp -- C++17 drop-in wrapper over the gfx950 engine's C ABI.
//
// Mirrors qpp's function set (namespace qpp_amd instead of qpp, same names,
// argument meaning and overloads) for the gate-application path:
//   apply, apply(A, Ks, ...), applyCTRL, applyCTRL_fan, ptrace, ptranspose,
//   syspermute, sample, measure_seq, QCircuit, QEngine / QNoisyEngine.
// Value semantics as in qpp: arguments are copied to the device, results
// copied back into new objects.  `State` keeps a state resident on the
// device for in-place work.  Errors throw qpp_amd::exception::* classes with
// qpp's names and what() text ("[WHERE: qpp::apply()] ...!").
//
// Matrices are qpp_amd::cmat (column-major std::complex<double>, Eigen's
// layout); when Eigen is available every function also accepts and returns
// Eigen::MatrixXcd / VectorXcd, so qpp code compiles after a namespace swap.
#ifndef QPP_AMD_HPP_
#define QPP_AMD_HPP_

#include <algorithm>
#include <complex>
#include <functional>
#include <cstdint>
#include <map>
#include <optional>
#include <random>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "qpp_amd.h"

#if defined(__has_include)
#if __has_include(<Eigen/Dense>)
#include <Eigen/Dense>
#define QPP_AMD_HAS_EIGEN 1
#endif
#endif

namespace qpp_amd {

using idx = std::size_t;
using cplx = std::complex<double>;

// ---------------------------------------------------------------------------
// dense column-major matrix (a ket is cols == 1)
// ---------------------------------------------------------------------------
struct cmat {
    idx rows_ = 0, cols_ = 0;
    std::vector<cplx> v;
    cmat() = default;
    cmat(idx r, idx c) : rows_(r), cols_(c), v(r * c, cplx(0, 0)) {}
    idx rows() const { return rows_; }
    idx cols() const { return cols_; }
    idx size() const { return v.size(); }
    cplx& operator()(idx i, idx j) { return v[i + j * rows_]; }
    const cplx& operator()(idx i, idx j) const { return v[i + j * rows_]; }
    cplx& operator()(idx i) { return v[i]; }
    const cplx& operator()(idx i) const { return v[i]; }
    const cplx* data() const { return v.data(); }
    cplx* data() { return v.data(); }
    static cmat Zero(idx r, idx c) { return cmat(r, c); }
    static cmat Identity(idx r, idx c) {
        cmat m(r, c);
        for (idx i = 0; i < r && i < c; ++i) m(i, i) = 1;
        return m;
    }
};
using ket = cmat;

// ---------------------------------------------------------------------------
// exceptions: one class per qs_status, named as qpp::exception::*
// ---------------------------------------------------------------------------
namespace exception {
struct Exception : std::runtime_error {
    int code;
    Exception(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
#define QPP_AMD_EXC(Name, Code)                                        \
    struct Name : Exception {                                          \
        explicit Name(const std::string& m) : Exception(Code, m) {}    \
    };
QPP_AMD_EXC(ZeroSize, QS_ERR_ZERO_SIZE)
QPP_AMD_EXC(MatrixNotSquare, QS_ERR_MATRIX_NOT_SQUARE)
QPP_AMD_EXC(MatrixNotCvector, QS_ERR_MATRIX_NOT_CVECTOR)
QPP_AMD_EXC(MatrixNotSquareNorCvector, QS_ERR_MATRIX_NOT_SQUARE_NOR_CVECTOR)
QPP_AMD_EXC(MatrixMismatchSubsys, QS_ERR_MATRIX_MISMATCH_SUBSYS)
QPP_AMD_EXC(DimsInvalid, QS_ERR_DIMS_INVALID)
QPP_AMD_EXC(DimsNotEqual, QS_ERR_DIMS_NOT_EQUAL)
QPP_AMD_EXC(DimsMismatchMatrix, QS_ERR_DIMS_MISMATCH_MATRIX)
QPP_AMD_EXC(DimsMismatchCvector, QS_ERR_DIMS_MISMATCH_CVECTOR)
QPP_AMD_EXC(SubsysMismatchDims, QS_ERR_SUBSYS_MISMATCH_DIMS)
QPP_AMD_EXC(PermInvalid, QS_ERR_PERM_INVALID)
QPP_AMD_EXC(PermMismatchDims, QS_ERR_PERM_MISMATCH_DIMS)
QPP_AMD_EXC(OutOfRange, QS_ERR_OUT_OF_RANGE)
QPP_AMD_EXC(TypeMismatch, QS_ERR_TYPE_MISMATCH)
QPP_AMD_EXC(SizeMismatch, QS_ERR_SIZE_MISMATCH)
QPP_AMD_EXC(QuditAlreadyMeasured, QS_ERR_QUDIT_ALREADY_MEASURED)
QPP_AMD_EXC(NotImplemented, QS_ERR_NOT_IMPLEMENTED)
QPP_AMD_EXC(InvalidConditional, QS_ERR_INVALID_CONDITIONAL)
QPP_AMD_EXC(DeviceError, QS_ERR_DEVICE)
#undef QPP_AMD_EXC
}  // namespace exception

inline void check(int rc) {
    if (rc == QS_OK) return;
    const std::string m = qs_last_error();
    switch (rc) {
        case QS_ERR_ZERO_SIZE: throw exception::ZeroSize(m);
        case QS_ERR_MATRIX_NOT_SQUARE: throw exception::MatrixNotSquare(m);
        case QS_ERR_MATRIX_NOT_CVECTOR: throw exception::MatrixNotCvector(m);
        case QS_ERR_MATRIX_NOT_SQUARE_NOR_CVECTOR: throw exception::MatrixNotSquareNorCvector(m);
        case QS_ERR_MATRIX_MISMATCH_SUBSYS: throw exception::MatrixMismatchSubsys(m);
        case QS_ERR_DIMS_INVALID: throw exception::DimsInvalid(m);
        case QS_ERR_DIMS_NOT_EQUAL: throw exception::DimsNotEqual(m);
        case QS_ERR_DIMS_MISMATCH_MATRIX: throw exception::DimsMismatchMatrix(m);
        case QS_ERR_DIMS_MISMATCH_CVECTOR: throw exception::DimsMismatchCvector(m);
        case QS_ERR_SUBSYS_MISMATCH_DIMS: throw exception::SubsysMismatchDims(m);
        case QS_ERR_PERM_INVALID: throw exception::PermInvalid(m);
        case QS_ERR_PERM_MISMATCH_DIMS: throw exception::PermMismatchDims(m);
        case QS_ERR_OUT_OF_RANGE: throw exception::OutOfRange(m);
        case QS_ERR_TYPE_MISMATCH: throw exception::TypeMismatch(m);
        case QS_ERR_SIZE_MISMATCH: throw exception::SizeMismatch(m);
        case QS_ERR_QUDIT_ALREADY_MEASURED: throw exception::QuditAlreadyMeasured(m);
        case QS_ERR_NOT_IMPLEMENTED: throw exception::NotImplemented(m);
        case QS_ERR_OVERFLOW: throw std::overflow_error("qpp::normalize(): Division by zero!");
        case QS_ERR_IO: throw std::runtime_error(m);
        case QS_ERR_INVALID_CONDITIONAL: throw exception::InvalidConditional(m);
        default: throw exception::DeviceError(m);
    }
}

// ---------------------------------------------------------------------------
// device-resident state (RAII over qs_state)
// ---------------------------------------------------------------------------
class State {
    qs_state* h_ = nullptr;

  public:
    State() = default;
    explicit State(qs_state* h) : h_(h) {}
    State(const std::vector<idx>& dims, qs_kind kind = QS_KET, qs_dtype dt = QS_C128) {
        std::vector<uint64_t> d(dims.begin(), dims.end());
        check(qs_state_create(dt, kind, static_cast<uint32_t>(d.size()), d.data(), &h_));
    }
    State(const cmat& A, const std::vector<idx>& dims) {
        const bool ket = A.cols() == 1;
        idx D = 1;
        for (idx x : dims) D *= x;
        if (dims.empty()) throw exception::DimsInvalid("[WHERE: qpp_amd] Invalid dimension(s)!");
        if ((ket && A.rows() != D) || (!ket && (A.rows() != A.cols() || A.rows() != D))) {
            if (!ket && A.rows() != A.cols())
                throw exception::MatrixNotSquareNorCvector(
                    "[WHERE: qpp_amd] Matrix is not square nor column vector!");
            if (ket) throw exception::DimsMismatchCvector("[WHERE: qpp_amd] Dimension(s) mismatch column vector size!");
            throw exception::DimsMismatchMatrix("[WHERE: qpp_amd] Dimension(s) mismatch matrix size!");
        }
        std::vector<uint64_t> d(dims.begin(), dims.end());
        check(qs_state_create(QS_C128, ket ? QS_KET : QS_RHO, static_cast<uint32_t>(d.size()),
                              d.data(), &h_));
        check(qs_state_upload(h_, A.data(), A.size() * sizeof(cplx)));
    }
    State(const State&) = delete;
    State& operator=(const State&) = delete;
    State(State&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    State& operator=(State&& o) noexcept {
        std::swap(h_, o.h_);
        return *this;
    }
    ~State() { qs_state_destroy(h_); }
    qs_state* get() const { return h_; }

    // qpp::save(A, os) (QS_FMT_TEXT) / qpp::obsolete::save (QS_FMT_BINARY)
    // of this state into a file (input_output.hpp:177-211, 318-345)
    const State& save(const std::string& path, qs_format fmt = QS_FMT_TEXT) const {
        check(qs_state_save(h_, path.c_str(), fmt));
        return *this;
    }
    // qpp::load<cmat>(is) / qpp::obsolete::load<cmat>(is) into a device state
    static State load(const std::string& path, const std::vector<idx>& dims,
                      qs_kind kind = QS_KET, qs_dtype dt = QS_C128,
                      qs_format fmt = QS_FMT_TEXT) {
        std::vector<uint64_t> d(dims.begin(), dims.end());
        qs_state* h = nullptr;
        check(qs_state_load(path.c_str(), fmt, dt, kind, static_cast<uint32_t>(d.size()), d.data(),
                            &h));
        return State(h);
    }

    cmat download() const {
        qs_dtype dt;
        qs_kind kd;
        uint32_t nd = 0;
        std::vector<uint64_t> dims(128);
        uint64_t ne = 0;
        check(qs_state_info(h_, &dt, &kd, &nd, dims.data(), &ne));
        idx D = 1;
        for (uint32_t i = 0; i < nd; ++i) D *= dims[i];
        cmat r(D, kd == QS_KET ? 1 : D);
        check(qs_state_download(h_, r.data(), r.size() * sizeof(cplx)));
        return r;
    }
    State& apply(const cmat& A, const std::vector<idx>& target) {
        std::vector<uint64_t> t(target.begin(), target.end());
        check(qs_apply(h_, reinterpret_cast<const double*>(A.data()), A.rows(),
                       static_cast<uint32_t>(t.size()), t.data()));
        return *this;
    }
    State& applyCTRL(const cmat& A, const std::vector<idx>& ctrl, const std::vector<idx>& target,
                     std::optional<std::vector<idx>> shift = std::nullopt, bool fan = false) {
        std::vector<uint64_t> c(ctrl.begin(), ctrl.end()), t(target.begin(), target.end());
        std::vector<uint64_t> s;
        if (shift) s.assign(shift->begin(), shift->end());
        auto f = fan ? qs_apply_ctrl_fan : qs_apply_ctrl;
        check(f(h_, reinterpret_cast<const double*>(A.data()), A.rows(),
                static_cast<uint32_t>(c.size()), c.data(), static_cast<uint32_t>(t.size()),
                t.data(), shift ? s.data() : nullptr));
        return *this;
    }
    State& apply(const std::vector<cmat>& Ks, const std::vector<idx>& target) {
        if (Ks.empty()) throw exception::ZeroSize("[WHERE: qpp::apply()] Object has zero size! [CONTEXT: Ks]");
        std::vector<cplx> buf;
        for (const cmat& K : Ks) {
            if (K.rows() != Ks[0].rows() || K.cols() != Ks[0].cols())
                throw exception::DimsNotEqual("[WHERE: qpp::apply()] Dimensions not equal! [CONTEXT: K]");
            buf.insert(buf.end(), K.v.begin(), K.v.end());
        }
        std::vector<uint64_t> t(target.begin(), target.end());
        check(qs_apply_kraus(h_, reinterpret_cast<const double*>(buf.data()),
                             static_cast<uint32_t>(Ks.size()), Ks[0].rows(),
                             static_cast<uint32_t>(t.size()), t.data()));
        return *this;
    }
};

namespace detail {
inline std::vector<idx> dims_of(const cmat& A, idx d) {
    if (d < 2) throw exception::DimsInvalid("[WHERE: qpp_amd] Invalid dimension(s)! [CONTEXT: d]");
    // internal/util.hpp:389-397, llround(log2(D) / log2(d))
    const idx n = static_cast<idx>(std::llround(std::log2(double(A.rows())) / std::log2(double(d))));
    return std::vector<idx>(n, d);
}
inline std::vector<uint64_t> u64(const std::vector<idx>& v) {
    return std::vector<uint64_t>(v.begin(), v.end());
}
}  // namespace detail

// ---------------------------------------------------------------------------
// qpp's free functions (operations.hpp, instruments.hpp)
// ---------------------------------------------------------------------------
inline cmat apply(const cmat& state, const cmat& A, const std::vector<idx>& target,
                  const std::vector<idx>& dims) {
    State s(state, dims);
    return s.apply(A, target).download();
}
inline cmat apply(const cmat& state, const cmat& A, const std::vector<idx>& target, idx d = 2) {
    return apply(state, A, target, detail::dims_of(state, d));
}
inline cmat apply(const cmat& rho, const std::vector<cmat>& Ks, const std::vector<idx>& target,
                  const std::vector<idx>& dims) {
    State s(rho, dims);
    return s.apply(Ks, target).download();
}
inline cmat apply(const cmat& rho, const std::vector<cmat>& Ks, const std::vector<idx>& target,
                  idx d = 2) {
    return apply(rho, Ks, target, detail::dims_of(rho, d));
}
// qpp::apply(A, Ks), operations.hpp:422-464: the channel on the whole
// system; the Kraus operators may be rectangular (Dout x D), the result is
// Dout x Dout
inline cmat apply(const cmat& rho, const std::vector<cmat>& Ks) {
    if (rho.size() == 0) throw exception::ZeroSize("[WHERE: qpp::apply()] Object has zero size! [CONTEXT: A]");
    if (rho.rows() != rho.cols())
        throw exception::MatrixNotSquare("[WHERE: qpp::apply()] Matrix is not square! [CONTEXT: A]");
    if (Ks.empty()) throw exception::ZeroSize("[WHERE: qpp::apply()] Object has zero size! [CONTEXT: Ks]");
    if (Ks[0].cols() != rho.rows())
        throw exception::DimsMismatchMatrix(
            "[WHERE: qpp::apply()] Dimension(s) mismatch matrix size! [CONTEXT: Ks[0]]");
    std::vector<cplx> buf;
    for (const cmat& K : Ks) {
        if (K.rows() != Ks[0].rows() || K.cols() != Ks[0].cols())
            throw exception::DimsNotEqual("[WHERE: qpp::apply()] Dimensions not equal! [CONTEXT: K]");
        buf.insert(buf.end(), K.v.begin(), K.v.end());
    }
    State s(rho, std::vector<idx>{rho.rows()});
    qs_state* out = nullptr;
    check(qs_apply_kraus_full(s.get(), reinterpret_cast<const double*>(buf.data()),
                              static_cast<uint32_t>(Ks.size()), Ks[0].rows(), Ks[0].cols(), &out));
    return State(out).download();
}
inline cmat applyCTRL(const cmat& state, const cmat& A, const std::vector<idx>& ctrl,
                      const std::vector<idx>& target, const std::vector<idx>& dims,
                      std::optional<std::vector<idx>> shift = std::nullopt) {
    State s(state, dims);
    return s.applyCTRL(A, ctrl, target, shift).download();
}
inline cmat applyCTRL(const cmat& state, const cmat& A, const std::vector<idx>& ctrl,
                      const std::vector<idx>& target, idx d = 2,
                      std::optional<std::vector<idx>> shift = std::nullopt) {
    return applyCTRL(state, A, ctrl, target, detail::dims_of(state, d), shift);
}
inline cmat applyCTRL_fan(const cmat& state, const cmat& A, const std::vector<idx>& ctrl,
                          const std::vector<idx>& target, const std::vector<idx>& dims,
                          std::optional<std::vector<idx>> shift = std::nullopt) {
    State s(state, dims);
    return s.applyCTRL(A, ctrl, target, shift, true).download();
}
inline cmat applyCTRL_fan(const cmat& state, const cmat& A, const std::vector<idx>& ctrl,
                          const std::vector<idx>& target, idx d = 2,
                          std::optional<std::vector<idx>> shift = std::nullopt) {
    return applyCTRL_fan(state, A, ctrl, target, detail::dims_of(state, d), shift);
}
inline cmat ptrace(const cmat& A, const std::vector<idx>& target, const std::vector<idx>& dims) {
    State s(A, dims);
    qs_state* o = nullptr;
    const auto t = detail::u64(target);
    check(qs_ptrace(s.get(), static_cast<uint32_t>(t.size()), t.data(), &o));
    return State(o).download();
}
inline cmat ptrace(const cmat& A, const std::vector<idx>& target, idx d = 2) {
    return ptrace(A, target, detail::dims_of(A, d));
}
inline cmat ptranspose(const cmat& A, const std::vector<idx>& target, const std::vector<idx>& dims,
                       qs_semantics sem = QS_SEM_DOCUMENTED) {
    State s(A, dims);
    qs_state* o = nullptr;
    const auto t = detail::u64(target);
    check(qs_ptranspose(s.get(), static_cast<uint32_t>(t.size()), t.data(), sem, &o));
    return State(o).download();
}
inline cmat ptranspose(const cmat& A, const std::vector<idx>& target, idx d = 2,
                       qs_semantics sem = QS_SEM_DOCUMENTED) {
    return ptranspose(A, target, detail::dims_of(A, d), sem);
}
inline cmat syspermute(const cmat& A, const std::vector<idx>& perm, const std::vector<idx>& dims,
                       qs_semantics sem = QS_SEM_DOCUMENTED) {
    if (perm.size() != dims.size())
        throw exception::PermMismatchDims(
            "[WHERE: qpp::syspermute()] Permutation mismatch dimensions! [CONTEXT: dims/perm]");
    State s(A, dims);
    qs_state* o = nullptr;
    const auto p = detail::u64(perm);
    check(qs_syspermute(s.get(), p.data(), sem, &o));
    return State(o).download();
}
inline cmat syspermute(const cmat& A, const std::vector<idx>& perm, idx d = 2,
                       qs_semantics sem = QS_SEM_DOCUMENTED) {
    return syspermute(A, perm, detail::dims_of(A, d), sem);
}

// qpp draws these uniforms from RandomDevices' mt19937 with
// std::generate_canonical<double, 53>; pass one explicitly or a generator.
inline std::tuple<std::vector<idx>, double> sample(const cmat& A, const std::vector<idx>& target,
                                                   const std::vector<idx>& dims, double u) {
    State s(A, dims);
    const auto t = detail::u64(target);
    std::vector<uint64_t> dig(t.size() ? t.size() : 1);
    double p = 0;
    uint64_t index = 0;
    check(qs_sample(s.get(), static_cast<uint32_t>(t.size()), t.data(), 1, &u, &index,
                    dig.data(), &p));
    return {std::vector<idx>(dig.begin(), dig.begin() + t.size()), p};
}
inline std::tuple<std::vector<idx>, double> sample(const cmat& A, const std::vector<idx>& target,
                                                   const std::vector<idx>& dims, std::mt19937& gen) {
    return sample(A, target, dims, std::generate_canonical<double, 53>(gen));
}
inline std::map<std::vector<idx>, idx> sample(idx num_samples, const cmat& A,
                                              const std::vector<idx>& target,
                                              const std::vector<idx>& dims, std::mt19937& gen) {
    State s(A, dims);
    std::vector<double> u(num_samples);
    for (auto& x : u) x = std::generate_canonical<double, 53>(gen);
    const auto t = detail::u64(target);
    std::vector<uint64_t> dig(num_samples * (t.size() ? t.size() : 1));
    check(qs_sample(s.get(), static_cast<uint32_t>(t.size()), t.data(), num_samples, u.data(),
                    nullptr, dig.data(), nullptr));
    std::map<std::vector<idx>, idx> r;
    for (idx i = 0; i < num_samples; ++i)
        ++r[std::vector<idx>(dig.begin() + i * t.size(), dig.begin() + (i + 1) * t.size())];
    return r;
}
inline std::tuple<std::vector<idx>, std::vector<double>, cmat> measure_seq(
    const cmat& A, const std::vector<idx>& target, const std::vector<idx>& dims, bool destructive,
    double u) {
    State s(A, dims);
    const auto t = detail::u64(target);
    std::vector<uint64_t> res(t.size() ? t.size() : 1);
    std::vector<double> ps(t.size() ? t.size() : 1);
    check(qs_measure_seq(s.get(), static_cast<uint32_t>(t.size()), t.data(), destructive ? 1 : 0,
                         u, res.data(), ps.data()));
    return {std::vector<idx>(res.begin(), res.begin() + t.size()),
            std::vector<double>(ps.begin(), ps.begin() + t.size()), s.download()};
}
// qpp::reset(A, target, dims) (instruments.hpp:1148-1196) with the uniform
// its measure_seq draws, or drawn from `gen` as qpp's RandomDevices would
inline cmat reset(const cmat& A, const std::vector<idx>& target, const std::vector<idx>& dims,
                  double u) {
    State s(A, dims);
    const auto t = detail::u64(target);
    check(qs_reset(s.get(), static_cast<uint32_t>(t.size()), t.data(), u));
    return s.download();
}
inline cmat reset(const cmat& A, const std::vector<idx>& target, const std::vector<idx>& dims,
                  std::mt19937& gen) {
    return reset(A, target, dims, std::generate_canonical<double, 53>(gen));
}

// ---------------------------------------------------------------------------
// QCircuit / QEngine (classes/qcircuit.hpp, classes/qengine.hpp)
// ---------------------------------------------------------------------------
class QCircuit {
  public:
    explicit QCircuit(idx nq, idx nc = 0, idx d = 2) : nq_(nq), nc_(nc), d_(d) {}
    QCircuit& gate(const cmat& U, idx i) { return add(QS_STEP_GATE, U, {}, {i}); }
    QCircuit& gate(const cmat& U, idx i, idx j) { return add(QS_STEP_GATE, U, {}, {i, j}); }
    QCircuit& gate(const cmat& U, idx i, idx j, idx k) {
        return add(QS_STEP_GATE, U, {}, {i, j, k});
    }
    QCircuit& gate_joint(const cmat& U, const std::vector<idx>& t) {
        return add(QS_STEP_GATE, U, {}, t);
    }
    QCircuit& gate_fan(const cmat& U, const std::vector<idx>& t) {
        return add(QS_STEP_FAN, U, {}, t);
    }
    QCircuit& CTRL(const cmat& U, const std::vector<idx>& ctrl, const std::vector<idx>& target,
                   std::optional<std::vector<idx>> shift = std::nullopt) {
        return add(QS_STEP_CTRL, U, ctrl, target, shift);
    }
    QCircuit& CTRL(const cmat& U, idx ctrl, idx target) {
        return add(QS_STEP_CTRL, U, {ctrl}, {target});
    }
    QCircuit& CTRL_fan(const cmat& U, const std::vector<idx>& ctrl, const std::vector<idx>& target,
                       std::optional<std::vector<idx>> shift = std::nullopt) {
        return add(QS_STEP_CTRL_FAN, U, ctrl, target, shift);
    }
    QCircuit& measure(idx target, idx c_reg, bool destructive = true) {
        return meas(QS_STEP_MEASURE, {target}, c_reg, destructive);
    }
    QCircuit& measure(const std::vector<idx>& target, idx c_reg = 0, bool destructive = true) {
        return meas(QS_STEP_MEASURE_MANY, target, c_reg, destructive);
    }
    // post_select (qcircuit.hpp:3366-3500): measure, keep the run only
    // when the results equal ps_vals (one value per target)
    QCircuit& post_select(idx target, idx ps_val, idx c_reg, bool destructive = true) {
        return post_select(std::vector<idx>{target}, std::vector<idx>{ps_val}, c_reg, destructive);
    }
    QCircuit& post_select(const std::vector<idx>& target, const std::vector<idx>& ps_vals,
                          idx c_reg, bool destructive = true) {
        if (ps_vals.size() != target.size())
            throw std::invalid_argument("qpp_amd::QCircuit::post_select(): one value per target");
        meas(QS_STEP_POST_SELECT, target, c_reg, destructive);
        steps_.back().ps_off = static_cast<uint32_t>(pool_.size());
        pool_.insert(pool_.end(), ps_vals.begin(), ps_vals.end());
        return *this;
    }
    // classically controlled gates (qcircuit.hpp:2604-2960): U^k on the
    // targets when every control dit, plus its shift mod d, equals k
    QCircuit& cCTRL(const cmat& U, const std::vector<idx>& ctrl_dits, const std::vector<idx>& target,
                    std::optional<std::vector<idx>> shift = std::nullopt) {
        return add(QS_STEP_CCTRL, U, ctrl_dits, target, shift);
    }
    QCircuit& cCTRL(const cmat& U, idx ctrl_dit, idx target) {
        return add(QS_STEP_CCTRL, U, {ctrl_dit}, {target});
    }
    QCircuit& cCTRL_fan(const cmat& U, const std::vector<idx>& ctrl_dits,
                        const std::vector<idx>& target,
                        std::optional<std::vector<idx>> shift = std::nullopt) {
        return add(QS_STEP_CCTRL_FAN, U, ctrl_dits, target, shift);
    }
    // runtime conditional block (qcircuit.hpp cond_if / cond_else /
    // cond_end): the condition is "dit dits[j] equals vals[j] for every j",
    // the form OpenQASM's if (c==n) takes
    QCircuit& cond_if(const std::vector<idx>& dits, const std::vector<idx>& vals) {
        if (dits.size() != vals.size())
            throw std::invalid_argument("qpp_amd::QCircuit::cond_if(): one value per dit");
        qs_step s{};
        s.type = QS_STEP_COND_IF;
        s.nctrl = static_cast<uint32_t>(dits.size());
        s.ctrl_off = static_cast<uint32_t>(pool_.size());
        pool_.insert(pool_.end(), dits.begin(), dits.end());
        s.ps_off = static_cast<uint32_t>(pool_.size());
        pool_.insert(pool_.end(), vals.begin(), vals.end());
        s.shift_off = UINT32_MAX;
        steps_.push_back(s);
        return *this;
    }
    // runtime while loop (cond_while ... cond_end), same condition form
    QCircuit& cond_while(const std::vector<idx>& dits, const std::vector<idx>& vals) {
        cond_if(dits, vals);
        steps_.back().type = QS_STEP_COND_WHILE;
        return *this;
    }
    // qpp's predicate form (cond_if(pred) / cond_while(pred),
    // qcircuit.hpp:968-1026): pred sees every classical dit, as qpp's dits
    // proxy does; it is stored as a truth table over their values 0..d-1
    // (d^nc entries, at most 2^20; a dit outside 0..d-1 when the condition
    // is evaluated is OutOfRange)
    QCircuit& cond_if(const std::function<bool(const std::vector<idx>&)>& pred) {
        return cond_table(QS_STEP_COND_IF_TABLE, pred);
    }
    QCircuit& cond_while(const std::function<bool(const std::vector<idx>&)>& pred) {
        return cond_table(QS_STEP_COND_WHILE_TABLE, pred);
    }
    QCircuit& cond_else() { return marker(QS_STEP_COND_ELSE); }
    QCircuit& cond_end() { return marker(QS_STEP_COND_END); }
    // reset / discard (qcircuit.hpp:3745-3940)
    QCircuit& reset(idx target) { return meas(QS_STEP_RESET, {target}, 0, false); }
    QCircuit& reset(const std::vector<idx>& target) { return meas(QS_STEP_RESET, target, 0, false); }
    QCircuit& discard(idx target) { return meas(QS_STEP_DISCARD, {target}, 0, true); }
    QCircuit& discard(const std::vector<idx>& target) {
        return meas(QS_STEP_DISCARD, target, 0, true);
    }
    QCircuit& nop() {
        qs_step s{};
        s.type = QS_STEP_NOP;
        s.shift_off = UINT32_MAX;
        steps_.push_back(s);
        return *this;
    }
    idx get_nq() const { return nq_; }
    idx get_nc() const { return nc_; }
    // the POD program (valid while the circuit lives and is not modified)
    const qs_circuit& program() const {
        c_.nq = static_cast<uint32_t>(nq_);
        c_.nc = static_cast<uint32_t>(nc_);
        c_.d = static_cast<uint32_t>(d_);
        c_.nsteps = static_cast<uint32_t>(steps_.size());
        c_.steps = steps_.data();
        c_.nmats = static_cast<uint32_t>(dim_.size());
        c_.mat_dim = dim_.data();
        c_.mat_off = off_.data();
        c_.mats = reinterpret_cast<const double*>(mats_.data());
        c_.pool = pool_.data();
        return c_;
    }

  private:
    QCircuit& marker(uint32_t type) {
        qs_step s{};
        s.type = type;
        s.shift_off = UINT32_MAX;
        steps_.push_back(s);
        return *this;
    }
    QCircuit& add(uint32_t type, const cmat& U, const std::vector<idx>& ctrl,
                  const std::vector<idx>& target,
                  std::optional<std::vector<idx>> shift = std::nullopt) {
        qs_step s{};
        s.type = type;
        s.gate = static_cast<uint32_t>(dim_.size());
        dim_.push_back(U.rows());
        off_.push_back(mats_.size());
        mats_.insert(mats_.end(), U.v.begin(), U.v.end());
        s.nctrl = static_cast<uint32_t>(ctrl.size());
        s.ntarget = static_cast<uint32_t>(target.size());
        s.ctrl_off = static_cast<uint32_t>(pool_.size());
        pool_.insert(pool_.end(), ctrl.begin(), ctrl.end());
        s.target_off = static_cast<uint32_t>(pool_.size());
        pool_.insert(pool_.end(), target.begin(), target.end());
        s.shift_off = UINT32_MAX;
        if (shift) {
            s.shift_off = static_cast<uint32_t>(pool_.size());
            pool_.insert(pool_.end(), shift->begin(), shift->end());
        }
        steps_.push_back(s);
        return *this;
    }
    QCircuit& meas(uint32_t type, const std::vector<idx>& target, idx c_reg, bool destructive) {
        qs_step s{};
        s.type = type;
        s.ntarget = static_cast<uint32_t>(target.size());
        s.target_off = static_cast<uint32_t>(pool_.size());
        pool_.insert(pool_.end(), target.begin(), target.end());
        s.shift_off = UINT32_MAX;
        s.c_reg = static_cast<uint32_t>(c_reg);
        s.destructive = destructive ? 1 : 0;
        steps_.push_back(s);
        return *this;
    }
    QCircuit& cond_table(uint32_t type, const std::function<bool(const std::vector<idx>&)>& pred) {
        idx len = 1;
        for (idx j = 0; j < nc_; ++j) {
            len *= d_;
            if (len > (idx{1} << 20))
                throw std::invalid_argument("qpp_amd::QCircuit::cond_if(): more than 2^20 dit values");
        }
        std::vector<idx> dits(nc_), table(len), v(nc_, 0);
        for (idx j = 0; j < nc_; ++j) dits[j] = j;
        for (idx i = 0; i < len; ++i) {  // v = the digits of i, dit 0 most significant
            idx r = i;
            for (idx j = nc_; j-- > 0;) {
                v[j] = r % d_;
                r /= d_;
            }
            table[i] = pred(v) ? 1 : 0;
        }
        return cond_table(type, dits, table);
    }

  public:
    // the truth-table step itself: table[i] over the values of `dits`,
    // i = sum_j value(dits[j]) d^(k-1-j) (QS_STEP_COND_IF_TABLE /
    // QS_STEP_COND_WHILE_TABLE)
    QCircuit& cond_table(uint32_t type, const std::vector<idx>& dits, const std::vector<idx>& table) {
        qs_step s{};
        s.type = type;
        s.nctrl = static_cast<uint32_t>(dits.size());
        s.ctrl_off = static_cast<uint32_t>(pool_.size());
        pool_.insert(pool_.end(), dits.begin(), dits.end());
        s.ps_off = static_cast<uint32_t>(pool_.size());
        pool_.insert(pool_.end(), table.begin(), table.end());
        s.shift_off = UINT32_MAX;
        steps_.push_back(s);
        return *this;
    }

  private:
    idx nq_, nc_, d_;
    std::vector<qs_step> steps_;
    std::vector<uint64_t> dim_, off_, pool_;
    std::vector<cplx> mats_;
    mutable qs_circuit c_{};
};

class QEngine {
  public:
    // noise_p < 0: QEngine; noise_p >= 0: QNoisyEngine with
    // QubitDepolarizingNoise(noise_p)
    explicit QEngine(const QCircuit& qc, std::uint64_t seed = 5489, double noise_p = -1,
                     qs_kind kind = QS_KET, qs_dtype dt = QS_C128)
        : nc_(qc.get_nc()), nq_(qc.get_nq()), nsteps_(qc.program().nsteps) {
        check(qs_engine_create(&qc.program(), kind, dt, noise_p, seed, &e_));
    }
    QEngine(const QEngine&) = delete;
    QEngine& operator=(const QEngine&) = delete;
    ~QEngine() { qs_engine_destroy(e_); }
    QEngine& execute(idx reps = 1) {
        check(qs_engine_execute(e_, reps));
        return *this;
    }
    std::map<std::vector<idx>, idx> get_stats() const {
        uint64_t n = 0;
        check(qs_engine_num_outcomes(e_, &n));
        std::vector<uint64_t> d(n * nc_ + 1), c(n + 1);
        check(qs_engine_get_stats(e_, d.data(), c.data()));
        std::map<std::vector<idx>, idx> r;
        for (uint64_t i = 0; i < n; ++i)
            r[std::vector<idx>(d.begin() + i * nc_, d.begin() + (i + 1) * nc_)] = c[i];
        return r;
    }
    std::vector<idx> get_dits() const {
        std::vector<uint64_t> d(nc_ + 1);
        check(qs_engine_get_dits(e_, d.data()));
        return std::vector<idx>(d.begin(), d.begin() + nc_);
    }
    std::vector<double> get_probs() const {
        std::vector<double> p(nc_ + 1);
        check(qs_engine_get_probs(e_, p.data()));
        p.resize(nc_);
        return p;
    }
    cmat get_state() const {
        qs_state* s = nullptr;
        check(qs_engine_get_state(e_, &s));
        return State(s).download();
    }
    // QEngineT::set_state (qengine.hpp:1195-1216): a state over the
    // non-measured qudits
    QEngine& set_state(const cmat& psi) {
        const idx D = psi.rows();
        State s(psi, std::vector<idx>{D});
        check(qs_engine_set_state(e_, s.get()));
        return *this;
    }
    idx get_dit(idx i) const {
        const auto d = get_dits();
        if (i >= d.size())
            throw exception::OutOfRange("[WHERE: qpp::QEngineT::get_dit()] Argument out of range! [CONTEXT: i]");
        return d[i];
    }
    QEngine& reset() {
        check(qs_engine_reset(e_));
        ensure_ = false;
        max_ps_ = UINT64_MAX;
        return *this;
    }
    // qengine.hpp:1177-1227
    QEngine& set_ensure_post_selection(bool val) {
        ensure_ = val;
        check(qs_engine_set_post_selection(e_, ensure_ ? 1 : 0, max_ps_));
        return *this;
    }
    QEngine& set_max_post_selection_reps(idx n) {
        max_ps_ = n;
        check(qs_engine_set_post_selection(e_, ensure_ ? 1 : 0, max_ps_));
        return *this;
    }
    bool get_ensure_post_selection() const { return ensure_; }
    idx get_max_post_selection_reps() const { return max_ps_; }
    // no qpp counterpart: where qpp runs for ever, stop with OutOfRange after
    // max_loop_reps runs of a while loop's body / max_ps_failures failures
    // of an unlimited post-selection (0: never, qpp's behaviour)
    QEngine& set_limits(idx max_loop_reps, idx max_ps_failures) {
        check(qs_engine_set_limits(e_, max_loop_reps, max_ps_failures));
        return *this;
    }
    // qengine.hpp:1035-1164: measured qudits, dits, statistics
    QEngine& set_dits(const std::vector<idx>& dits) {
        std::vector<uint64_t> d(dits.begin(), dits.end());
        check(qs_engine_set_dits(e_, d.data(), d.size()));
        return *this;
    }
    QEngine& set_dit(idx i, idx value) {
        auto d = get_dits();
        if (i >= d.size())
            throw exception::OutOfRange("[WHERE: qpp::QEngineT::set_dit()] Argument out of range! [CONTEXT: i]");
        d[i] = value;
        return set_dits(d);
    }
    std::vector<idx> get_measured_d() const { return measured_where(true); }
    std::vector<idx> get_non_measured_d() const { return measured_where(false); }
    bool was_measured_d(idx i) const {
        const auto m = get_measured_d();
        return std::find(m.begin(), m.end(), i) != m.end();
    }
    QEngine& reset_stats() {
        check(qs_engine_reset_stats(e_));
        return *this;
    }
    // QNoisyEngineT::get_noise_results (qnoisy_engine.hpp:174-176)
    std::vector<std::vector<idx>> get_noise_results() const {
        std::vector<std::vector<idx>> r(nsteps_);
        for (idx i = 0; i < nsteps_; ++i) {
            uint64_t n = 0;
            check(qs_engine_noise_results(e_, i, nullptr, &n));
            std::vector<uint64_t> v(n);
            check(qs_engine_noise_results(e_, i, v.data(), &n));
            r[i].assign(v.begin(), v.end());
        }
        return r;
    }
    bool post_select_ok() const {
        int ok = 0;
        check(qs_engine_post_select_ok(e_, &ok));
        return ok != 0;
    }

  private:
    std::vector<idx> measured_where(bool want) const {
        std::vector<uint8_t> m(nq_ + 1);
        check(qs_engine_get_measured(e_, m.data()));
        std::vector<idx> r;
        for (idx i = 0; i < nq_; ++i)
            if ((m[i] != 0) == want) r.push_back(i);
        return r;
    }
    bool ensure_ = false;
    idx max_ps_ = UINT64_MAX;
    idx nc_;
    idx nq_;
    idx nsteps_;
    qs_engine* e_ = nullptr;
};

#ifdef QPP_AMD_HAS_EIGEN
// Eigen overloads: qpp code keeps its types after the namespace swap
namespace detail {
template <typename Derived>
cmat from_eigen(const Eigen::MatrixBase<Derived>& A) {
    const Eigen::MatrixXcd M = A.template cast<cplx>();
    cmat r(M.rows(), M.cols());
    for (idx j = 0; j < r.cols(); ++j)
        for (idx i = 0; i < r.rows(); ++i) r(i, j) = M(i, j);
    return r;
}
inline Eigen::MatrixXcd to_eigen(const cmat& A) {
    Eigen::MatrixXcd M(A.rows(), A.cols());
    for (idx j = 0; j < A.cols(); ++j)
        for (idx i = 0; i < A.rows(); ++i) M(i, j) = A(i, j);
    return M;
}
}  // namespace detail
template <typename D1, typename D2>
Eigen::MatrixXcd apply(const Eigen::MatrixBase<D1>& state, const Eigen::MatrixBase<D2>& A,
                       const std::vector<idx>& target, const std::vector<idx>& dims) {
    return detail::to_eigen(apply(detail::from_eigen(state), detail::from_eigen(A), target, dims));
}
template <typename D1, typename D2>
Eigen::MatrixXcd applyCTRL(const Eigen::MatrixBase<D1>& state, const Eigen::MatrixBase<D2>& A,
                           const std::vector<idx>& ctrl, const std::vector<idx>& target,
                           const std::vector<idx>& dims,
                           std::optional<std::vector<idx>> shift = std::nullopt) {
    return detail::to_eigen(applyCTRL(detail::from_eigen(state), detail::from_eigen(A), ctrl,
                                      target, dims, shift));
}
template <typename D1>
Eigen::MatrixXcd ptrace(const Eigen::MatrixBase<D1>& A, const std::vector<idx>& target,
                        const std::vector<idx>& dims) {
    return detail::to_eigen(ptrace(detail::from_eigen(A), target, dims));
}
#endif

}  // namespace qpp_amd

#endif  // QPP_AMD_HPP_

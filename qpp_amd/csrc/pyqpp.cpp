// pyqpp.cpp -- a pybind11 module with pyqpp's names and signatures over the
// C ABI (through include/qpp_amd.hpp): the free functions of
// pyqpp/include/pyqpp/operations_bind.hpp:37-293 and instruments, and the
// QCircuit / QEngine classes of classes/qcircuit_bind.hpp and
// classes/qengine_bind.hpp:52-160.  Host code only; every state operation
// runs on the GPU through libqpp_amd.so.
//
// Matrices cross as numpy arrays: a 1-D array is a column vector, results
// come back 2-D (a ket as D x 1), as Eigen matrices do in pyqpp.  Random
// draws come from one std::mt19937 of this module (pyqpp: RandomDevices'
// generator), read through std::generate_canonical<double, 53>; seed it with
// pyqpp.seed(s).  Engines take their seed from the same generator.
#include <pybind11/pybind11.h>
#include <pybind11/numpy.h>
#include <pybind11/stl.h>

#include <cmath>
#include <complex>
#include <map>
#include <memory>
#include <optional>
#include <random>
#include <string>
#include <vector>

#include "../../include/qpp_amd.hpp"

namespace py = pybind11;
using namespace qpp_amd;
using carr = py::array_t<std::complex<double>, py::array::forcecast>;

namespace {

std::mt19937& prng() {
    static std::mt19937 g{std::random_device{}()};
    return g;
}

double draw() { return std::generate_canonical<double, 53>(prng()); }

cmat to_cmat(const py::handle& h) {
    const carr a = carr::ensure(h);
    if (!a) throw py::type_error("expected a complex matrix or vector");
    if (a.ndim() == 1) {
        auto r = a.unchecked<1>();
        cmat m(static_cast<idx>(r.shape(0)), 1);
        for (py::ssize_t i = 0; i < r.shape(0); ++i) m(static_cast<idx>(i)) = r(i);
        return m;
    }
    if (a.ndim() != 2) throw py::value_error("expected a 1-D or 2-D array");
    auto r = a.unchecked<2>();
    cmat m(static_cast<idx>(r.shape(0)), static_cast<idx>(r.shape(1)));
    for (py::ssize_t j = 0; j < r.shape(1); ++j)
        for (py::ssize_t i = 0; i < r.shape(0); ++i)
            m(static_cast<idx>(i), static_cast<idx>(j)) = r(i, j);
    return m;
}

py::array to_np(const cmat& m) {
    py::array_t<std::complex<double>> out({static_cast<py::ssize_t>(m.rows()),
                                           static_cast<py::ssize_t>(m.cols())});
    auto w = out.mutable_unchecked<2>();
    for (idx j = 0; j < m.cols(); ++j)
        for (idx i = 0; i < m.rows(); ++i) w(i, j) = m(i, j);
    return out;
}

std::vector<idx> idx_list(const py::handle& h) {
    if (py::isinstance<py::int_>(h)) return {h.cast<idx>()};
    return h.cast<std::vector<idx>>();
}

// pyqpp binds each function twice, with `dims` and with `d` (default 2)
std::vector<idx> dims_for(const cmat& A, const py::object& dims, idx d) {
    if (!dims.is_none()) return dims.cast<std::vector<idx>>();
    return detail::dims_of(A, d);
}

// qpp runs its qubit kernels (the as-built results) when every dim is 2
// and the generic code otherwise (operations.hpp:2207, 2437)
qs_semantics sem_for(const std::vector<idx>& dims) {
    for (idx x : dims)
        if (x != 2) return QS_SEM_DOCUMENTED;
    return QS_SEM_AS_BUILT;
}

std::optional<std::vector<idx>> opt_list(const py::object& o) {
    if (o.is_none()) return std::nullopt;
    return idx_list(o);
}

py::dict stats_dict(const std::map<std::vector<idx>, idx>& m) {
    py::dict d;
    for (const auto& kv : m) d[py::tuple(py::cast(kv.first))] = kv.second;
    return d;
}

// one engine class for pyqpp's QEngine / QDensityEngine / QNoisyEngine
// `seed`: the engine's std::mt19937 seed; None takes the next output of the
// module's generator (pyqpp's engines share RandomDevices' generator)
std::uint64_t seed_of(const py::object& seed) {
    return seed.is_none() ? std::uint64_t{prng()()} : seed.cast<std::uint64_t>();
}

// Polymorphic, so that a method returning the engine by reference (execute,
// set_state, ...) on a QDensityEngine / QNoisyEngine finds the existing
// Python object through the dynamic type instead of wrapping the pointer in
// a new, non-owning QEngine object that outlives its owner.
struct PyEngine {
    std::unique_ptr<QEngine> e;
    PyEngine(const QCircuit& qc, double noise_p, qs_kind kind, const py::object& seed)
        : e(new QEngine(qc, seed_of(seed), noise_p, kind)) {}
    virtual ~PyEngine() = default;
};

struct PyDensityEngine : PyEngine {
    PyDensityEngine(const QCircuit& qc, const py::object& seed) : PyEngine(qc, -1.0, QS_RHO, seed) {}
};

struct PyNoisyEngine : PyEngine {
    PyNoisyEngine(const QCircuit& qc, double p, const py::object& seed)
        : PyEngine(qc, p, QS_KET, seed) {}
};

struct DepolarizingNoise {
    double p;
};

}  // namespace

PYBIND11_MODULE(pyqpp, m) {
    m.doc() = "pyqpp's operations, instruments, QCircuit and QEngine on the GPU (qpp_amd)";

    m.def("seed", [](std::uint32_t s) { prng().seed(s); },
          "Seeds the module's std::mt19937 (RandomDevices::get_prng() in pyqpp)", py::arg("seed"));

    // operations (operations_bind.hpp:37-293) ---------------------------------
    m.def(
        "apply",
        [](const py::object& state, const py::object& A, const py::object& target,
           const py::object& dims, idx d) {
            const cmat s = to_cmat(state);
            if (py::isinstance<py::list>(A) || py::isinstance<py::tuple>(A)) {
                std::vector<cmat> Ks;
                for (auto k : A) Ks.push_back(to_cmat(k));
                if (target.is_none())  // apply(A, Ks): the whole system
                    return to_np(qpp_amd::apply(s, Ks));
                return to_np(apply(s, Ks, idx_list(target), dims_for(s, dims, d)));
            }
            return to_np(apply(s, to_cmat(A), idx_list(target), dims_for(s, dims, d)));
        },
        "Applies the gate A (or the channel with Kraus operators A) to the part "
        "target of the multi-partite state",
        py::arg("state"), py::arg("A"), py::arg("target") = py::none(),
        py::arg("dims") = py::none(), py::arg("d") = 2);
    for (const char* name : {"applyCTRL", "applyCTRL_fan"}) {
        const bool fan = std::string(name) == "applyCTRL_fan";
        m.def(
            name,
            [fan](const py::object& state, const py::object& A, const py::object& ctrl,
                  const py::object& target, const py::object& dims, idx d,
                  const py::object& shift) {
                const cmat s = to_cmat(state);
                const auto dd = dims_for(s, dims, d);
                return to_np(fan ? applyCTRL_fan(s, to_cmat(A), idx_list(ctrl), idx_list(target),
                                                 dd, opt_list(shift))
                                 : applyCTRL(s, to_cmat(A), idx_list(ctrl), idx_list(target), dd,
                                             opt_list(shift)));
            },
            "Applies the controlled gate A", py::arg("state"), py::arg("A"), py::arg("ctrl"),
            py::arg("target"), py::arg("dims") = py::none(), py::arg("d") = 2,
            py::arg("shift") = py::none());
    }
    m.def(
        "ptrace",
        [](const py::object& A, const py::object& target, const py::object& dims, idx d) {
            const cmat a = to_cmat(A);
            return to_np(ptrace(a, idx_list(target), dims_for(a, dims, d)));
        },
        "Partial trace", py::arg("A"), py::arg("target"), py::arg("dims") = py::none(),
        py::arg("d") = 2);
    m.def(
        "ptranspose",
        [](const py::object& A, const py::object& target, const py::object& dims, idx d) {
            const cmat a = to_cmat(A);
            const auto dd = dims_for(a, dims, d);
            return to_np(ptranspose(a, idx_list(target), dd, sem_for(dd)));
        },
        "Partial transpose (qpp's results for qubits, operations.hpp:2167-2335)", py::arg("A"),
        py::arg("target"), py::arg("dims") = py::none(), py::arg("d") = 2);
    m.def(
        "syspermute",
        [](const py::object& A, const py::object& perm, const py::object& dims, idx d) {
            const cmat a = to_cmat(A);
            const auto dd = dims_for(a, dims, d);
            return to_np(syspermute(a, idx_list(perm), dd, sem_for(dd)));
        },
        "Subsystem permutation (qpp's results for qubits, operations.hpp:2386-2514)", py::arg("A"),
        py::arg("perm"), py::arg("dims") = py::none(), py::arg("d") = 2);

    // instruments ----------------------------------------------------------------
    m.def(
        "sample",
        [](const py::object& A, const py::object& target, const py::object& dims, idx d) {
            const cmat a = to_cmat(A);
            return sample(a, idx_list(target), dims_for(a, dims, d), draw());
        },
        "Samples the target qudits (instruments.hpp:746-929)", py::arg("A"), py::arg("target"),
        py::arg("dims") = py::none(), py::arg("d") = 2);
    m.def(
        "measure_seq",
        [](const py::object& A, const py::object& target, const py::object& dims, idx d,
           bool destructive) {
            const cmat a = to_cmat(A);
            auto [r, p, s] = measure_seq(a, idx_list(target), dims_for(a, dims, d), destructive,
                                         draw());
            return py::make_tuple(r, p, to_np(s));
        },
        "Sequential Z measurements (instruments.hpp:985-1099)", py::arg("A"), py::arg("target"),
        py::arg("dims") = py::none(), py::arg("d") = 2, py::arg("destructive") = true);
    m.def(
        "reset",
        [](const py::object& A, const py::object& target, const py::object& dims, idx d) {
            const cmat a = to_cmat(A);
            return to_np(reset(a, idx_list(target), dims_for(a, dims, d), draw()));
        },
        "Resets the target qudits to |0> (instruments.hpp:1148-1196)", py::arg("A"),
        py::arg("target"), py::arg("dims") = py::none(), py::arg("d") = 2);

    // QCircuit ----------------------------------------------------------------------
    py::class_<QCircuit>(m, "QCircuit")
        .def(py::init<idx, idx, idx>(), py::arg("nq") = 1, py::arg("nc") = 0, py::arg("d") = 2)
        .def("get_nq", &QCircuit::get_nq)
        .def("get_nc", &QCircuit::get_nc)
        .def(
            "gate",
            [](QCircuit& c, const py::object& U, const py::args& targets) -> QCircuit& {
                std::vector<idx> t;
                for (auto x : targets)
                    if (py::isinstance<py::int_>(x)) t.push_back(x.cast<idx>());
                if (t.empty() || t.size() > 3) throw py::value_error("gate(U, i[, j[, k]])");
                return c.gate_joint(to_cmat(U), t);
            },
            py::return_value_policy::reference)
        .def(
            "gate_joint",
            [](QCircuit& c, const py::object& U, const py::object& t) -> QCircuit& {
                return c.gate_joint(to_cmat(U), idx_list(t));
            },
            py::arg("U"), py::arg("target"), py::return_value_policy::reference)
        .def(
            "gate_fan",
            [](QCircuit& c, const py::object& U, const py::object& t) -> QCircuit& {
                return c.gate_fan(to_cmat(U), idx_list(t));
            },
            py::arg("U"), py::arg("target"), py::return_value_policy::reference)
        .def(
            "CTRL",
            [](QCircuit& c, const py::object& U, const py::object& ctrl, const py::object& t,
               const py::object& shift) -> QCircuit& {
                return c.CTRL(to_cmat(U), idx_list(ctrl), idx_list(t), opt_list(shift));
            },
            py::arg("U"), py::arg("ctrl"), py::arg("target"), py::arg("shift") = py::none(),
            py::return_value_policy::reference)
        .def(
            "CTRL_fan",
            [](QCircuit& c, const py::object& U, const py::object& ctrl, const py::object& t,
               const py::object& shift) -> QCircuit& {
                return c.CTRL_fan(to_cmat(U), idx_list(ctrl), idx_list(t), opt_list(shift));
            },
            py::arg("U"), py::arg("ctrl"), py::arg("target"), py::arg("shift") = py::none(),
            py::return_value_policy::reference)
        .def(
            "cCTRL",
            [](QCircuit& c, const py::object& U, const py::object& ctrl, const py::object& t,
               const py::object& shift) -> QCircuit& {
                return c.cCTRL(to_cmat(U), idx_list(ctrl), idx_list(t), opt_list(shift));
            },
            py::arg("U"), py::arg("ctrl_dits"), py::arg("target"), py::arg("shift") = py::none(),
            py::return_value_policy::reference)
        .def(
            "cCTRL_fan",
            [](QCircuit& c, const py::object& U, const py::object& ctrl, const py::object& t,
               const py::object& shift) -> QCircuit& {
                return c.cCTRL_fan(to_cmat(U), idx_list(ctrl), idx_list(t), opt_list(shift));
            },
            py::arg("U"), py::arg("ctrl_dits"), py::arg("target"), py::arg("shift") = py::none(),
            py::return_value_policy::reference)
        .def(
            "measure",
            [](QCircuit& c, const py::object& t, idx c_reg, bool destructive) -> QCircuit& {
                if (py::isinstance<py::int_>(t)) return c.measure(t.cast<idx>(), c_reg, destructive);
                return c.measure(idx_list(t), c_reg, destructive);
            },
            py::arg("target"), py::arg("c_reg") = 0, py::arg("destructive") = true,
            py::return_value_policy::reference)
        .def(
            "post_select",
            [](QCircuit& c, const py::object& t, const py::object& ps, idx c_reg,
               bool destructive) -> QCircuit& {
                return c.post_select(idx_list(t), idx_list(ps), c_reg, destructive);
            },
            py::arg("target"), py::arg("ps_vals"), py::arg("c_reg") = 0,
            py::arg("destructive") = true, py::return_value_policy::reference)
        .def(
            "reset",
            [](QCircuit& c, const py::object& t) -> QCircuit& { return c.reset(idx_list(t)); },
            py::arg("target"), py::return_value_policy::reference)
        .def(
            "discard",
            [](QCircuit& c, const py::object& t) -> QCircuit& { return c.discard(idx_list(t)); },
            py::arg("target"), py::return_value_policy::reference)
        .def(
            "cond_if",
            [](QCircuit& c, const py::object& dits, const py::object& vals) -> QCircuit& {
                return c.cond_if(idx_list(dits), idx_list(vals));
            },
            "Runtime conditional: true when dit dits[j] equals vals[j] for every j",
            py::arg("dits"), py::arg("vals"), py::return_value_policy::reference)
        .def(
            "cond_if",
            [](QCircuit& c, const py::function& pred) -> QCircuit& {
                return c.cond_if([&](const std::vector<idx>& v) { return pred(v).cast<bool>(); });
            },
            "Runtime conditional with pyqpp's predicate form: pred(dits) over the list of "
            "every classical dit (stored as a truth table over the values 0..d-1)",
            py::arg("pred"), py::return_value_policy::reference)
        .def(
            "cond_while",
            [](QCircuit& c, const py::object& dits, const py::object& vals) -> QCircuit& {
                return c.cond_while(idx_list(dits), idx_list(vals));
            },
            "Runtime while loop with the cond_if condition form", py::arg("dits"),
            py::arg("vals"), py::return_value_policy::reference)
        .def(
            "cond_while",
            [](QCircuit& c, const py::function& pred) -> QCircuit& {
                return c.cond_while([&](const std::vector<idx>& v) { return pred(v).cast<bool>(); });
            },
            "Runtime while loop with pyqpp's predicate form", py::arg("pred"),
            py::return_value_policy::reference)
        .def("cond_else", &QCircuit::cond_else, py::return_value_policy::reference)
        .def("cond_end", &QCircuit::cond_end, py::return_value_policy::reference)
        .def("nop", &QCircuit::nop, py::return_value_policy::reference);

    // engines (qengine_bind.hpp:52-160) ------------------------------------------
    py::class_<DepolarizingNoise>(m, "QubitDepolarizingNoise")
        .def(py::init([](double p) { return DepolarizingNoise{p}; }), py::arg("p"));
    auto eng = py::class_<PyEngine>(m, "QEngine");
    eng.def(py::init([](const QCircuit& qc, const py::object& seed) {
                return new PyEngine(qc, -1.0, QS_KET, seed);
            }),
            py::arg("qc"), py::arg("seed") = py::none())
        .def(
            "execute", [](PyEngine& e, idx reps) -> PyEngine& { e.e->execute(reps); return e; },
            py::arg("reps") = 1, py::return_value_policy::reference)
        .def("get_dit", [](const PyEngine& e, idx i) { return e.e->get_dit(i); }, py::arg("i"))
        .def("get_dits", [](const PyEngine& e) { return e.e->get_dits(); })
        .def("get_probs", [](const PyEngine& e) { return e.e->get_probs(); })
        .def("get_state", [](const PyEngine& e) { return to_np(e.e->get_state()); })
        .def(
            "set_state",
            [](PyEngine& e, const py::object& s) -> PyEngine& {
                e.e->set_state(to_cmat(s));
                return e;
            },
            py::arg("state"), py::return_value_policy::reference)
        .def("get_stats", [](const PyEngine& e) { return stats_dict(e.e->get_stats()); })
        .def("post_select_ok", [](const PyEngine& e) { return e.e->post_select_ok(); })
        .def("get_noise_results", [](const PyEngine& e) { return e.e->get_noise_results(); })
        .def("get_measured_d", [](const PyEngine& e) { return e.e->get_measured_d(); })
        .def("get_non_measured_d", [](const PyEngine& e) { return e.e->get_non_measured_d(); })
        .def("was_measured_d", [](const PyEngine& e, idx i) { return e.e->was_measured_d(i); },
             py::arg("i"))
        .def(
            "set_dit",
            [](PyEngine& e, idx i, idx v) -> PyEngine& { e.e->set_dit(i, v); return e; },
            py::arg("i"), py::arg("value"), py::return_value_policy::reference)
        .def(
            "set_dits",
            [](PyEngine& e, const std::vector<idx>& d) -> PyEngine& { e.e->set_dits(d); return e; },
            py::arg("dits"), py::return_value_policy::reference)
        .def(
            "reset_stats", [](PyEngine& e) -> PyEngine& { e.e->reset_stats(); return e; },
            py::return_value_policy::reference)
        .def(
            "set_ensure_post_selection",
            [](PyEngine& e, bool v) -> PyEngine& { e.e->set_ensure_post_selection(v); return e; },
            py::arg("val"), py::return_value_policy::reference)
        .def("get_ensure_post_selection",
             [](const PyEngine& e) { return e.e->get_ensure_post_selection(); })
        .def(
            "set_max_post_selection_reps",
            [](PyEngine& e, idx n) -> PyEngine& { e.e->set_max_post_selection_reps(n); return e; },
            py::arg("max_post_selection_reps"), py::return_value_policy::reference)
        .def("get_max_post_selection_reps",
             [](const PyEngine& e) { return e.e->get_max_post_selection_reps(); })
        .def(
            "reset", [](PyEngine& e) -> PyEngine& { e.e->reset(); return e; },
            py::return_value_policy::reference);
    m.attr("QKetEngine") = eng;
    py::class_<PyDensityEngine, PyEngine>(m, "QDensityEngine")
        .def(py::init([](const QCircuit& qc, const py::object& seed) {
                 return new PyDensityEngine(qc, seed);
             }),
             py::arg("qc"), py::arg("seed") = py::none());
    auto noisy = py::class_<PyNoisyEngine, PyEngine>(m, "QNoisyEngine");
    noisy.def(py::init([](const QCircuit& qc, const DepolarizingNoise& n, const py::object& seed) {
                  return new PyNoisyEngine(qc, n.p, seed);
              }),
              py::arg("qc"), py::arg("noise"), py::arg("seed") = py::none());
    m.attr("QKetNoisyEngine") = noisy;

    // gates and states (pyqpp.gates / pyqpp.states) --------------------------------
    auto gt = m.def_submodule("gates", "qpp::Gates constants");
    const double s2 = 1 / std::sqrt(2.0);
    const std::complex<double> I(0, 1);
    auto mk = [](std::initializer_list<std::complex<double>> rm, idx D) {
        cmat M(D, D);
        idx k = 0;
        for (auto z : rm) {
            M(k / D, k % D) = z;
            ++k;
        }
        return to_np(M);
    };
    gt.attr("Id2") = mk({1, 0, 0, 1}, 2);
    gt.attr("H") = mk({s2, s2, s2, -s2}, 2);
    gt.attr("X") = mk({0, 1, 1, 0}, 2);
    gt.attr("Y") = mk({0, -I, I, 0}, 2);
    gt.attr("Z") = mk({1, 0, 0, -1}, 2);
    gt.attr("S") = mk({1, 0, 0, I}, 2);
    gt.attr("T") = mk({1, 0, 0, std::exp(I * (M_PI / 4))}, 2);
    gt.attr("CNOT") = mk({1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0}, 4);
    gt.attr("CZ") = mk({1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, -1}, 4);
    gt.attr("SWAP") = mk({1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 1}, 4);
    auto st = m.def_submodule("states", "qpp::States constants");
    st.attr("z0") = to_np([] { cmat k(2, 1); k(0) = 1; return k; }());
    st.attr("z1") = to_np([] { cmat k(2, 1); k(1) = 1; return k; }());
    st.attr("x0") = to_np([&] { cmat k(2, 1); k(0) = s2; k(1) = s2; return k; }());
    st.attr("x1") = to_np([&] { cmat k(2, 1); k(0) = s2; k(1) = -s2; return k; }());

    // qasm (qasm.hpp:794-823): the Python reader's circuit as a QCircuit here
    auto qasm = m.def_submodule("qasm", "OpenQASM 2.0 reader");
    auto convert = [](const py::object& pc) {
        QCircuit c(pc.attr("nq").cast<idx>(), pc.attr("nc").cast<idx>(), pc.attr("d").cast<idx>());
        const py::list mats = pc.attr("mats");
        const py::dict ps = pc.attr("ps_vals");
        idx i = 0;
        for (auto h : pc.attr("steps")) {
            const py::tuple s = h.cast<py::tuple>();
            const uint32_t type = s[0].cast<uint32_t>();
            const auto ctrl = s[2].cast<std::vector<idx>>();
            const auto tgt = s[3].cast<std::vector<idx>>();
            const auto shift = opt_list(s[4]);
            const idx c_reg = s[5].cast<idx>();
            const bool dest = s[6].cast<int>() != 0;
            auto U = [&] { return to_cmat(mats[s[1].cast<size_t>()]); };
            switch (type) {
                case QS_STEP_GATE: c.gate_joint(U(), tgt); break;
                case QS_STEP_FAN: c.gate_fan(U(), tgt); break;
                case QS_STEP_CTRL: c.CTRL(U(), ctrl, tgt, shift); break;
                case QS_STEP_CTRL_FAN: c.CTRL_fan(U(), ctrl, tgt, shift); break;
                case QS_STEP_CCTRL: c.cCTRL(U(), ctrl, tgt, shift); break;
                case QS_STEP_CCTRL_FAN: c.cCTRL_fan(U(), ctrl, tgt, shift); break;
                case QS_STEP_MEASURE: c.measure(tgt[0], c_reg, dest); break;
                case QS_STEP_MEASURE_MANY: c.measure(tgt, c_reg, dest); break;
                case QS_STEP_RESET: c.reset(tgt); break;
                case QS_STEP_DISCARD: c.discard(tgt); break;
                case QS_STEP_POST_SELECT:
                    c.post_select(tgt, ps[py::int_(i)].cast<std::vector<idx>>(), c_reg, dest);
                    break;
                case QS_STEP_COND_IF: c.cond_if(ctrl, ps[py::int_(i)].cast<std::vector<idx>>()); break;
                case QS_STEP_COND_WHILE:
                    c.cond_while(ctrl, ps[py::int_(i)].cast<std::vector<idx>>());
                    break;
                case QS_STEP_COND_IF_TABLE:
                case QS_STEP_COND_WHILE_TABLE:
                    c.cond_table(type, ctrl, ps[py::int_(i)].cast<std::vector<idx>>());
                    break;
                case QS_STEP_COND_ELSE: c.cond_else(); break;
                case QS_STEP_COND_END: c.cond_end(); break;
                default: c.nop(); break;
            }
            ++i;
        }
        return c;
    };
    qasm.def(
        "read_from_file",
        [convert](const std::string& path) {
            return convert(py::module_::import("qpp_amd.qasm").attr("read_from_file")(path));
        },
        py::arg("fname"));
    qasm.def(
        "read_from_string",
        [convert](const std::string& text) {
            return convert(py::module_::import("qpp_amd.qasm").attr("read_from_string")(text));
        },
        py::arg("qasm_string"));

    // qpp's exceptions surface as RuntimeError with qpp's what() text
    py::register_exception_translator([](std::exception_ptr p) {
        try {
            if (p) std::rethrow_exception(p);
        } catch (const exception::Exception& e) {
            PyErr_SetString(PyExc_RuntimeError, e.what());
        }
    });
}

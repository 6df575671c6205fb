// Drop-in check of include/qpp_amd.hpp: qpp-style calls through the C++
// wrapper (value semantics, exceptions, QCircuit/QEngine), each compared with
// the CPU oracle on the same inputs.  Exit status 0 = all checks passed.
// Test program; links liboracle.so as the checker.
#include <cmath>
#include <complex>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <random>
#include <vector>

#include <unistd.h>

#include "../../include/qpp_amd.hpp"
#include "../../oracle/oracle.h"

using namespace qpp_amd;

static int failures = 0;
#define CHECK(cond, what)                                              \
    do {                                                               \
        if (!(cond)) {                                                 \
            std::printf("FAIL %s (line %d)\n", what, __LINE__);        \
            ++failures;                                                \
        } else {                                                       \
            std::printf("ok   %s\n", what);                            \
        }                                                              \
    } while (0)

static cmat haar(idx D, std::mt19937& g) {
    std::normal_distribution<double> nd;
    cmat A(D, D);
    for (auto& x : A.v) x = cplx(nd(g), nd(g));
    // Gram-Schmidt on columns
    for (idx j = 0; j < D; ++j) {
        for (idx k = 0; k < j; ++k) {
            cplx s = 0;
            for (idx i = 0; i < D; ++i) s += std::conj(A(i, k)) * A(i, j);
            for (idx i = 0; i < D; ++i) A(i, j) -= s * A(i, k);
        }
        double nrm = 0;
        for (idx i = 0; i < D; ++i) nrm += std::norm(A(i, j));
        nrm = std::sqrt(nrm);
        for (idx i = 0; i < D; ++i) A(i, j) /= nrm;
    }
    return A;
}

static cmat randket(idx D, std::mt19937& g) {
    std::normal_distribution<double> nd;
    cmat k(D, 1);
    double s = 0;
    for (auto& x : k.v) {
        x = cplx(nd(g), nd(g));
        s += std::norm(x);
    }
    for (auto& x : k.v) x /= std::sqrt(s);
    return k;
}

static cmat randrho(idx D, std::mt19937& g) {
    cmat k = randket(D, g), r(D, D);
    cmat l = randket(D, g);
    for (idx i = 0; i < D; ++i)
        for (idx j = 0; j < D; ++j)
            r(i, j) = 0.6 * k(i) * std::conj(k(j)) + 0.4 * l(i) * std::conj(l(j));
    return r;
}

static double maxdiff(const cmat& a, const std::vector<cplx>& b) {
    if (a.size() != b.size()) return 1e300;
    double m = 0;
    for (idx i = 0; i < b.size(); ++i) m = std::max(m, std::abs(a.v[i] - b[i]));
    return m;
}

static const double* dp(const cmat& a) { return reinterpret_cast<const double*>(a.data()); }

int main() {
    std::mt19937 g(12345);
    const idx n = 7;
    const std::vector<idx> dims(n, 2);
    const std::vector<uint64_t> dims64(n, 2);
    const idx D = 1u << n;
    cmat psi = randket(D, g), rho = randrho(D, g);
    cmat U1 = haar(2, g), U2 = haar(4, g);

    {  // apply on a ket and a density matrix
        std::vector<idx> t{5, 2};
        std::vector<uint64_t> t64{5, 2};
        std::vector<cplx> ref(D), refr(D * D);
        orc_apply(dp(psi), 0, n, dims64.data(), dp(U2), 2, t64.data(),
                  reinterpret_cast<double*>(ref.data()));
        orc_apply(dp(rho), 1, n, dims64.data(), dp(U2), 2, t64.data(),
                  reinterpret_cast<double*>(refr.data()));
        CHECK(maxdiff(apply(psi, U2, t, 2), ref) <= 1e-12, "apply ket");
        CHECK(maxdiff(apply(rho, U2, t, dims), refr) <= 1e-12, "apply rho");
    }
    {  // applyCTRL with shift, applyCTRL_fan
        std::vector<idx> c{0, 6}, t{3};
        std::vector<uint64_t> c64{0, 6}, t64{3}, s64{1, 0};
        std::vector<cplx> ref(D), reff(D);
        orc_apply_ctrl(dp(psi), 0, n, dims64.data(), dp(U1), 2, c64.data(), 1, t64.data(),
                       s64.data(), reinterpret_cast<double*>(ref.data()));
        CHECK(maxdiff(applyCTRL(psi, U1, c, t, 2, std::vector<idx>{1, 0}), ref) <= 1e-12,
              "applyCTRL shift");
        std::vector<idx> tf{1, 4};
        std::vector<uint64_t> tf64{1, 4};
        orc_apply_ctrl_fan(dp(psi), 0, n, dims64.data(), dp(U1), 2, c64.data(), 2, tf64.data(),
                           nullptr, reinterpret_cast<double*>(reff.data()));
        CHECK(maxdiff(applyCTRL_fan(psi, U1, c, tf), reff) <= 1e-12, "applyCTRL_fan");
    }
    {  // Kraus channel (amplitude damping) on a density matrix
        const double gam = 0.3;
        cmat K0(2, 2), K1(2, 2);
        K0(0, 0) = 1;
        K0(1, 1) = std::sqrt(1 - gam);
        K1(0, 1) = std::sqrt(gam);
        std::vector<cplx> ks(K0.v);
        ks.insert(ks.end(), K1.v.begin(), K1.v.end());
        std::vector<uint64_t> t64{4};
        std::vector<cplx> ref(D * D);
        orc_apply_kraus(dp(rho), n, dims64.data(), 2, reinterpret_cast<double*>(ks.data()), 1,
                        t64.data(), reinterpret_cast<double*>(ref.data()));
        CHECK(maxdiff(apply(rho, std::vector<cmat>{K0, K1}, {4}, 2), ref) <= 1e-12, "apply Kraus");
    }
    {  // ptrace / ptranspose / syspermute
        std::vector<uint64_t> t64{1, 5};
        std::vector<cplx> ref(32 * 32);
        orc_ptrace(dp(rho), 1, n, dims64.data(), 2, t64.data(),
                   reinterpret_cast<double*>(ref.data()));
        CHECK(maxdiff(ptrace(rho, {1, 5}, dims), ref) <= 1e-12, "ptrace rho");
        orc_ptrace(dp(psi), 0, n, dims64.data(), 2, t64.data(),
                   reinterpret_cast<double*>(ref.data()));
        CHECK(maxdiff(ptrace(psi, {1, 5}, dims), ref) <= 1e-12, "ptrace ket");
        std::vector<cplx> refpt(D * D);
        orc_ptranspose(dp(rho), 1, n, dims64.data(), 2, t64.data(), 0,
                       reinterpret_cast<double*>(refpt.data()));
        CHECK(maxdiff(ptranspose(rho, {1, 5}, dims), refpt) == 0, "ptranspose");
        std::vector<uint64_t> p64{3, 0, 6, 1, 5, 2, 4};
        std::vector<idx> p(p64.begin(), p64.end());
        std::vector<cplx> refsp(D * D);
        orc_syspermute(dp(rho), 1, n, dims64.data(), p64.data(), 0,
                       reinterpret_cast<double*>(refsp.data()));
        CHECK(maxdiff(syspermute(rho, p, dims), refsp) == 0, "syspermute");
    }
    {  // sample / measure_seq with explicit uniforms
        std::vector<uint64_t> t64{2, 0, 6};
        const double u = 0.61803398875;
        std::vector<uint64_t> dig(3);
        double pr = 0;
        uint64_t sidx = 0;
        orc_sample(dp(psi), 0, n, dims64.data(), 3, t64.data(), u, &sidx, dig.data(), &pr);
        auto [res, p] = sample(psi, {2, 0, 6}, dims, u);
        CHECK(res == std::vector<idx>(dig.begin(), dig.end()) && p == pr, "sample");
        std::vector<uint64_t> r3(3);
        std::vector<double> p3(3);
        std::vector<cplx> post(D);
        uint64_t plen = 0;
        orc_measure_seq(dp(psi), 0, n, dims64.data(), 3, t64.data(), 1, u, r3.data(), p3.data(),
                        reinterpret_cast<double*>(post.data()), &plen);
        post.resize(plen);
        auto [mr, mp, mst] = measure_seq(psi, {2, 0, 6}, dims, true, u);
        // outcomes bit-exact; probabilities and the post-measurement state
        // come from reductions and agree to 1e-12
        double pd = 0;
        for (size_t i = 0; i < p3.size(); ++i) pd = std::max(pd, std::abs(mp[i] - p3[i]));
        CHECK(mr == std::vector<idx>(r3.begin(), r3.end()) && pd <= 1e-12 &&
                  maxdiff(mst, post) <= 1e-12,
              "measure_seq");
    }
    {  // error behaviour: qpp's exception types
        bool threw = false;
        try {
            apply(psi, U2, {3, 3}, 2);
        } catch (const exception::Exception& e) {
            threw = true;
            std::printf("     %s\n", e.what());
        }
        CHECK(threw, "duplicate targets throw");
        threw = false;
        try {
            ptrace(rho, {0}, std::vector<idx>{2, 2, 2});
        } catch (const exception::DimsMismatchMatrix&) {
            threw = true;
        }
        CHECK(threw, "DimsMismatchMatrix");
    }
    {  // qpp::save / load and obsolete::save / load through device states
        const std::string base = "/tmp/qpp_amd_dropin_" + std::to_string(::getpid());
        bool same = true;
        for (qs_format fmt : {QS_FMT_TEXT, QS_FMT_BINARY}) {
            const std::string path = base + (fmt == QS_FMT_TEXT ? ".txt" : ".dat");
            State(rho, dims).save(path, fmt);
            same = same && maxdiff(State::load(path, dims, QS_RHO, QS_C128, fmt).download(),
                                   rho.v) == 0.0;
            State(psi, dims).save(path, fmt);
            same = same && maxdiff(State::load(path, dims, QS_KET, QS_C128, fmt).download(),
                                   psi.v) == 0.0;
            std::remove(path.c_str());
        }
        CHECK(same, "save/load round trips (text, binary)");
        bool threw = false;
        try {
            State::load(base + ".missing", dims);
        } catch (const std::runtime_error& e) {
            threw = std::string(e.what()) == "qpp::load(): Error opening input stream!";
        }
        CHECK(threw, "load of a missing file throws qpp's runtime_error");
    }
    {  // QCircuit + QEngine against the oracle engine
        QCircuit qc(5, 5);
        for (idx q = 0; q < 5; ++q) qc.gate(haar(2, g), q);
        cmat X(2, 2);
        X(0, 1) = X(1, 0) = 1;
        qc.CTRL(X, 0, 3).CTRL(X, 2, 4).gate(haar(4, g), 1, 3);
        qc.measure({0, 1, 2, 3, 4}, 0);
        QEngine e(qc, 777);
        e.execute(200);
        auto stats = e.get_stats();
        const auto& prog = qc.program();
        uint64_t nout = 0;
        std::vector<uint64_t> od(64 * 5), oc(64), dits(5);
        std::vector<double> probs(5);
        std::vector<double> st(2 * 32);
        uint64_t st_len = 0;
        orc_engine_run(&prog, 0, -1.0, 777, 200, 64, &nout, od.data(), oc.data(), dits.data(),
                       probs.data(), st.data(), &st_len);
        std::map<std::vector<idx>, idx> ref;
        for (uint64_t i = 0; i < nout; ++i)
            ref[std::vector<idx>(od.begin() + i * 5, od.begin() + (i + 1) * 5)] = oc[i];
        CHECK(stats == ref, "QEngine statistics");
        CHECK(e.get_dits() == std::vector<idx>(dits.begin(), dits.end()), "QEngine dits");
    }
    {  // qpp::reset, and circuits with cCTRL, post-selection and if/else
        std::vector<cplx> ref(D);
        std::vector<uint64_t> t64{3, 6};
        orc_reset(dp(psi), 0, n, dims64.data(), 2, t64.data(), 0.37,
                  reinterpret_cast<double*>(ref.data()));
        CHECK(maxdiff(reset(psi, {3, 6}, dims, 0.37), ref) <= 1e-12, "reset");
        cmat X(2, 2), Z(2, 2), H(2, 2);
        X(0, 1) = X(1, 0) = 1;
        Z(0, 0) = 1;
        Z(1, 1) = -1;
        H(0, 0) = H(0, 1) = H(1, 0) = 1 / std::sqrt(2.0);
        H(1, 1) = -1 / std::sqrt(2.0);
        QCircuit qc(3, 3);
        qc.gate(haar(2, g), 0).gate(H, 1).CTRL(X, 1, 2).CTRL(X, 0, 1).gate(H, 0);
        qc.measure(0, 0).measure(1, 1).cCTRL(X, 1, 2).cCTRL(Z, 0, 2);
        qc.gate(H, 2).post_select(2, 1, 2, false);
        qc.cond_if({2}, {1}).gate(X, 2).cond_else().gate(Z, 2).cond_end();
        const auto& prog = qc.program();
        for (idx reps : {idx{1}, idx{40}}) {
            QEngine e(qc, 99);
            e.set_ensure_post_selection(true).set_max_post_selection_reps(3);
            e.execute(reps);
            uint64_t nout = 0;
            std::vector<uint64_t> od(64 * 3), oc(64), dits(3);
            std::vector<double> probs(3), st(2 * 2);
            uint64_t st_len = 0;
            int ok = 0;
            orc_engine_run_ps(&prog, 0, -1.0, 99, reps, 1, 3, 64, &nout, od.data(), oc.data(),
                              dits.data(), probs.data(), st.data(), &st_len, &ok);
            std::map<std::vector<idx>, idx> ref_stats;
            for (uint64_t i = 0; i < nout; ++i)
                ref_stats[std::vector<idx>(od.begin() + i * 3, od.begin() + (i + 1) * 3)] = oc[i];
            const cmat fin = e.get_state();
            std::vector<cplx> fin_ref(reinterpret_cast<cplx*>(st.data()),
                                      reinterpret_cast<cplx*>(st.data()) + st_len);
            CHECK(e.get_stats() == ref_stats && e.post_select_ok() == (ok != 0) &&
                      e.get_dits() == std::vector<idx>(dits.begin(), dits.end()) &&
                      maxdiff(fin, fin_ref) <= 1e-12,
                  reps == 1 ? "engine: cCTRL, post-selection, if/else (reps 1)"
                            : "engine: cCTRL, post-selection, if/else (reps 40)");
        }
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}

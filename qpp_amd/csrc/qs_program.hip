// qs_program.hip -- C ABI of fused programs: lowering of circuit gate steps
// (qpp::QCircuit gate / CTRL / FAN steps, classes/qcircuit.hpp:1166-2372,
// executed by QEngineT::visit_gate_step_, classes/qengine.hpp:533-579) to
// bit-level gates, compiled into tile passes by qs_fused.hip.
#include <hip/hip_runtime.h>

#include <complex>
#include <vector>

#include "qs_fused.hpp"
#include "qs_internal.hpp"

struct qs_program {
    FusedProgram P;
    uint64_t ngates = 0;
    uint32_t grid_cap = 2048;
    // the pass launches captured as a HIP graph for one state buffer on one
    // stream (qs_program_capture); runs on that buffer and stream replay it
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    void* gd = nullptr;
    hipStream_t gs = nullptr;
};

namespace {

void drop_graph(qs_program* p) {
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->gexec = nullptr;
    p->graph = nullptr;
    p->gd = nullptr;
    p->gs = nullptr;
}

}  // namespace

namespace {

using qs::fail;
using C = std::complex<double>;

std::vector<C> mat_of(const qs_circuit* qc, uint32_t g, bool conj) {
    const uint64_t d = qc->mat_dim[g];
    const double* p = qc->mats + 2 * qc->mat_off[g];
    std::vector<C> m(d * d);
    for (uint64_t i = 0; i < d * d; ++i) m[i] = C(p[2 * i], conj ? -p[2 * i + 1] : p[2 * i + 1]);
    return m;
}

}  // namespace

namespace qs {

// Lowers gate steps to bit gates: a ket over n qubits uses bit n-1-q; a
// density matrix applies U on the row bits and conj(U) on the column bits
// (bit 2n-1-q), each under the controls of its own index.
int lower_steps(const qs_circuit* qc, uint32_t b, uint32_t e, bool rho,
                std::vector<BitGate>& out) {
    const uint32_t n = qc->nq;
    auto bit = [&](uint64_t q, bool col) { return static_cast<uint32_t>((col ? 2 * n : n) - 1 - q); };
    for (uint32_t s = b; s < e; ++s) {
        const qs_step& st = qc->steps[s];
        if (st.type == QS_STEP_NOP) continue;
        if (st.type == QS_STEP_MEASURE || st.type == QS_STEP_MEASURE_MANY ||
            st.type == QS_STEP_POST_SELECT || st.type == QS_STEP_RESET ||
            st.type == QS_STEP_DISCARD)
            return fail(QS_ERR_NOT_IMPLEMENTED, "qpp_amd::program", "measurement step in a program");
        if (st.type > QS_STEP_CTRL_FAN)
            return fail(QS_ERR_NOT_IMPLEMENTED, "qpp_amd::program", "runtime or classically controlled step");
        const uint64_t* tg = qc->pool + st.target_off;
        const uint64_t* cg = qc->pool + st.ctrl_off;
        const uint64_t* sh = st.shift_off == UINT32_MAX ? nullptr : qc->pool + st.shift_off;
        const bool fan = st.type == QS_STEP_FAN || st.type == QS_STEP_CTRL_FAN;
        const bool ctrl = st.type == QS_STEP_CTRL || st.type == QS_STEP_CTRL_FAN;
        std::vector<std::vector<uint64_t>> groups;
        if (fan)
            for (uint32_t i = 0; i < st.ntarget; ++i) groups.push_back({tg[i]});
        else
            groups.push_back(std::vector<uint64_t>(tg, tg + st.ntarget));
        for (const auto& g : groups) {
            if (g.size() > 2)
                return fail(QS_ERR_NOT_IMPLEMENTED, "qpp_amd::program", "gate on more than 2 qubits");
            for (int pass = 0; pass < (rho ? 2 : 1); ++pass) {
                const bool col = pass == 1;
                BitGate bg;
                for (uint64_t q : g) bg.bits.push_back(bit(q, col));
                bg.m = mat_of(qc, st.gate, col);
                if (ctrl)
                    for (uint32_t c = 0; c < st.nctrl; ++c) {
                        const uint64_t m = uint64_t{1} << bit(cg[c], col);
                        bg.cmask |= m;
                        if (!sh || sh[c] == 0) bg.cval |= m;
                    }
                out.push_back(bg);
            }
        }
    }
    return QS_OK;
}

}  // namespace qs

extern "C" {

int qs_program_create(const qs_circuit* qc, uint32_t step_begin, uint32_t step_end,
                      qs_kind kind, qs_dtype dtype, const qs_fuse_opts* opts,
                      qs_program** out) {
    constexpr const char* W = "qpp_amd::program";
    if (!qc || !out) return fail(QS_ERR_INVALID_HANDLE, W, "circuit");
    *out = nullptr;
    if (qc->d != 2) return fail(QS_ERR_NOT_IMPLEMENTED, W, "qudit circuit");
    if (step_end > qc->nsteps || step_begin > step_end)
        return fail(QS_ERR_OUT_OF_RANGE, W, "steps");
    if (int rc = qs::ensure_device()) return rc;
    std::vector<BitGate> gates;
    if (int rc = qs::lower_steps(qc, step_begin, step_end, kind == QS_RHO, gates)) return rc;
    FusedOptions o;
    uint32_t cap = 2048;  // resident workgroups walk the tiles
    if (opts) {
        if (opts->reg_bits) o.reg_bits = opts->reg_bits;
        if (opts->low_bits) o.low_bits = opts->low_bits;
        if (opts->max_ops) o.max_ops = opts->max_ops;
        if (opts->grid_cap) cap = opts->grid_cap;
    }
    if (o.reg_bits != 3 && o.reg_bits != 4) return fail(QS_ERR_OUT_OF_RANGE, W, "reg_bits");
    auto* p = new qs_program;
    p->ngates = gates.size();
    p->grid_cap = cap;
    const uint32_t nb = (kind == QS_RHO ? 2 : 1) * qc->nq;
    if (int rc = qs::fused_compile(gates, nb, o, p->P)) {
        delete p;
        return rc;
    }
    if (int rc = qs::fused_upload(p->P, dtype)) {
        delete p;
        return rc;
    }
    *out = p;
    return QS_OK;
}

int qs_program_create_bits(uint32_t nb, uint32_t ngates, const qs_bitgate* gates,
                           const double* mats, qs_dtype dtype, const qs_fuse_opts* opts,
                           qs_program** out) {
    constexpr const char* W = "qpp_amd::program";
    if (!out || (ngates && (!gates || !mats))) return fail(QS_ERR_INVALID_HANDLE, W);
    *out = nullptr;
    if (nb == 0 || nb > 48) return fail(QS_ERR_OUT_OF_RANGE, W, "nb");
    std::vector<BitGate> bg(ngates);
    for (uint32_t i = 0; i < ngates; ++i) {
        const qs_bitgate& g = gates[i];
        if (g.nbits < 1 || g.nbits > 2) return fail(QS_ERR_NOT_IMPLEMENTED, W, "gate width");
        uint64_t used = 0;
        for (uint32_t k = 0; k < g.nbits; ++k) {
            if (g.bits[k] >= nb || ((used >> g.bits[k]) & 1))
                return fail(QS_ERR_SUBSYS_MISMATCH_DIMS, W, "bits");
            used |= uint64_t{1} << g.bits[k];
        }
        if ((g.cmask & used) || (nb < 64 && (g.cmask >> nb)) || (g.cval & ~g.cmask))
            return fail(QS_ERR_SUBSYS_MISMATCH_DIMS, W, "control");
        bg[i].bits.assign(g.bits, g.bits + g.nbits);
        const uint64_t d = uint64_t{1} << g.nbits;
        bg[i].m.resize(d * d);
        for (uint64_t e = 0; e < d * d; ++e)
            bg[i].m[e] = C(mats[2 * (g.mat_off + e)], mats[2 * (g.mat_off + e) + 1]);
        bg[i].cmask = g.cmask;
        bg[i].cval = g.cval;
    }
    if (int rc = qs::ensure_device()) return rc;
    FusedOptions o;
    uint32_t cap = 2048;
    if (opts) {
        if (opts->reg_bits) o.reg_bits = opts->reg_bits;
        if (opts->low_bits) o.low_bits = opts->low_bits;
        if (opts->max_ops) o.max_ops = opts->max_ops;
        if (opts->grid_cap) cap = opts->grid_cap;
    }
    if (o.reg_bits != 3 && o.reg_bits != 4) return fail(QS_ERR_OUT_OF_RANGE, W, "reg_bits");
    auto* p = new qs_program;
    p->ngates = bg.size();
    p->grid_cap = cap;
    if (int rc = qs::fused_compile(bg, nb, o, p->P)) {
        delete p;
        return rc;
    }
    if (int rc = qs::fused_upload(p->P, dtype)) {
        delete p;
        return rc;
    }
    *out = p;
    return QS_OK;
}

// The program's passes as specialised kernels (generated source compiled
// with hipRTC); later runs on states of the program's dtype use them.
int qs_program_compile_native(qs_program* p) {
    if (!p) return fail(QS_ERR_INVALID_HANDLE, "qpp_amd::program");
    qs::clear_error();
    // a captured graph holds the old kernels (and native_build unloads
    // earlier specialised ones): drop it
    if (p->gexec && p->gs) (void)hipStreamSynchronize(p->gs);
    drop_graph(p);
    return qs::native_build(p->P, p->P.dtype);
}

int qs_program_capture(qs_program* p, qs_state* st) {
    constexpr const char* W = "qpp_amd::program";
    if (!p || !st || !st->d) return fail(QS_ERR_INVALID_HANDLE, W);
    if (st->dtype != p->P.dtype) return fail(QS_ERR_TYPE_MISMATCH, W, "dtype");
    if (!st->qubits || st->nbits != p->P.nb) return fail(QS_ERR_DIMS_MISMATCH_CVECTOR, W, "state/program");
    if (p->gexec && p->gs) (void)hipStreamSynchronize(p->gs);
    drop_graph(p);
    QS_HIP(hipStreamBeginCapture(st->stream, hipStreamCaptureModeThreadLocal), W);
    const int rc = qs::fused_run(p->P, st, p->grid_cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st->stream, &g);  // always ends the capture
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : qs::fail_hip(e, W);
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        return qs::fail_hip(ei, W);
    }
    p->graph = g;
    p->gexec = x;
    p->gd = st->d;
    p->gs = st->stream;
    // a capture records the launches without running them: run them once
    QS_HIP(hipGraphLaunch(x, st->stream), W);
    return QS_OK;
}

int qs_program_run(qs_program* p, qs_state* st) {
    if (!p || !st || !st->d) return fail(QS_ERR_INVALID_HANDLE, "qpp_amd::program");
    if (p->gexec && st->d == p->gd && st->stream == p->gs && st->dtype == p->P.dtype && st->qubits &&
        st->nbits == p->P.nb) {
        QS_HIP(hipGraphLaunch(p->gexec, st->stream), "qpp_amd::program");
        return QS_OK;
    }
    return qs::fused_run(p->P, st, p->grid_cap);
}

int qs_program_run_timed(qs_program* p, qs_state* st, double* pass_ms) {
    constexpr const char* W = "qpp_amd::program";
    if (!p || !st || !st->d || !pass_ms) return fail(QS_ERR_INVALID_HANDLE, W);
    const size_t np = p->P.passes.size();
    std::vector<hipEvent_t> ev(np + 1, nullptr);
    int rc = QS_OK;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = fail(QS_ERR_DEVICE, W, "hipEventCreate");
    if (rc == QS_OK) rc = qs::fused_run(p->P, st, p->grid_cap, ev.data());
    if (rc == QS_OK && hipEventSynchronize(ev[np]) != hipSuccess)
        rc = fail(QS_ERR_DEVICE, W, "hipEventSynchronize");
    for (size_t i = 0; rc == QS_OK && i < np; ++i) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) != hipSuccess)
            rc = fail(QS_ERR_DEVICE, W, "hipEventElapsedTime");
        pass_ms[i] = ms;
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

int qs_program_info(const qs_program* p, uint64_t* num_gates, uint64_t* num_passes,
                    uint64_t* num_phases) {
    if (!p) return fail(QS_ERR_INVALID_HANDLE, "qpp_amd::program");
    if (num_gates) *num_gates = p->ngates;
    if (num_passes) *num_passes = p->P.passes.size();
    if (num_phases) *num_phases = p->P.phases.size();
    return QS_OK;
}

int qs_program_native_stats(const qs_program* p, uint64_t counts[5]) {
    if (!p) return fail(QS_ERR_INVALID_HANDLE, "qpp_amd::program");
    if (!counts) return fail(QS_ERR_SIZE_MISMATCH, "qpp_amd::program", "counts is null");
    for (int i = 0; i < 5; ++i) counts[i] = p->P.native_stats[i];
    return QS_OK;
}

int qs_program_destroy(qs_program* p) {
    if (!p) return QS_OK;
    if (p->gexec && p->gs) (void)hipStreamSynchronize(p->gs);
    drop_graph(p);
    qs::native_free(p->P);
    qs::fused_free(p->P);
    delete p;
    return QS_OK;
}

}  // extern "C"

// qs_engine.hip -- QEngine / QNoisyEngine execute loop over device states.
//
// Restates the step interpreter of qpp's engines (classes/qengine.hpp,
// classes/qnoisy_engine.hpp) on a resident device state:
//   execute(reps)        qengine.hpp:1341-1405 (canonical form for reps > 1,
//                        steps before the first measurement run once, then
//                        sampling or per-rep re-execution)
//   sampling tail        qengine.hpp:374-481 (rep 0 draws and discards one
//                        sample, executes the measurements; later reps
//                        sample the pre-measurement state)
//   retries              qengine.hpp:316-345 (state restored between reps,
//                        the first rep's state kept)
//   destructive measure  qengine.hpp:489-503, 514-526 (relative positions)
//   noise                qnoisy_engine.hpp:104-140, noise.hpp:131-160
//                        (before every step, one draw per live qubit)
//   reset / discard      qengine.hpp:804-821 (instruments.hpp:1148-1196)
//   post-selection       qengine.hpp:650-733 (re-measured while ensured),
//                        240-294 (a failed run stops), 374-481 (sampling
//                        with post-selected qudits)
// Every uniform comes from the engine's std::mt19937 through
// std::generate_canonical<double, 53>, in qpp's order.  Runs of gate steps
// between measurements are compiled into tile-fused programs (cached) when
// the state is a qubit state of at least 12 index bits, and applied with the
// gate kernels otherwise.  A depolarizing event applies its Pauli operator;
// the identity event and the renormalisation after a Pauli (noise.hpp:159,
// a division by a norm equal to 1 up to rounding) are not launched.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "qs_internal.hpp"

namespace {

using qs::fail;
constexpr uint64_t kMeasured = std::numeric_limits<uint64_t>::max();
// Retrying a post-selection without limit (qpp's default, and the first
// successful rep of the sampling tail) never ends when the selected result
// has probability 0, e.g. after a measurement that contradicts it; qpp
// then loops for ever.  Here such a step gives up with an error after
// ps_give_up failed attempts, 2^16 unless set (a result of probability p
// fails that often with probability (1-p)^65536, below 1e-28 for p >=
// 1e-3).  The same for a runtime while loop whose condition never turns
// false: loop_give_up repetitions of its body in one run, 2^24 unless set.
// qs_engine_set_limits changes both; 0 means no limit (qpp's behaviour).
constexpr uint64_t kUnlimited = std::numeric_limits<uint64_t>::max();
constexpr uint64_t kPsGiveUp = uint64_t{1} << 16;
constexpr uint64_t kLoopGiveUp = uint64_t{1} << 24;

struct EStep {
    uint32_t type = 0, gate = 0;
    std::vector<uint64_t> ctrl, target, shift;
    bool has_shift = false;
    uint64_t c_reg = 0;
    bool destructive = true;
    std::vector<uint64_t> ps;  // POST_SELECT values, COND_IF dit values or truth table
    bool table = false;        // COND_IF / COND_WHILE: ps is a truth table over the ctrl dits
    size_t jump = 0;           // COND_IF / COND_WHILE: its COND_ELSE or else COND_END;
                               // COND_ELSE: its COND_END; COND_END of a while: the COND_WHILE
    bool loop_end = false;     // a COND_END that closes a COND_WHILE
};

bool is_cctrl(const EStep& s) { return s.type == QS_STEP_CCTRL || s.type == QS_STEP_CCTRL_FAN; }

bool is_runtime(const EStep& s) {
    return s.type == QS_STEP_COND_IF || s.type == QS_STEP_COND_ELSE ||
           s.type == QS_STEP_COND_END || s.type == QS_STEP_COND_WHILE;
}

// is_measurement, classes/qcircuit.hpp:7008-7025: measurements and
// post-selections; reset and discard are not measurements
bool is_meas(const EStep& s) {
    return s.type == QS_STEP_MEASURE || s.type == QS_STEP_MEASURE_MANY ||
           s.type == QS_STEP_POST_SELECT;
}

bool is_reset_or_discard(const EStep& s) {
    return s.type == QS_STEP_RESET || s.type == QS_STEP_DISCARD;
}

std::vector<uint64_t> c_regs_of(const EStep& s) {
    std::vector<uint64_t> c;
    if (is_meas(s)) {
        c.resize(s.target.size());
        std::iota(c.begin(), c.end(), s.c_reg);
    } else if (is_cctrl(s)) {
        c = s.ctrl;  // its control dits (qcircuit.hpp:7420-7427)
    }
    return c;
}

// can_swap, classes/qcircuit.hpp:7465-7492, on the lists
// extract_ctrl_target_ps_vals_c_reg (:7371-7449) gives: a reset or discard
// step yields no controls, targets or dits, so it swaps with anything
bool can_swap(const EStep& a, const EStep& b) {
    auto overlap = [](const std::vector<uint64_t>& x, const std::vector<uint64_t>& y) {
        return std::find_first_of(x.begin(), x.end(), y.begin(), y.end()) != x.end();
    };
    const std::vector<uint64_t> none;
    const auto& ta = is_reset_or_discard(a) ? none : a.target;
    const auto& tb = is_reset_or_discard(b) ? none : b.target;
    const auto& ca = is_cctrl(a) ? none : a.ctrl;  // a cCTRL's controls are dits
    const auto& cb = is_cctrl(b) ? none : b.ctrl;
    return !overlap(ca, cb) && !overlap(ta, tb) && !overlap(c_regs_of(a), c_regs_of(b));
}

// canonical_form, classes/qcircuit.hpp:7556-7612 (a circuit with runtime
// steps is left as it is, :7563-7567)
std::vector<EStep> canonical_form(const std::vector<EStep>& steps) {
    if (std::any_of(steps.begin(), steps.end(), is_runtime)) return steps;
    auto first = std::find_if(steps.begin(), steps.end(), is_meas);
    std::vector<EStep> before(steps.begin(), first), rest(first, steps.end());
    // cCTRL steps first (:7580-7595), then measurements
    for (std::ptrdiff_t ri = static_cast<std::ptrdiff_t>(rest.size()) - 1; ri >= 0; --ri) {
        if (!is_cctrl(rest[ri])) continue;
        for (size_t after = ri + 1; after < rest.size(); ++after) {
            if (can_swap(rest[after - 1], rest[after]) && !is_cctrl(rest[after]))
                std::swap(rest[after - 1], rest[after]);
            else
                break;
        }
    }
    for (std::ptrdiff_t ri = static_cast<std::ptrdiff_t>(rest.size()) - 1; ri >= 0; --ri) {
        if (!is_meas(rest[ri])) continue;
        for (size_t after = ri + 1; after < rest.size(); ++after) {
            if (can_swap(rest[after - 1], rest[after]) && !is_meas(rest[after]))
                std::swap(rest[after - 1], rest[after]);
            else
                break;
        }
    }
    before.insert(before.end(), rest.begin(), rest.end());
    return before;
}

// one gate application in relative positions; gate < 0 selects a Pauli
// (-1 X, -2 Y, -3 Z) of a noise event
struct GateApp {
    uint32_t type;
    int gate;
    std::vector<uint64_t> ctrl, target, shift;
    bool has_shift;
    std::vector<double> own;  // the matrix, when it is a power U^k (k > 1) of gate
};

const double kPauli[3][8] = {
    {0, 0, 1, 0, 1, 0, 0, 0},    // X, column-major (re, im)
    {0, 0, 0, 1, 0, -1, 0, 0},   // Y = [[0, -i], [i, 0]]
    {1, 0, 0, 0, 0, 0, -1, 0}};  // Z

}  // namespace

struct qs_engine {
    uint32_t nq = 0, nc = 0, d = 2;
    qs_kind kind = QS_KET;
    qs_dtype dtype = QS_C128;
    double noise_p = -1;
    std::vector<EStep> steps;
    std::vector<uint64_t> mat_dim, mat_off;
    std::vector<double> mats;
    std::mt19937 gen;
    qs_state* st = nullptr;
    std::vector<uint64_t> dits, subsys;
    std::vector<double> probs;
    std::map<std::vector<uint64_t>, uint64_t> stats;
    std::map<std::string, qs_program*> progs;
    // QNoisyEngineT::noise_results_ (qnoisy_engine.hpp:62, 104-121): per
    // step, the Kraus index drawn for each live qubit at each visit
    std::vector<std::vector<uint64_t>> noise_results;
    // internal/classes/qengine_state.hpp:66-72
    bool ps_ok = true, ensure_ps = false;
    uint64_t max_ps = std::numeric_limits<uint64_t>::max();
    uint64_t ps_give_up = kPsGiveUp, loop_give_up = kLoopGiveUp;  // 0: none

    double canonical() { return std::generate_canonical<double, 53>(gen); }

    bool was_measured(uint64_t i) const { return subsys[i] == kMeasured; }

    // get_relative_pos_, qengine.hpp:514-526: a step on a destructively
    // measured qudit is QuditAlreadyMeasured
    // 1 true, 0 false, -1 a dit outside [0, d) under a truth table
    int cond_true(const EStep& c) const {
        if (c.table) {
            uint64_t i = 0;
            for (uint64_t q : c.ctrl) {
                if (dits[q] >= d) return -1;
                i = i * d + dits[q];
            }
            return c.ps[i] != 0;
        }
        for (size_t j = 0; j < c.ctrl.size(); ++j)
            if (dits[c.ctrl[j]] != c.ps[j]) return 0;
        return 1;
    }

    bool on_measured(const EStep& x) const {
        if (is_runtime(x)) return false;
        const std::vector<uint64_t> none;
        for (const auto* v : {&x.target, is_cctrl(x) ? &none : &x.ctrl})
            for (uint64_t q : *v)
                if (was_measured(q)) return true;
        return false;
    }

    // the steps before the failing one run first, as in qpp
    int already_measured(const std::vector<GateApp>& pending) {
        run_gates(pending, false);
        return fail(QS_ERR_QUDIT_ALREADY_MEASURED, "qpp::QEngineT::get_relative_pos_()", "v[i]");
    }

    std::vector<uint64_t> rel(const std::vector<uint64_t>& v) const {
        std::vector<uint64_t> r = v;
        for (auto& x : r) x = subsys[x];
        return r;
    }

    std::vector<uint64_t> live() const {
        std::vector<uint64_t> l;
        for (uint64_t i = 0; i < nq; ++i)
            if (!was_measured(i)) l.push_back(i);
        return l;
    }

    uint64_t live_dim() const {
        uint64_t D = 1;
        for (uint64_t i = 0; i < nq; ++i)
            if (!was_measured(i)) D *= d;
        return D;
    }

    // set_measured_destructively_, qengine.hpp:489-503
    void set_measured(uint64_t i) {
        subsys[i] = kMeasured;
        for (uint64_t m = i; m < nq; ++m)
            if (!was_measured(m)) --subsys[m];
    }

    const double* mat(int g, uint64_t& dim) const {
        if (g < 0) {
            dim = 2;
            return kPauli[-g - 1];
        }
        dim = mat_dim[g];
        return mats.data() + 2 * mat_off[g];
    }

    int run_gate_by_gate(const std::vector<GateApp>& gl) {
        for (const GateApp& a : gl) {
            uint64_t dim;
            const double* m = mat(a.gate, dim);
            if (!a.own.empty()) m = a.own.data();
            const uint64_t* sh = a.has_shift ? a.shift.data() : nullptr;
            int rc = QS_OK;
            switch (a.type) {
                case QS_STEP_GATE:
                    rc = qs_apply(st, m, dim, uint32_t(a.target.size()), a.target.data());
                    break;
                case QS_STEP_FAN:
                    for (uint64_t t : a.target)
                        if ((rc = qs_apply(st, m, dim, 1, &t))) break;
                    break;
                case QS_STEP_CTRL:
                    rc = qs_apply_ctrl(st, m, dim, uint32_t(a.ctrl.size()), a.ctrl.data(),
                                       uint32_t(a.target.size()), a.target.data(), sh);
                    break;
                case QS_STEP_CTRL_FAN:
                    rc = qs_apply_ctrl_fan(st, m, dim, uint32_t(a.ctrl.size()), a.ctrl.data(),
                                           uint32_t(a.target.size()), a.target.data(), sh);
                    break;
                default:
                    break;
            }
            if (rc) return rc;
        }
        return QS_OK;
    }

    // a run of gate applications: one fused program when the state allows
    int run_gates(const std::vector<GateApp>& gl, bool cacheable) {
        if (gl.empty()) return QS_OK;
        const uint64_t nl = live().size();
        const uint64_t nbits = (kind == QS_RHO ? 2 : 1) * nl;
        bool fusible = d == 2 && nbits >= 12;
        for (const GateApp& a : gl) {
            const bool fan = a.type == QS_STEP_FAN || a.type == QS_STEP_CTRL_FAN;
            if (!fan && a.target.size() > 2) fusible = false;
            if (!a.own.empty()) fusible = false;
        }
        if (!fusible) return run_gate_by_gate(gl);
        std::string key;
        if (cacheable) {
            key.reserve(gl.size() * 16);
            for (const GateApp& a : gl) {
                key += std::to_string(a.type) + ":" + std::to_string(a.gate) + ":";
                for (uint64_t c : a.ctrl) key += std::to_string(c) + ",";
                key += "|";
                for (uint64_t t : a.target) key += std::to_string(t) + ",";
                key += "|";
                if (a.has_shift)
                    for (uint64_t s : a.shift) key += std::to_string(s) + ",";
                key += ";";
            }
            key += "n" + std::to_string(nl);
            auto it = progs.find(key);
            if (it != progs.end()) return qs_program_run(it->second, st);
        }
        // lower to a qs_circuit over the live qudits
        std::vector<qs_step> ss(gl.size());
        std::vector<uint64_t> pool, mdim = mat_dim, moff = mat_off;
        std::vector<double> mm = mats;
        for (int p = 0; p < 3; ++p) {
            mdim.push_back(2);
            moff.push_back(mm.size() / 2);
            mm.insert(mm.end(), kPauli[p], kPauli[p] + 8);
        }
        const uint32_t npre = static_cast<uint32_t>(mat_dim.size());
        for (size_t i = 0; i < gl.size(); ++i) {
            const GateApp& a = gl[i];
            qs_step& s = ss[i];
            std::memset(&s, 0, sizeof(s));
            s.type = a.type;
            s.gate = a.gate >= 0 ? uint32_t(a.gate) : npre + uint32_t(-a.gate - 1);
            s.nctrl = uint32_t(a.ctrl.size());
            s.ntarget = uint32_t(a.target.size());
            s.ctrl_off = uint32_t(pool.size());
            pool.insert(pool.end(), a.ctrl.begin(), a.ctrl.end());
            s.target_off = uint32_t(pool.size());
            pool.insert(pool.end(), a.target.begin(), a.target.end());
            if (a.has_shift) {
                s.shift_off = uint32_t(pool.size());
                pool.insert(pool.end(), a.shift.begin(), a.shift.end());
            } else {
                s.shift_off = UINT32_MAX;
            }
        }
        pool.push_back(0);
        qs_circuit qc{};
        qc.nq = uint32_t(nl);
        qc.nc = nc;
        qc.d = d;
        qc.nsteps = uint32_t(ss.size());
        qc.steps = ss.data();
        qc.nmats = uint32_t(mdim.size());
        qc.mat_dim = mdim.data();
        qc.mat_off = moff.data();
        qc.mats = mm.data();
        qc.pool = pool.data();
        qs_program* prog = nullptr;
        int rc = qs_program_create(&qc, 0, qc.nsteps, kind, dtype, nullptr, &prog);
        if (rc == QS_ERR_NOT_IMPLEMENTED) return run_gate_by_gate(gl);
        if (rc) return rc;
        rc = qs_program_run(prog, st);
        if (cacheable && !rc)
            progs[key] = prog;
        else
            qs_program_destroy(prog);
        return rc;
    }

    // cCTRL / cCTRL_FAN, qengine.hpp:581-630: every control dit, plus its
    // shift mod d, must equal the first one's value k; then U^k on the
    // targets (each target for the fan form).  Nothing is applied when the
    // dits disagree or k = 0 (U^0 = I leaves the state as it is).
    void cctrl_apps(const EStep& s, std::vector<GateApp>& out) const {
        auto val = [&](size_t m) {
            return s.has_shift ? (dits[s.ctrl[m]] + s.shift[m]) % d : dits[s.ctrl[m]];
        };
        const uint64_t k = val(0);
        for (size_t m = 1; m < s.ctrl.size(); ++m)
            if (val(m) != k) return;
        if (k == 0) return;
        GateApp a;
        a.type = s.type == QS_STEP_CCTRL_FAN ? QS_STEP_FAN : QS_STEP_GATE;
        a.gate = int(s.gate);
        a.target = rel(s.target);
        a.has_shift = false;
        if (k > 1) {
            // powm(U, k) by repeated products, column-major complex
            const uint64_t D = mat_dim[s.gate];
            const std::complex<double>* U =
                reinterpret_cast<const std::complex<double>*>(mats.data() + 2 * mat_off[s.gate]);
            std::vector<std::complex<double>> P(U, U + D * D), T(D * D);
            for (uint64_t r = 1; r < k; ++r) {
                for (uint64_t j = 0; j < D; ++j)
                    for (uint64_t i = 0; i < D; ++i) {
                        std::complex<double> acc = 0;
                        for (uint64_t l = 0; l < D; ++l) acc += P[i + l * D] * U[l + j * D];
                        T[i + j * D] = acc;
                    }
                P.swap(T);
            }
            const double* p = reinterpret_cast<const double*>(P.data());
            a.own.assign(p, p + 2 * D * D);
        }
        out.push_back(std::move(a));
    }

    GateApp app_of(const EStep& s) const {
        GateApp a;
        a.type = s.type;
        a.gate = int(s.gate);
        a.ctrl = rel(s.ctrl);
        a.target = rel(s.target);
        a.shift = s.shift;
        a.has_shift = s.has_shift;
        return a;
    }

    // visit_measurement_step_, qengine.hpp:639-734, with its draw u (or,
    // when `given`, the basis index that draw selects, already sampled)
    int measure(const EStep& s, double u, const uint64_t* given = nullptr) {
        const std::vector<uint64_t> t = rel(s.target);
        std::vector<uint64_t> res(t.size());
        std::vector<double> ps(t.size());
        const int dst = s.destructive ? 1 : 0;
        if (int rc = given ? qs::measure_seq_given(st, uint32_t(t.size()), t.data(), dst, *given,
                                                   res.data(), ps.data())
                           : qs_measure_seq(st, uint32_t(t.size()), t.data(), dst, u, res.data(),
                                            ps.data()))
            return rc;
        for (size_t i = 0; i < t.size(); ++i) {
            dits[s.c_reg + i] = res[i];
            probs[s.c_reg + i] = ps[i];
        }
        if (s.destructive)
            for (uint64_t q : s.target) set_measured(q);
        return QS_OK;
    }

    // POST_SELECT(_MANY)(_ND), qengine.hpp:650-733.  Each attempt draws and
    // samples the basis index measure_seq would collapse to; a result that
    // differs from the values is drawn again while ensure_ps is set and
    // fewer than max_ps attempts were made.  Out of attempts the step
    // leaves state and dits as they were; without ensure_ps the failed
    // result is measured and kept.  Either failure clears ps_ok.
    int post_select(const EStep& s) {
        const std::vector<uint64_t> t = rel(s.target);
        std::vector<uint64_t> dig(t.size());
        uint64_t attempts = 0;
        for (;;) {
            double u = live_dim() >= 2 ? canonical() : 0.0;
            uint64_t index = 0;
            if (int rc = qs_sample(st, uint32_t(t.size()), t.data(), 1, &u, &index, dig.data(),
                                   nullptr))
                return rc;
            if (dig != s.ps) {
                ++attempts;
                if (!ensure_ps) {
                    ps_ok = false;
                } else if (attempts < max_ps) {
                    if (max_ps == kUnlimited && ps_give_up && attempts >= ps_give_up)
                        return fail(QS_ERR_OUT_OF_RANGE, "qpp::QEngineT::execute()",
                                    "post-selection failed the give-up count of times with unlimited retries");
                    continue;
                } else {
                    ps_ok = false;
                    return QS_OK;
                }
            }
            return measure(s, 0.0, &index);
        }
    }

    // RESET / RESET_MANY, qengine.hpp:804-808: qpp::reset on the relative
    // positions, one draw for its measure_seq
    int reset_step(const EStep& s) {
        const std::vector<uint64_t> t = rel(s.target);
        const double u = live_dim() >= 2 ? canonical() : 0.0;
        return qs_reset(st, uint32_t(t.size()), t.data(), u);
    }

    // DISCARD / DISCARD_MANY, qengine.hpp:810-821: destructive measure_seq
    // with the results dropped
    int discard_step(const EStep& s) {
        const std::vector<uint64_t> t = rel(s.target);
        std::vector<uint64_t> res(t.size());
        std::vector<double> ps(t.size());
        const double u = live_dim() >= 2 ? canonical() : 0.0;
        if (int rc = qs_measure_seq(st, uint32_t(t.size()), t.data(), 1, u, res.data(), ps.data()))
            return rc;
        for (uint64_t q : s.target) set_measured(q);
        return QS_OK;
    }

    // noise before a step (qnoisy_engine.hpp:108-116): one draw per live
    // qubit in ascending order, k = lower_bound(cp{1-p, p/3, p/3, p/3}, u)
    void noise_apps(std::vector<GateApp>& pending, size_t ip) {
        const std::vector<double> w{1 - noise_p, noise_p / 3, noise_p / 3, noise_p / 3};
        const double sum = std::accumulate(w.begin(), w.end(), 0.0);
        std::vector<double> q(4), cp(4);
        for (int i = 0; i < 4; ++i) q[i] = w[i] / sum;
        std::partial_sum(q.begin(), q.end(), cp.begin());
        cp.back() = 1.0;
        for (uint64_t qb : live()) {
            const double u = canonical();
            const int k = int(std::lower_bound(cp.begin(), cp.end(), u) - cp.begin());
            noise_results[ip].push_back(uint64_t(k));
            if (k == 0) continue;
            GateApp a;
            a.type = QS_STEP_GATE;
            a.gate = -k;
            a.target = {subsys[qb]};
            a.has_shift = false;
            pending.push_back(a);
        }
    }

    // execute_steps_once_, qengine.hpp:240-294: steps [pos, end) once,
    // noise before every step when noisy; stops after a step whose
    // post-selection failed; a run that measured and got to the end adds
    // its dits to the statistics.  `ensure` forces ensure_ps on meanwhile.
    int execute_steps_once(const std::vector<EStep>& s, size_t pos, bool ensure) {
        const bool saved = ensure_ps;
        if (ensure) ensure_ps = true;
        const int rc = steps_once(s, pos);
        ensure_ps = saved;
        return rc;
    }

    int steps_once(const std::vector<EStep>& s, size_t pos) {
        bool measured = false;
        uint64_t loops = 0;
        std::vector<GateApp> pending;
        bool cacheable = noise_p < 0;
        for (size_t i = pos; i < s.size(); ++i) {
            if (noise_p >= 0) noise_apps(pending, i);
            const EStep& x = s[i];
            if (on_measured(x)) return already_measured(pending);
            if (x.type == QS_STEP_COND_IF || x.type == QS_STEP_COND_WHILE) {
                // visit_runtime_step_, qengine.hpp:837-860: on false, go on
                // after the matching ELSE (or END)
                const int c = cond_true(x);
                if (c < 0) {
                    run_gates(pending, false);
                    return fail(QS_ERR_OUT_OF_RANGE, "qpp::QEngineT::execute()", "condition dit outside [0, d)");
                }
                if (!c) i = x.jump;
            } else if (x.type == QS_STEP_COND_ELSE) {
                i = x.jump;  // the true branch ends here (:862-866)
            } else if (x.type == QS_STEP_COND_END) {
                // ENDWHILE (:869-880): while the loop's condition holds, go
                // on at the first step of its body.  A loop whose condition
                // never turns false would run for ever (as in qpp); it stops
                // here once its body has run loop_give_up times in one run
                // and the condition still holds.
                const int c = x.loop_end ? cond_true(s[x.jump]) : 0;
                if (c < 0) {
                    run_gates(pending, false);
                    return fail(QS_ERR_OUT_OF_RANGE, "qpp::QEngineT::execute()", "condition dit outside [0, d)");
                }
                if (c) {
                    if (++loops >= loop_give_up && loop_give_up) {
                        run_gates(pending, false);
                        return fail(QS_ERR_OUT_OF_RANGE, "qpp::QEngineT::execute()",
                                    "while loop repeated the give-up count of times");
                    }
                    i = x.jump;
                }
            } else if (is_meas(x) || is_reset_or_discard(x)) {
                if (int rc = run_gates(pending, cacheable && !pending.empty())) return rc;
                pending.clear();
                int rc = QS_OK;
                if (x.type == QS_STEP_POST_SELECT) {
                    rc = post_select(x);
                } else if (x.type == QS_STEP_RESET) {
                    rc = reset_step(x);
                } else if (x.type == QS_STEP_DISCARD) {
                    rc = discard_step(x);
                } else {
                    const double u = live_dim() >= 2 ? canonical() : 0.0;
                    rc = measure(x, u);
                }
                if (rc) return rc;
                if (is_meas(x)) measured = true;
            } else if (is_cctrl(x)) {
                cctrl_apps(x, pending);
            } else if (x.type != QS_STEP_NOP) {
                pending.push_back(app_of(x));
            }
            if (!ps_ok) return run_gates(pending, false);
        }
        if (int rc = run_gates(pending, cacheable)) return rc;
        if (measured) ++stats[dits];
        return QS_OK;
    }

    struct Snap {
        qs_state* st = nullptr;
        std::vector<uint64_t> dits, subsys;
        std::vector<double> probs;
        bool ps_ok = true;
    };

    int snapshot(Snap& sn) const {
        sn.dits = dits;
        sn.subsys = subsys;
        sn.probs = probs;
        sn.ps_ok = ps_ok;
        return qs_state_clone(st, &sn.st);
    }

    // The engine's state becomes the snapshot's (taken over, or cloned).  A
    // failed clone leaves the current state in place and reports why.
    int restore(Snap& sn, bool take) {
        if (take) {
            qs_state_destroy(st);
            st = sn.st;
            sn.st = nullptr;
        } else {
            qs_state* c = nullptr;
            if (int rc = qs_state_clone(sn.st, &c)) return rc;
            qs_state_destroy(st);
            st = c;
        }
        dits = sn.dits;
        subsys = sn.subsys;
        probs = sn.probs;
        ps_ok = sn.ps_ok;
        return QS_OK;
    }

    // execute_steps_with_postselection_retries_, qengine.hpp:316-345: every
    // rep starts from the saved engine state; the engine ends in the first
    // rep whose post-selection held, else in the last rep's state
    int execute_with_retries(const std::vector<EStep>& s, size_t pos, uint64_t reps) {
        Snap copy, ok;
        if (reps > 1)
            if (int rc = snapshot(copy)) return rc;
        bool found = false;
        int rc = QS_OK;
        for (uint64_t rep = 0; rep < reps && !rc; ++rep) {
            rc = execute_steps_once(s, pos, ensure_ps);
            if (!rc && ps_ok && !found) {
                found = true;
                if (rep + 1 < reps) rc = snapshot(ok);
            }
            if (!rc && rep + 1 < reps) rc = restore(copy, false);
        }
        if (!rc && ok.st) rc = restore(ok, true);
        qs_state_destroy(ok.st);
        qs_state_destroy(copy.st);
        return rc;
    }

    // execute_measurement_steps_with_sampling_and_postselection_retries_,
    // qengine.hpp:374-481, and build_sampling_maps_, :150-213.  The sampled
    // qudits (ascending) fill the classical dits in ascending c_reg order.
    // A qudit measured into two dits would make qpp read past its sample
    // (one digit per qudit); such circuits run rep by rep here.
    int execute_sampling(const std::vector<EStep>& s, size_t pos, uint64_t reps) {
        std::map<uint64_t, uint64_t> c_q, q_ps;
        bool any = false, has_ps = false;
        for (size_t i = pos; i < s.size(); ++i)
            if (is_meas(s[i])) {
                any = true;
                const auto cr = c_regs_of(s[i]);
                for (size_t q = 0; q < s[i].target.size(); ++q) {
                    const uint64_t t = s[i].target[q];
                    c_q[cr[q]] = t;
                    if (s[i].type != QS_STEP_POST_SELECT) continue;
                    has_ps = true;
                    auto it = q_ps.find(t);
                    if (it == q_ps.end()) {
                        q_ps[t] = s[i].ps[q];
                    } else if (it->second != s[i].ps[q]) {
                        ps_ok = false;  // incompatible post-selections
                        return QS_OK;
                    }
                }
            }
        if (!any) return QS_OK;
        std::set<uint64_t> sset;
        for (auto& kv : c_q) sset.insert(kv.second);
        const std::vector<uint64_t> from(sset.begin(), sset.end());
        if (c_q.size() != from.size()) return execute_with_retries(s, pos, reps);
        if (has_ps) return execute_sampling_ps(s, pos, reps, c_q, q_ps, from);
        const uint64_t D = live_dim();
        if (s.size() == pos + 1) {
            // A single measurement step: its draw and every later rep's draw
            // sample the same pre-measurement state (sample() reads the full
            // basis index, instruments.hpp:783-799), so all of them come from
            // one sampling call and no copy of the state is needed.
            if (D >= 2) (void)canonical();  // rep 0's discarded sample
            std::vector<double> u;
            for (uint64_t rep = 0; rep < reps; ++rep) u.push_back(D >= 2 ? canonical() : 0.0);
            std::vector<uint64_t> idx(u.size()), dig(u.size() * from.size());
            if (int rc = qs_sample(st, uint32_t(from.size()), from.data(), u.size(), u.data(),
                                   idx.data(), dig.data(), nullptr))
                return rc;
            if (int rc = measure(s[pos], 0.0, &idx[0])) return rc;
            ++stats[dits];
            for (size_t r = 1; r < u.size(); ++r) {
                std::vector<uint64_t> res = dits;
                size_t i = 0;
                for (auto& kv : c_q) res[kv.first] = dig[r * from.size() + i++];
                ++stats[res];
            }
            return QS_OK;
        }
        qs_state* pre = nullptr;
        if (int rc = qs_state_clone(st, &pre)) return rc;
        // rep 0: one sample drawn and discarded, then the measurement steps
        if (D >= 2) (void)canonical();
        int rc = execute_steps_once(s, pos, true);
        // reps 1..: one draw each, sampled from the pre-measurement state
        std::vector<double> u;
        for (uint64_t rep = 1; rep < reps && !rc; ++rep) u.push_back(D >= 2 ? canonical() : 0.0);
        if (!rc && !u.empty()) {
            const std::vector<uint64_t> frel = from;  // every qudit is live at pos
            std::vector<uint64_t> dig(u.size() * frel.size());
            rc = qs_sample(pre, uint32_t(frel.size()), frel.data(), u.size(), u.data(), nullptr,
                           dig.data(), nullptr);
            for (size_t r = 0; r < u.size() && !rc; ++r) {
                std::vector<uint64_t> res = dits;
                size_t i = 0;
                for (auto& kv : c_q) res[kv.first] = dig[r * frel.size() + i++];
                ++stats[res];
            }
        }
        qs_state_destroy(pre);
        return rc;
    }

    // The sampling tail with post-selected qudits (qengine.hpp:414-473).
    // Every attempt draws once and samples the pre-measurement state; an
    // attempt whose post-selected qudits differ from their values fails
    // its rep, or with ensure_ps is drawn again while fewer than max_ps
    // attempts were made.  The first successful rep runs the steps once
    // (ensure on, no attempt limit), which fixes the engine's state; later
    // successes add statistics only.  Until that first success attempts go
    // one by one, since the steps' own draws follow it; after it no other
    // draw happens, so the draws come in batches of the remaining rep count
    // (each remaining rep takes at least one) and are sampled in one call.
    int execute_sampling_ps(const std::vector<EStep>& s, size_t pos, uint64_t reps,
                            const std::map<uint64_t, uint64_t>& c_q,
                            const std::map<uint64_t, uint64_t>& q_ps,
                            const std::vector<uint64_t>& from) {
        std::vector<std::pair<size_t, uint64_t>> chk;  // (position in `from`, value)
        for (size_t q = 0; q < from.size(); ++q) {
            auto it = q_ps.find(from[q]);
            if (it != q_ps.end()) chk.push_back({q, it->second});
        }
        auto failed = [&](const uint64_t* dig) {
            for (const auto& c : chk)
                if (dig[c.first] != c.second) return true;
            return false;
        };
        const uint64_t D = live_dim(), nf = from.size();
        qs_state* pre = nullptr;
        if (int rc = qs_state_clone(st, &pre)) return rc;
        bool found = false;
        uint64_t rep = 0, attempts = 0;
        int rc = QS_OK;
        std::vector<uint64_t> dig(nf);
        while (!rc && rep < reps && !found) {
            ++attempts;
            double u = D >= 2 ? canonical() : 0.0;
            rc = qs_sample(pre, uint32_t(nf), from.data(), 1, &u, nullptr, dig.data(), nullptr);
            if (rc) break;
            if (failed(dig.data())) {
                ps_ok = false;
                if (!(ensure_ps && attempts < max_ps)) {
                    ++rep;
                    attempts = 0;
                } else if (max_ps == kUnlimited && ps_give_up && attempts >= ps_give_up) {
                    rc = fail(QS_ERR_OUT_OF_RANGE, "qpp::QEngineT::execute()",
                              "post-selection failed the give-up count of times with unlimited retries");
                }
                continue;
            }
            ps_ok = true;
            found = true;
            const uint64_t m0 = max_ps;
            max_ps = std::numeric_limits<uint64_t>::max();
            rc = execute_steps_once(s, pos, true);
            max_ps = m0;
            ++rep;
            attempts = 0;
        }
        while (!rc && rep < reps) {
            const uint64_t B = reps - rep;
            std::vector<double> u(B);
            for (auto& x : u) x = D >= 2 ? canonical() : 0.0;
            std::vector<uint64_t> dg(B * nf);
            rc = qs_sample(pre, uint32_t(nf), from.data(), B, u.data(), nullptr, dg.data(), nullptr);
            for (uint64_t b = 0; b < B && !rc; ++b) {
                ++attempts;
                const uint64_t* d = dg.data() + b * nf;
                if (failed(d)) {
                    ps_ok = false;
                    if (!(ensure_ps && attempts < max_ps)) {
                        ++rep;
                        attempts = 0;
                    } else if (max_ps == kUnlimited && ps_give_up && attempts >= ps_give_up) {
                        rc = fail(QS_ERR_OUT_OF_RANGE, "qpp::QEngineT::execute()",
                                  "post-selection failed the give-up count of times with unlimited retries");
                    }
                    continue;
                }
                ps_ok = true;
                std::vector<uint64_t> res = dits;
                size_t i = 0;
                for (auto& kv : c_q) res[kv.first] = d[i++];
                ++stats[res];
                ++rep;
                attempts = 0;
            }
        }
        if (found) ps_ok = true;
        qs_state_destroy(pre);
        return rc;
    }

    void reset_bookkeeping() {
        dits.assign(nc, 0);
        probs.assign(nc, 0.0);
        subsys.resize(nq);
        std::iota(subsys.begin(), subsys.end(), uint64_t{0});
    }

    // set by qs::engine_execute_prefixed: the state after the gate prefix,
    // which execute() takes instead of running the prefix
    qs_state* prefixed = nullptr;

    int execute(uint64_t reps) {
        if (noise_p >= 0) return execute_with_retries(steps, 0, reps);
        // reset(qstate, false): keep the state and the statistics
        if (!live().empty() && live().size() != nq)
            return fail(QS_ERR_DIMS_NOT_EQUAL, "qpp::QEngineT::execute()", "state");
        reset_bookkeeping();
        ps_ok = true;  // ensure_ps and max_ps are kept (qengine.hpp:1354-1361)
        const std::vector<EStep> s = reps > 1 ? canonical_form(steps) : steps;
        if (s.empty()) return QS_OK;
        // compute_optimize_up_to_pos_ and can_sample_from_, qengine.hpp:87-143
        const size_t pos = std::find_if(s.begin(), s.end(), [](const EStep& x) {
                               return is_meas(x) || is_reset_or_discard(x) || is_runtime(x);
                           }) - s.begin();
        bool can_sample = true;
        for (size_t i = pos; i < s.size(); ++i)
            if (!is_meas(s[i])) can_sample = false;
        std::vector<GateApp> pre;
        for (size_t i = 0; i < pos; ++i)
            if (s[i].type != QS_STEP_NOP) {
                if (on_measured(s[i])) return already_measured(pre);
                if (is_cctrl(s[i]))
                    cctrl_apps(s[i], pre);
                else
                    pre.push_back(app_of(s[i]));
            }
        if (prefixed) {
            qs_state_destroy(st);
            st = prefixed;
            prefixed = nullptr;
        } else if (int rc = run_gates(pre, true)) {
            return rc;
        }
        if (reps > 1 && can_sample) return execute_sampling(s, pos, reps);
        return execute_with_retries(s, pos, reps);
    }

    ~qs_engine() {
        if (prefixed) qs_state_destroy(prefixed);
        for (auto& kv : progs) qs_program_destroy(kv.second);
        if (st) qs_state_destroy(st);
    }
};

extern "C" {

int qs_engine_create(const qs_circuit* qc, qs_kind kind, qs_dtype dtype, double noise_p,
                     uint64_t seed, qs_engine** out) {
    constexpr const char* W = "qpp::QEngineT::QEngineT()";
    if (!qc || !out) return fail(QS_ERR_INVALID_HANDLE, W, "circuit");
    *out = nullptr;
    if (noise_p > 1) return fail(QS_ERR_OUT_OF_RANGE, "qpp::QubitDepolarizingNoise()", "p");
    if (noise_p >= 0 && qc->d != 2)
        return fail(QS_ERR_NOT_IMPLEMENTED, W, "depolarizing noise on qudits");
    auto* e = new qs_engine;
    e->nq = qc->nq;
    e->nc = qc->nc;
    e->d = qc->d;
    e->kind = kind;
    e->dtype = dtype;
    e->noise_p = noise_p;
    e->gen.seed(static_cast<std::mt19937::result_type>(seed));
    for (uint32_t m = 0; m < qc->nmats; ++m) {
        e->mat_dim.push_back(qc->mat_dim[m]);
        e->mat_off.push_back(e->mats.size() / 2);
        const double* p = qc->mats + 2 * qc->mat_off[m];
        e->mats.insert(e->mats.end(), p, p + 2 * qc->mat_dim[m] * qc->mat_dim[m]);
    }
    for (uint32_t i = 0; i < qc->nsteps; ++i) {
        const qs_step& s = qc->steps[i];
        EStep t;
        t.type = s.type;
        t.gate = s.gate;
        t.ctrl.assign(qc->pool + s.ctrl_off, qc->pool + s.ctrl_off + s.nctrl);
        t.target.assign(qc->pool + s.target_off, qc->pool + s.target_off + s.ntarget);
        if (s.shift_off != UINT32_MAX) {
            t.has_shift = true;
            t.shift.assign(qc->pool + s.shift_off, qc->pool + s.shift_off + s.nctrl);
        }
        t.c_reg = s.c_reg;
        t.destructive = s.destructive != 0;
        if (s.type > QS_STEP_COND_WHILE_TABLE) {
            delete e;
            return fail(QS_ERR_OUT_OF_RANGE, W, "step type");
        }
        if (s.type == QS_STEP_COND_IF_TABLE || s.type == QS_STEP_COND_WHILE_TABLE) {
            // a truth-table condition runs as COND_IF / COND_WHILE
            t.type = s.type == QS_STEP_COND_IF_TABLE ? QS_STEP_COND_IF : QS_STEP_COND_WHILE;
            t.table = true;
        }
        if (is_meas(t) && t.c_reg + t.target.size() > qc->nc) {
            delete e;
            return fail(QS_ERR_OUT_OF_RANGE, "qpp::QCircuit::measure()", "c_reg");
        }
        bool in_range = true;
        const bool dit_ctrl = t.type == QS_STEP_COND_IF || t.type == QS_STEP_COND_WHILE || is_cctrl(t);
        const uint64_t ctrl_lim = dit_ctrl ? qc->nc : qc->nq;  // dits or qudits
        if (is_cctrl(t) && t.ctrl.empty()) {
            delete e;
            return fail(QS_ERR_ZERO_SIZE, "qpp::QCircuit::cCTRL()", "ctrl_dits");
        }
        for (uint64_t q : t.ctrl) in_range = in_range && q < ctrl_lim;
        for (uint64_t q : t.target) in_range = in_range && q < qc->nq;
        if (t.type == QS_STEP_COND_IF || t.type == QS_STEP_COND_WHILE) {
            uint64_t len = s.nctrl;
            if (t.table) {
                len = 1;
                for (uint32_t j = 0; j < s.nctrl && len <= (uint64_t{1} << 20); ++j) len *= qc->d;
                if (len > (uint64_t{1} << 20)) {
                    delete e;
                    return fail(QS_ERR_OUT_OF_RANGE, "qpp::QCircuit::cond_if()", "truth table above 2^20 entries");
                }
            }
            t.ps.assign(qc->pool + s.ps_off, qc->pool + s.ps_off + len);
        }
        if (!in_range) {
            delete e;
            return fail(QS_ERR_OUT_OF_RANGE, W, "ctrl/target");
        }
        if (t.type == QS_STEP_POST_SELECT) {
            t.ps.assign(qc->pool + s.ps_off, qc->pool + s.ps_off + s.ntarget);
            for (uint64_t v : t.ps)
                if (v >= qc->d) {
                    delete e;
                    return fail(QS_ERR_OUT_OF_RANGE, "qpp::QCircuit::post_select()", "ps_vals");
                }
        }
        e->steps.push_back(t);
    }
    // match COND_IF / COND_WHILE / COND_ELSE / COND_END
    // (QCircuit::cond_* and validate_conditionals, qcircuit.hpp:940-1130)
    std::vector<size_t> open;
    int in_while = 0;
    for (size_t i = 0; i < e->steps.size(); ++i) {
        EStep& t = e->steps[i];
        const bool destroys = ((is_meas(t) && t.destructive) || t.type == QS_STEP_DISCARD);
        if (in_while && destroys) {
            delete e;
            return fail(QS_ERR_INVALID_CONDITIONAL, "qpp::QCircuit::measure()",
                        "destructive step inside a while");
        }
        if (t.type == QS_STEP_COND_IF) {
            open.push_back(i);
        } else if (t.type == QS_STEP_COND_WHILE) {
            open.push_back(i);
            ++in_while;
        } else if (t.type == QS_STEP_COND_ELSE) {
            if (open.empty() || e->steps[open.back()].type != QS_STEP_COND_IF) {
                delete e;
                return fail(QS_ERR_INVALID_CONDITIONAL, "qpp::QCircuit::cond_else()",
                            "No matching cond_if()");
            }
            e->steps[open.back()].jump = i;
            open.back() = i;  // the ELSE now waits for the END
        } else if (t.type == QS_STEP_COND_END) {
            if (open.empty()) {
                delete e;
                return fail(QS_ERR_INVALID_CONDITIONAL, "qpp::QCircuit::cond_end()",
                            "No matching cond_if()");
            }
            EStep& top = e->steps[open.back()];
            top.jump = i;
            if (top.type == QS_STEP_COND_WHILE) {
                t.jump = open.back();
                t.loop_end = true;
                --in_while;
            }
            open.pop_back();
        }
    }
    if (!open.empty()) {
        delete e;
        return fail(QS_ERR_INVALID_CONDITIONAL, "qpp::QEngineT::execute()",
                    "Missing QCircuit::cond_end() delimiter");
    }
    e->noise_results.assign(e->steps.size(), {});
    std::vector<uint64_t> dims(qc->nq, qc->d);
    if (int rc = qs_state_create(dtype, kind, qc->nq, dims.data(), &e->st)) {
        delete e;
        return rc;
    }
    e->reset_bookkeeping();
    *out = e;
    return QS_OK;
}

int qs_engine_destroy(qs_engine* e) {
    delete e;
    return QS_OK;
}

int qs_engine_execute(qs_engine* e, uint64_t reps) {
    if (!e || !e->st) return fail(QS_ERR_INVALID_HANDLE, "qpp::QEngineT::execute()");
    qs::clear_error();
    if (reps == 0) return fail(QS_ERR_OUT_OF_RANGE, "qpp::QEngineT::execute()", "reps");
    return e->execute(reps);
}

int qs_engine_num_outcomes(const qs_engine* e, uint64_t* n) {
    if (!e || !n) return fail(QS_ERR_INVALID_HANDLE, "qpp::QEngineT::get_stats()");
    *n = e->stats.size();
    return QS_OK;
}

int qs_engine_get_stats(const qs_engine* e, uint64_t* dits, uint64_t* counts) {
    if (!e) return fail(QS_ERR_INVALID_HANDLE, "qpp::QEngineT::get_stats()");
    size_t k = 0;
    for (const auto& kv : e->stats) {
        if (dits) std::copy(kv.first.begin(), kv.first.end(), dits + k * e->nc);
        if (counts) counts[k] = kv.second;
        ++k;
    }
    return QS_OK;
}

int qs_engine_get_dits(const qs_engine* e, uint64_t* dits) {
    if (!e || !dits) return fail(QS_ERR_INVALID_HANDLE, "qpp::QEngineT::get_dits()");
    std::copy(e->dits.begin(), e->dits.end(), dits);
    return QS_OK;
}

int qs_engine_get_probs(const qs_engine* e, double* probs) {
    if (!e || !probs) return fail(QS_ERR_INVALID_HANDLE, "qpp::QEngineT::get_probs()");
    std::copy(e->probs.begin(), e->probs.end(), probs);
    return QS_OK;
}

int qs_engine_get_state(const qs_engine* e, qs_state** out) {
    if (!e || !out || !e->st) return fail(QS_ERR_INVALID_HANDLE, "qpp::QEngineT::get_state()");
    return qs_state_clone(e->st, out);
}

int qs_engine_set_state(qs_engine* e, const qs_state* st) {
    if (!e || !st) return fail(QS_ERR_INVALID_HANDLE, "qpp::QEngineT::set_state()");
    qs_dtype dt;
    qs_kind kd;
    uint32_t nd = 0;
    uint64_t ne = 0;
    std::vector<uint64_t> dims(128);
    if (int rc = qs_state_info(st, &dt, &kd, &nd, dims.data(), &ne)) return rc;
    uint64_t Dl = e->live_dim();
    const uint64_t D = kd == QS_KET ? ne : static_cast<uint64_t>(std::llround(std::sqrt(double(ne))));
    if (dt != e->dtype || kd != e->kind || D != Dl)
        return fail(QS_ERR_DIMS_NOT_EQUAL, "qpp::QEngineT::set_state()", "state");
    qs_state* c = nullptr;
    if (int rc = qs_state_clone(st, &c)) return rc;
    qs_state_destroy(e->st);
    e->st = c;
    return QS_OK;
}

int qs_engine_reset(qs_engine* e) {
    if (!e) return fail(QS_ERR_INVALID_HANDLE, "qpp::QEngineT::reset()");
    std::vector<uint64_t> dims(e->nq, e->d);
    qs_state* s = nullptr;
    {
        const qs::StreamScope on(e->st ? e->st->stream : qs::current_stream());
        if (int rc = qs_state_create(e->dtype, e->kind, e->nq, dims.data(), &s)) return rc;
    }
    qs_state_destroy(e->st);
    e->st = s;
    e->reset_bookkeeping();
    e->stats.clear();
    e->ps_ok = true;
    e->ensure_ps = false;
    e->max_ps = std::numeric_limits<uint64_t>::max();
    // qnoisy_engine.hpp:152-158 empties the list and the next execute then
    // indexes it; here every step's list is emptied instead
    for (auto& v : e->noise_results) v.clear();
    return QS_OK;
}

int qs_engine_set_dits(qs_engine* e, const uint64_t* dits, uint64_t n) {
    constexpr const char* W = "qpp::QEngineT::set_dits()";
    if (!e || (n && !dits)) return fail(QS_ERR_INVALID_HANDLE, W);
    if (n != e->dits.size()) return fail(QS_ERR_SIZE_MISMATCH, W, "dits");
    std::copy(dits, dits + n, e->dits.begin());
    return QS_OK;
}

int qs_engine_get_measured(const qs_engine* e, uint8_t* measured) {
    if (!e || !measured) return fail(QS_ERR_INVALID_HANDLE, "qpp::QEngineT::get_measured_d()");
    for (uint64_t i = 0; i < e->nq; ++i) measured[i] = e->was_measured(i) ? 1 : 0;
    return QS_OK;
}

int qs_engine_reset_stats(qs_engine* e) {
    if (!e) return fail(QS_ERR_INVALID_HANDLE, "qpp::QEngineT::reset_stats()");
    e->stats.clear();
    return QS_OK;
}

int qs_engine_noise_results(const qs_engine* e, uint64_t step, uint64_t* out, uint64_t* count) {
    constexpr const char* W = "qpp::QNoisyEngineT::get_noise_results()";
    if (!e || !count) return fail(QS_ERR_INVALID_HANDLE, W);
    if (step >= e->noise_results.size()) return fail(QS_ERR_OUT_OF_RANGE, W, "step");
    const auto& v = e->noise_results[step];
    *count = v.size();
    if (out) std::copy(v.begin(), v.end(), out);
    return QS_OK;
}

int qs_engine_set_post_selection(qs_engine* e, int ensure, uint64_t max_reps) {
    if (!e) return fail(QS_ERR_INVALID_HANDLE, "qpp::QEngineT::set_ensure_post_selection()");
    e->ensure_ps = ensure != 0;
    e->max_ps = max_reps;
    return QS_OK;
}

// The give-up counts of unlimited post-selection retries and of runtime
// while loops (2^16 and 2^24 by default; 0 = no limit, as in qpp).
int qs_engine_set_limits(qs_engine* e, uint64_t max_loop_reps, uint64_t max_ps_failures) {
    if (!e) return fail(QS_ERR_INVALID_HANDLE, "qpp::QEngineT::set_limits()");
    e->loop_give_up = max_loop_reps;
    e->ps_give_up = max_ps_failures;
    return QS_OK;
}

int qs_engine_post_select_ok(const qs_engine* e, int* ok) {
    if (!e || !ok) return fail(QS_ERR_INVALID_HANDLE, "qpp::QEngineT::post_select_ok()");
    *ok = e->ps_ok ? 1 : 0;
    return QS_OK;
}

}  // extern "C"

int qs::engine_execute_prefixed(qs_engine* e, qs_state* prefixed, uint64_t reps) {
    constexpr const char* W = "qpp::QEngineT::execute()";
    if (!e || !prefixed || !e->st) {
        qs_state_destroy(prefixed);
        return fail(QS_ERR_INVALID_HANDLE, W);
    }
    if (e->noise_p >= 0 || prefixed->dtype != e->dtype || prefixed->kind != e->kind ||
        prefixed->nelem != e->st->nelem) {
        qs_state_destroy(prefixed);
        return fail(QS_ERR_DIMS_NOT_EQUAL, W, "state");
    }
    bool meas = false, ok = reps > 0;
    for (const EStep& s : e->steps) {
        const bool gate = s.type == QS_STEP_GATE || s.type == QS_STEP_FAN ||
                          s.type == QS_STEP_CTRL || s.type == QS_STEP_CTRL_FAN ||
                          s.type == QS_STEP_NOP;
        if (is_meas(s) && s.type != QS_STEP_POST_SELECT) meas = true;
        else if (!gate || meas) ok = false;
    }
    if (!ok) {
        qs_state_destroy(prefixed);
        return fail(QS_ERR_NOT_IMPLEMENTED, W, "a prefixed run needs gates, then measurements");
    }
    qs::clear_error();
    e->prefixed = prefixed;
    const int rc = e->execute(reps);
    if (e->prefixed) {  // execute() returned before taking it
        qs_state_destroy(e->prefixed);
        e->prefixed = nullptr;
    }
    return rc;
}

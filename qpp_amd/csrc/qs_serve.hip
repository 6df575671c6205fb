// qs_serve.hip -- request server: continuous admission of circuits onto one
// GPU, states resident for the life of a request, overlapping execution.
//
// A request is QEngine(circuit, seed).execute(shots) (classes/qengine.hpp:
// 1341-1405) in complex128 or complex64.  The server keeps `lanes` worker
// threads, each bound to its own HIP stream (default priority); a worker
// takes the oldest queued request whose device footprint fits the free part
// of the HBM budget (first fit: a large request waiting for memory does not
// block small ones behind it), runs it to completion on its stream, releases
// its memory and takes the next.  Requests that arrive while others run are
// admitted as soon as a lane and memory are free, so work from several
// requests is on the device at once.  All device memory of a request is
// stream-ordered (hipMallocAsync from a caching pool), so admitting or
// retiring a request never synchronises the device.
//
// Small requests run batched: a request of kBatchMinQ..kBatchMaxQ qubits
// whose circuit is gates (dense 1-qubit or Clifford on <= 2 qubits)
// followed by one measurement of every qubit is "batchable".  A lane that
// takes one also takes every other queued batchable request of the same
// dtype that fits (up to kMaxBatch), and their gate steps run in ONE
// persistent launch (qs_batch_run_states: the k_traj work queue, one
// workgroup per request, each writing its request's final state).  Each
// request's QEngine::execute(shots) then continues from that state (the
// sampling tail, qs::engine_execute_prefixed).  A batchable request always
// takes this path, alone or in a mix, so its result does not depend on the
// mix: every request's computation is its own workgroup's, in the same
// order of operations.  Above kBatchMaxQ a single workgroup streaming a
// whole state is slower than the per-pass launches that spread it over the
// chip, so larger requests run alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "qs_internal.hpp"

namespace {

using qs::fail;
using Clock = std::chrono::steady_clock;

// owned copy of a qs_circuit
struct CircuitCopy {
    std::vector<qs_step> steps;
    std::vector<uint64_t> mat_dim, mat_off, pool;
    std::vector<double> mats;
    qs_circuit c{};
    void take(const qs_circuit* q) {
        steps.assign(q->steps, q->steps + q->nsteps);
        mat_dim.assign(q->mat_dim, q->mat_dim + q->nmats);
        mat_off.assign(q->mat_off, q->mat_off + q->nmats);
        uint64_t nm = 0, np = 0;
        for (uint32_t i = 0; i < q->nmats; ++i)
            nm = std::max(nm, 2 * (q->mat_off[i] + q->mat_dim[i] * q->mat_dim[i]));
        mats.assign(q->mats, q->mats + nm);
        for (uint32_t s = 0; s < q->nsteps; ++s) {
            const qs_step& st = q->steps[s];
            np = std::max<uint64_t>(np, uint64_t(st.ctrl_off) + st.nctrl);
            np = std::max<uint64_t>(np, uint64_t(st.target_off) + st.ntarget);
            if (st.shift_off != UINT32_MAX) np = std::max<uint64_t>(np, uint64_t(st.shift_off) + st.nctrl);
        }
        pool.assign(q->pool, q->pool + np);
        c = *q;
        c.steps = steps.data();
        c.mat_dim = mat_dim.data();
        c.mat_off = mat_off.data();
        c.mats = mats.data();
        c.pool = pool.data();
    }
};

struct Request {
    uint64_t id = 0;
    CircuitCopy circ;
    qs_dtype dtype = QS_C128;
    uint64_t shots = 1, seed = 0;
    uint64_t footprint = 0;
    Clock::time_point t_submit, t_start, t_done;
    bool done = false;
    int status = QS_OK;
    std::string error;
    std::vector<uint64_t> stat_dits, stat_counts;  // outcomes (n x nc) and counts
    std::vector<uint64_t> dits;
    uint32_t lane = 0;
    bool batchable = false;
    uint32_t batch_size = 0;  // requests in the launch that ran it (1: alone)
};

constexpr uint32_t kBatchMinQ = 12;  // one 2^12-amplitude tile
constexpr uint32_t kBatchMaxQ = 14;
constexpr size_t kMaxBatch = 64;

// What qs_batch accepts, checked on the steps: qubits, gate steps on at most
// two qubits (controls included), then one measurement of every qubit.
bool batchable(const qs_circuit* c) {
    if (c->d != 2 || c->nq < kBatchMinQ || c->nq > kBatchMaxQ || c->nsteps == 0) return false;
    for (uint32_t i = 0; i + 1 < c->nsteps; ++i) {
        const qs_step& st = c->steps[i];
        if (st.type == QS_STEP_NOP) continue;
        if (st.type != QS_STEP_GATE && st.type != QS_STEP_CTRL) return false;
        if (st.nctrl + st.ntarget > 2 || st.ntarget == 0) return false;
        if (st.shift_off != UINT32_MAX) return false;
    }
    const qs_step& m = c->steps[c->nsteps - 1];
    return m.type == QS_STEP_MEASURE_MANY && m.ntarget == c->nq;
}

// Device bytes a request can hold at once, over every path QEngine's
// execute(shots) takes (qengine.hpp:316-481), with D the dimension and s the
// bytes per amplitude:
//   the state                                               s D
//   post-selection retries: the saved state and the first
//   successful rep's state (execute_with_retries)           2 s D
//   a destructive measurement's contracted output           s D / 2
//   the sampler's workspace: FP64 probabilities and the
//   exact cumulative table (2 x 8 D), block tables and
//   per-shot slots (qs_measure.hip)                         16 D + 64 D / 4096 + 24 shots
// plus 2 MiB for the compiled passes' tables and small buffers.  The
// sampling-tail path holds less (the state, its pre-measurement copy and the
// workspace), so this bound covers it.
uint64_t footprint_of(const qs_circuit* c, qs_dtype dt, uint64_t shots) {
    double D = 1;
    for (uint32_t i = 0; i < c->nq; ++i) D *= c->d;
    const double s = dt == QS_C128 ? 16 : 8;
    const double ws = 16 * D + 64 * (D / 4096 + 1) + 24 * double(shots) + 12 * 256;
    return static_cast<uint64_t>(3.5 * s * D + ws + (2 << 20));
}

}  // namespace

struct qs_server {
    uint64_t budget = 0, in_use = 0;
    uint32_t lanes = 0;
    std::vector<hipStream_t> streams;
    std::vector<std::thread> workers;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> all;
    uint64_t next_id = 1;
    bool stop = false;
    Clock::time_point t0 = Clock::now();
    uint64_t completed = 0, peak_in_use = 0, peak_running = 0, running = 0;

    void work(uint32_t lane) {
        qs::set_thread_stream(streams[lane]);
        std::vector<Request*> group;
        for (;;) {
            Request* r = nullptr;
            {
                std::unique_lock<std::mutex> lk(m);
                // wake for a request that fits, or to leave once stopped
                // and drained (a queued request always fits once nothing
                // runs: submit rejects footprints above the budget)
                cv_work.wait(lk, [&] {
                    if (stop && queue.empty()) return true;
                    for (Request* q : queue)
                        if (in_use + q->footprint <= budget) return true;
                    return false;
                });
                if (queue.empty()) return;
                auto it = std::find_if(queue.begin(), queue.end(),
                                       [&](Request* q) { return in_use + q->footprint <= budget; });
                if (it == queue.end()) continue;
                r = *it;
                queue.erase(it);
                group.assign(1, r);
                in_use += r->footprint;
                if (r->batchable)  // the other batchable requests that fit
                    for (auto jt = queue.begin(); jt != queue.end() && group.size() < kMaxBatch;) {
                        Request* q = *jt;
                        if (q->batchable && q->dtype == r->dtype && in_use + q->footprint <= budget) {
                            in_use += q->footprint;
                            group.push_back(q);
                            jt = queue.erase(jt);
                        } else {
                            ++jt;
                        }
                    }
                peak_in_use = std::max(peak_in_use, in_use);
                running += group.size();
                peak_running = std::max(peak_running, running);
            }
            const auto t_start = Clock::now();
            for (Request* q : group) {
                q->lane = lane;
                q->t_start = t_start;
            }
            if (r->batchable)
                run_batch(group);
            else
                run(r);
            const auto t_done = Clock::now();
            {
                std::lock_guard<std::mutex> lk(m);
                for (Request* q : group) {
                    q->t_done = t_done;
                    in_use -= q->footprint;
                    --running;
                    q->done = true;
                    ++completed;
                }
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }

    void run(Request* r) {
        qs_engine* e = nullptr;
        int rc = qs_engine_create(&r->circ.c, QS_KET, r->dtype, -1.0, r->seed, &e);
        if (!rc) rc = qs_engine_execute(e, r->shots);
        r->batch_size = 1;
        finish(r, e, rc);
    }

    // The gate steps of every request of `g` in one launch, then each one's
    // sampling tail from its final state.  If the batch cannot be built, the
    // requests run one by one.
    void run_batch(const std::vector<Request*>& g) {
        std::vector<const qs_circuit*> circs;
        for (Request* q : g) circs.push_back(&q->circ.c);
        qs_batch* b = nullptr;
        std::vector<qs_state*> sts(g.size(), nullptr);
        int rc = qs_batch_create(circs.data(), uint32_t(circs.size()), g[0]->dtype, nullptr, &b);
        for (size_t i = 0; i < g.size() && !rc; ++i) {
            const std::vector<uint64_t> dims(g[i]->circ.c.nq, 2);
            rc = qs_state_create(g[i]->dtype, QS_KET, uint32_t(dims.size()), dims.data(), &sts[i]);
        }
        if (!rc) {
            std::vector<uint32_t> circuit_of(g.size());
            std::iota(circuit_of.begin(), circuit_of.end(), 0u);
            rc = qs_batch_run_states(b, g.size(), circuit_of.data(), sts.data());
        }
        qs_batch_destroy(b);
        if (rc) {  // not batchable after all: each request alone
            for (qs_state* st : sts) qs_state_destroy(st);
            for (Request* q : g) run(q);
            return;
        }
        for (size_t i = 0; i < g.size(); ++i) {
            Request* q = g[i];
            qs_engine* e = nullptr;
            int rq = qs_engine_create(&q->circ.c, QS_KET, q->dtype, -1.0, q->seed, &e);
            if (!rq) {
                rq = qs::engine_execute_prefixed(e, sts[i], q->shots);
            } else {
                qs_state_destroy(sts[i]);
            }
            sts[i] = nullptr;
            q->batch_size = uint32_t(g.size());
            finish(q, e, rq);
        }
    }

    // the request's results from its engine (rc: the run's status)
    void finish(Request* r, qs_engine* e, int rc) {
        if (!rc) {
            uint64_t n = 0;
            rc = qs_engine_num_outcomes(e, &n);
            const uint32_t nc = r->circ.c.nc;
            if (!rc) {
                r->stat_dits.assign(n * nc + 1, 0);
                r->stat_counts.assign(n + 1, 0);
                rc = qs_engine_get_stats(e, r->stat_dits.data(), r->stat_counts.data());
                r->stat_dits.resize(n * nc);
                r->stat_counts.resize(n);
            }
            if (!rc) {
                r->dits.assign(nc + 1, 0);
                rc = qs_engine_get_dits(e, r->dits.data());
                r->dits.resize(nc);
            }
        }
        if (rc) r->error = qs_last_error();
        if (e) qs_engine_destroy(e);
        // the request's memory is back in the pool in stream order; wait so
        // the budget the scheduler sees is the memory actually free
        (void)hipStreamSynchronize(qs::current_stream());
        r->status = rc;
    }
};

extern "C" {

int qs_server_create(uint32_t lanes, uint64_t hbm_budget, qs_server** out) {
    constexpr const char* W = "qpp_amd::server";
    if (!out) return fail(QS_ERR_INVALID_HANDLE, W);
    *out = nullptr;
    if (lanes == 0 || lanes > 64) return fail(QS_ERR_OUT_OF_RANGE, W, "lanes");
    if (int rc = qs::ensure_device()) return rc;
    auto s = std::make_unique<qs_server>();
    if (hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        QS_HIP(hipMemGetInfo(&fr, &tot), W);
        hbm_budget = static_cast<uint64_t>(0.9 * double(fr));
    }
    s->budget = hbm_budget;
    s->lanes = lanes;
    s->streams.resize(lanes);
    for (uint32_t i = 0; i < lanes; ++i)
        QS_HIP(hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking), W);
    qs_server* sp = s.release();
    for (uint32_t i = 0; i < lanes; ++i) sp->workers.emplace_back([sp, i] { sp->work(i); });
    *out = sp;
    return QS_OK;
}

// The device bytes a request is admitted against (footprint_of).
int qs_server_footprint(const qs_circuit* qc, qs_dtype dtype, uint64_t shots, uint64_t* bytes) {
    if (!qc || !bytes) return fail(QS_ERR_INVALID_HANDLE, "qpp_amd::server");
    if (dtype != QS_C128 && dtype != QS_C64) return fail(QS_ERR_TYPE_MISMATCH, "qpp_amd::server", "dtype");
    *bytes = footprint_of(qc, dtype, shots);
    return QS_OK;
}

// Queues QEngine(circuit, seed).execute(shots); the circuit is copied.
int qs_server_submit(qs_server* s, const qs_circuit* qc, qs_dtype dtype, uint64_t shots,
                     uint64_t seed, uint64_t* ticket) {
    constexpr const char* W = "qpp_amd::server";
    if (!s || !qc || !ticket) return fail(QS_ERR_INVALID_HANDLE, W);
    if (dtype != QS_C128 && dtype != QS_C64) return fail(QS_ERR_TYPE_MISMATCH, W, "dtype");
    if (shots == 0) return fail(QS_ERR_OUT_OF_RANGE, W, "shots");
    auto r = std::make_unique<Request>();
    r->circ.take(qc);
    r->dtype = dtype;
    r->shots = shots;
    r->seed = seed;
    r->footprint = footprint_of(qc, dtype, shots);
    r->batchable = batchable(qc);
    if (r->footprint > s->budget) return fail(QS_ERR_OUT_OF_MEMORY, W, "request exceeds the HBM budget");
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->m);
        r->id = s->next_id++;
        *ticket = r->id;
        s->queue.push_back(r.get());
        s->all.emplace(r->id, std::move(r));
    }
    s->cv_work.notify_one();
    return QS_OK;
}

// Blocks until the request is done.  times_ms (optional, 3): queued time,
// run time, submit-to-done latency.  Then qs_server_stats reads the results.
int qs_server_wait(qs_server* s, uint64_t ticket, double* times_ms) {
    constexpr const char* W = "qpp_amd::server";
    if (!s) return fail(QS_ERR_INVALID_HANDLE, W);
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(QS_ERR_OUT_OF_RANGE, W, "ticket");
    Request* r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->done; });
    if (times_ms) {
        using ms = std::chrono::duration<double, std::milli>;
        times_ms[0] = ms(r->t_start - r->t_submit).count();
        times_ms[1] = ms(r->t_done - r->t_start).count();
        times_ms[2] = ms(r->t_done - r->t_submit).count();
    }
    if (r->status) return fail(r->status, W, r->error.c_str());
    return QS_OK;
}

// Statistics of a finished request (QEngine::get_stats order) and its
// classical dits; call with dits/counts NULL to get n_outcomes first.
int qs_server_stats(qs_server* s, uint64_t ticket, uint64_t* n_outcomes, uint64_t* dits,
                    uint64_t* counts, uint64_t* last_dits) {
    constexpr const char* W = "qpp_amd::server";
    if (!s) return fail(QS_ERR_INVALID_HANDLE, W);
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->all.find(ticket);
    if (it == s->all.end() || !it->second->done) return fail(QS_ERR_OUT_OF_RANGE, W, "ticket");
    const Request* r = it->second.get();
    if (n_outcomes) *n_outcomes = r->stat_counts.size();
    if (dits) std::copy(r->stat_dits.begin(), r->stat_dits.end(), dits);
    if (counts) std::copy(r->stat_counts.begin(), r->stat_counts.end(), counts);
    if (last_dits) std::copy(r->dits.begin(), r->dits.end(), last_dits);
    return QS_OK;
}

// How a finished request ran: *batch_size = the number of requests whose
// gate steps ran in the same launch as its own (1: alone, or not batchable).
int qs_server_batch_size(qs_server* s, uint64_t ticket, uint32_t* batch_size) {
    constexpr const char* W = "qpp_amd::server";
    if (!s || !batch_size) return fail(QS_ERR_INVALID_HANDLE, W);
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->all.find(ticket);
    if (it == s->all.end() || !it->second->done) return fail(QS_ERR_OUT_OF_RANGE, W, "ticket");
    *batch_size = it->second->batch_size;
    return QS_OK;
}

// Forgets a finished request (its results).
int qs_server_release(qs_server* s, uint64_t ticket) {
    if (!s) return fail(QS_ERR_INVALID_HANDLE, "qpp_amd::server");
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->all.find(ticket);
    if (it != s->all.end() && it->second->done) s->all.erase(it);
    return QS_OK;
}

// info[6]: completed, queued, running, peak running, bytes in use, peak bytes
int qs_server_info(qs_server* s, uint64_t* info) {
    if (!s || !info) return fail(QS_ERR_INVALID_HANDLE, "qpp_amd::server");
    std::lock_guard<std::mutex> lk(s->m);
    info[0] = s->completed;
    info[1] = s->queue.size();
    info[2] = s->running;
    info[3] = s->peak_running;
    info[4] = s->in_use;
    info[5] = s->peak_in_use;
    return QS_OK;
}

// Finishes queued work, stops the workers and frees the server.
int qs_server_destroy(qs_server* s) {
    if (!s) return QS_OK;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    for (hipStream_t st : s->streams) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    delete s;
    return QS_OK;
}

}  // extern "C"

/*
 * oracle.h -- CPU restatement of qpp 7.0.3's gate-application path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU engine is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (qpp_amd/) never calls it.
 *
 * qpp itself cannot be compiled here (it needs Eigen >= 5.0, fetched at
 * configure time, cmake/qpp_eigen.cmake:9-43), so this is a restatement in
 * plain C++17 + OpenMP of the reference's algorithms, each function citing
 * the file:line it follows (paths relative to the qpp source tree).
 *
 * Storage: interleaved complex doubles, column-major (Eigen's layout);
 * rho(i, j) at element i + j*D.  Index vectors are uint64_t.
 * Return value: 0 on success, otherwise a qs_status code (qpp_amd.h).
 */
#ifndef QPP_ORACLE_H_
#define QPP_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* apply(state, A, target, dims): operations.hpp:72-372 dispatch;
 * qubit kernels apply.hpp:57-1167 when all dims are 2. */
int orc_apply(const double* state, int is_rho, uint32_t ndims,
              const uint64_t* dims, const double* A, uint32_t ntarget,
              const uint64_t* target, double* out);

/* applyCTRL: operations.hpp:616-887; apply_ctrl.hpp:63-1124. */
int orc_apply_ctrl(const double* state, int is_rho, uint32_t ndims,
                   const uint64_t* dims, const double* A, uint32_t nctrl,
                   const uint64_t* ctrl, uint32_t ntarget,
                   const uint64_t* target, const uint64_t* shift,
                   double* out);

/* applyCTRL_fan: operations.hpp:961-1234; apply_ctrl_fan.hpp:62-303. */
int orc_apply_ctrl_fan(const double* state, int is_rho, uint32_t ndims,
                       const uint64_t* dims, const double* A, uint32_t nctrl,
                       const uint64_t* ctrl, uint32_t ntarget,
                       const uint64_t* target, const uint64_t* shift,
                       double* out);

/* apply(A, Ks): operations.hpp:422-464; Ks are Dout x D (rectangular
 * allowed), column-major, back to back; out is Dout x Dout. */
int orc_apply_kraus_full(const double* rho, uint64_t D, uint32_t nK, const double* Ks,
                         uint64_t Dout, double* out);

/* apply(rho, Ks, target, dims): operations.hpp:476-560. */
int orc_apply_kraus(const double* rho, uint32_t ndims, const uint64_t* dims,
                    uint32_t nK, const double* Ks, uint32_t ntarget,
                    const uint64_t* target, double* out);

/* ptrace: operations.hpp:1908-2114; ptrace.hpp:61-257.  out: Dbar x Dbar. */
int orc_ptrace(const double* A, int is_rho, uint32_t ndims,
               const uint64_t* dims, uint32_t ntarget, const uint64_t* target,
               double* out);

/* ptranspose: operations.hpp:2167-2335 (sem 0); ptranspose.hpp:59-219
 * (sem 1, qubits only).  out: D x D. */
int orc_ptranspose(const double* A, int is_rho, uint32_t ndims,
                   const uint64_t* dims, uint32_t ntarget,
                   const uint64_t* target, int sem, double* out);

/* syspermute: operations.hpp:2386-2514 (sem 0); syspermute.hpp:58-180
 * (sem 1, qubits only). */
int orc_syspermute(const double* A, int is_rho, uint32_t ndims,
                   const uint64_t* dims, const uint64_t* perm, int sem,
                   double* out);

/* std::norm per element / Re diag: functions.hpp:1756-1800,
 * instruments.hpp:783-792. */
int orc_probs(const double* A, int is_rho, uint64_t D, double* out);

/* libstdc++ discrete_distribution<size_t>(p, p+N) evaluated at uniform u
 * (random.tcc:2656-2712): normalise by the sequential sum, sequential
 * partial_sum, last := 1, lower_bound.  Returns the index. */
uint64_t orc_discrete_draw(const double* p, uint64_t N, double u);
/* Same, also returning the cumulative table (N doubles) it built. */
void orc_discrete_table(const double* p, uint64_t N, double* cp);

/* sample(A, target, dims) with the given uniform: instruments.hpp:746-809. */
int orc_sample(const double* A, int is_rho, uint32_t ndims,
               const uint64_t* dims, uint32_t ntarget, const uint64_t* target,
               double u, uint64_t* out_index, uint64_t* out_digits,
               double* out_prob);

/* measure_seq(A, target, dims, destructive) with the given uniform:
 * instruments.hpp:985-1099.  out_state receives the post-measurement state
 * (capacity: the input size), *out_len its number of complex elements. */
int orc_measure_seq(const double* A, int is_rho, uint32_t ndims,
                    const uint64_t* dims, uint32_t ntarget,
                    const uint64_t* target, int destructive, double u,
                    uint64_t* results, double* probs, double* out_state,
                    uint64_t* out_len);

/* reset(A, target, dims) with the given uniform: instruments.hpp:1148-1196.
 * out: same size as A. */
int orc_reset(const double* A, int is_rho, uint32_t ndims, const uint64_t* dims,
              uint32_t ntarget, const uint64_t* target, double u, double* out);

/* std::mt19937 + std::generate_canonical<double, 53> (the draws qpp's
 * RandomDevices feeds to every distribution; random.h:164-192). */
void* orc_rng_new(uint64_t seed);
double orc_rng_canonical(void* rng);
void orc_rng_free(void* rng);

/* CPU baseline: runs `ngates` gates on |0..0> of n qubits with the restated
 * qubit kernels apply_psi_1q (kind 0, apply.hpp:57-126) and
 * apply_ctrl_psi_1q (kind 1, apply_ctrl.hpp:63-177), value semantics and
 * OpenMP placement as in qpp, and writes each gate's wall time (seconds,
 * std::chrono::steady_clock as qpp::Timer, classes/timer.hpp:51) to
 * seconds[].  mats: 4 complex (column-major) per gate.  Returns the thread
 * count OpenMP used. */
int orc_time_gates(uint32_t n, uint32_t ngates, const uint32_t* kind,
                   const uint32_t* target, const uint32_t* ctrl,
                   const double* mats, double* seconds);

/* QEngineT / QNoisyEngineT execute(reps) over a qs_circuit program
 * (qengine.hpp:240-481, 533-823, 1341-1405; qnoisy_engine.hpp:104-140;
 * noise.hpp:131-160, 388-408).  noise_p < 0 means no noise.  is_rho
 * selects QEngineT<cmat>.  Returns statistics (outcomes in std::map order),
 * the final classical dits/probs and the final state.
 * stats_dits: capacity max_outcomes * nc; returns count in *n_outcomes. */
struct qs_circuit;
int orc_engine_run(const struct qs_circuit* qc, int is_rho, double noise_p,
                   uint64_t seed, uint64_t reps, uint64_t max_outcomes,
                   uint64_t* n_outcomes, uint64_t* stats_dits,
                   uint64_t* stats_counts, uint64_t* dits, double* probs,
                   double* out_state, uint64_t* out_len);
/* The same with set_ensure_post_selection(ensure_ps) and
 * set_max_post_selection_reps(max_ps) (qengine.hpp:1177-1227) before the
 * run; *ps_ok (may be NULL) receives post_select_ok() after it.  Reset,
 * discard and post-selection steps: qengine.hpp:639-823.  Returns 0, or the
 * qs_status code of the error the run stopped at: QS_ERR_QUDIT_ALREADY_MEASURED
 * for a step on a destructively measured qudit (qengine.hpp:514-526),
 * QS_ERR_OUT_OF_RANGE for a post-selection retried without limit that
 * failed 2^16 times. */
/* give-up counts of the following engine runs: a while loop whose body has
 * run max_loop_reps times with its condition true, an unlimited
 * post-selection that failed max_ps_failures times (2^24, 2^16 by
 * default; 0 = never, qpp's behaviour) */
int orc_set_engine_limits(uint64_t max_loop_reps, uint64_t max_ps_failures);
int orc_engine_run_ps(const struct qs_circuit* qc, int is_rho, double noise_p,
                      uint64_t seed, uint64_t reps, int ensure_ps,
                      uint64_t max_ps, uint64_t max_outcomes,
                      uint64_t* n_outcomes, uint64_t* stats_dits,
                      uint64_t* stats_counts, uint64_t* dits, double* probs,
                      double* out_state, uint64_t* out_len, int* ps_ok);

/* QNoisyEngineT::get_noise_results() of the engine the last
 * orc_engine_run(_ps) call ran: step `step`'s Kraus indices (out may be
 * NULL; *count gets the number). */
int orc_last_noise_results(uint64_t step, uint64_t* out, uint64_t* count);

#ifdef __cplusplus
}
#endif

#endif /* QPP_ORACLE_H_ */

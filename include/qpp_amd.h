/*
 * qpp_amd.h -- C ABI of the MI355X (gfx950) gate-application engine.
 *
 * The engine replaces the hot path of Quantum++ 7.0.3 (qpp): qpp::apply /
 * applyCTRL / applyCTRL_fan on kets and density matrices, apply(rho, Kraus),
 * ptrace, ptranspose, syspermute, sample / measure_seq and the QEngine /
 * QNoisyEngine execute loop.  Every entry point below names the qpp function
 * it stands in for (paths relative to the qpp source tree).
 *
 * Conventions (identical to qpp):
 *   - subsystem 0 is the most significant digit of a basis index
 *     (include/qpp/internal/util.hpp:58-116);
 *   - a gate acting on target[0..k-1] has target[0] as the most significant
 *     digit of its row/column index (internal/kernels/qubit/apply.hpp:527-534);
 *   - matrices, kets and density matrices are column-major arrays of
 *     interleaved (re, im) doubles, exactly Eigen's storage of qpp::cmat /
 *     qpp::ket, so a caller passes `U.data()` unchanged; rho(i, j) lives at
 *     element i + j * D;
 *   - index vectors (target, ctrl, shift, perm, dims) are uint64_t arrays,
 *     i.e. std::vector<qpp::idx>::data().
 *
 * States live in device memory behind an opaque handle.  Operations on a
 * handle are in place and ordered on the handle's HIP stream; the qpp-style
 * value-semantics wrappers (copy in, compute, copy out) sit on top in
 * qpp_amd.hpp.  Every call validates its arguments on the host, in the order
 * qpp does, before anything is launched, and returns a qs_status whose
 * non-zero values correspond one-to-one to qpp::exception classes
 * (include/qpp/classes/exception.hpp:129-667).  qs_last_error() returns the
 * message qpp's what() would print ("[WHERE: qpp::apply()] ...!").
 *
 * Nothing here allocates on a hidden RNG: every random draw is a uniform
 * double in [0, 1) supplied by the caller, or drawn from a std::mt19937 the
 * caller seeds (the generator qpp's RandomDevices holds,
 * include/qpp/classes/random_devices.hpp:55-91).
 */
#ifndef QPP_AMD_H_
#define QPP_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes: one per qpp::exception class used on the path.              */
/* ------------------------------------------------------------------------ */
typedef enum qs_status {
    QS_OK = 0,
    QS_ERR_ZERO_SIZE = 1,                /* exception::ZeroSize */
    QS_ERR_MATRIX_NOT_SQUARE = 2,        /* exception::MatrixNotSquare */
    QS_ERR_MATRIX_NOT_CVECTOR = 3,       /* exception::MatrixNotCvector */
    QS_ERR_MATRIX_NOT_SQUARE_NOR_CVECTOR = 4, /* MatrixNotSquareNorCvector */
    QS_ERR_MATRIX_MISMATCH_SUBSYS = 5,   /* exception::MatrixMismatchSubsys */
    QS_ERR_DIMS_INVALID = 6,             /* exception::DimsInvalid */
    QS_ERR_DIMS_NOT_EQUAL = 7,           /* exception::DimsNotEqual */
    QS_ERR_DIMS_MISMATCH_MATRIX = 8,     /* exception::DimsMismatchMatrix */
    QS_ERR_DIMS_MISMATCH_CVECTOR = 9,    /* exception::DimsMismatchCvector */
    QS_ERR_SUBSYS_MISMATCH_DIMS = 10,    /* exception::SubsysMismatchDims */
    QS_ERR_PERM_INVALID = 11,            /* exception::PermInvalid */
    QS_ERR_PERM_MISMATCH_DIMS = 12,      /* exception::PermMismatchDims */
    QS_ERR_OUT_OF_RANGE = 13,            /* exception::OutOfRange */
    QS_ERR_TYPE_MISMATCH = 14,           /* exception::TypeMismatch */
    QS_ERR_SIZE_MISMATCH = 15,           /* exception::SizeMismatch */
    QS_ERR_QUDIT_ALREADY_MEASURED = 16,  /* exception::QuditAlreadyMeasured */
    QS_ERR_NOT_IMPLEMENTED = 17,         /* exception::NotImplemented */
    QS_ERR_OVERFLOW = 18,                /* std::overflow_error (normalize) */
    QS_ERR_IO = 19,                      /* std::runtime_error (load/save streams) */
    QS_ERR_INVALID_CONDITIONAL = 20,     /* exception::InvalidConditional */
    /* engine-side conditions with no qpp counterpart */
    QS_ERR_INVALID_HANDLE = 100,
    QS_ERR_OUT_OF_MEMORY = 101,          /* device memory exhausted */
    QS_ERR_DEVICE = 102,                 /* HIP runtime error */
    QS_ERR_NO_DEVICE = 103               /* no gfx950 device / library not built for it */
} qs_status;

typedef enum qs_dtype { QS_C128 = 0, QS_C64 = 1 } qs_dtype;
typedef enum qs_kind { QS_KET = 0, QS_RHO = 1 } qs_kind;

/* ptranspose / syspermute semantics.  qpp's generic (documented) path and
 * its qubit kernels disagree for most targets / permutations
 * (internal/kernels/qubit/ptranspose.hpp:101-119, syspermute.hpp:70-93
 * against operations.hpp:2246-2331, 2480-2495); both are offered. */
typedef enum qs_semantics {
    QS_SEM_DOCUMENTED = 0, /* generic path: operations.hpp */
    QS_SEM_AS_BUILT = 1    /* QPP_QUBIT_OPTIMIZATIONS qubit kernels */
} qs_semantics;

typedef struct qs_state qs_state; /* opaque, device resident */

/* Human-readable message of the last failing call on this thread. */
const char* qs_last_error(void);
/* Library / device information (device name, CU count, HBM bytes). */
int qs_device_info(char* name, size_t name_len, int* num_cu,
                   uint64_t* hbm_bytes, uint64_t* hbm_free);
/* Synchronise the device (all streams the library uses). */
int qs_synchronize(void);
/* Timing on the stream the library launches its kernels on: begin records a
 * HIP event, end records a second one, waits for it and returns the elapsed
 * milliseconds between the two. */
int qs_profile_begin(void);
int qs_profile_end(double* ms);

/* ------------------------------------------------------------------------ */
/* State handles                                                             */
/* ------------------------------------------------------------------------ */
/* Allocate a state over subsystems of dimensions dims[0..ndims-1]: a ket of
 * D = prod(dims) amplitudes or a D x D density matrix.  Contents are |0..0>
 * (ket) or |0..0><0..0| (rho), as QEngineState's initial state
 * (internal/classes/qengine_state.hpp:137-143). */
int qs_state_create(qs_dtype dtype, qs_kind kind, uint32_t ndims,
                    const uint64_t* dims, qs_state** out);
int qs_state_destroy(qs_state* st);
int qs_state_clone(const qs_state* st, qs_state** out);
int qs_state_info(const qs_state* st, qs_dtype* dtype, qs_kind* kind,
                  uint32_t* ndims, uint64_t* dims_out /* >= ndims */,
                  uint64_t* num_elems);
/* Host <-> device copies of the whole state.  `host` holds num_elems
 * interleaved complex values of the state's dtype (complex128: 16 B each,
 * complex64: 8 B each), column-major. */
int qs_state_upload(qs_state* st, const void* host, uint64_t nbytes);
int qs_state_download(const qs_state* st, void* host, uint64_t nbytes);
/* Files in qpp's formats (input_output.hpp), so states travel between qpp
 * and the engine.  A ket is written as a D x 1 matrix, a density matrix as
 * D x D; complex64 states are written as complex128 (qpp's cplx).
 *   QS_FMT_TEXT: qpp::save(A, os) / qpp::load<cmat>(is), input_output.hpp:
 *     177-211 / 233-262: "rows cols\n", then one line per row of "(re,im)"
 *     entries separated by single spaces, each real with max_digits10 (17)
 *     significant digits (internal/util.hpp:440-446).
 *   QS_FMT_BINARY: qpp::obsolete::save / load, input_output.hpp:318-345 /
 *     357-390: the 19 bytes "TYPE::Eigen::Matrix", rows and cols as 8-byte
 *     size_t, then the column-major complex<double> data.
 * qs_state_load creates a state with the given dims from a file whose
 * rows x cols must be D x 1 (QS_KET) or D x D (QS_RHO).  Stream errors
 * return QS_ERR_IO with qpp's std::runtime_error text. */
typedef enum qs_format { QS_FMT_TEXT = 0, QS_FMT_BINARY = 1 } qs_format;
int qs_state_save(const qs_state* st, const char* path, qs_format fmt);
int qs_state_load(const char* path, qs_format fmt, qs_dtype dtype, qs_kind kind,
                  uint32_t ndims, const uint64_t* dims, qs_state** out);
/* Basis state |index> or |index><index|. */
int qs_state_set_basis(qs_state* st, uint64_t index);
/* Raw device pointer and HIP stream of a handle (for zero-copy callers). */
int qs_state_device_ptr(const qs_state* st, void** dptr);

/* ------------------------------------------------------------------------ */
/* Gate application (in place).  A: column-major (2^k..) x (..) complex128.   */
/* ------------------------------------------------------------------------ */
/* qpp::apply(state, A, target, dims)  operations.hpp:72-372;
 * qubit kernels apply_psi_{1q,2q,3q,kq}, apply_rho_{1q,2q,3q,kq}
 * (internal/kernels/qubit/apply.hpp:57-1167). */
int qs_apply(qs_state* st, const double* A, uint64_t dimA, uint32_t ntarget,
             const uint64_t* target);

/* qpp::applyCTRL(state, A, ctrl, target, dims, shift)  operations.hpp:616-887;
 * qubit kernels apply_ctrl_{psi,rho}_{1q,2q,kq} (apply_ctrl.hpp:63-1124).
 * shift may be NULL (std::nullopt). */
int qs_apply_ctrl(qs_state* st, const double* A, uint64_t dimA, uint32_t nctrl,
                  const uint64_t* ctrl, uint32_t ntarget,
                  const uint64_t* target, const uint64_t* shift);

/* qpp::applyCTRL_fan(state, A, ctrl, target, dims, shift)
 * operations.hpp:961-1234; apply_ctrl_fan_{psi,rho}
 * (internal/kernels/qubit/apply_ctrl_fan.hpp:62-303).  A acts on one
 * subsystem and is applied to every target under the same controls. */
int qs_apply_ctrl_fan(qs_state* st, const double* A, uint64_t dimA,
                      uint32_t nctrl, const uint64_t* ctrl, uint32_t ntarget,
                      const uint64_t* target, const uint64_t* shift);

/* qpp::apply(rho, Ks, target, dims)  operations.hpp:476-560: the channel
 * sum_K K rho K^dagger, fused into one pass (one term per operator when the
 * targets span more than 32 basis states).  Ks: nK matrices, each
 * dimK x dimK, stored back to back.  A ket fails with
 * QS_ERR_MATRIX_NOT_SQUARE, as in qpp. */
int qs_apply_kraus(qs_state* rho, const double* Ks, uint32_t nK, uint64_t dimK,
                   uint32_t ntarget, const uint64_t* target);

/* qpp::apply(A, Ks)  operations.hpp:422-464: sum_K K rho K^dagger over the
 * whole system.  The Kraus operators may be rectangular: each is
 * rowsK x colsK (colsK = the dimension of rho), stored back to back
 * column-major; the result is a new rowsK x rowsK density matrix (one
 * subsystem of dimension rowsK) in *out.  Errors in qpp's order:
 * MatrixNotSquare (a ket), ZeroSize (no operators), DimsMismatchMatrix
 * (colsK != dimension of rho). */
int qs_apply_kraus_full(const qs_state* rho, const double* Ks, uint32_t nK,
                        uint64_t rowsK, uint64_t colsK, qs_state** out);

/* ------------------------------------------------------------------------ */
/* Reductions and index maps (out-of-place: create a new handle).            */
/* ------------------------------------------------------------------------ */
/* qpp::ptrace(A, target, dims)  operations.hpp:1908-2114;
 * ptrace_{psi,rho}_kq (internal/kernels/qubit/ptrace.hpp:61-257).
 * Result: density matrix over the kept subsystems. */
int qs_ptrace(const qs_state* st, uint32_t ntarget, const uint64_t* target,
              qs_state** out);
/* qpp::ptranspose(A, target, dims)  operations.hpp:2167-2335;
 * ptranspose_{psi,rho}_kq (internal/kernels/qubit/ptranspose.hpp:59-219).
 * Result: density matrix. */
int qs_ptranspose(const qs_state* st, uint32_t ntarget, const uint64_t* target,
                  qs_semantics sem, qs_state** out);
/* qpp::syspermute(A, perm, dims)  operations.hpp:2386-2514;
 * syspermute_{psi,rho}_kq (internal/kernels/qubit/syspermute.hpp:58-180). */
int qs_syspermute(const qs_state* st, const uint64_t* perm, qs_semantics sem,
                  qs_state** out);

/* Squared 2-norm (ket) or real trace (rho), FP64 accumulation
 * (qpp::norm functions.hpp:282-294, trace). */
int qs_norm2(const qs_state* st, double* out);
/* qpp::normalize(A)  functions.hpp:304-333: divide by the norm (ket) or
 * by the trace (rho); QS_ERR_OVERFLOW on zero. */
int qs_normalize(qs_state* st);

/* ------------------------------------------------------------------------ */
/* Precision (complex64 path)                                                */
/* ------------------------------------------------------------------------ */
/* No single qpp counterpart: qpp computes in one dtype.  These report how
 * far a complex64 result is from a complex128 one, accumulated in FP64:
 * out[5] = { Re<a|b>, Im<a|b>, <a|a>, <b|b>, max_i |a_i - b_i| }
 * (for density matrices <a|b> = tr(a^dagger b)).  The infidelity of pure
 * states is 1 - |<a|b>|^2 / (<a|a><b|b>) and their trace distance
 * sqrt(infidelity). */
int qs_compare(const qs_state* a, const qs_state* b, double* out);
/* A copy of st stored as dtype (complex128 <-> complex64). */
int qs_state_convert(const qs_state* st, qs_dtype dtype, qs_state** out);
/* dst := src (same shape and dtype), asynchronous on dst's stream, with a
 * plain streaming kernel. */
int qs_state_copy(qs_state* dst, const qs_state* src);

/* ------------------------------------------------------------------------ */
/* Measurement                                                               */
/* ------------------------------------------------------------------------ */
/* Z-basis probabilities |psi_i|^2 (ket, re*re + im*im as std::norm) or
 * Re rho(i,i) (rho), written to host memory (num basis states doubles).
 * qpp::abssq functions.hpp:1756-1800; instruments.hpp:783-792. */
int qs_probs(const qs_state* st, double* host_out, uint64_t n_out);

/* qpp::sample(A, target, dims) repeated nshots times with caller uniforms
 * u[0..nshots-1]: each shot picks basis index
 *   lower_bound(partial_sum(p / sum(p)) with last := 1, u)
 * exactly as libstdc++'s discrete_distribution does
 * (instruments.hpp:746-809; /usr/include/c++/11/bits/random.tcc:2656-2712).
 * Outputs (host, any may be NULL): full basis index per shot, target digits
 * per shot (nshots x ntarget, row per shot), probability p[index] per shot.
 * The chosen indices are bit-identical to the sequential reference. */
int qs_sample(const qs_state* st, uint32_t ntarget, const uint64_t* target,
              uint64_t nshots, const double* u, uint64_t* out_index,
              uint64_t* out_digits, double* out_prob);

/* No counterpart in the reference.  qs_sample for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): same arguments, validation, results and
 * errors as qs_sample, but the call allocates nothing.  Its workspace is a
 * fixed array in device memory, one per GPU (about 16.4 MiB, held until the
 * process ends), so a call takes at most 2^20 basis states and 16384 shots
 * (larger ones fail with QS_ERR_OUT_OF_MEMORY before anything runs).  Every
 * shot is searched in the exact cumulative table.  Calls that overlap on one
 * GPU share that array and race: serialise them. */
int qs_sample_noalloc(const qs_state* st, uint32_t ntarget, const uint64_t* target,
                      uint64_t nshots, const double* u, uint64_t* out_index,
                      uint64_t* out_digits, double* out_prob);

/* qpp::measure_seq(A, target, dims, destructive)  instruments.hpp:985-1099,
 * in place, with one caller uniform for the sample.  results/probs: ntarget
 * entries in target order.  A destructive measurement removes the measured
 * subsystems from the handle (its dims shrink). */
int qs_measure_seq(qs_state* st, uint32_t ntarget, const uint64_t* target,
                   int destructive, double u, uint64_t* results,
                   double* probs);

/* qpp::reset(A, target, dims)  instruments.hpp:1148-1196, in place: a
 * non-destructive measure_seq of the targets with the caller uniform u,
 * then Xd(dims[i])^(dims[i] - result[i]) on target[i].  As in qpp the
 * correction's dimension is dims[i] (the i-th subsystem's), not
 * dims[target[i]]: where they differ qpp's apply() throws
 * MatrixMismatchSubsys, which is reported here before anything runs. */
int qs_reset(qs_state* st, uint32_t ntarget, const uint64_t* target, double u);

/* ------------------------------------------------------------------------ */
/* Circuits and engines                                                      */
/* ------------------------------------------------------------------------ */
/* A circuit is a flat program of steps (the VarStep list of qpp::QCircuit,
 * classes/qcircuit.hpp:139-143, lowered to POD).  Gate steps reference a
 * matrix table; index lists live in one pool. */
typedef enum qs_step_type {
    QS_STEP_NOP = 0,
    QS_STEP_GATE = 1,        /* SINGLE/TWO/THREE/JOINT: apply(U, target) */
    QS_STEP_FAN = 2,         /* FAN: apply(U, {t}) for every target */
    QS_STEP_CTRL = 3,        /* CTRL: applyCTRL(U, ctrl, target, shift) */
    QS_STEP_CTRL_FAN = 4,    /* CTRL_FAN: applyCTRL_fan */
    QS_STEP_MEASURE = 5,     /* MEASURE / MEASURE_ND (flag) on target[0] */
    QS_STEP_MEASURE_MANY = 6, /* MEASURE_MANY / _ND on all targets */
    QS_STEP_RESET = 7,        /* RESET / RESET_MANY: reset(state, target, d),
                                 qengine.hpp:804-808 */
    QS_STEP_DISCARD = 8,      /* DISCARD / DISCARD_MANY: destructive
                                 measure_seq, results dropped, :810-821 */
    QS_STEP_POST_SELECT = 9,  /* POST_SELECT(_MANY) / _ND (flag): measure_seq
                                 of all targets, written from c_reg; the run
                                 goes on only if the results equal the
                                 post-selection values, :650-733 */
    QS_STEP_COND_IF = 10,     /* runtime if (qcircuit.hpp cond_if,
                                 qengine.hpp:837-860): true when dit
                                 pool[ctrl_off + j] equals pool[ps_off + j]
                                 for j < nctrl; false jumps past the matching
                                 COND_ELSE, or else COND_END */
    QS_STEP_COND_ELSE = 11,   /* reached from the true branch: jumps past
                                 the matching COND_END */
    QS_STEP_COND_END = 12,    /* end of the conditional block */
    QS_STEP_CCTRL = 13,       /* cCTRL (qengine.hpp:581-630): the nctrl dits
                                 at ctrl_off, each plus its shift mod d, must
                                 all equal one value k; then U^k on the
                                 targets */
    QS_STEP_CCTRL_FAN = 14,   /* cCTRL_FAN: the same, U^k on each target */
    QS_STEP_COND_WHILE = 15,  /* runtime while (cond_while, qengine.hpp:837-880):
                                 the COND_IF condition; false jumps past the
                                 matching COND_END, whose visit re-evaluates it
                                 and, while true, goes back to the body.
                                 Destructive measurements, post-selections and
                                 discards inside a while are rejected
                                 (InvalidConditional, as QCircuit does) */
    /* No qpp form of their own: COND_IF and COND_WHILE with the condition
     * given as a truth table, the C-ABI form of qpp's predicate lambdas
     * (cond_if(pred), qcircuit.hpp:1026; visit_runtime_step_,
     * qengine.hpp:837-880) over the nctrl dits at ctrl_off: true when
     * pool[ps_off + i] != 0, i = sum_j dit(ctrl[j]) * d^(nctrl-1-j).  The
     * table has d^nctrl entries (at most 2^20); a dit outside [0, d) when the
     * condition is evaluated is OutOfRange. */
    QS_STEP_COND_IF_TABLE = 16,
    QS_STEP_COND_WHILE_TABLE = 17
} qs_step_type;

typedef struct qs_step {
    uint32_t type;        /* qs_step_type */
    uint32_t gate;        /* matrix index (gate steps) */
    uint32_t nctrl;       /* number of controls */
    uint32_t ntarget;     /* number of targets */
    uint32_t ctrl_off;    /* offset of the controls in the index pool */
    uint32_t target_off;  /* offset of the targets in the index pool */
    uint32_t shift_off;   /* offset of the shifts, UINT32_MAX = no shift */
    uint32_t c_reg;       /* first classical dit written by a measurement */
    uint32_t destructive; /* measurement: 1 = destructive (MEASURE), 0 = _ND */
    uint32_t ps_off;      /* POST_SELECT: offset of ntarget values in the pool */
} qs_step;

typedef struct qs_circuit {
    uint32_t nq;              /* number of qudits */
    uint32_t nc;              /* number of classical dits */
    uint32_t d;               /* local dimension (2 for qubits) */
    uint32_t nsteps;
    const qs_step* steps;
    uint32_t nmats;
    const uint64_t* mat_dim;  /* nmats: matrix dimension */
    const uint64_t* mat_off;  /* nmats: offset in complex elements */
    const double* mats;       /* interleaved complex, column-major */
    const uint64_t* pool;     /* index pool */
} qs_circuit;

/* Fused programs: the gate steps [step_begin, step_end) of a qubit circuit
 * compiled into tile passes (each pass streams the state through HBM once
 * and applies many gates from registers).  Gate steps only (GATE with 1 or
 * 2 targets, FAN, CTRL / CTRL_FAN with one target per application).  The
 * same gates applied one qs_apply / qs_apply_ctrl at a time give the same
 * state up to rounding (the pass order only swaps gates on disjoint
 * qubits). */
typedef struct qs_program qs_program;
typedef struct qs_fuse_opts {
    uint32_t reg_bits; /* log2(amplitudes per thread): 3 or 4 (tile 2^11 or 2^12) */
    uint32_t low_bits; /* low index bits kept in every tile (coalescing) */
    uint32_t max_ops;  /* gates per pass at most */
    uint32_t grid_cap; /* workgroups per pass launch at most (0 = default) */
} qs_fuse_opts;
int qs_program_create(const qs_circuit* qc, uint32_t step_begin,
                      uint32_t step_end, qs_kind kind, qs_dtype dtype,
                      const qs_fuse_opts* opts /* NULL: defaults */,
                      qs_program** out);
/* A program of bit-level operations on a 2^nb-element state (a ket over nb
 * qubits, or vec(rho) over 2n bits: rho(i, j) at i + j*2^n, so row qubit q
 * is bit n-1-q and column qubit q bit 2n-1-q).  Each operation is a dense
 * 2^nbits x 2^nbits column-major matrix at mats + 2*mat_off (complex128,
 * interleaved) on index bits `bits` (bits[0] = most significant digit of
 * the matrix index), applied where (index & cmask) == cval.  A unitary on
 * rho is U on the row bit and conj(U) on the column bit; a 1-qubit channel
 * {K} on rho is the superoperator sum_K conj(K) (x) K on (column, row) bits
 * (the form qs_apply_kraus uses; operations.hpp:476-560).  No single qpp
 * counterpart: it is what a fused sequence of apply / apply(rho, Ks) calls
 * computes. */
typedef struct qs_bitgate {
    uint32_t nbits;   /* 1 or 2 */
    uint32_t bits[2];
    uint32_t pad;
    uint64_t cmask, cval;
    uint64_t mat_off; /* complex elements */
} qs_bitgate;
int qs_program_create_bits(uint32_t nb, uint32_t ngates, const qs_bitgate* gates,
                           const double* mats, qs_dtype dtype,
                           const qs_fuse_opts* opts, qs_program** out);
int qs_program_run(qs_program* p, qs_state* st);
/* No qpp counterpart: capture the program's pass launches for this state's
 * buffer on its stream as a HIP graph (the passes run once, as in
 * qs_program_run).  Later qs_program_run calls on the same buffer and
 * stream replay the graph (one launch); other states run the passes
 * directly.  A later capture or qs_program_compile_native drops it. */
int qs_program_capture(qs_program* p, qs_state* st);
/* No qpp counterpart: generate the program's passes as specialised kernels
 * (matrices and layouts as literals, register-only X gates as renamings, no
 * op dispatch), compile them with hipRTC for gfx950 and use them in later
 * runs (reg_bits 4 programs).  Takes seconds; a failure leaves the program
 * on the general tile kernel. */
int qs_program_compile_native(qs_program* p);
/* No qpp counterpart: diagnostics of the specialised kernels built by
 * qs_program_compile_native.  counts[0..4]: layout changes moved whole,
 * split on a wave bit, split on a lane bit, split on a register bit, and
 * passes whose kernels need half the LDS (all their changes split).  All
 * zero before a successful compile. */
int qs_program_native_stats(const qs_program* p, uint64_t counts[5]);
/* qs_program_run with a HIP event recorded on the state's stream before the
 * first pass and after each pass; waits for the last one and writes each
 * pass's elapsed milliseconds to pass_ms[0 .. num_passes-1] (the kernel
 * durations, launches being back to back).  No qpp counterpart:
 * measurement. */
int qs_program_run_timed(qs_program* p, qs_state* st, double* pass_ms);
int qs_program_info(const qs_program* p, uint64_t* num_gates,
                    uint64_t* num_passes, uint64_t* num_phases);
int qs_program_destroy(qs_program* p);

typedef struct qs_engine qs_engine;

/* Engine over a circuit (a deep copy is taken).  kind QS_KET is
 * QEngine (QEngineT<ket>), QS_RHO is QEngineT<cmat>.  noise_p < 0: no noise;
 * noise_p >= 0: QNoisyEngine<QubitDepolarizingNoise(noise_p)>
 * (classes/qengine.hpp:77-1516, classes/qnoisy_engine.hpp:59-178,
 * classes/noise.hpp:388-408).  The engine's std::mt19937 is seeded with
 * `seed`; it draws uniforms exactly as qpp's engines do (same number, same
 * order, same use). */
int qs_engine_create(const qs_circuit* qc, qs_kind kind, qs_dtype dtype,
                     double noise_p, uint64_t seed, qs_engine** out);
int qs_engine_destroy(qs_engine* e);
/* QEngineT::execute(reps)  qengine.hpp:1341-1405 (QNoisyEngine: 129-140). */
int qs_engine_execute(qs_engine* e, uint64_t reps);
/* Statistics: number of distinct outcomes, then (outcome dits, count)
 * pairs in lexicographic order (QEngineStatistics, std::map order). */
int qs_engine_num_outcomes(const qs_engine* e, uint64_t* n);
int qs_engine_get_stats(const qs_engine* e, uint64_t* dits /* n x nc */,
                        uint64_t* counts /* n */);
int qs_engine_get_dits(const qs_engine* e, uint64_t* dits /* nc */);
int qs_engine_get_probs(const qs_engine* e, double* probs /* nc */);
/* Current quantum state (over the non-measured qudits). */
int qs_engine_get_state(const qs_engine* e, qs_state** out);
int qs_engine_set_state(qs_engine* e, const qs_state* st);
/* QEngine::reset(): |0..0>, stats cleared. */
int qs_engine_reset(qs_engine* e);
/* QEngineT::set_ensure_post_selection / set_max_post_selection_reps
 * (qengine.hpp:1177-1227): with ensure set, a failed POST_SELECT step is
 * measured again (a new draw) up to max_reps attempts in all.  Defaults:
 * ensure 0, max_reps UINT64_MAX (internal/classes/qengine_state.hpp:66-72). */
int qs_engine_set_post_selection(qs_engine* e, int ensure, uint64_t max_reps);
/* QEngineT::set_dits (qengine.hpp:1155-1164): SizeMismatch unless n == nc.
 * execute() starts from zero dits, as qpp's does. */
int qs_engine_set_dits(qs_engine* e, const uint64_t* dits, uint64_t n);
/* QEngineT::get_measured_d (qengine.hpp:1052-1062): measured[i] = 1 for a
 * destructively measured qudit i (nq entries). */
int qs_engine_get_measured(const qs_engine* e, uint8_t* measured);
/* QEngineT::reset_stats: the statistics emptied, state and dits kept. */
int qs_engine_reset_stats(qs_engine* e);
/* QEngineT::post_select_ok(): 0 after a run whose post-selection failed. */
int qs_engine_post_select_ok(const qs_engine* e, int* ok);
/* No qpp counterpart: where qpp would run for ever, execute stops with
 * QS_ERR_OUT_OF_RANGE -- a runtime while loop whose body has run
 * max_loop_reps times in one run with its condition still true (default
 * 2^24), and a post-selection retried without limit (max_reps unset, see
 * qs_engine_set_post_selection) that has failed max_ps_failures times
 * (default 2^16).  0 removes a limit (qpp's behaviour: qengine.hpp:414-473,
 * 837-880). */
int qs_engine_set_limits(qs_engine* e, uint64_t max_loop_reps, uint64_t max_ps_failures);
/* QNoisyEngineT::get_noise_results (qnoisy_engine.hpp:104-121, 174-176):
 * the Kraus indices drawn before circuit step `step` (0 the identity, 1, 2, 3
 * X, Y, Z of the depolarizing channel), one per live qubit, in ascending
 * qubit order, for every visit of the step over all runs since creation or
 * the last qs_engine_reset.  out may be NULL; *count gets the number. */
int qs_engine_noise_results(const qs_engine* e, uint64_t step, uint64_t* out, uint64_t* count);

/* ------------------------------------------------------------------------ */
/* Batched noisy trajectories (one persistent launch)                        */
/* ------------------------------------------------------------------------ */
/* Trajectory i of a run computes what
 *   QNoisyEngine<QubitDepolarizingNoise(noise_p)>(circuits[circuit_of[i]],
 *                                                 seed = seeds[i]).execute(1)
 * computes (classes/qnoisy_engine.hpp:104-140, classes/noise.hpp:131-160,
 * 388-408; noise_p < 0: QEngine(..).execute(1), qengine.hpp:1341-1405): the
 * same std::mt19937 draws in the same order, and the final measurement
 * outcome drawn from the trajectory's last uniform.  Circuits: qubits, gate
 * steps whose gates are dense 1-qubit gates or Clifford gates on at most 2
 * qubits (CNOT, CZ, X, Z, ...), ending in one MEASURE_MANY of every qubit.
 * Circuits may have different qubit counts (>= 8 + reg_bits).
 * out_index[i]: basis index of the outcome (measured digit of target j =
 * bit (n-1-target[j])); out_prob[i]: its probability. */
typedef struct qs_batch qs_batch;
int qs_batch_create(const qs_circuit* const* circuits, uint32_t ncircuits,
                    qs_dtype dtype, const qs_fuse_opts* opts, qs_batch** out);
int qs_batch_run(qs_batch* b, uint64_t ntraj, const uint32_t* circuit_of,
                 const uint64_t* seeds, double noise_p, uint64_t* out_index,
                 double* out_prob);
/* Program size of one circuit; statistics of the last run (shots decided by
 * the exact sequential replay, host planning and kernel times in ms). */
int qs_batch_info(const qs_batch* b, uint32_t circuit, uint64_t* num_passes,
                  uint64_t* num_ops, uint64_t* last_exact, double* last_host_ms,
                  double* last_device_ms);
/* Host only: the Pauli-frame plan of one trajectory -- the circuit steps
 * whose 1-qubit matrix is replaced (U X^x Z^z, column-major, 8 doubles
 * each; up to cap entries), the frame left at the measurement
 * (i^phase X^xmask Z^zmask on index bits) and the measurement uniform. */
int qs_batch_plan(const qs_batch* b, uint32_t circuit, uint64_t seed,
                  double noise_p, uint64_t cap, uint64_t* n_ovr,
                  uint32_t* ovr_step, double* ovr_mats, uint64_t* xmask,
                  uint64_t* zmask, uint32_t* phase, double* u);
/* No qpp counterpart: the gate steps of circuit circuit_of[i] (everything
 * before its measurement, noiseless) applied to |0..0> and left in out[i],
 * a ket of that circuit's qubits in the batch's dtype, for all i in one
 * persistent launch (one workgroup per state).  What QEngineT::execute runs
 * before its sampling tail (qengine.hpp:1341-1405), for many circuits at
 * once; the request server's path for small requests. */
int qs_batch_run_states(qs_batch* b, uint64_t nstates, const uint32_t* circuit_of,
                        qs_state* const* out);
int qs_batch_destroy(qs_batch* b);

/* ------------------------------------------------------------------------ */
/* Serving                                                                   */
/* ------------------------------------------------------------------------ */
/* No qpp counterpart (qpp runs one engine per caller thread): a request
 * server that admits circuits continuously onto the GPU.  Each request is
 * QEngine(qc, seed).execute(shots) (classes/qengine.hpp:1341-1405) in
 * complex128 or complex64, with the engine's statistics as result.  `lanes`
 * worker threads each own a HIP stream; a free lane takes the oldest queued
 * request whose device footprint (qs_server_footprint) fits what is left of
 * `hbm_budget` bytes (0 = 90% of
 * the free device memory), so requests run concurrently and new ones are
 * admitted as lanes and memory free up.  Small requests (12-14 qubits,
 * gates then one measurement of every qubit) run batched: a lane takes
 * every queued one of the same dtype that fits and runs their gate steps
 * in one launch (qs_batch_run_states), then each request's sampling tail.
 * A request's result does not depend on the mix it ran in. */
typedef struct qs_server qs_server;
int qs_server_create(uint32_t lanes, uint64_t hbm_budget, qs_server** out);
/* Non-blocking; the circuit is copied.  QS_ERR_OUT_OF_MEMORY if the request
 * can never fit the budget. */
int qs_server_submit(qs_server* s, const qs_circuit* qc, qs_dtype dtype,
                     uint64_t shots, uint64_t seed, uint64_t* ticket);
/* The device bytes a request is admitted against: the most it can hold at
 * once on any path of execute(shots) (the state, the two saved states of
 * post-selection retries, a measurement's contracted output, and the
 * sampler's workspace: 3.5 s D + 16 D + small terms, s the bytes per
 * amplitude). */
int qs_server_footprint(const qs_circuit* qc, qs_dtype dtype, uint64_t shots,
                        uint64_t* bytes);
/* Blocks until done; times_ms[3] (optional): queued, running, total. */
int qs_server_wait(qs_server* s, uint64_t ticket, double* times_ms);
/* Results of a finished request: outcomes (n x nc dits, QEngine::get_stats
 * order), their counts and the last classical dits (nc).  Pass NULL arrays
 * to read n_outcomes first. */
int qs_server_stats(qs_server* s, uint64_t ticket, uint64_t* n_outcomes,
                    uint64_t* dits, uint64_t* counts, uint64_t* last_dits);
/* How a finished request ran: the number of requests whose gate steps ran
 * in one launch with it (1: alone or not batchable). */
int qs_server_batch_size(qs_server* s, uint64_t ticket, uint32_t* batch_size);
int qs_server_release(qs_server* s, uint64_t ticket);
/* info[6]: completed, queued, running, peak running, bytes held, peak bytes */
int qs_server_info(qs_server* s, uint64_t* info);
/* Runs what is queued to completion, then stops the lanes. */
int qs_server_destroy(qs_server* s);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* QPP_AMD_H_ */

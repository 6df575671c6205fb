// qs_internal.hpp -- shared host/device declarations of the gfx950 engine.
//
// Storage model.  A state is an array of N complex elements (complex128 as
// double2, complex64 as float2) in device memory.  A ket over n qubits has
// N = 2^n and qubit q is bit (n-1-q) of the element index (qpp's big-endian
// order).  A density matrix is stored column-major, rho(i, j) at i + j*D, so
// the same array is a "ket" over 2n bits: row qubit q is bit n-1-q and
// column qubit q is bit 2n-1-q.  Every density-matrix operation is therefore
// a bit-addressed operation on that 2n-bit index:
//   U rho U^dagger  = U on the row bits, conj(U) on the column bits,
//   sum_K K rho K^dagger = (sum_K conj(K) (x) K) on (column bits, row bits).
// Qudit states (any dimension != 2) use mixed-radix strides instead.
#ifndef QPP_AMD_QS_INTERNAL_HPP_
#define QPP_AMD_QS_INTERNAL_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/qpp_amd.h"

struct qs_state {
    qs_dtype dtype = QS_C128;
    qs_kind kind = QS_KET;
    std::vector<uint64_t> dims;  // subsystem dimensions
    uint64_t D = 1;              // prod(dims)
    uint64_t nelem = 1;          // D (ket) or D*D (rho)
    bool qubits = true;          // all dims == 2
    uint32_t nbits = 0;          // log2(nelem) when qubits
    void* d = nullptr;           // device buffer
    hipStream_t stream = nullptr;
    int device = 0;
    uint64_t elem_bytes() const { return dtype == QS_C128 ? 16 : 8; }
    uint64_t bytes() const { return nelem * elem_bytes(); }
};

namespace qs {

// ---------------------------------------------------------------------------
// errors: status + qpp-style message, per thread
// ---------------------------------------------------------------------------
int fail(int code, const char* where, const char* context = nullptr);
int fail_hip(hipError_t e, const char* where);
// an error whose message is a std::runtime_error text (no WHERE/CONTEXT form)
int fail_runtime(int code, const std::string& text);
void clear_error();
const char* description(int code);

#define QS_HIP(call, where)                                \
    do {                                                   \
        hipError_t qs_e_ = (call);                         \
        if (qs_e_ != hipSuccess) return qs::fail_hip(qs_e_, where); \
    } while (0)

// default stream / device bring-up (returns QS_OK or an error code)
int ensure_device();
hipStream_t default_stream();
// stream new states are created on: the calling thread's override if set
// (a serving lane), else the library stream
hipStream_t current_stream();
void set_thread_stream(hipStream_t s);
// test hook: QPP_AMD_FORCE_EXACT=1 searches every sampled shot in the exact
// table (the reference's sequential sums) instead of accepting the
// error-bounded parallel search
bool force_exact_sampling();
// measurement hook: QPP_AMD_GATEK_VALU=1 applies 4- and 5-qubit dense gates
// on complex64 states with the VALU kernel instead of the MFMA one
bool gatek_valu_forced();
// qpp::measure_seq with the sampled basis index given (qs_measure.hip)
int measure_seq_given(qs_state* st, uint32_t ntarget, const uint64_t* target,
                      int destructive, uint64_t index, uint64_t* results, double* probs);
// sum_K K rho K^dagger on the targets, one term per operator (qs_kraus.hip);
// validation done by the caller
int kraus_per_operator(qs_state* st, const double* Ks, uint32_t nK, uint64_t DK,
                       uint32_t ntarget, const uint64_t* target);
// creates states on `s` for the scope (outputs derived from a state live on
// that state's stream)
struct StreamScope {
    hipStream_t old;
    explicit StreamScope(hipStream_t s);
    ~StreamScope();
};

// ---------------------------------------------------------------------------
// kernel launchers (defined in the .hip files).  `m` is a column-major
// complex128 matrix (interleaved re/im).  Bit positions are LSB numbering of
// the element index.  A control is (cmask, cval): the operation acts on an
// element group iff (index & cmask) == cval.
// ---------------------------------------------------------------------------

// One-bit dense gate on bit `bit`, optionally controlled.
int launch_gate1(qs_state* st, uint32_t bit, const double* m, uint64_t cmask,
                 uint64_t cval, bool conj);

// k-bit dense gate, bits[0] the most significant digit of the gate index.
int launch_gatek(qs_state* st, const uint32_t* bits, uint32_t k,
                 const double* m, uint64_t cmask, uint64_t cval, bool conj);

// Mixed-radix (qudit) dense gate: element strides and digit ranges of the
// gate's subsystems, spectator subsystems enumerated by the kernel.  Used
// for any state with a dimension != 2.  `rows_only` restricts the gate to
// the row index of a density matrix (cols when !rows); for kets rows = true.
struct QuditGate {
    uint32_t ntg = 0;             // gate subsystems
    uint64_t tg_stride[16];       // element stride of each gate digit
    uint64_t tg_dim[16];          // its dimension
    uint32_t nsp = 0;             // spectator digits
    uint64_t sp_stride[64];
    uint64_t sp_dim[64];
    // controls: digit (c + shift) mod d must all equal r; A^r applied
    uint32_t nctrl = 0;
    uint64_t c_stride[64];
    uint64_t c_shift[64];
    uint64_t c_dim = 2;
};
int launch_qudit_gate(qs_state* st, const QuditGate& g, const double* mats,
                      uint32_t nmats, uint64_t DA, bool conj);

// QEngine::execute(reps) on an engine whose gate prefix (the steps before
// the first measurement) has already been applied elsewhere: `prefixed`
// holds that state and is taken over by the engine (destroyed on error
// too).  Valid for circuits whose steps are gates followed by measurements
// only (the prefix is then every gate step, in order, for any reps).
int engine_execute_prefixed(qs_engine* e, qs_state* prefixed, uint64_t reps);

}  // namespace qs

#endif  // QPP_AMD_QS_INTERNAL_HPP_

// qs_fused.hpp -- tables of the tile-fused pass engine (see qs_fused.hip).
#ifndef QPP_AMD_QS_FUSED_HPP_
#define QPP_AMD_QS_FUSED_HPP_

#include <complex>
#include <cstdint>
#include <set>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/qpp_amd.h"

enum FOpKind : int {
    OP_U1 = 0,      // 2x2 dense on register bit j0
    OP_X = 1,       // swap on register bit j0 (X / CNOT target)
    OP_DIAG_R = 2,  // diag(d0, d1) on register bit j0
    OP_DIAG_T = 3,  // diag(d0, d1) on a thread-number bit (tbit)
    OP_DIAG_O = 4,  // diag(d0, d1) on a bit outside the tile (dbit)
    OP_U2 = 5,      // 4x4 dense on register bits j0 < j1 (j1 = gate MSB)
    OP_U1B = 6,     // uncontrolled 2x2 dense on each register bit k set in
                    // tbit, matrices at mats + moff + 4k (commuting ops
                    // grouped so the kernel applies them in one straight run)
    // specialised kernels only (made by native_fuse, never uploaded): a
    // product on register bits that holds an X under a lane condition
    // (tmask, tval) or a tile condition (omask, oval); mats + moff holds the
    // product where the condition fails, then the one where it holds
    OP_U1S = 7,     // 2x2 on j0: mats + moff, + 4
    OP_U2S = 8      // 4x4 on j0 < j1: mats + moff, + 16
};

// One gate of a phase.  Its control (global mask, value) is split by where
// the control bits live in the phase's layout:
//   register bits -> regok: bit rho set when register group rho satisfies it
//   thread bits   -> (t & tmask) == tval on the thread number t
//   other bits    -> (tile base & omask) == oval, uniform per tile
// The kernels dispatch on `code` alone (loaded with the matrix and the other
// fields of the 96-byte head); the other fields are read only by the cases
// that need them.  `kind`, `j0`, `j1`, `ctrl` are the host's description,
// packed into `code` by seal_op().
struct FOp {
    // --- the 96-byte head (OpHead): everything the dispatch and the
    // common cases read, fetched with two scalar loads
    uint32_t code;      // kind | j0 << 4 | j1 << 6 | kCodeCtrl | kCodeOut
    uint32_t regok;
    uint32_t tmask, tval;
    uint32_t moff;      // OP_U2 / OP_U1B: matrix offset (complex elements)
    uint32_t tbit;      // OP_DIAG_T: thread bit; OP_U1B: register bits
    uint32_t pad0, pad1;
    double m[8];        // U1: u00 u10 u01 u11 (re, im); DIAG: d0 d1
    // --- the rest, read only by the cases that need it
    uint64_t omask, oval;
    uint64_t dbit;      // OP_DIAG_O: global bit
    int kind;
    int j0, j1;
    int ctrl;           // 1 when regok / tmask restrict the op per element
};

struct OpHead {
    uint32_t code, regok, tmask, tval, moff, tbit, pad0, pad1;
    double m[8];
};
static_assert(sizeof(OpHead) == 96, "OpHead is the first 96 bytes of FOp");

constexpr uint32_t kCodeCtrl = 1u << 8;  // regok / tmask apply
constexpr uint32_t kCodeOut = 1u << 9;   // omask / oval apply

inline void seal_op(FOp& o) {
    o.code = uint32_t(o.kind) | uint32_t(o.j0) << 4 | uint32_t(o.j1) << 6 |
             (o.ctrl ? kCodeCtrl : 0u) | (o.omask ? kCodeOut : 0u);
}

constexpr int kMaxNR = 16;

struct FPhase {
    uint64_t g_thr[8];            // global index bit of thread-number bit j
    uint64_t g_reg[kMaxNR];       // global index offset of register group rho
    uint32_t s_thr[8];            // swizzled LDS index bit of thread-number bit j
    uint32_t s_reg[kMaxNR];       // swizzled LDS index of register group rho
    uint8_t regbits[8];           // tile bits held in registers
    uint8_t tbits[16];            // tile bits indexed by the thread number
    // per register bit k: its global index bit and swizzled LDS offset; both
    // offsets are XOR-linear in rho, so g_reg[rho] / s_reg[rho] are the XOR
    // of these over the set bits of rho (what the kernels use: 4 uniform
    // values per phase instead of 2 x 16 table loads)
    uint64_t g_rb[4];
    uint32_t s_rb[4];
    int op_begin, op_end;
    int coalesced;                // lanes on tile bits 0..5: loads/stores coalesce
    int pad;
};

struct FPass {
    int m;               // tile bits
    int nrest;           // bits outside the tile
    uint8_t S[16];       // tile bit i -> global bit S[i] (ascending)
    uint8_t rest[64];    // global bits outside the tile (ascending)
    // the same as runs of consecutive bits: tile counter bits [src, src+len)
    // go to global bits [dst, dst+len); run = src | dst << 8 | len << 16
    int nruns;
    uint32_t runs[24];
    uint64_t ntiles;
    int phase_begin, phase_end;
};

// A gate at bit level: dense matrix on bits (bits[0] = MSB of the gate
// index), column-major, with a control (mask, value) on global bits.
struct BitGate {
    std::vector<uint32_t> bits;
    std::vector<std::complex<double>> m;
    uint64_t cmask = 0, cval = 0;
};

struct FusedOptions {
    uint32_t reg_bits = 4;  // log2(elements per thread): tile = 2^(8 + reg_bits)
    uint32_t low_bits = 6;  // low index bits every tile keeps (coalescing)
    size_t max_ops = 64;    // gates per pass at most (balances VALU vs HBM)
    bool bundle = true;     // group commuting 1-bit ops of a phase (OP_U1B)
    bool improve_sets = true;  // hill-climb each pass's tile-bit set (more gates per pass)
};

struct FusedProgram {
    uint32_t nb = 0;
    uint32_t reg_bits = 4;
    std::vector<FPass> passes;
    std::vector<FPhase> phases;
    std::vector<FOp> ops;
    std::vector<std::complex<double>> mats;  // OP_U2 matrices
    std::vector<size_t> gates_per_pass;
    std::vector<int> gate_op;  // gate index -> index of its op in `ops`
    qs_dtype dtype = QS_C128;
    hipStream_t stream = nullptr;       // stream of the upload (and of the free)
    mutable hipStream_t last = nullptr;  // stream of the last run
    void* dblob = nullptr;
    FPass* d_passes = nullptr;
    FPhase* d_phases = nullptr;
    FOp* d_ops = nullptr;
    void* d_mats = nullptr;
    // specialised pass kernels (qs_native.hip): hipModule_t per group of
    // passes and one hipFunction_t per pass, for states of native_dtype
    std::vector<void*> native_modules;
    std::vector<void*> native_fn;
    qs_dtype native_dtype = QS_C128;
    void* native_tab = nullptr;          // the passes' constants (device)
    std::vector<uint64_t> native_off;   // each pass's first constant
    std::vector<uint32_t> native_lds;   // each pass kernel's LDS bytes per workgroup
    // layout changes moved whole / split on a wave, lane, register bit;
    // passes at half the LDS
    uint64_t native_stats[5] = {0, 0, 0, 0, 0};
};

namespace qs {
int fused_compile(const std::vector<BitGate>& gates, uint32_t nb, const FusedOptions& opt,
                  FusedProgram& P);
int fused_upload(FusedProgram& P, qs_dtype dtype);
// The layout of a phase with register tile bits R (tile bits 0..m-1 at
// global bits Sv; lanes 0..lc-1 on tile bits 0..lc-1 when R allows); thread
// bits j >= 4 that are thread bits of `hint` too keep their position there.
FPhase layout_phase(const std::set<uint32_t>& R, uint32_t m, uint32_t RB,
                    const std::vector<uint32_t>& Sv, uint32_t lc, const FPhase* hint);
void fused_free(FusedProgram& P);
// ev: nullptr, or passes + 1 events recorded on the state's stream before
// the first pass and after each pass (per-pass kernel timing)
int fused_run(const FusedProgram& P, qs_state* st, uint32_t grid_cap, hipEvent_t* ev = nullptr);
// specialised pass kernels generated from the program and compiled with
// hipRTC (qs_native.hip); fused_run uses them once built
std::string native_source(const FusedProgram& P, qs_dtype dt, std::vector<double>* tab = nullptr,
                          std::vector<uint64_t>* off = nullptr,
                          std::vector<std::string>* pass_src = nullptr,
                          std::string* preamble = nullptr,
                          std::vector<uint32_t>* lds = nullptr, uint64_t* stats = nullptr);
int native_build(FusedProgram& P, qs_dtype dt);
// the program with runs of ops on one or two register bits of a phase
// multiplied into one op where that costs fewer FP64 instructions
FusedProgram native_fuse(const FusedProgram& P);
void native_free(FusedProgram& P);
int native_launch(const FusedProgram& P, qs_state* st, int pass, uint32_t grid_cap);
// circuit gate steps [b, e) -> bit gates (qs_program.hip)
int lower_steps(const qs_circuit* qc, uint32_t b, uint32_t e, bool rho,
                std::vector<BitGate>& out);
}  // namespace qs

#endif  // QPP_AMD_QS_FUSED_HPP_

"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker / CPU baseline.  Never imported
by the product package.  Arrays follow qpp's conventions: kets are 1-D
complex128, density matrices 2-D complex128 (stored column-major at the C
boundary, as Eigen does).
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u64p = POINTER(c_uint64)
        dp = POINTER(c_double)
        L.orc_apply.argtypes = [dp, c_int, c_uint32, u64p, dp, c_uint32, u64p, dp]
        L.orc_apply_ctrl.argtypes = [dp, c_int, c_uint32, u64p, dp, c_uint32, u64p,
                                     c_uint32, u64p, u64p, dp]
        L.orc_apply_ctrl_fan.argtypes = L.orc_apply_ctrl.argtypes
        L.orc_apply_kraus.argtypes = [dp, c_uint32, u64p, c_uint32, dp, c_uint32, u64p, dp]
        L.orc_apply_kraus_full.argtypes = [dp, c_uint64, c_uint32, dp, c_uint64, dp]
        L.orc_ptrace.argtypes = [dp, c_int, c_uint32, u64p, c_uint32, u64p, dp]
        L.orc_ptranspose.argtypes = [dp, c_int, c_uint32, u64p, c_uint32, u64p, c_int, dp]
        L.orc_syspermute.argtypes = [dp, c_int, c_uint32, u64p, u64p, c_int, dp]
        L.orc_probs.argtypes = [dp, c_int, c_uint64, dp]
        L.orc_discrete_draw.argtypes = [dp, c_uint64, c_double]
        L.orc_discrete_draw.restype = c_uint64
        L.orc_discrete_table.argtypes = [dp, c_uint64, dp]
        L.orc_discrete_table.restype = None
        L.orc_sample.argtypes = [dp, c_int, c_uint32, u64p, c_uint32, u64p, c_double,
                                 u64p, u64p, dp]
        L.orc_measure_seq.argtypes = [dp, c_int, c_uint32, u64p, c_uint32, u64p, c_int,
                                      c_double, u64p, dp, dp, u64p]
        L.orc_rng_new.argtypes = [c_uint64]
        L.orc_rng_new.restype = c_void_p
        L.orc_rng_canonical.argtypes = [c_void_p]
        L.orc_rng_canonical.restype = c_double
        L.orc_rng_free.argtypes = [c_void_p]
        L.orc_rng_free.restype = None
        L.orc_engine_run.argtypes = [c_void_p, c_int, c_double, c_uint64, c_uint64, c_uint64,
                                     u64p, u64p, u64p, u64p, dp, dp, u64p]
        L.orc_engine_run_ps.argtypes = [c_void_p, c_int, c_double, c_uint64, c_uint64, c_int,
                                        c_uint64, c_uint64, u64p, u64p, u64p, u64p, dp, dp, u64p,
                                        POINTER(c_int)]
        L.orc_reset.argtypes = [dp, c_int, c_uint32, u64p, c_uint32, u64p, c_double, dp]
        L.orc_last_noise_results.argtypes = [c_uint64, u64p, u64p]
        L.orc_set_engine_limits.argtypes = [c_uint64, c_uint64]
        _lib = L
    return _lib


def _u64(seq):
    a = np.ascontiguousarray(np.asarray(list(seq), dtype=np.uint64))
    return a, a.ctypes.data_as(POINTER(c_uint64))


def _buf(A):
    """(flat column-major complex128 buffer, pointer, is_rho)."""
    A = np.asarray(A, dtype=np.complex128)
    if A.ndim == 2 and A.shape[1] == 1:
        A = A[:, 0]
    is_rho = A.ndim == 2
    flat = np.asfortranarray(A).reshape(-1, order="F") if is_rho else np.ascontiguousarray(A)
    flat = np.ascontiguousarray(flat)
    return flat, flat.ctypes.data_as(POINTER(c_double)), is_rho


def _out(n):
    o = np.zeros(n, dtype=np.complex128)
    return o, o.ctypes.data_as(POINTER(c_double))


def _shape(o, is_rho, D):
    return o.reshape((D, D), order="F") if is_rho else o


def _dims(state, dims):
    D = np.asarray(state).shape[0]
    if isinstance(dims, (int, np.integer)):
        d = int(dims)
        n = int(round(np.log2(D) / np.log2(d)))
        return [d] * n
    return list(dims)


def apply(state, A, target, dims=2):
    dims = _dims(state, dims)
    s, sp, is_rho = _buf(state)
    a, ap, _ = _buf(np.asarray(A, dtype=np.complex128))
    d, dp_ = _u64(dims)
    t, tp = _u64(target)
    D = int(np.prod(dims))
    o, op = _out(D * D if is_rho else D)
    lib().orc_apply(sp, int(is_rho), len(d), dp_, ap, len(t), tp, op)
    return _shape(o, is_rho, D)


def applyCTRL(state, A, ctrl, target, dims=2, shift=None, fan=False):
    dims = _dims(state, dims)
    s, sp, is_rho = _buf(state)
    a, ap, _ = _buf(np.asarray(A, dtype=np.complex128))
    d, dp_ = _u64(dims)
    t, tp = _u64(target)
    c, cp = _u64(ctrl)
    sh, shp = _u64(shift) if shift is not None else (None, None)
    D = int(np.prod(dims))
    o, op = _out(D * D if is_rho else D)
    f = lib().orc_apply_ctrl_fan if fan else lib().orc_apply_ctrl
    f(sp, int(is_rho), len(d), dp_, ap, len(c), cp, len(t), tp, shp, op)
    return _shape(o, is_rho, D)


def applyCTRL_fan(state, A, ctrl, target, dims=2, shift=None):
    return applyCTRL(state, A, ctrl, target, dims, shift, fan=True)


def apply_kraus(rho, Ks, target, dims=2):
    dims = _dims(rho, dims)
    s, sp, _ = _buf(rho)
    kb = np.concatenate([np.asfortranarray(np.asarray(K, dtype=np.complex128)).reshape(-1, order="F")
                         for K in Ks])
    kb = np.ascontiguousarray(kb)
    d, dp_ = _u64(dims)
    t, tp = _u64(target)
    D = int(np.prod(dims))
    o, op = _out(D * D)
    lib().orc_apply_kraus(sp, len(d), dp_, len(Ks), kb.ctypes.data_as(POINTER(c_double)),
                          len(t), tp, op)
    return _shape(o, True, D)


def apply_kraus_full(rho, Ks):
    """apply(A, Ks), operations.hpp:422-464 (rectangular Kraus operators
    allowed): sum_K K rho K^dagger."""
    s, sp, _ = _buf(rho)
    D = np.asarray(rho).shape[0]
    Dout = np.asarray(Ks[0]).shape[0]
    kb = np.ascontiguousarray(np.concatenate(
        [np.asfortranarray(np.asarray(K, dtype=np.complex128)).reshape(-1, order="F") for K in Ks]))
    o, op = _out(Dout * Dout)
    lib().orc_apply_kraus_full(sp, D, len(Ks), kb.ctypes.data_as(POINTER(c_double)), Dout, op)
    return _shape(o, True, Dout)


def ptrace(A, target, dims=2):
    dims = _dims(A, dims)
    s, sp, is_rho = _buf(A)
    d, dp_ = _u64(dims)
    t, tp = _u64(target)
    Dbar = int(np.prod([dims[i] for i in range(len(dims)) if i not in set(target)]))
    o, op = _out(Dbar * Dbar)
    lib().orc_ptrace(sp, int(is_rho), len(d), dp_, len(t), tp, op)
    return _shape(o, True, Dbar)


def ptranspose(A, target, dims=2, semantics=0):
    dims = _dims(A, dims)
    s, sp, is_rho = _buf(A)
    d, dp_ = _u64(dims)
    t, tp = _u64(target)
    D = int(np.prod(dims))
    o, op = _out(D * D)
    lib().orc_ptranspose(sp, int(is_rho), len(d), dp_, len(t), tp, semantics, op)
    return _shape(o, True, D)


def syspermute(A, perm, dims=2, semantics=0):
    dims = _dims(A, dims)
    s, sp, is_rho = _buf(A)
    d, dp_ = _u64(dims)
    p, pp = _u64(perm)
    D = int(np.prod(dims))
    o, op = _out(D * D if is_rho else D)
    lib().orc_syspermute(sp, int(is_rho), len(d), dp_, pp, semantics, op)
    return _shape(o, is_rho, D)


def probs(A):
    s, sp, is_rho = _buf(A)
    D = np.asarray(A).shape[0]
    o = np.zeros(D)
    lib().orc_probs(sp, int(is_rho), D, o.ctypes.data_as(POINTER(c_double)))
    return o


def discrete_draw(p, u):
    p = np.ascontiguousarray(np.asarray(p, dtype=np.float64))
    return int(lib().orc_discrete_draw(p.ctypes.data_as(POINTER(c_double)), len(p), float(u)))


def discrete_table(p):
    p = np.ascontiguousarray(np.asarray(p, dtype=np.float64))
    cp = np.zeros(len(p))
    lib().orc_discrete_table(p.ctypes.data_as(POINTER(c_double)), len(p),
                             cp.ctypes.data_as(POINTER(c_double)))
    return cp


def sample(A, target, dims=2, u=0.5):
    dims = _dims(A, dims)
    s, sp, is_rho = _buf(A)
    d, dp_ = _u64(dims)
    t, tp = _u64(target)
    idx = np.zeros(1, dtype=np.uint64)
    dig = np.zeros(max(1, len(t)), dtype=np.uint64)
    pr = np.zeros(1)
    lib().orc_sample(sp, int(is_rho), len(d), dp_, len(t), tp, float(u),
                     idx.ctypes.data_as(POINTER(c_uint64)), dig.ctypes.data_as(POINTER(c_uint64)),
                     pr.ctypes.data_as(POINTER(c_double)))
    return int(idx[0]), [int(x) for x in dig[: len(t)]], float(pr[0])


def measure_seq(A, target, dims=2, destructive=True, u=0.5):
    dims = _dims(A, dims)
    s, sp, is_rho = _buf(A)
    d, dp_ = _u64(dims)
    t, tp = _u64(target)
    res = np.zeros(max(1, len(t)), dtype=np.uint64)
    pr = np.zeros(max(1, len(t)))
    out = np.zeros(len(s), dtype=np.complex128)
    olen = np.zeros(1, dtype=np.uint64)
    lib().orc_measure_seq(sp, int(is_rho), len(d), dp_, len(t), tp, int(destructive), float(u),
                          res.ctypes.data_as(POINTER(c_uint64)), pr.ctypes.data_as(POINTER(c_double)),
                          out.ctypes.data_as(POINTER(c_double)), olen.ctypes.data_as(POINTER(c_uint64)))
    n = int(olen[0])
    st = out[:n]
    if is_rho:
        Dn = int(round(np.sqrt(n)))
        st = st.reshape((Dn, Dn), order="F")
    return [int(x) for x in res[: len(t)]], [float(x) for x in pr[: len(t)]], st


def reset(A, target, dims=2, u=0.5):
    """qpp::reset(A, target, dims) with the uniform u."""
    dims = _dims(A, dims)
    s, sp, is_rho = _buf(A)
    d, dp_ = _u64(dims)
    t, tp = _u64(target)
    out = np.zeros(len(s), dtype=np.complex128)
    lib().orc_reset(sp, int(is_rho), len(d), dp_, len(t), tp, float(u),
                    out.ctypes.data_as(POINTER(c_double)))
    if is_rho:
        D = int(round(np.sqrt(len(s))))
        return out.reshape((D, D), order="F")
    return out


def engine_run(qc, reps=1, noise_p=-1.0, seed=5489, is_rho=False, max_outcomes=1 << 16,
               ensure_post_selection=False, max_post_selection_reps=None, with_ps_ok=False,
               max_loop_reps=1 << 24, max_ps_failures=1 << 16):
    """QEngine / QNoisyEngine execute(reps) on a qpp_amd.QCircuit; returns
    (stats dict, dits, probs, final state), plus post_select_ok() when
    with_ps_ok.  max_loop_reps / max_ps_failures: where qpp runs for ever,
    give up with OutOfRange after that many while-loop body runs or failed
    unlimited post-selections (0: never)."""
    lib().orc_set_engine_limits(int(max_loop_reps), int(max_ps_failures))
    cs = qc.struct()
    nc = qc.nc
    n_out = np.zeros(1, dtype=np.uint64)
    sd = np.zeros(max(1, max_outcomes * nc), dtype=np.uint64)
    sc = np.zeros(max_outcomes, dtype=np.uint64)
    dits = np.zeros(max(1, nc), dtype=np.uint64)
    probs = np.zeros(max(1, nc))
    D = qc.d ** qc.nq
    st = np.zeros(D * D if is_rho else D, dtype=np.complex128)
    olen = np.zeros(1, dtype=np.uint64)
    ok = c_int(1)
    mx = (1 << 64) - 1 if max_post_selection_reps is None else int(max_post_selection_reps)
    rc = lib().orc_engine_run_ps(ctypes.cast(ctypes.byref(cs), c_void_p), int(is_rho), float(noise_p),
                            int(seed), int(reps), int(bool(ensure_post_selection)), mx,
                            max_outcomes,
                            n_out.ctypes.data_as(POINTER(c_uint64)), sd.ctypes.data_as(POINTER(c_uint64)),
                            sc.ctypes.data_as(POINTER(c_uint64)), dits.ctypes.data_as(POINTER(c_uint64)),
                            probs.ctypes.data_as(POINTER(c_double)), st.ctypes.data_as(POINTER(c_double)),
                            olen.ctypes.data_as(POINTER(c_uint64)), ctypes.byref(ok))
    lib().orc_set_engine_limits(1 << 24, 1 << 16)
    if rc != 0:
        err = RuntimeError(f"oracle engine stopped with qs_status {rc}")
        err.code = rc
        raise err
    n = int(n_out[0])
    stats = {tuple(int(x) for x in sd[i * nc:(i + 1) * nc]): int(sc[i]) for i in range(min(n, max_outcomes))}
    L_ = int(olen[0])
    state = st[:L_]
    if is_rho:
        Dn = int(round(np.sqrt(L_)))
        state = state.reshape((Dn, Dn), order="F")
    out = (stats, [int(x) for x in dits[:nc]], [float(x) for x in probs[:nc]], state)
    return out + (bool(ok.value),) if with_ps_ok else out


def last_noise_results(nsteps):
    """get_noise_results() of the engine the last engine_run call ran."""
    out = []
    for i in range(nsteps):
        n = np.zeros(1, dtype=np.uint64)
        lib().orc_last_noise_results(i, None, n.ctypes.data_as(POINTER(c_uint64)))
        buf = np.zeros(max(1, int(n[0])), dtype=np.uint64)
        lib().orc_last_noise_results(i, buf.ctypes.data_as(POINTER(c_uint64)),
                                     n.ctypes.data_as(POINTER(c_uint64)))
        out.append([int(x) for x in buf[: int(n[0])]])
    return out


class Rng:
    """The oracle's std::mt19937 + std::generate_canonical<double, 53>."""

    def __init__(self, seed):
        self.h = lib().orc_rng_new(seed)

    def canonical(self):
        return lib().orc_rng_canonical(self.h)

    def __del__(self):
        try:
            lib().orc_rng_free(self.h)
        except Exception:
            pass

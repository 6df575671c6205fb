"""qpp_amd -- MI355X (gfx950) engine for Quantum++'s gate-application path.

Python mirror of the qpp functions the engine replaces, over the C ABI of
libqpp_amd.so (include/qpp_amd.h).  Names, argument meaning and error
behaviour follow qpp 7.0.3:

    apply(state, A, target, dims_or_d)          operations.hpp:72-409
    apply(rho, Ks, target, dims_or_d)           operations.hpp:476-593
    applyCTRL(state, A, ctrl, target, dims_or_d, shift=None)
                                                operations.hpp:616-936
    applyCTRL_fan(state, A, ctrl, target, dims_or_d, shift=None)
                                                operations.hpp:961-1284
    ptrace(A, target, dims_or_d)                operations.hpp:1908-2152
    ptranspose(A, target, dims_or_d, semantics) operations.hpp:2167-2372
    syspermute(A, perm, dims_or_d, semantics)   operations.hpp:2386-2550
    sample(A, target, dims_or_d, u)             instruments.hpp:746-929
    measure_seq(A, target, dims_or_d, destructive, u)
                                                instruments.hpp:985-1135

Value semantics as in qpp: numpy arrays in, new numpy arrays out (a ket is a
1-D complex array, a density matrix a 2-D one).  For device-resident work
use `State` (in-place operations on a device handle).  Random draws are
explicit uniforms `u` in [0, 1) (what qpp draws from its mt19937 through
std::generate_canonical); `Rng` reproduces that generator.

Errors raise `QppError` subclasses named after qpp::exception classes.  The
library has no CPU fallback: without a gfx950 device every compute call
raises `NoDevice`.
"""
import ctypes
import itertools
import math
import os
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_size_t, c_uint32, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqpp_amd.so")

C128, C64 = 0, 1
TEXT, BINARY = 0, 1  # qs_format: qpp::save/load text, qpp::obsolete binary
KET, RHO = 0, 1
DOCUMENTED, AS_BUILT = 0, 1

# ---------------------------------------------------------------------------
# exceptions (include/qpp/classes/exception.hpp)
# ---------------------------------------------------------------------------


class QppError(Exception):
    """Base of every error raised by the engine (qpp::exception::Exception)."""

    code = -1


def _mk(name, code):
    return type(name, (QppError,), {"code": code})


ZeroSize = _mk("ZeroSize", 1)
MatrixNotSquare = _mk("MatrixNotSquare", 2)
MatrixNotCvector = _mk("MatrixNotCvector", 3)
MatrixNotSquareNorCvector = _mk("MatrixNotSquareNorCvector", 4)
MatrixMismatchSubsys = _mk("MatrixMismatchSubsys", 5)
DimsInvalid = _mk("DimsInvalid", 6)
DimsNotEqual = _mk("DimsNotEqual", 7)
DimsMismatchMatrix = _mk("DimsMismatchMatrix", 8)
DimsMismatchCvector = _mk("DimsMismatchCvector", 9)
SubsysMismatchDims = _mk("SubsysMismatchDims", 10)
PermInvalid = _mk("PermInvalid", 11)
PermMismatchDims = _mk("PermMismatchDims", 12)
OutOfRange = _mk("OutOfRange", 13)
TypeMismatch = _mk("TypeMismatch", 14)
SizeMismatch = _mk("SizeMismatch", 15)
QuditAlreadyMeasured = _mk("QuditAlreadyMeasured", 16)
NotImplemented_ = _mk("NotImplemented", 17)
OverflowError_ = _mk("OverflowError", 18)
RuntimeError_ = _mk("RuntimeError", 19)  # std::runtime_error (load/save streams)
InvalidConditional = _mk("InvalidConditional", 20)
InvalidHandle = _mk("InvalidHandle", 100)
OutOfMemory = _mk("OutOfMemory", 101)
DeviceError = _mk("DeviceError", 102)
NoDevice = _mk("NoDevice", 103)

_BY_CODE = {c.code: c for c in [
    ZeroSize, MatrixNotSquare, MatrixNotCvector, MatrixNotSquareNorCvector,
    MatrixMismatchSubsys, DimsInvalid, DimsNotEqual, DimsMismatchMatrix,
    DimsMismatchCvector, SubsysMismatchDims, PermInvalid, PermMismatchDims,
    OutOfRange, TypeMismatch, SizeMismatch, QuditAlreadyMeasured,
    NotImplemented_, OverflowError_, RuntimeError_, InvalidConditional, InvalidHandle, OutOfMemory,
    DeviceError, NoDevice]}

# ---------------------------------------------------------------------------
# library
# ---------------------------------------------------------------------------
_lib = None


def lib():
    """Load libqpp_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run python qpp_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        u64p = POINTER(c_uint64)
        dp = POINTER(c_double)
        vp = c_void_p
        sigs = {
            "qs_last_error": (c_char_p, []),
            "qs_device_info": (c_int, [c_char_p, c_size_t, POINTER(c_int), u64p, u64p]),
            "qs_synchronize": (c_int, []),
            "qs_state_create": (c_int, [c_int, c_int, c_uint32, u64p, POINTER(vp)]),
            "qs_state_destroy": (c_int, [vp]),
            "qs_state_clone": (c_int, [vp, POINTER(vp)]),
            "qs_state_info": (c_int, [vp, POINTER(c_int), POINTER(c_int), POINTER(c_uint32), u64p, u64p]),
            "qs_state_upload": (c_int, [vp, vp, c_uint64]),
            "qs_state_download": (c_int, [vp, vp, c_uint64]),
            "qs_state_set_basis": (c_int, [vp, c_uint64]),
            "qs_state_device_ptr": (c_int, [vp, POINTER(vp)]),
            "qs_apply": (c_int, [vp, dp, c_uint64, c_uint32, u64p]),
            "qs_apply_ctrl": (c_int, [vp, dp, c_uint64, c_uint32, u64p, c_uint32, u64p, u64p]),
            "qs_apply_ctrl_fan": (c_int, [vp, dp, c_uint64, c_uint32, u64p, c_uint32, u64p, u64p]),
            "qs_apply_kraus": (c_int, [vp, dp, c_uint32, c_uint64, c_uint32, u64p]),
            "qs_ptrace": (c_int, [vp, c_uint32, u64p, POINTER(vp)]),
            "qs_ptranspose": (c_int, [vp, c_uint32, u64p, c_int, POINTER(vp)]),
            "qs_syspermute": (c_int, [vp, u64p, c_int, POINTER(vp)]),
            "qs_norm2": (c_int, [vp, dp]),
            "qs_normalize": (c_int, [vp]),
            "qs_compare": (c_int, [vp, vp, dp]),
            "qs_state_convert": (c_int, [vp, c_int, POINTER(vp)]),
            "qs_state_save": (c_int, [vp, c_char_p, c_int]),
            "qs_state_load": (c_int, [c_char_p, c_int, c_int, c_int, c_uint32, u64p, POINTER(vp)]),
            "qs_probs": (c_int, [vp, dp, c_uint64]),
            "qs_sample": (c_int, [vp, c_uint32, u64p, c_uint64, dp, u64p, u64p, dp]),
            "qs_measure_seq": (c_int, [vp, c_uint32, u64p, c_int, c_double, u64p, dp]),
            "qs_reset": (c_int, [vp, c_uint32, u64p, c_double]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().qs_last_error().decode(errors="replace")
        raise _BY_CODE.get(rc, QppError)(msg)


def _u64(seq):
    a = np.ascontiguousarray(np.asarray(list(seq), dtype=np.uint64))
    return a, a.ctypes.data_as(POINTER(c_uint64))


def _mat(A):
    """Column-major complex128 buffer of a square matrix (Eigen layout)."""
    A = np.asarray(A, dtype=np.complex128)
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise MatrixNotSquare("[WHERE: qpp_amd] Matrix is not square!")
    buf = np.asfortranarray(A)
    return buf, buf.ctypes.data_as(POINTER(c_double)), A.shape[0]


def device_info():
    name = ctypes.create_string_buffer(256)
    cu = c_int()
    tot = c_uint64()
    free = c_uint64()
    _check(lib().qs_device_info(name, 256, byref(cu), byref(tot), byref(free)))
    return {"name": name.value.decode(), "cu": cu.value, "hbm_bytes": tot.value,
            "hbm_free": free.value}


def synchronize():
    _check(lib().qs_synchronize())


# ---------------------------------------------------------------------------
# device-resident state
# ---------------------------------------------------------------------------


class State:
    """A ket or density matrix in device memory (qs_state handle)."""

    def __init__(self, dims, kind=KET, dtype=C128, _handle=None):
        self._h = c_void_p()
        if _handle is not None:
            self._h = _handle
        else:
            arr, p = _u64(dims)
            _check(lib().qs_state_create(dtype, kind, len(arr), p, byref(self._h)))

    @classmethod
    def _wrap(cls, handle):
        return cls(None, _handle=handle)

    @classmethod
    def from_array(cls, A, dims, dtype=C128):
        A = np.asarray(A)
        kind = RHO if A.ndim == 2 and A.shape[1] != 1 else KET
        if A.ndim == 2 and A.shape[1] == 1:
            A = A[:, 0]
        st = cls(dims, kind, dtype)
        st.upload(A)
        return st

    def close(self):
        if self._h and self._h.value:
            lib().qs_state_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def info(self):
        dt, kd, nd = c_int(), c_int(), c_uint32()
        dims = (c_uint64 * 128)()
        ne = c_uint64()
        _check(lib().qs_state_info(self._h, byref(dt), byref(kd), byref(nd), dims, byref(ne)))
        return dt.value, kd.value, [dims[i] for i in range(nd.value)], ne.value

    @property
    def dims(self):
        return self.info()[2]

    @property
    def kind(self):
        return self.info()[1]

    @property
    def dtype(self):
        return self.info()[0]

    def _np_dtype(self):
        return np.complex128 if self.dtype == C128 else np.complex64

    def upload(self, A):
        dt, kd, dims, ne = self.info()
        npdt = np.complex128 if dt == C128 else np.complex64
        A = np.asarray(A, dtype=npdt)
        buf = np.asfortranarray(A).reshape(-1, order="F") if A.ndim == 2 else np.ascontiguousarray(A)
        _check(lib().qs_state_upload(self._h, buf.ctypes.data_as(c_void_p), buf.nbytes))

    def download(self):
        dt, kd, dims, ne = self.info()
        npdt = np.complex128 if dt == C128 else np.complex64
        out = np.empty(ne, dtype=npdt)
        _check(lib().qs_state_download(self._h, out.ctypes.data_as(c_void_p), out.nbytes))
        if kd == RHO:
            D = int(round(math.sqrt(ne)))
            return out.reshape((D, D), order="F")
        return out

    def clone(self):
        h = c_void_p()
        _check(lib().qs_state_clone(self._h, byref(h)))
        return State._wrap(h)

    def set_basis(self, index):
        _check(lib().qs_state_set_basis(self._h, index))

    def save(self, path, fmt=TEXT):
        """Write the state to `path` in qpp's format (TEXT: qpp::save,
        BINARY: qpp::obsolete::save; input_output.hpp:177-211, 318-345)."""
        _check(lib().qs_state_save(self._h, os.fsencode(path), int(fmt)))
        return self

    @classmethod
    def load(cls, path, dims, kind=KET, dtype=C128, fmt=TEXT):
        """A state with `dims` read from a file in qpp's format (TEXT:
        qpp::load<cmat>, BINARY: qpp::obsolete::load<cmat>)."""
        d, dp = _u64(dims)
        h = c_void_p()
        _check(lib().qs_state_load(os.fsencode(path), int(fmt), int(dtype), int(kind), len(d), dp,
                                   byref(h)))
        return cls(list(dims), kind, dtype, _handle=h)

    # in-place operations -------------------------------------------------
    def apply(self, A, target):
        buf, p, dim = _mat(A)
        t, tp = _u64(target)
        _check(lib().qs_apply(self._h, p, dim, len(t), tp))
        return self

    def apply_ctrl(self, A, ctrl, target, shift=None):
        buf, p, dim = _mat(A)
        c, cp = _u64(ctrl)
        t, tp = _u64(target)
        s, sp = _u64(shift) if shift is not None else (None, None)
        _check(lib().qs_apply_ctrl(self._h, p, dim, len(c), cp, len(t), tp, sp))
        return self

    def apply_ctrl_fan(self, A, ctrl, target, shift=None):
        buf, p, dim = _mat(A)
        c, cp = _u64(ctrl)
        t, tp = _u64(target)
        s, sp = _u64(shift) if shift is not None else (None, None)
        _check(lib().qs_apply_ctrl_fan(self._h, p, dim, len(c), cp, len(t), tp, sp))
        return self

    def apply_kraus(self, Ks, target):
        Ks = [np.asarray(K, dtype=np.complex128) for K in Ks]
        if not Ks:
            raise ZeroSize("[WHERE: qpp::apply()] Object has zero size! [CONTEXT: Ks]")
        dim = Ks[0].shape[0]
        for K in Ks:
            if K.shape != Ks[0].shape:
                raise DimsNotEqual("[WHERE: qpp::apply()] Dimensions not equal! [CONTEXT: K]")
        buf = np.concatenate([np.asfortranarray(K).reshape(-1, order="F") for K in Ks])
        t, tp = _u64(target)
        _check(lib().qs_apply_kraus(self._h, buf.ctypes.data_as(POINTER(c_double)),
                                    len(Ks), dim, len(t), tp))
        return self

    def apply_kraus_full(self, Ks):
        """qpp::apply(rho, Ks) over the whole system (operations.hpp:422-464):
        a new density matrix of dimension Ks[0].shape[0] (one subsystem)."""
        Ks = [np.asarray(K, dtype=np.complex128) for K in Ks]
        if not Ks:
            raise ZeroSize("[WHERE: qpp::apply()] Object has zero size! [CONTEXT: Ks]")
        for K in Ks:
            if K.shape != Ks[0].shape:
                raise DimsNotEqual("[WHERE: qpp::apply()] Dimensions not equal! [CONTEXT: K]")
        buf = np.ascontiguousarray(np.concatenate([np.asfortranarray(K).reshape(-1, order="F")
                                                   for K in Ks]))
        L = lib()
        L.qs_apply_kraus_full.argtypes = [c_void_p, POINTER(c_double), c_uint32, c_uint64,
                                          c_uint64, POINTER(c_void_p)]
        out = c_void_p()
        _check(L.qs_apply_kraus_full(self._h, buf.ctypes.data_as(POINTER(c_double)), len(Ks),
                                     Ks[0].shape[0], Ks[0].shape[1], byref(out)))
        return State._wrap(out)

    def ptrace(self, target):
        t, tp = _u64(target)
        h = c_void_p()
        _check(lib().qs_ptrace(self._h, len(t), tp, byref(h)))
        return State._wrap(h)

    def ptranspose(self, target, semantics=DOCUMENTED):
        t, tp = _u64(target)
        h = c_void_p()
        _check(lib().qs_ptranspose(self._h, len(t), tp, semantics, byref(h)))
        return State._wrap(h)

    def syspermute(self, perm, semantics=DOCUMENTED):
        p, pp = _u64(perm)
        h = c_void_p()
        _check(lib().qs_syspermute(self._h, pp, semantics, byref(h)))
        return State._wrap(h)

    def norm2(self):
        v = c_double()
        _check(lib().qs_norm2(self._h, byref(v)))
        return v.value

    def normalize(self):
        _check(lib().qs_normalize(self._h))
        return self

    def astype(self, dtype):
        """A device copy stored as C128 or C64."""
        h = c_void_p()
        _check(lib().qs_state_convert(self._h, dtype, byref(h)))
        return State._wrap(h)

    def compare(self, other):
        """FP64 comparison with another state of the same shape (any dtype):
        overlap, norms, max-abs amplitude difference and, for kets, the
        infidelity 1 - |<a|b>|^2/(<a|a><b|b>) and pure-state trace distance
        sqrt(infidelity)."""
        out = (c_double * 5)()
        _check(lib().qs_compare(self._h, other._h, out))
        ov = complex(out[0], out[1])
        r = {"overlap": ov, "norm2_a": out[2], "norm2_b": out[3], "max_abs": out[4]}
        if self.kind == KET:
            inf = max(0.0, 1.0 - abs(ov) ** 2 / (out[2] * out[3]))
            r["infidelity"] = inf
            r["trace_distance"] = math.sqrt(inf)
        else:
            # Hilbert-Schmidt distance; trace distance <= sqrt(rank) / 2 times it
            r["hs_distance"] = math.sqrt(max(0.0, out[2] + out[3] - 2 * out[0]))
        return r

    def probs(self):
        dt, kd, dims, ne = self.info()
        D = int(np.prod(dims, dtype=np.uint64))
        out = np.empty(D, dtype=np.float64)
        _check(lib().qs_probs(self._h, out.ctypes.data_as(POINTER(c_double)), D))
        return out

    def sample(self, target, u):
        """Indices, target digits and probabilities of len(u) shots."""
        u = np.ascontiguousarray(np.atleast_1d(np.asarray(u, dtype=np.float64)))
        t, tp = _u64(target)
        n = len(u)
        idx = np.empty(n, dtype=np.uint64)
        dig = np.empty(max(1, n * len(t)), dtype=np.uint64)
        pr = np.empty(n, dtype=np.float64)
        _check(lib().qs_sample(self._h, len(t), tp, n, u.ctypes.data_as(POINTER(c_double)),
                               idx.ctypes.data_as(POINTER(c_uint64)),
                               dig.ctypes.data_as(POINTER(c_uint64)),
                               pr.ctypes.data_as(POINTER(c_double))))
        return idx, dig[: n * len(t)].reshape(n, len(t)), pr

    def measure_seq(self, target, u, destructive=True):
        t, tp = _u64(target)
        res = np.empty(max(1, len(t)), dtype=np.uint64)
        pr = np.empty(max(1, len(t)), dtype=np.float64)
        _check(lib().qs_measure_seq(self._h, len(t), tp, int(bool(destructive)), float(u),
                                    res.ctypes.data_as(POINTER(c_uint64)),
                                    pr.ctypes.data_as(POINTER(c_double))))
        return [int(x) for x in res[: len(t)]], [float(x) for x in pr[: len(t)]]

    def reset(self, target, u):
        """qpp::reset in place (instruments.hpp:1148-1196) with uniform u."""
        t, tp = _u64(target)
        _check(lib().qs_reset(self._h, len(t), tp, float(u)))
        return self


# ---------------------------------------------------------------------------
# qpp-style value-semantics functions
# ---------------------------------------------------------------------------


def get_num_subsys(D, d):
    """internal/util.hpp:389-397: llround(log2(D) / log2(d))."""
    return int(round(math.log2(D) / math.log2(d)))


def _dims_of(state, dims_or_d):
    state = np.asarray(state)
    D = state.shape[0]
    if np.isscalar(dims_or_d) or isinstance(dims_or_d, (int, np.integer)):
        d = int(dims_or_d)
        if d < 2:
            raise DimsInvalid("[WHERE: qpp_amd] Invalid dimension(s)! [CONTEXT: d]")
        return [d] * get_num_subsys(D, d)
    return [int(x) for x in dims_or_d]


def _as_state(state, dims, dtype=C128):
    A = np.asarray(state)
    if A.size == 0:
        raise ZeroSize("[WHERE: qpp_amd] Object has zero size! [CONTEXT: state]")
    if A.ndim == 2 and A.shape[1] == 1:
        A = A[:, 0]
    is_rho = A.ndim == 2
    if is_rho and A.shape[0] != A.shape[1]:
        raise MatrixNotSquareNorCvector(
            "[WHERE: qpp_amd] Matrix is not square nor column vector! [CONTEXT: state]")
    D = int(np.prod(dims)) if dims else 0
    if not dims or any(x == 0 for x in dims):
        raise DimsInvalid("[WHERE: qpp_amd] Invalid dimension(s)! [CONTEXT: dims]")
    if D != A.shape[0]:
        if is_rho:
            raise DimsMismatchMatrix("[WHERE: qpp_amd] Dimension(s) mismatch matrix size! [CONTEXT: dims/state]")
        raise DimsMismatchCvector("[WHERE: qpp_amd] Dimension(s) mismatch column vector size! [CONTEXT: dims/state]")
    return State.from_array(A, dims, dtype)


def apply(state, A, target=None, dims=2):
    """qpp::apply(state, A, target, dims|d) -- or, when A is a list of Kraus
    operators, qpp::apply(rho, Ks, target, dims|d), and without a target
    qpp::apply(rho, Ks) on the whole system (rectangular Ks allowed)."""
    if isinstance(A, (list, tuple)) and target is None:
        rho = np.asarray(state)
        if rho.ndim != 2 or rho.shape[0] != rho.shape[1]:
            raise MatrixNotSquare("[WHERE: qpp::apply()] Matrix is not square! [CONTEXT: A]")
        st = State.from_array(rho, [rho.shape[0]])
        return st.apply_kraus_full(A).download()
    if target is None:
        raise TypeError("apply() needs a target for a gate")
    if isinstance(A, (list, tuple)):
        dims = _dims_of(state, dims)
        st = _as_state(state, dims)
        st.apply_kraus(A, target)
        return st.download()
    dims = _dims_of(state, dims)
    st = _as_state(state, dims)
    st.apply(A, target)
    return st.download()


def applyCTRL(state, A, ctrl, target, dims=2, shift=None):
    dims = _dims_of(state, dims)
    st = _as_state(state, dims)
    st.apply_ctrl(A, ctrl, target, shift)
    return st.download()


def applyCTRL_fan(state, A, ctrl, target, dims=2, shift=None):
    dims = _dims_of(state, dims)
    st = _as_state(state, dims)
    st.apply_ctrl_fan(A, ctrl, target, shift)
    return st.download()


def ptrace(A, target, dims=2):
    dims = _dims_of(A, dims)
    st = _as_state(A, dims)
    return st.ptrace(target).download()


def ptranspose(A, target, dims=2, semantics=DOCUMENTED):
    dims = _dims_of(A, dims)
    st = _as_state(A, dims)
    return st.ptranspose(target, semantics).download()


def syspermute(A, perm, dims=2, semantics=DOCUMENTED):
    dims = _dims_of(A, dims)
    if len(perm) != len(dims):
        if sorted(perm) != list(range(len(perm))):
            raise PermInvalid("[WHERE: qpp::syspermute()] Invalid permutation! [CONTEXT: perm]")
        raise PermMismatchDims("[WHERE: qpp::syspermute()] Permutation mismatch dimensions! [CONTEXT: dims/perm]")
    st = _as_state(A, dims)
    return st.syspermute(perm, semantics).download()


def sample(A, target, dims=2, u=None):
    """qpp::sample(A, target, dims): (target digits, probability) for one
    uniform u, or for an array u the per-shot digits."""
    dims = _dims_of(A, dims)
    st = _as_state(A, dims)
    idx, dig, pr = st.sample(target, u)
    if np.ndim(u) == 0:
        return [int(x) for x in dig[0]], float(pr[0])
    return dig, pr


def save(A, path, dims=2, fmt=TEXT):
    """qpp::save(A, os) (TEXT) or qpp::obsolete::save (BINARY) of a ket or
    density matrix, written through the device state."""
    dims = _dims_of(A, dims)
    _as_state(A, dims).save(path, fmt)


def load(path, dims, kind=KET, fmt=TEXT):
    """qpp::load<cmat>(is) (TEXT) or qpp::obsolete::load (BINARY) as a numpy
    array (D for a ket, D x D for a density matrix)."""
    return State.load(path, dims, kind, C128, fmt).download()


def reset(A, target, dims=2, u=0.5):
    """qpp::reset(A, target, dims) (instruments.hpp:1148-1196) -> the reset
    state; u is the uniform its measure_seq draws."""
    dims = _dims_of(A, dims)
    st = _as_state(A, dims)
    st.reset(target, u)
    return st.download()


def measure_seq(A, target, dims=2, destructive=True, u=None):
    """qpp::measure_seq -> (results, probs, post-measurement state)."""
    dims = _dims_of(A, dims)
    st = _as_state(A, dims)
    res, pr = st.measure_seq(target, u, destructive)
    return res, pr, st.download()


class Rng:
    """std::mt19937 + std::generate_canonical<double, 53> (the uniforms qpp's
    RandomDevices feeds its distributions; bits/random.h:164-192)."""

    def __init__(self, seed=5489):
        self.mt = [0] * 624
        self.idx = 624
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xFFFFFFFF

    def _twist(self):
        mt = self.mt
        for i in range(624):
            y = (mt[i] & 0x80000000) | (mt[(i + 1) % 624] & 0x7FFFFFFF)
            mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
        self.idx = 0

    def next_u32(self):
        if self.idx >= 624:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF

    def canonical(self):
        # generate_canonical<double, 53> with a 32-bit engine: k = 2 draws,
        # sum = x0 + x1 * 2^32, r = sum / 2^64, clamped below 1
        x0 = self.next_u32()
        x1 = self.next_u32()
        s = float(x0) + float(x1) * 4294967296.0
        r = s / 18446744073709551616.0
        if r >= 1.0:
            r = math.nextafter(1.0, 0.0)
        return r


# ---------------------------------------------------------------------------
# circuits (qpp::QCircuit lowered to the qs_circuit program format)
# ---------------------------------------------------------------------------
(STEP_NOP, STEP_GATE, STEP_FAN, STEP_CTRL, STEP_CTRL_FAN, STEP_MEASURE, STEP_MEASURE_MANY,
 STEP_RESET, STEP_DISCARD, STEP_POST_SELECT, STEP_COND_IF, STEP_COND_ELSE,
 STEP_COND_END, STEP_CCTRL, STEP_CCTRL_FAN, STEP_COND_WHILE, STEP_COND_IF_TABLE,
 STEP_COND_WHILE_TABLE) = range(18)


class _Step(ctypes.Structure):
    _fields_ = [(f, c_uint32) for f in ("type", "gate", "nctrl", "ntarget", "ctrl_off",
                                         "target_off", "shift_off", "c_reg", "destructive",
                                         "ps_off")]


class _Circuit(ctypes.Structure):
    _fields_ = [("nq", c_uint32), ("nc", c_uint32), ("d", c_uint32), ("nsteps", c_uint32),
                ("steps", POINTER(_Step)), ("nmats", c_uint32),
                ("mat_dim", POINTER(c_uint64)), ("mat_off", POINTER(c_uint64)),
                ("mats", POINTER(c_double)), ("pool", POINTER(c_uint64))]


class QCircuit:
    """Step list of a qpp::QCircuit (classes/qcircuit.hpp): gate / gate_fan /
    gate_joint / CTRL / CTRL_fan / measure / post_select / reset / discard /
    nop, in program order."""

    def __init__(self, nq, nc=0, d=2):
        self.nq, self.nc, self.d = int(nq), int(nc), int(d)
        self.steps = []       # tuples (type, gate, ctrl, target, shift, c_reg, destructive)
        self.ps_vals = {}     # step index -> post-selection values
        self.mats = []
        self._mat_index = {}

    def _mat(self, U):
        U = np.asarray(U, dtype=np.complex128)
        key = U.tobytes()
        if key not in self._mat_index:
            self._mat_index[key] = len(self.mats)
            self.mats.append(U)
        return self._mat_index[key]

    def gate(self, U, *targets):
        t = list(targets[0]) if len(targets) == 1 and not np.isscalar(targets[0]) else list(targets)
        self.steps.append((STEP_GATE, self._mat(U), [], t, None, 0, 0))
        return self

    gate_joint = gate

    def gate_fan(self, U, targets):
        self.steps.append((STEP_FAN, self._mat(U), [], list(targets), None, 0, 0))
        return self

    def CTRL(self, U, ctrl, target, shift=None):
        c = [ctrl] if np.isscalar(ctrl) else list(ctrl)
        t = [target] if np.isscalar(target) else list(target)
        self.steps.append((STEP_CTRL, self._mat(U), c, t, None if shift is None else list(shift), 0, 0))
        return self

    def CTRL_fan(self, U, ctrl, target, shift=None):
        c = [ctrl] if np.isscalar(ctrl) else list(ctrl)
        self.steps.append((STEP_CTRL_FAN, self._mat(U), c, list(target),
                           None if shift is None else list(shift), 0, 0))
        return self

    def measure(self, target, c_reg=None, destructive=True):
        if np.isscalar(target):
            t = [int(target)]
            kind = STEP_MEASURE
        else:
            t = list(target)
            kind = STEP_MEASURE_MANY
        c = t[0] if c_reg is None else int(c_reg)
        self.steps.append((kind, 0, [], t, None, c, int(bool(destructive))))
        return self

    def cCTRL(self, U, ctrl_dits, target, shift=None):
        """QCircuit::cCTRL (qcircuit.hpp:2604-2960; qengine.hpp:581-630):
        U^k on the target(s) when every control dit (plus its shift mod d)
        equals k."""
        c = [int(ctrl_dits)] if np.isscalar(ctrl_dits) else list(ctrl_dits)
        t = [int(target)] if np.isscalar(target) else list(target)
        self.steps.append((STEP_CCTRL, self._mat(U), c, t, None if shift is None else list(shift),
                           0, 0))
        return self

    def cCTRL_fan(self, U, ctrl_dits, targets, shift=None):
        """QCircuit::cCTRL_fan: U^k on each target."""
        c = [int(ctrl_dits)] if np.isscalar(ctrl_dits) else list(ctrl_dits)
        self.steps.append((STEP_CCTRL_FAN, self._mat(U), c, list(targets),
                           None if shift is None else list(shift), 0, 0))
        return self

    def post_select(self, target, ps_vals, c_reg=None, destructive=True):
        """QCircuit::post_select (qcircuit.hpp:3366-3500): measure the
        targets into c_reg.. and keep the run only when the results equal
        ps_vals (one value per target)."""
        t = [int(target)] if np.isscalar(target) else list(target)
        v = [int(ps_vals)] if np.isscalar(ps_vals) else list(ps_vals)
        if len(v) != len(t):
            raise ValueError("post_select: one value per target")
        c = t[0] if c_reg is None else int(c_reg)
        self.ps_vals[len(self.steps)] = v
        self.steps.append((STEP_POST_SELECT, 0, [], t, None, c, int(bool(destructive))))
        return self

    def reset(self, target):
        """QCircuit::reset (qcircuit.hpp:3862-3940): the targets back to |0>."""
        t = [int(target)] if np.isscalar(target) else list(target)
        self.steps.append((STEP_RESET, 0, [], t, None, 0, 0))
        return self

    def discard(self, target):
        """QCircuit::discard (qcircuit.hpp:3745-3830): measure the targets
        destructively and drop the results."""
        t = [int(target)] if np.isscalar(target) else list(target)
        self.steps.append((STEP_DISCARD, 0, [], t, None, 0, 1))
        return self

    def _cond(self, name, plain, table, dits, values):
        d = [int(dits)] if np.isscalar(dits) else list(dits)
        if callable(values):
            # qpp's predicate lambda (cond_if(pred), qcircuit.hpp:1026) as a
            # truth table over the values 0..d-1 of the dits it reads:
            # entry i = sum_j v_j d^(k-1-j) holds pred([v_0, ..., v_{k-1}])
            if self.d ** len(d) > 1 << 20:
                raise ValueError(f"{name}: a predicate reads at most 2^20 value combinations")
            self.ps_vals[len(self.steps)] = [
                int(bool(values(list(v)))) for v in itertools.product(range(self.d), repeat=len(d))]
            self.steps.append((table, 0, d, [], None, 0, 0))
            return self
        v = [int(values)] if np.isscalar(values) else list(values)
        if len(v) != len(d):
            raise ValueError(f"{name}: one value per dit")
        self.ps_vals[len(self.steps)] = v
        self.steps.append((plain, 0, d, [], None, 0, 0))
        return self

    def cond_if(self, dits, values):
        """QCircuit::cond_if (runtime conditional, qengine.hpp:837-860).
        values: the condition "dit dits[j] == values[j] for every j" (the
        form OpenQASM's if (c==n) takes), or a predicate pred(v) over the
        list v of those dits' values (qpp's cond_if(pred); the dits must
        hold values in 0..d-1 when the condition is evaluated, else the run
        is OutOfRange).  Close with cond_end(), optionally after
        cond_else()."""
        return self._cond("cond_if", STEP_COND_IF, STEP_COND_IF_TABLE, dits, values)

    def cond_while(self, dits, values):
        """QCircuit::cond_while (qengine.hpp:837-880): repeat the steps up
        to cond_end() while the condition holds (values as in cond_if).
        Destructive measurements and discards are not allowed inside."""
        return self._cond("cond_while", STEP_COND_WHILE, STEP_COND_WHILE_TABLE, dits, values)

    def cond_else(self):
        self.steps.append((STEP_COND_ELSE, 0, [], [], None, 0, 0))
        return self

    def cond_end(self):
        self.steps.append((STEP_COND_END, 0, [], [], None, 0, 0))
        return self

    def nop(self):
        self.steps.append((STEP_NOP, 0, [], [], None, 0, 0))
        return self

    @classmethod
    def from_gates(cls, n, gates):
        """From a workloads gate list (kind, targets, controls, matrix)."""
        qc = cls(n)
        for kind, target, ctrl, U in gates:
            if ctrl:
                qc.CTRL(U, ctrl, target)
            else:
                qc.gate(U, target)
        return qc

    def struct(self):
        """Build the C structure (arrays kept alive on the returned object)."""
        pool = []
        steps = (_Step * max(1, len(self.steps)))()
        for i, (tp, g, ctrl, tg, shift, creg, dest) in enumerate(self.steps):
            s = steps[i]
            s.type, s.gate, s.nctrl, s.ntarget = tp, g, len(ctrl), len(tg)
            s.ctrl_off = len(pool)
            pool += ctrl
            s.target_off = len(pool)
            pool += tg
            if shift is None:
                s.shift_off = 0xFFFFFFFF
            else:
                s.shift_off = len(pool)
                pool += shift
            s.c_reg, s.destructive, s.ps_off = creg, dest, 0
            if i in self.ps_vals:
                s.ps_off = len(pool)
                pool += self.ps_vals[i]
        pool_a = np.ascontiguousarray(np.asarray(pool + [0], dtype=np.uint64))
        dims = np.ascontiguousarray(np.asarray([m.shape[0] for m in self.mats] + [0], dtype=np.uint64))
        offs, off = [], 0
        for m in self.mats:
            offs.append(off)
            off += m.size
        offs_a = np.ascontiguousarray(np.asarray(offs + [0], dtype=np.uint64))
        flat = np.ascontiguousarray(np.concatenate(
            [np.asfortranarray(m).reshape(-1, order="F") for m in self.mats] + [np.zeros(1, complex)]))
        c = _Circuit()
        c.nq, c.nc, c.d, c.nsteps = self.nq, self.nc, self.d, len(self.steps)
        c.steps = steps
        c.nmats = len(self.mats)
        c.mat_dim = dims.ctypes.data_as(POINTER(c_uint64))
        c.mat_off = offs_a.ctypes.data_as(POINTER(c_uint64))
        c.mats = flat.ctypes.data_as(POINTER(c_double))
        c.pool = pool_a.ctypes.data_as(POINTER(c_uint64))
        c._keep = (steps, pool_a, dims, offs_a, flat)
        return c


class _FuseOpts(ctypes.Structure):
    _fields_ = [("reg_bits", c_uint32), ("low_bits", c_uint32), ("max_ops", c_uint32),
                ("grid_cap", c_uint32)]


def _native_stats(h):
    L = lib()
    L.qs_program_native_stats.argtypes = [c_void_p, POINTER(c_uint64)]
    c = (c_uint64 * 5)()
    _check(L.qs_program_native_stats(h, c))
    return {"whole": c[0], "wave_split": c[1], "lane_split": c[2], "register_split": c[3],
            "half_lds_passes": c[4]}


class Program:
    """Gate steps of a circuit compiled into tile-fused passes."""

    def __init__(self, n_or_circuit, gates=None, kind=KET, dtype=C128, reg_bits=4,
                 low_bits=5, max_ops=64, grid_cap=0, step_begin=0, step_end=None):
        qc = n_or_circuit if isinstance(n_or_circuit, QCircuit) else QCircuit.from_gates(n_or_circuit, gates)
        self.circuit = qc
        self._cs = qc.struct()
        L = lib()
        L.qs_program_create.argtypes = [POINTER(_Circuit), c_uint32, c_uint32, c_int, c_int,
                                        POINTER(_FuseOpts), POINTER(c_void_p)]
        L.qs_program_run.argtypes = [c_void_p, c_void_p]
        L.qs_program_info.argtypes = [c_void_p, POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint64)]
        L.qs_program_destroy.argtypes = [c_void_p]
        o = _FuseOpts(reg_bits, low_bits, max_ops, grid_cap)
        self._dtype, self._reg_bits = dtype, reg_bits
        self._h = c_void_p()
        end = len(qc.steps) if step_end is None else step_end
        _check(L.qs_program_create(ctypes.byref(self._cs), step_begin, end, kind, dtype,
                                   ctypes.byref(o), ctypes.byref(self._h)))

    def run(self, st):
        _check(lib().qs_program_run(self._h, st.handle))

    def compile_native(self):
        """Generate this program's passes as specialised kernels and compile
        them with hipRTC (qs_program_compile_native; seconds).  Later runs
        use them."""
        L = lib()
        L.qs_program_compile_native.argtypes = [c_void_p]
        _check(L.qs_program_compile_native(self._h))
        self._native = True
        return self

    def capture(self, st):
        """Run once while capturing the pass launches as a HIP graph for this
        state's buffer and stream (qs_program_capture); later runs on the
        same state replay the graph."""
        L = lib()
        L.qs_program_capture.argtypes = [c_void_p, c_void_p]
        _check(L.qs_program_capture(self._h, st.handle))
        return self

    def native_stats(self):
        """How the specialised kernels move their layout changes through LDS
        (qs_program_native_stats): counts of changes moved whole, split on a
        wave bit, a lane bit or a register bit, and of passes at half the
        LDS."""
        return _native_stats(self._h)

    def run_timed(self, st):
        """Run once; the milliseconds of each pass, from HIP events recorded
        on the state's stream between the pass launches."""
        L = lib()
        L.qs_program_run_timed.argtypes = [c_void_p, c_void_p, POINTER(c_double)]
        ms = np.zeros(max(1, self.num_passes()))
        _check(L.qs_program_run_timed(self._h, st.handle, ms.ctypes.data_as(POINTER(c_double))))
        return ms[: self.num_passes()]

    def info(self):
        g, p, h = c_uint64(), c_uint64(), c_uint64()
        _check(lib().qs_program_info(self._h, byref(g), byref(p), byref(h)))
        return {"gates": g.value, "passes": p.value, "phases": h.value}

    def num_passes(self):
        return self.info()["passes"]

    def kernel_name(self):
        if getattr(self, "_native", False):
            return ("qs_pass_c128_<i>" if self._dtype == C128 else "qs_pass_c64_<i>") + \
                " (specialised tile passes, one kernel per pass)"
        r = "double" if self._dtype == C128 else "float"
        return f"k_tile<{r},{16 if self._reg_bits == 4 else 8}>"

    def __del__(self):
        try:
            if self._h:
                lib().qs_program_destroy(self._h)
        except Exception:
            pass


class _BitGate(ctypes.Structure):
    _fields_ = [("nbits", c_uint32), ("bits", c_uint32 * 2), ("pad", c_uint32),
                ("cmask", c_uint64), ("cval", c_uint64), ("mat_off", c_uint64)]


class BitProgram:
    """A fused program of bit-level operations (qs_program_create_bits).
    On a density matrix of n qubits (vec(rho): row qubit q = bit n-1-q,
    column qubit q = bit 2n-1-q) `rho_apply`, `rho_ctrl` and `rho_channel`
    add what qpp's apply(rho, U, ...), applyCTRL(rho, ...) and
    apply(rho, Ks, {q}) compute; all of them run in the same tile passes."""

    def __init__(self, nb):
        self.nb = int(nb)
        self.ops = []  # (bits, matrix, cmask, cval)
        self._h = None

    def gate(self, U, bits, cmask=0, cval=0):
        self.ops.append((list(bits), np.asarray(U, dtype=np.complex128), int(cmask), int(cval)))
        return self

    def rho_apply(self, U, target, n):
        """U rho U^dagger on qubits `target` (1 or 2)."""
        U = np.asarray(U, dtype=np.complex128)
        self.gate(U, [n - 1 - t for t in target])
        return self.gate(U.conj(), [2 * n - 1 - t for t in target])

    def rho_ctrl(self, U, ctrl, target, n, shift=None):
        """applyCTRL on rho (qubits): U on target where every control is |1>
        (|0> where shift is 1), on rows and, conjugated, on columns."""
        shift = shift or [0] * len(ctrl)
        for col in (False, True):
            off = 2 * n - 1 if col else n - 1
            cm = cv = 0
            for c, s in zip(ctrl, shift):
                b = 1 << (off - c)
                cm |= b
                if not s:
                    cv |= b
            M = np.asarray(U, dtype=np.complex128)
            self.gate(M.conj() if col else M, [off - t for t in target], cm, cv)
        return self

    def rho_channel(self, Ks, q, n):
        """sum_K K rho K^dagger on qubit q: the 4x4 superoperator
        sum_K conj(K) (x) K on (column bit, row bit)."""
        S = sum(np.kron(np.asarray(K).conj(), np.asarray(K)) for K in Ks)
        return self.gate(S, [2 * n - 1 - q, n - 1 - q])

    def compile(self, dtype=C128, reg_bits=4, low_bits=5, max_ops=48, grid_cap=0):
        L = lib()
        L.qs_program_create_bits.argtypes = [c_uint32, c_uint32, POINTER(_BitGate),
                                             POINTER(c_double), c_int, POINTER(_FuseOpts),
                                             POINTER(c_void_p)]
        L.qs_program_run.argtypes = [c_void_p, c_void_p]
        L.qs_program_info.argtypes = [c_void_p, POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint64)]
        L.qs_program_destroy.argtypes = [c_void_p]
        arr = (_BitGate * max(1, len(self.ops)))()
        mats = []
        off = 0
        for i, (bits, M, cm, cv) in enumerate(self.ops):
            g = arr[i]
            g.nbits = len(bits)
            for k, b in enumerate(bits):
                g.bits[k] = b
            g.cmask, g.cval, g.mat_off = cm, cv, off
            flat = np.asfortranarray(M).reshape(-1, order="F")
            mats.append(flat)
            off += len(flat)
        m = np.ascontiguousarray(np.concatenate(mats) if mats else np.zeros(1, dtype=complex))
        o = _FuseOpts(reg_bits, low_bits, max_ops, grid_cap)
        if self._h:
            L.qs_program_destroy(self._h)
        self._h = c_void_p()
        _check(L.qs_program_create_bits(self.nb, len(self.ops), arr,
                                        m.ctypes.data_as(POINTER(c_double)), dtype,
                                        ctypes.byref(o), ctypes.byref(self._h)))
        return self

    def run(self, st):
        _check(lib().qs_program_run(self._h, st.handle))

    def compile_native(self):
        """Specialised pass kernels for the compiled program (see
        Program.compile_native)."""
        L = lib()
        L.qs_program_compile_native.argtypes = [c_void_p]
        _check(L.qs_program_compile_native(self._h))
        return self

    def native_stats(self):
        """See Program.native_stats."""
        return _native_stats(self._h)

    def run_timed(self, st):
        """See Program.run_timed."""
        L = lib()
        L.qs_program_run_timed.argtypes = [c_void_p, c_void_p, POINTER(c_double)]
        np_ = self.info()["passes"]
        ms = np.zeros(max(1, np_))
        _check(L.qs_program_run_timed(self._h, st.handle, ms.ctypes.data_as(POINTER(c_double))))
        return ms[:np_]

    def info(self):
        g, p, h = c_uint64(), c_uint64(), c_uint64()
        _check(lib().qs_program_info(self._h, byref(g), byref(p), byref(h)))
        return {"gates": g.value, "passes": p.value, "phases": h.value}

    def __del__(self):
        try:
            if self._h:
                lib().qs_program_destroy(self._h)
        except Exception:
            pass


class QEngine:
    """qpp::QEngineT / QNoisyEngineT (classes/qengine.hpp,
    classes/qnoisy_engine.hpp) over a device-resident state.  noise_p >= 0
    selects QNoisyEngine<QubitDepolarizingNoise(noise_p)>.  Draws come from
    an std::mt19937 seeded with `seed`, in qpp's order."""

    def __init__(self, qc, kind=KET, dtype=C128, noise_p=-1.0, seed=5489):
        self.circuit = qc
        self._cs = qc.struct()
        L = lib()
        L.qs_engine_create.argtypes = [POINTER(_Circuit), c_int, c_int, c_double, c_uint64,
                                       POINTER(c_void_p)]
        L.qs_engine_destroy.argtypes = [c_void_p]
        L.qs_engine_execute.argtypes = [c_void_p, c_uint64]
        L.qs_engine_num_outcomes.argtypes = [c_void_p, POINTER(c_uint64)]
        L.qs_engine_get_stats.argtypes = [c_void_p, POINTER(c_uint64), POINTER(c_uint64)]
        L.qs_engine_get_dits.argtypes = [c_void_p, POINTER(c_uint64)]
        L.qs_engine_get_probs.argtypes = [c_void_p, POINTER(c_double)]
        L.qs_engine_get_state.argtypes = [c_void_p, POINTER(c_void_p)]
        L.qs_engine_set_state.argtypes = [c_void_p, c_void_p]
        L.qs_engine_reset.argtypes = [c_void_p]
        self._h = c_void_p()
        _check(L.qs_engine_create(ctypes.byref(self._cs), kind, dtype, float(noise_p), int(seed),
                                  ctypes.byref(self._h)))

    def execute(self, reps=1):
        _check(lib().qs_engine_execute(self._h, int(reps)))
        return self

    def get_stats(self):
        n = c_uint64()
        _check(lib().qs_engine_num_outcomes(self._h, byref(n)))
        nc = self.circuit.nc
        d = np.zeros(max(1, n.value * nc), dtype=np.uint64)
        c = np.zeros(max(1, n.value), dtype=np.uint64)
        _check(lib().qs_engine_get_stats(self._h, d.ctypes.data_as(POINTER(c_uint64)),
                                         c.ctypes.data_as(POINTER(c_uint64))))
        return {tuple(int(x) for x in d[i * nc:(i + 1) * nc]): int(c[i]) for i in range(n.value)}

    def get_dits(self):
        d = np.zeros(max(1, self.circuit.nc), dtype=np.uint64)
        _check(lib().qs_engine_get_dits(self._h, d.ctypes.data_as(POINTER(c_uint64))))
        return [int(x) for x in d[: self.circuit.nc]]

    def get_probs(self):
        p = np.zeros(max(1, self.circuit.nc))
        _check(lib().qs_engine_get_probs(self._h, p.ctypes.data_as(POINTER(c_double))))
        return [float(x) for x in p[: self.circuit.nc]]

    def get_state(self):
        h = c_void_p()
        _check(lib().qs_engine_get_state(self._h, byref(h)))
        return State._wrap(h)

    def set_state(self, st):
        _check(lib().qs_engine_set_state(self._h, st.handle))
        return self

    def reset(self):
        _check(lib().qs_engine_reset(self._h))
        self._ensure, self._max_ps = False, (1 << 64) - 1
        return self

    def set_ensure_post_selection(self, val):
        """QEngineT::set_ensure_post_selection (qengine.hpp:1177-1181)."""
        self._ensure = bool(val)
        return self._set_ps()

    def set_max_post_selection_reps(self, n):
        """QEngineT::set_max_post_selection_reps (qengine.hpp:1222-1226)."""
        self._max_ps = int(n)
        return self._set_ps()

    def set_limits(self, max_loop_reps=1 << 24, max_ps_failures=1 << 16):
        """Where qpp runs for ever, stop with OutOfRange: after a while
        loop's body has run max_loop_reps times in one run with its
        condition still true, and after max_ps_failures failures of a
        post-selection retried without limit.  0 removes a limit (qpp's
        behaviour)."""
        L = lib()
        L.qs_engine_set_limits.argtypes = [c_void_p, c_uint64, c_uint64]
        _check(L.qs_engine_set_limits(self._h, int(max_loop_reps), int(max_ps_failures)))
        return self

    def _set_ps(self):
        L = lib()
        L.qs_engine_set_post_selection.argtypes = [c_void_p, c_int, c_uint64]
        _check(L.qs_engine_set_post_selection(self._h, int(getattr(self, "_ensure", False)),
                                              getattr(self, "_max_ps", (1 << 64) - 1)))
        return self

    def set_dits(self, dits):
        """QEngineT::set_dits (qengine.hpp:1155-1164)."""
        L = lib()
        L.qs_engine_set_dits.argtypes = [c_void_p, POINTER(c_uint64), c_uint64]
        d, dp = _u64(dits)
        _check(L.qs_engine_set_dits(self._h, dp, len(d)))
        return self

    def set_dit(self, i, value):
        d = self.get_dits()
        if i >= len(d):
            raise OutOfRange("[WHERE: qpp::QEngineT::set_dit()] Argument out of range! [CONTEXT: i]")
        d[i] = int(value)
        return self.set_dits(d)

    def get_dit(self, i):
        d = self.get_dits()
        if i >= len(d):
            raise OutOfRange("[WHERE: qpp::QEngineT::get_dit()] Argument out of range! [CONTEXT: i]")
        return d[i]

    def _measured(self):
        L = lib()
        L.qs_engine_get_measured.argtypes = [c_void_p, POINTER(ctypes.c_uint8)]
        m = (ctypes.c_uint8 * max(1, self.circuit.nq))()
        _check(L.qs_engine_get_measured(self._h, m))
        return [bool(m[i]) for i in range(self.circuit.nq)]

    def get_measured_d(self):
        """QEngineT::get_measured_d (qengine.hpp:1052-1062)."""
        return [i for i, x in enumerate(self._measured()) if x]

    def get_non_measured_d(self):
        return [i for i, x in enumerate(self._measured()) if not x]

    def was_measured_d(self, i):
        return self._measured()[i]

    def reset_stats(self):
        L = lib()
        L.qs_engine_reset_stats.argtypes = [c_void_p]
        _check(L.qs_engine_reset_stats(self._h))
        return self

    def get_noise_results(self):
        """QNoisyEngineT::get_noise_results (qnoisy_engine.hpp:174-176): per
        circuit step, the Kraus index (0 I, 1 X, 2 Y, 3 Z) drawn for each
        live qubit at each visit."""
        L = lib()
        L.qs_engine_noise_results.argtypes = [c_void_p, c_uint64, POINTER(c_uint64), POINTER(c_uint64)]
        out = []
        for i in range(len(self.circuit.steps)):
            n = c_uint64()
            _check(L.qs_engine_noise_results(self._h, i, None, byref(n)))
            buf = (c_uint64 * max(1, n.value))()
            _check(L.qs_engine_noise_results(self._h, i, buf, byref(n)))
            out.append([int(buf[j]) for j in range(n.value)])
        return out

    def post_select_ok(self):
        L = lib()
        L.qs_engine_post_select_ok.argtypes = [c_void_p, POINTER(c_int)]
        ok = c_int()
        _check(L.qs_engine_post_select_ok(self._h, byref(ok)))
        return bool(ok.value)

    def __del__(self):
        try:
            if self._h:
                lib().qs_engine_destroy(self._h)
        except Exception:
            pass


class Batch:
    """Batched noisy trajectories in one persistent launch.  Trajectory i is
    QNoisyEngine<QubitDepolarizingNoise(noise_p)>(circuits[circuit_of[i]],
    seed=seeds[i]).execute(1) (noise_p < 0: QEngine(...).execute(1));
    every circuit ends in one measurement of all its qubits."""

    def __init__(self, circuits, dtype=C64, reg_bits=4, low_bits=5, max_ops=48):
        if isinstance(circuits, QCircuit):
            circuits = [circuits]
        self.circuits = list(circuits)
        self._cs = [qc.struct() for qc in self.circuits]
        L = lib()
        L.qs_batch_create.argtypes = [POINTER(POINTER(_Circuit)), c_uint32, c_int,
                                      POINTER(_FuseOpts), POINTER(c_void_p)]
        L.qs_batch_run.argtypes = [c_void_p, c_uint64, POINTER(c_uint32), POINTER(c_uint64),
                                   c_double, POINTER(c_uint64), POINTER(c_double)]
        L.qs_batch_info.argtypes = [c_void_p, c_uint32, POINTER(c_uint64), POINTER(c_uint64),
                                    POINTER(c_uint64), POINTER(c_double), POINTER(c_double)]
        L.qs_batch_plan.argtypes = [c_void_p, c_uint32, c_uint64, c_double, c_uint64,
                                    POINTER(c_uint64), POINTER(c_uint32), POINTER(c_double),
                                    POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint32),
                                    POINTER(c_double)]
        L.qs_batch_destroy.argtypes = [c_void_p]
        arr = (POINTER(_Circuit) * len(self._cs))(*[ctypes.pointer(c) for c in self._cs])
        o = _FuseOpts(reg_bits, low_bits, max_ops, 0)
        self._dtype = dtype
        self._h = c_void_p()
        _check(L.qs_batch_create(arr, len(self._cs), dtype, ctypes.byref(o), ctypes.byref(self._h)))

    def run(self, circuit_of, seeds, noise_p):
        """Outcome basis indices and their probabilities, one per trajectory."""
        c = np.ascontiguousarray(circuit_of, dtype=np.uint32)
        s = np.ascontiguousarray(seeds, dtype=np.uint64)
        if c.shape != s.shape:
            raise SizeMismatch("[WHERE: qpp_amd::batch] Size mismatch! [CONTEXT: seeds]")
        idx = np.zeros(max(1, len(c)), dtype=np.uint64)
        pr = np.zeros(max(1, len(c)))
        _check(lib().qs_batch_run(self._h, len(c), c.ctypes.data_as(POINTER(c_uint32)),
                                  s.ctypes.data_as(POINTER(c_uint64)), float(noise_p),
                                  idx.ctypes.data_as(POINTER(c_uint64)),
                                  pr.ctypes.data_as(POINTER(c_double))))
        return idx[: len(c)], pr[: len(c)]

    def run_states(self, circuit_of, dtype=None):
        """The gate steps of circuits[circuit_of[i]] applied to |0..0>, all in
        one launch (qs_batch_run_states): one State per entry, in the batch's
        dtype; no measurement is taken."""
        L = lib()
        L.qs_batch_run_states.argtypes = [c_void_p, c_uint64, POINTER(c_uint32), POINTER(c_void_p)]
        c = np.ascontiguousarray(circuit_of, dtype=np.uint32)
        dt = self._dtype if dtype is None else dtype
        states = [State([2] * self.circuits[int(k)].nq, dtype=dt) for k in c]
        hs = (c_void_p * max(1, len(states)))(*[st.handle for st in states])
        _check(L.qs_batch_run_states(self._h, len(c), c.ctypes.data_as(POINTER(c_uint32)), hs))
        return states

    def info(self, circuit=0):
        v = [c_uint64() for _ in range(3)]
        h, d = c_double(), c_double()
        _check(lib().qs_batch_info(self._h, circuit, byref(v[0]), byref(v[1]), byref(v[2]),
                                   byref(h), byref(d)))
        return {"passes": v[0].value, "ops": v[1].value, "exact_replays": v[2].value,
                "host_ms": h.value, "device_ms": d.value}

    def plan(self, circuit, seed, noise_p, cap=4096):
        n = c_uint64()
        st = np.zeros(cap, dtype=np.uint32)
        m = np.zeros(8 * cap)
        xm, zm, ph, u = c_uint64(), c_uint64(), c_uint32(), c_double()
        _check(lib().qs_batch_plan(self._h, circuit, int(seed), float(noise_p), cap, byref(n),
                                   st.ctypes.data_as(POINTER(c_uint32)),
                                   m.ctypes.data_as(POINTER(c_double)), byref(xm), byref(zm),
                                   byref(ph), byref(u)))
        k = min(n.value, cap)
        mats = [np.array([m[8 * i + 2 * j] + 1j * m[8 * i + 2 * j + 1] for j in range(4)])
                .reshape(2, 2, order="F") for i in range(k)]
        return {"steps": [int(x) for x in st[:k]], "mats": mats, "xmask": xm.value,
                "zmask": zm.value, "phase": ph.value, "u": u.value}

    def dits(self, circuit, index):
        """The engine's classical dits for an outcome index (the final
        MEASURE_MANY writes target j's digit to c_reg + j)."""
        qc = self.circuits[circuit]
        tp, g, ctrl, tg, shift, creg, dest = [s for s in qc.steps if s[0] == STEP_MEASURE_MANY][-1]
        d = [0] * qc.nc
        n = qc.nq
        for j, q in enumerate(tg):
            d[creg + j] = (int(index) >> (n - 1 - q)) & 1
        return tuple(d)

    def __del__(self):
        try:
            if self._h:
                lib().qs_batch_destroy(self._h)
        except Exception:
            pass


class Server:
    """Request server (qs_server_*): each request is QEngine(qc, seed)
    .execute(shots) in complex128 or complex64, admitted onto the GPU by
    device footprint as lanes and memory free up; `lanes` lanes run
    concurrently on their own HIP streams (4 by default: the device's
    hardware queues).  Small requests (12-14 qubits, gates then one
    measurement of every qubit) queued together run their gate steps in one
    launch."""

    def __init__(self, lanes=4, hbm_budget=0):
        L = lib()
        L.qs_server_create.argtypes = [c_uint32, c_uint64, POINTER(c_void_p)]
        L.qs_server_submit.argtypes = [c_void_p, POINTER(_Circuit), c_int, c_uint64, c_uint64,
                                       POINTER(c_uint64)]
        L.qs_server_wait.argtypes = [c_void_p, c_uint64, POINTER(c_double)]
        L.qs_server_stats.argtypes = [c_void_p, c_uint64, POINTER(c_uint64), POINTER(c_uint64),
                                      POINTER(c_uint64), POINTER(c_uint64)]
        L.qs_server_release.argtypes = [c_void_p, c_uint64]
        L.qs_server_info.argtypes = [c_void_p, POINTER(c_uint64)]
        L.qs_server_destroy.argtypes = [c_void_p]
        self._h = c_void_p()
        self._nc = {}
        _check(L.qs_server_create(int(lanes), int(hbm_budget), byref(self._h)))

    @staticmethod
    def footprint(qc, dtype=C128, shots=1024):
        """Device bytes a request is admitted against (qs_server_footprint)."""
        L = lib()
        L.qs_server_footprint.argtypes = [POINTER(_Circuit), c_int, c_uint64, POINTER(c_uint64)]
        cs = qc.struct()
        b = c_uint64()
        _check(L.qs_server_footprint(ctypes.byref(cs), dtype, shots, byref(b)))
        return b.value

    def submit(self, qc, dtype=C128, shots=1024, seed=5489):
        cs = qc.struct()
        t = c_uint64()
        _check(lib().qs_server_submit(self._h, byref(cs), dtype, int(shots), int(seed), byref(t)))
        self._nc[t.value] = qc.nc
        return t.value

    def wait(self, ticket):
        """(queued_ms, run_ms, total_ms) of a finished request."""
        tm = (c_double * 3)()
        _check(lib().qs_server_wait(self._h, int(ticket), tm))
        return tuple(tm)

    def stats(self, ticket):
        """(statistics dict as QEngine.get_stats, last dits)."""
        n = c_uint64()
        _check(lib().qs_server_stats(self._h, int(ticket), byref(n), None, None, None))
        nc = self._nc[ticket]
        d = np.zeros(max(1, n.value * nc), dtype=np.uint64)
        c = np.zeros(max(1, n.value), dtype=np.uint64)
        last = np.zeros(max(1, nc), dtype=np.uint64)
        _check(lib().qs_server_stats(self._h, int(ticket), byref(n),
                                     d.ctypes.data_as(POINTER(c_uint64)),
                                     c.ctypes.data_as(POINTER(c_uint64)),
                                     last.ctypes.data_as(POINTER(c_uint64))))
        st = {tuple(int(x) for x in d[i * nc:(i + 1) * nc]): int(c[i]) for i in range(n.value)}
        return st, [int(x) for x in last[:nc]]

    def batch_size(self, ticket):
        """Requests whose gate steps ran in one launch with this finished
        request (1: alone, or not batchable)."""
        L = lib()
        L.qs_server_batch_size.argtypes = [c_void_p, c_uint64, POINTER(c_uint32)]
        v = c_uint32()
        _check(L.qs_server_batch_size(self._h, int(ticket), byref(v)))
        return v.value

    def release(self, ticket):
        _check(lib().qs_server_release(self._h, int(ticket)))
        self._nc.pop(ticket, None)

    def info(self):
        v = (c_uint64 * 6)()
        _check(lib().qs_server_info(self._h, v))
        keys = ["completed", "queued", "running", "peak_running", "bytes_in_use", "peak_bytes"]
        return dict(zip(keys, [int(x) for x in v]))

    def close(self):
        if self._h and self._h.value:
            lib().qs_server_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


from . import qasm  # noqa: E402  (OpenQASM 2.0 reader; needs QCircuit)

"""Shared test helpers: gates, basis states, seeded Haar unitaries and an
independent numpy tensor-contraction reference.

The tensor reference reshapes a state into one axis per subsystem and uses
numpy.tensordot / transpose, so it shares no index arithmetic with the
oracle (which restates qpp's bit/stride loops) and pins it.
"""
import numpy as np

SQ2 = 1 / np.sqrt(2.0)
I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
H = np.array([[SQ2, SQ2], [SQ2, -SQ2]], dtype=complex)
CNOT = np.eye(4, dtype=complex)
CNOT[2:, 2:] = X
CZ = np.diag([1, 1, 1, -1]).astype(complex)
TOF = np.eye(8, dtype=complex)
TOF[6:, 6:] = X


def Xd(d):
    """qpp Gates::Xd: |j> -> |j+1 mod d>."""
    M = np.zeros((d, d), dtype=complex)
    for j in range(d):
        M[(j + 1) % d, j] = 1
    return M


def mket(digits, dims):
    """qpp::mket: basis ket of the multi-index (subsystem 0 most significant)."""
    if isinstance(dims, int):
        dims = [dims] * len(digits)
    D = int(np.prod(dims))
    v = np.zeros(D, dtype=complex)
    v[np.ravel_multi_index(digits, dims)] = 1
    return v


def kets(bits):
    """|b0 b1 ...> for a bit string."""
    return mket([int(c) for c in bits], 2)


def prj(psi):
    psi = np.asarray(psi)
    return np.outer(psi, psi.conj()) / np.vdot(psi, psi).real


def haar(d, rng):
    """Haar unitary: QR of a complex Gaussian with the phases of R's diagonal
    divided out (as qpp::randU, random.hpp:257-283)."""
    Zm = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / np.sqrt(2)
    Q, R = np.linalg.qr(Zm)
    ph = np.diag(R) / np.abs(np.diag(R))
    return Q * ph


def randket(D, rng):
    v = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    return v / np.linalg.norm(v)


def randrho(D, rng):
    A = rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D))
    R = A @ A.conj().T
    return R / np.trace(R).real


# ---------------------------------------------------------------------------
# tensor reference
# ---------------------------------------------------------------------------

def t_apply_ket(psi, U, target, dims):
    """U on the target axes (target[0] = most significant gate digit)."""
    n = len(dims)
    T = np.asarray(psi).reshape(dims)
    Ut = np.asarray(U).reshape([dims[t] for t in target] * 2)
    k = len(target)
    out = np.tensordot(Ut, T, axes=(list(range(k, 2 * k)), list(target)))
    # out axes: gate axes (target order) then the remaining axes in order
    rest = [q for q in range(n) if q not in target]
    order = list(target) + rest
    inv = np.argsort(order)
    return out.transpose(inv).reshape(-1)


def t_apply_rho(rho, U, target, dims):
    """U rho U^dagger via axis contractions on the row and column axes."""
    n = len(dims)
    D = int(np.prod(dims))
    T = np.asarray(rho).reshape(list(dims) * 2)
    k = len(target)
    Ut = np.asarray(U).reshape([dims[t] for t in target] * 2)
    out = np.tensordot(Ut, T, axes=(list(range(k, 2 * k)), list(target)))
    rest = [q for q in range(2 * n) if q not in target]
    inv = np.argsort(list(target) + rest)
    T = out.transpose(inv)
    cols = [n + t for t in target]
    out = np.tensordot(Ut.conj(), T, axes=(list(range(k, 2 * k)), cols))
    rest = [q for q in range(2 * n) if q not in cols]
    inv = np.argsort(cols + rest)
    return out.transpose(inv).reshape(D, D)


def t_apply(state, U, target, dims):
    state = np.asarray(state)
    return t_apply_rho(state, U, target, dims) if state.ndim == 2 else t_apply_ket(state, U, target, dims)


def t_ctrl_op(U, ctrl, target, dims, shift=None):
    """Full D x D operator of qpp's applyCTRL: on each basis component whose
    (control digit + shift) mod d equals the same r for every control, U^r
    acts on the targets (operations.hpp:605-615 doc, :777-886)."""
    n = len(dims)
    D = int(np.prod(dims))
    d = dims[ctrl[0]]
    shift = shift or [0] * len(ctrl)
    powers = [np.linalg.matrix_power(U, r) for r in range(d)]
    M = np.zeros((D, D), dtype=complex)
    for col in range(D):
        e = np.zeros(D, dtype=complex)
        e[col] = 1
        mid = np.unravel_index(col, dims)
        rs = {(mid[c] + s) % d for c, s in zip(ctrl, shift)}
        if len(rs) == 1:
            r = rs.pop()
            M[:, col] = t_apply_ket(e, powers[r], target, dims)
        else:
            M[:, col] = e
    return M


def t_ptrace(state, target, dims):
    state = np.asarray(state)
    n = len(dims)
    rho = np.outer(state, state.conj()) if state.ndim == 1 else state
    T = rho.reshape(list(dims) * 2)
    keep = [q for q in range(n) if q not in target]
    letters = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"
    row = [letters[q] for q in range(n)]
    col = [letters[n + q] if q in keep else letters[q] for q in range(n)]
    outs = "".join(letters[q] for q in keep) + "".join(letters[n + q] for q in keep)
    res = np.einsum("".join(row) + "".join(col) + "->" + outs, T)
    Db = int(np.prod([dims[q] for q in keep])) if keep else 1
    return res.reshape(Db, Db)


def t_ptranspose(state, target, dims):
    state = np.asarray(state)
    n = len(dims)
    D = int(np.prod(dims))
    rho = np.outer(state, state.conj()) if state.ndim == 1 else state
    T = rho.reshape(list(dims) * 2)
    axes = list(range(2 * n))
    for t in target:
        axes[t], axes[n + t] = axes[n + t], axes[t]
    return T.transpose(axes).reshape(D, D)


def t_syspermute(state, perm, dims):
    """Output subsystem k is input subsystem perm[k]."""
    state = np.asarray(state)
    n = len(dims)
    if state.ndim == 1:
        return state.reshape(dims).transpose(perm).reshape(-1)
    D = int(np.prod(dims))
    T = state.reshape(list(dims) * 2)
    return T.transpose(list(perm) + [n + p for p in perm]).reshape(D, D)

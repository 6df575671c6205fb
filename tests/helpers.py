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


# ---------------------------------------------------------------------------
# golden fixtures (tests/golden/known_answers.json)
# ---------------------------------------------------------------------------
GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


def golden_known_answers():
    import json
    import os
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        return json.load(f)["cases"]


def golden_gate(name):
    return {"X": X, "Y": Y, "Z": Z, "H": H, "CNOT": CNOT, "TOF": TOF,
            "XZ": np.kron(X, Z), "ZZ": np.kron(Z, Z), "Xd3": Xd(3),
            "Xd3Xd3": np.kron(Xd(3), Xd(3))}[name]


def golden_state(spec, dims):
    """A fixture state: a list of ket terms [digits, re, im], or
    {"prj_mix": [[w, ket terms], ...]} for sum_i w_i prj(ket_i)."""
    def ket(terms):
        v = np.zeros(int(np.prod(dims)), dtype=complex)
        for digits, re, im in terms:
            v += complex(re, im) * mket(digits, dims)
        return v
    if isinstance(spec, dict):
        return sum(w * prj(ket(t)) for w, t in spec["prj_mix"])
    return ket(spec)


def golden_run(mod, case, u=None):
    """Runs one fixture case through `mod` (the oracle or qpp_amd, which share
    the qpp names) and returns what the case compares."""
    dims = case["dims"]
    A = golden_state(case["in"], dims)
    fn = case["fn"]
    if fn == "apply":
        return mod.apply(A, golden_gate(case["gate"]), case["target"], dims)
    if fn == "applyCTRL":
        return mod.applyCTRL(A, golden_gate(case["gate"]), case["ctrl"], case["target"], dims,
                             shift=case["shift"])
    if fn == "applyCTRL_fan":
        return mod.applyCTRL_fan(A, golden_gate(case["gate"]), case["ctrl"], case["target"], dims,
                                 shift=case["shift"])
    if fn == "measure_seq":
        return mod.measure_seq(A, case["target"], dims, case["destructive"], u)
    raise ValueError(fn)


def golden_check(mod, case):
    if case["fn"] == "measure_seq":
        exp = golden_state(case["out"], case["out_dims"])
        for u in case["u"]:
            ms, ps, st = golden_run(mod, case, u)
            assert [int(m) for m in ms] == case["results"], (case["ref"], u)
            assert np.allclose(ps, case["probs"], atol=1e-12, rtol=0), (case["ref"], u)
            assert np.max(np.abs(np.asarray(st) - exp)) <= case["tol"], (case["ref"], u)
        return
    out = golden_run(mod, case)
    exp = golden_state(case["out"], case["dims"])
    assert np.max(np.abs(np.asarray(out) - exp)) <= case.get("tol", 1e-12), case["ref"]

"""Seeded synthetic workloads (the benchmark configurations).

A circuit is a list of gates (kind, targets, controls, matrix):
  ("u",  [q],    [],  U 2x2)      dense 1-qubit gate, qpp::apply(psi, U, {q})
  ("cx", [t],    [c], X)          CNOT, qpp::applyCTRL(psi, X, {c}, {t})
  ("cz", [t],    [c], Z)          CZ,   qpp::applyCTRL(rho, Z, {c}, {t})

Haar unitaries are QR decompositions of complex Gaussian matrices with the
phases of R's diagonal divided out (the construction of qpp::randU,
random.hpp:257-283), drawn from numpy's PCG64 with a fixed seed.
"""
import hashlib
import struct

import numpy as np

X = np.array([[0, 1], [1, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)


def haar(d, rng):
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def random_layers(n, depth, seed):
    """Random circuit: `depth` layers of a Haar 1-qubit gate on every qubit
    (q = 0..n-1) followed by CNOTs on a random perfect pairing
    (perm[2i] -> perm[2i+1]; with odd n one qubit stays unpaired)."""
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(depth):
        for q in range(n):
            gates.append(("u", [q], [], haar(2, rng)))
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            gates.append(("cx", [int(perm[i + 1])], [int(perm[i])], X))
    return gates


def dm_layers(n, depth, seed):
    """Density-matrix circuit: `depth` layers of Haar 1-qubit gates on every
    qubit followed by CZ on adjacent pairs (2i, 2i+1)."""
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(depth):
        for q in range(n):
            gates.append(("u", [q], [], haar(2, rng)))
        for i in range(0, n - 1, 2):
            gates.append(("cz", [i + 1], [i], Z))
    return gates


def gate_bytes(gate, n, elem_bytes=16, rho=False):
    """Algorithmic HBM bytes of one gate applied on its own: every amplitude
    a dense gate touches is read once and written once; a controlled gate
    touches the 2^(n-c) control-satisfying amplitudes."""
    nb = 2 * n if rho else n
    kind, target, ctrl, _ = gate
    c = len(ctrl) * (2 if rho else 1)
    if rho and ctrl:
        # rows pass and columns pass each touch the control-satisfying half
        return 2 * 2 * elem_bytes * 2 ** (nb - len(ctrl))
    if rho:
        return 2 * 2 * elem_bytes * 2 ** nb  # U on rows, conj(U) on columns
    return 2 * elem_bytes * 2 ** (nb - c)


def apply_gates(st, gates):
    """Apply a gate list gate by gate on a qpp_amd.State."""
    for kind, target, ctrl, U in gates:
        if ctrl:
            st.apply_ctrl(U, ctrl, target)
        else:
            st.apply(U, target)


def ladder_layers(n, depth, seed):
    """Trajectory circuit: `depth` layers of a Haar 1-qubit gate on every
    qubit followed by a CNOT ladder (i -> i+1, i = 0..n-2)."""
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(depth):
        for q in range(n):
            gates.append(("u", [q], [], haar(2, rng)))
        for i in range(n - 1):
            gates.append(("cx", [i + 1], [i], X))
    return gates


def trajectory_circuit(qa, n, depth, seed):
    """A qpp_amd.QCircuit of ladder_layers ending in one destructive
    measurement of every qubit into dits 0..n-1 (the noisy-trajectory
    workload)."""
    qc = qa.QCircuit.from_gates(n, ladder_layers(n, depth, seed))
    qc.nc = n
    qc.measure(list(range(n)), 0)
    return qc


def measured_circuit(qa, n, depth, seed):
    """random_layers ending in one measurement of every qubit (the serving
    workload: an odd n leaves one qubit unpaired per layer)."""
    qc = qa.QCircuit.from_gates(n, random_layers(n, depth, seed))
    qc.nc = n
    qc.measure(list(range(n)), 0)
    return qc


def depolarizing_kraus(p):
    """Kraus operators of QubitDepolarizingNoise(p) (noise.hpp:388-408):
    sqrt(1-p) I, sqrt(p/3) X, sqrt(p/3) Y, sqrt(p/3) Z."""
    Y = np.array([[0, -1j], [1j, 0]])
    return [np.sqrt(1 - p) * np.eye(2, dtype=complex), np.sqrt(p / 3) * X,
            np.sqrt(p / 3) * Y, np.sqrt(p / 3) * Z]


def dm_pipeline(n, depth, seed, p=1e-3):
    """Density-matrix workload: rho = 0.9 |psi><psi| + 0.1 I / 2^n (psi a
    seeded random ket); `depth` layers of a Haar 1-qubit gate on every qubit
    and CZ on the pairs (2i, 2i+1), each gate followed by the depolarizing
    channel on every qubit it touched; then ptrace over the upper half of
    the qubits, ptranspose on qubits {0, 1, 2} and syspermute by a seeded
    permutation.  Returns (rho, ops, trace_out, pt_target, perm) with ops
    ("u", [q], U), ("cz", [c, t], Z), ("kraus", [q], Ks)."""
    rng = np.random.default_rng(seed)
    D = 2 ** n
    psi = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    psi /= np.linalg.norm(psi)
    rho = 0.9 * np.outer(psi, psi.conj()) + 0.1 * np.eye(D) / D
    Ks = depolarizing_kraus(p)
    ops = []
    for _ in range(depth):
        for q in range(n):
            ops.append(("u", [q], haar(2, rng)))
            ops.append(("kraus", [q], Ks))
        for i in range(0, n - 1, 2):
            ops.append(("cz", [i, i + 1], Z))
            ops.append(("kraus", [i], Ks))
            ops.append(("kraus", [i + 1], Ks))
    trace_out = list(range(n // 2, n))
    perm = [int(x) for x in rng.permutation(n)]
    return rho, ops, trace_out, [0, 1, 2][: min(3, n)], perm


def dm_program(qa, n, ops):
    """The gates and channels of dm_pipeline as one fused BitProgram."""
    bp = qa.BitProgram(2 * n)
    for kind, q, M in ops:
        if kind == "u":
            bp.rho_apply(M, q, n)
        elif kind == "cz":
            bp.rho_ctrl(M, [q[0]], [q[1]], n)
        else:
            bp.rho_channel(M, q[0], n)
    return bp


def dm_bytes(n, ops, elem_bytes=16):
    """Algorithmic bytes of the ops applied one by one: every gate and every
    channel reads and writes rho once (2 s 4^n)."""
    return len(ops) * 2 * elem_bytes * 4 ** n


def serve_request(k, rng):
    """One serving request's parameters drawn from the numpy Generator
    `rng`: n ~ U{10..26}, depth ~ U{10..100}, complex64 with probability 1/2
    ("c64" True), an engine seed, and gseed = 7000 + k, the seed of the
    request's measured_circuit."""
    n = int(rng.integers(10, 27))
    depth = int(rng.integers(10, 101))
    c64 = bool(rng.random() < 0.5)
    return {"n": n, "depth": depth, "c64": c64, "seed": int(rng.integers(1, 2 ** 31)),
            "gseed": 7000 + k}


def serve_requests(nreq, seed, first=0):
    """`nreq` serve_request draws (k = first..first+nreq-1) from PCG64 with
    `seed`."""
    rng = np.random.default_rng(seed)
    return [serve_request(first + k, rng) for k in range(nreq)]


_KIND = {"u": 0, "cx": 1, "cz": 2, "kraus": 3}


def _mats(M):
    if isinstance(M, (list, tuple)):
        return b"".join(np.ascontiguousarray(K, dtype="<c16").tobytes() for K in M)
    return np.ascontiguousarray(M, dtype="<c16").tobytes()


def gates_digest(gates):
    """SHA-256 of a gate list (random_layers / dm_layers / ladder_layers):
    per gate the kind code (u 0, cx 1, cz 2), the target and control lists
    (counts as one byte, qubits as little-endian int32) and the matrix as
    little-endian complex128, row-major."""
    h = hashlib.sha256()
    for kind, target, ctrl, U in gates:
        h.update(struct.pack("<BB", _KIND[kind], len(target)))
        h.update(struct.pack(f"<{len(target)}i", *target))
        h.update(struct.pack("<B", len(ctrl)))
        h.update(struct.pack(f"<{len(ctrl)}i", *ctrl))
        h.update(_mats(U))
    return h.hexdigest()


def dm_digest(rho, ops, trace_out, pt_target, perm):
    """SHA-256 of a dm_pipeline workload: rho's first row (it fixes the
    seeded pure state up to a phase), then per op the kind code (u 0, cz 2,
    kraus 3), its qubits and its matrices (little-endian complex128), then
    the traced-out qubits, the ptranspose targets and the permutation."""
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(rho[0], dtype="<c16").tobytes())
    for kind, q, M in ops:
        h.update(struct.pack(f"<BB{len(q)}i", _KIND[kind], len(q), *q))
        h.update(_mats(M))
    for lst in (trace_out, pt_target, perm):
        h.update(struct.pack(f"<B{len(lst)}i", len(lst), *lst))
    return h.hexdigest()


def requests_digest(reqs):
    """SHA-256 of serving requests' parameters (n, depth, c64, seed, gseed per
    request, as little-endian int64); each request's circuit follows from
    random_layers(n, depth, gseed)."""
    h = hashlib.sha256()
    for q in reqs:
        h.update(struct.pack("<5q", q["n"], q["depth"], int(q["c64"]), q["seed"], q["gseed"]))
    return h.hexdigest()


def save_gates(path, gates):
    """Write a gate list of 1-target gates with at most one control as an
    .npz of plain arrays (kind codes, targets, controls with -1 for none,
    the 2x2 matrices as complex128); load_gates reads it back."""
    kinds = np.array([_KIND[g[0]] for g in gates], dtype=np.uint8)
    tg = np.array([g[1][0] for g in gates], dtype=np.int32)
    ct = np.array([g[2][0] if g[2] else -1 for g in gates], dtype=np.int32)
    mats = np.array([g[3] for g in gates], dtype=np.complex128)
    np.savez_compressed(path, kinds=kinds, targets=tg, controls=ct, mats=mats)


def load_gates(path):
    """The gate list saved by save_gates, in the generators' form."""
    names = {v: k for k, v in _KIND.items()}
    with np.load(path) as f:
        kinds, tg, ct, mats = f["kinds"], f["targets"], f["controls"], f["mats"]
    return [(names[int(k)], [int(t)], [int(c)] if c >= 0 else [], mats[i].copy())
            for i, (k, t, c) in enumerate(zip(kinds, tg, ct))]

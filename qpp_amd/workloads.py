"""Seeded synthetic workloads (the benchmark configurations).

A circuit is a list of gates (kind, targets, controls, matrix):
  ("u",  [q],    [],  U 2x2)      dense 1-qubit gate, qpp::apply(psi, U, {q})
  ("cx", [t],    [c], X)          CNOT, qpp::applyCTRL(psi, X, {c}, {t})
  ("cz", [t],    [c], Z)          CZ,   qpp::applyCTRL(rho, Z, {c}, {t})

Haar unitaries are QR decompositions of complex Gaussian matrices with the
phases of R's diagonal divided out (the construction of qpp::randU,
random.hpp:257-283), drawn from numpy's PCG64 with a fixed seed.
"""
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

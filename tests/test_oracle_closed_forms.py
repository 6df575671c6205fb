"""Closed forms and brute-force reductions that pin the oracle where the
reference holds no vectors: each check is one a plausible slip (dropped
term, wrong bit order, transposed operand, missing conjugate) would fail.

Brute force builds the full 2^n x 2^n operator by Kronecker products
(np.kron, subsystem 0 leftmost = most significant), an index convention
independent of both the oracle's bit loops and the tensordot helpers."""
import functools

import numpy as np
import pytest

from helpers import CNOT, CZ, H, I2, TOF, X, Y, Z, haar, kets, prj, randket, randrho

SWAP = np.eye(4, dtype=complex)[[0, 2, 1, 3]]


def kron_all(ops):
    return functools.reduce(np.kron, ops)


def full_1q(U, q, n):
    return kron_all([U if k == q else I2 for k in range(n)])


def full_ctrl(U, ctrl, target, n, shift=None):
    """sum over control patterns: |c><c| (x) (U if all controls on else I)."""
    D = 2 ** n
    M = np.zeros((D, D), dtype=complex)
    shift = shift or [0] * len(ctrl)
    Ufull = full_gate(U, target, n)
    for i in range(D):
        bits = [(i >> (n - 1 - q)) & 1 for q in range(n)]
        on = all((bits[c] + s) % 2 == 1 for c, s in zip(ctrl, shift))
        col = Ufull[:, i] if on else np.eye(D)[:, i]
        M[:, i] = col
    return M


def full_gate(U, target, n):
    """Dense operator of U on `target` (target[0] = most significant bit of
    U's index) by an explicit basis sweep."""
    D, k = 2 ** n, len(target)
    M = np.zeros((D, D), dtype=complex)
    for col in range(D):
        bits = [(col >> (n - 1 - q)) & 1 for q in range(n)]
        a = 0
        for t in target:
            a = (a << 1) | bits[t]
        for r in range(2 ** k):
            out = list(bits)
            for j, t in enumerate(target):
                out[t] = (r >> (k - 1 - j)) & 1
            row = 0
            for b in out:
                row = (row << 1) | b
            M[row, col] += U[r, a]
    return M


def test_full_gate_matches_kron():
    rng = np.random.default_rng(0)
    U = haar(2, rng)
    for q in range(4):
        assert np.allclose(full_gate(U, [q], 4), full_1q(U, q, 4))
    V = haar(4, rng)
    assert np.allclose(full_gate(V, [0, 1], 3), np.kron(V, I2))
    assert np.allclose(full_gate(V, [1, 0], 2), SWAP @ V @ SWAP)


@pytest.mark.parametrize("n", [1, 2, 5, 8])
def test_hadamard_layer_uniform(orc, n):
    psi = kets("0" * n)
    for q in range(n):
        psi = orc.apply(psi, H, [q], [2] * n)
    assert np.max(np.abs(psi - 2 ** (-n / 2))) < 1e-14


@pytest.mark.parametrize("n", [2, 3, 6, 9])
def test_ghz_ladder(orc, n):
    psi = orc.apply(kets("0" * n), H, [0], [2] * n)
    for q in range(n - 1):
        psi = orc.applyCTRL(psi, X, [q], [q + 1], [2] * n)
    ref = (kets("0" * n) + kets("1" * n)) / np.sqrt(2)
    assert np.max(np.abs(psi - ref)) < 1e-14


@pytest.mark.parametrize("n", [3, 5, 7])
def test_apply_vs_kron_bruteforce(orc, n):
    rng = np.random.default_rng(n)
    psi, rho = randket(2 ** n, rng), randrho(2 ** n, rng)
    for k in (1, 2, 3):
        tgt = [int(t) for t in rng.permutation(n)[:k]]
        U = haar(2 ** k, rng)
        M = full_gate(U, tgt, n)
        assert np.max(np.abs(orc.apply(psi, U, tgt, [2] * n) - M @ psi)) < 1e-12
        got = orc.apply(rho, U, tgt, [2] * n)
        assert np.max(np.abs(got - M @ rho @ M.conj().T)) < 1e-12


@pytest.mark.parametrize("n", [3, 5, 6])
def test_applyctrl_vs_kron_bruteforce(orc, n):
    rng = np.random.default_rng(10 + n)
    psi, rho = randket(2 ** n, rng), randrho(2 ** n, rng)
    perm = [int(t) for t in rng.permutation(n)]
    ctrl, tgt = perm[:2], perm[2:3]
    U = haar(2, rng)
    for shift in (None, [1, 0], [1, 1]):
        M = full_ctrl(U, ctrl, tgt, n, shift)
        got = orc.applyCTRL(psi, U, ctrl, tgt, [2] * n, shift=shift)
        assert np.max(np.abs(got - M @ psi)) < 1e-12
        got = orc.applyCTRL(rho, U, ctrl, tgt, [2] * n, shift=shift)
        assert np.max(np.abs(got - M @ rho @ M.conj().T)) < 1e-12


def test_cnot_cz_toffoli_matrices_are_controlled_gates(orc):
    n = 3
    psi = randket(8, np.random.default_rng(3))
    assert np.allclose(orc.apply(psi, CNOT, [0, 2], [2] * n), full_ctrl(X, [0], [2], n) @ psi)
    assert np.allclose(orc.apply(psi, CZ, [2, 1], [2] * n), full_ctrl(Z, [1], [2], n) @ psi)
    assert np.allclose(orc.apply(psi, TOF, [1, 2, 0], [2] * n), full_ctrl(X, [1, 2], [0], n) @ psi)


def test_prj_commutes_with_apply(orc):
    rng = np.random.default_rng(7)
    n = 5
    psi = randket(2 ** n, rng)
    U = haar(4, rng)
    a = orc.apply(prj(psi), U, [3, 1], [2] * n)
    b = prj(orc.apply(psi, U, [3, 1], [2] * n))
    assert np.max(np.abs(a - b)) < 1e-12


def test_ptrace_of_product_is_factor(orc):
    rng = np.random.default_rng(8)
    a, b, c = randrho(2, rng), randrho(4, rng), randrho(2, rng)
    rho = kron_all([a, b, c])
    assert np.max(np.abs(orc.ptrace(rho, [0, 3], [2] * 4) - b)) < 1e-13
    assert np.max(np.abs(orc.ptrace(rho, [1, 2], [2] * 4) - np.kron(a, c))) < 1e-13
    psi = np.kron(kets("1"), randket(8, rng))
    assert np.max(np.abs(orc.ptrace(psi, [0], [2] * 4) - prj(psi[8:]))) < 1e-13


def test_ptranspose_bell_projector_is_swap_over_two(orc):
    bell = (kets("00") + kets("11")) / np.sqrt(2)
    P = prj(bell)
    for tgt in ([0], [1]):
        assert np.max(np.abs(orc.ptranspose(P, tgt, [2, 2]) - SWAP / 2)) == 0
    assert np.max(np.abs(orc.ptranspose(P, [0, 1], [2, 2]) - P.T)) == 0


def test_syspermute_inverse_roundtrip(orc):
    rng = np.random.default_rng(9)
    n = 6
    rho = randrho(2 ** n, rng)
    for sem in (0, 1):
        perm = [int(p) for p in rng.permutation(n)]
        inv = [0] * n
        for i, p in enumerate(perm):
            inv[p] = i
        once = orc.syspermute(rho, perm, [2] * n, semantics=sem)
        back = orc.syspermute(once, inv, [2] * n, semantics=sem)
        assert np.array_equal(back, rho)


def test_depolarizing_channel_trace_and_fixed_point(orc):
    """The Kraus set of QubitDepolarizingNoise (noise.hpp:396-400) is
    trace preserving and leaves I/2^n fixed."""
    p = 0.3
    Ks = [np.sqrt(1 - p) * I2, np.sqrt(p / 3) * X, np.sqrt(p / 3) * Y, np.sqrt(p / 3) * Z]
    n = 4
    rho = randrho(16, np.random.default_rng(4))
    out = orc.apply_kraus(rho, Ks, [2], [2] * n)
    assert abs(np.trace(out) - 1) < 1e-13
    mixed = np.eye(16) / 16
    assert np.max(np.abs(orc.apply_kraus(mixed, Ks, [1], [2] * n) - mixed)) < 1e-15

"""Pins the CPU oracle before it is trusted as the GPU engine's checker.

Three kinds of checks (none retypes the oracle's formulas):
  * known answers transcribed from qpp's own unit tests (file:line cited);
  * the numpy tensor-contraction reference in helpers.py (axis contractions,
    no bit/stride arithmetic) on seeded random inputs, qubits and qudits;
  * closed forms: depolarizing channel, Appendix-A style index maps for the
    as-built ptranspose / syspermute semantics, Bell-state partial
    transposes, projector identities.
"""
import itertools

import numpy as np
import pytest

from helpers import (CNOT, CZ, H, I2, TOF, X, Xd, Y, Z, golden_check,
                     golden_known_answers, haar, kets, mket, prj, randket,
                     randrho, t_apply, t_ctrl_op, t_ptrace, t_ptranspose,
                     t_syspermute)

TOL = 1e-12


def close(a, b, tol=TOL):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) <= tol


# ---------------------------------------------------------------------------
# known answers transcribed from qpp's unit tests, kept as data in
# tests/golden/known_answers.json (the same cases run on the GPU engine in
# test_gpu_parity.py)
# ---------------------------------------------------------------------------
_GOLDEN = golden_known_answers()


@pytest.mark.parametrize("case", _GOLDEN, ids=[f"{c['fn']}-{i}" for i, c in enumerate(_GOLDEN)])
def test_known_answers_oracle(orc, case):
    golden_check(orc, case)


# known answers: unit_tests/tests/operations.cpp:16-53
def test_apply_known_answers_qubits(orc):
    psi = kets("1")
    assert close(orc.apply(psi, X, [0], [2]), kets("0"))
    assert close(orc.apply(psi, Y, [0], [2]), -1j * kets("0"))
    assert close(orc.apply(psi, Z, [0], [2]), -kets("1"))
    assert close(orc.apply(psi, H, [0], [2]), (kets("0") - kets("1")) / np.sqrt(2))
    psi = 0.8 * kets("00") + 0.6 * kets("11")
    assert close(orc.apply(psi, X, [1], [2, 2]), 0.8 * kets("01") + 0.6 * kets("10"))
    assert close(orc.apply(psi, Y, [1], [2, 2]), 0.8j * kets("01") - 0.6j * kets("10"))
    assert close(orc.apply(psi, Z, [1], [2, 2]), 0.8 * kets("00") - 0.6 * kets("11"))
    assert close(orc.apply(psi, CNOT, [0, 1], [2, 2]), 0.8 * kets("00") + 0.6 * kets("10"))
    assert close(orc.apply(psi, CNOT, [1, 0], [2, 2]), 0.8 * kets("00") + 0.6 * kets("01"))
    assert close(orc.apply(psi, np.kron(Z, Z), [0, 1], [2, 2]), psi)
    psi = 0.8 * kets("0000") + 0.6 * kets("1111")
    XZ = np.kron(X, Z)
    assert close(orc.apply(psi, XZ, [1, 2], [2] * 4), 0.8 * kets("0100") - 0.6 * kets("1011"))
    assert close(orc.apply(psi, XZ, [2, 1], [2] * 4), 0.8 * kets("0010") - 0.6 * kets("1101"))
    assert close(orc.apply(psi, TOF, [1, 2, 0], [2] * 4), 0.8 * kets("0000") + 0.6 * kets("0111"))


def test_apply_known_answers_qutrits(orc):
    d = 3
    psi = mket([0, 0], d) + mket([1, 0], d) / np.sqrt(2)
    exp = mket([1, 0], d) + mket([2, 0], d) / np.sqrt(2)
    assert close(orc.apply(psi, Xd(d), [0], [d, d]), exp)
    rho = 0.6 * prj(mket([0, 0], d) + mket([1, 0], d) / np.sqrt(2)) + \
        0.4 * prj(mket([0, 1], d) + mket([1, 1], d) / np.sqrt(2))
    exp = 0.6 * prj(mket([1, 0], d) + mket([2, 0], d) / np.sqrt(2)) + \
        0.4 * prj(mket([1, 1], d) + mket([2, 1], d) / np.sqrt(2))
    assert close(orc.apply(rho, Xd(d), [0], [d, d]), exp)


# unit_tests/tests/operations.cpp:101-162 (deterministic parts)
def test_applyctrl_known_answers_qutrits(orc):
    d, dims = 3, [3] * 4
    ctrl, target = [0, 2], [1, 3]
    U = np.kron(Xd(3), Xd(3))
    assert close(orc.applyCTRL(mket([2, 1, 2, 1], d), U, ctrl, target, dims), mket([2, 0, 2, 0], d))
    assert close(orc.applyCTRL(mket([2, 1, 1, 1], d), U, ctrl, target, dims, shift=[1, 2]),
                 mket([2, 1, 1, 1], d))
    assert close(orc.applyCTRL(prj(mket([2, 0, 2, 0], d)), U, ctrl, target, dims),
                 prj(mket([2, 2, 2, 2], d)))
    assert close(orc.applyCTRL(prj(mket([1, 0, 0, 0], d)), U, ctrl, target, dims, shift=[1, 2]),
                 prj(mket([1, 2, 0, 2], d)))


# unit_tests/tests/operations.cpp:176-230 (deterministic parts)
def test_applyctrl_fan_known_answers_qutrits(orc):
    d, dims = 3, [3] * 4
    ctrl, target = [0, 2], [1, 3]
    assert close(orc.applyCTRL_fan(mket([2, 1, 2, 1], d), Xd(3), ctrl, target, dims), mket([2, 0, 2, 0], d))
    assert close(orc.applyCTRL_fan(mket([2, 1, 1, 1], d), Xd(3), ctrl, target, dims, shift=[1, 2]),
                 mket([2, 1, 1, 1], d))
    assert close(orc.applyCTRL_fan(prj(mket([2, 0, 2, 0], d)), Xd(3), ctrl, target, dims),
                 prj(mket([2, 2, 2, 2], d)))
    assert close(orc.applyCTRL_fan(prj(mket([1, 0, 0, 0], d)), Xd(3), ctrl, target, dims, shift=[1, 2]),
                 prj(mket([1, 2, 0, 2], d)))


# unit_tests/tests/operations.cpp:237-397 (qubit fan kernels, 1e-12)
def test_applyctrl_fan_known_answers_qubits(orc):
    n, m = 6, 3
    ctrl, target = [0, 1, 2], [3, 4, 5]
    dims = [2] * n
    midx = [1, 1, 1, 0, 0, 0]
    out = orc.applyCTRL_fan(mket(midx, 2), X, ctrl, target, dims, shift=[0] * m)
    assert close(out, mket([1, 1, 1, 1, 1, 1], 2))
    shift = [1, 0, 1]
    midx = [0, 1, 0, 0, 0, 0]
    out = orc.applyCTRL_fan(mket(midx, 2), X, ctrl, target, dims, shift=shift)
    assert close(out, mket([0, 1, 0, 1, 1, 1], 2))
    broken = [0, 0, 0, 0, 0, 0]
    out = orc.applyCTRL_fan(mket(broken, 2), X, ctrl, target, dims, shift=shift)
    assert close(out, mket(broken, 2))
    out = orc.applyCTRL_fan(prj(mket([1, 1, 1, 0, 0, 0], 2)), X, ctrl, target, dims, shift=[0] * m)
    assert close(out, prj(mket([1] * 6, 2)))
    out = orc.applyCTRL_fan(prj(mket(midx, 2)), X, ctrl, target, dims, shift=shift)
    assert close(out, prj(mket([0, 1, 0, 1, 1, 1], 2)))
    rng = np.random.default_rng(7)
    psi = randket(64, rng)
    U = haar(2, rng)
    A = orc.applyCTRL_fan(psi, U, ctrl, target, dims, shift=[0] * m)
    B = orc.applyCTRL_fan(np.outer(psi, psi.conj()), U, ctrl, target, dims, shift=[0] * m)
    assert close(np.outer(A, A.conj()), B, 1e-10)


# unit_tests/tests/instruments.cpp:102-160
def test_measure_seq_known_answers(orc):
    dims = [2, 2, 3, 2, 3, 2]
    mask = [0, 1, 2, 1, 2, 0]
    psi = mket(mask, dims)
    rho = prj(psi)
    # u = 0 exactly would select index 0 (lower_bound over the zero prefix of
    # the table); generate_canonical yields it with probability 2^-64
    for u in (1e-9, 0.3, 0.999):
        ms, ps, st = orc.measure_seq(psi, [4, 2, 3, 0], dims, True, u)
        assert ms == [2, 2, 1, 0] and ps == [1, 1, 1, 1]
        assert close(st, mket([1, 0], [2, 2]), 1e-5)
        ms, ps, st = orc.measure_seq(rho, [4, 2, 3, 0], dims, True, u)
        assert ms == [2, 2, 1, 0] and ps == [1, 1, 1, 1]
        assert close(st, prj(mket([1, 0], [2, 2])), 1e-5)
        ms, ps, st = orc.measure_seq(psi, [4, 2, 3, 0], dims, False, u)
        assert ms == [2, 2, 1, 0] and ps == [1, 1, 1, 1] and close(st, psi, 1e-5)
        ms, ps, st = orc.measure_seq(psi, [3, 0, 4, 2], dims, False, u)
        assert ms == [1, 0, 2, 2] and ps == [1, 1, 1, 1] and close(st, psi, 1e-5)
        ms, ps, st = orc.measure_seq(rho, [3, 0, 4, 2], dims, False, u)
        assert ms == [1, 0, 2, 2] and ps == [1, 1, 1, 1] and close(st, rho, 1e-5)


# unit_tests/tests/entanglement.cpp:107-112: syspermute(kron(rho, sigma),
# {0, 2, 1, 3}) of two 2-qubit states interleaves their qubits
def test_syspermute_known_answer(orc):
    rng = np.random.default_rng(3)
    r, s = randrho(4, rng), randrho(4, rng)
    out = orc.syspermute(np.kron(r, s), [0, 2, 1, 3], [2] * 4)
    assert close(out, t_syspermute(np.kron(r, s), [0, 2, 1, 3], [2] * 4))
    for sem in (0, 1):  # reversal-symmetric perm: both semantics agree
        assert close(orc.syspermute(np.kron(r, s), [0, 2, 1, 3], [2] * 4, sem), out)


# ---------------------------------------------------------------------------
# tensor reference on seeded random inputs
# ---------------------------------------------------------------------------
DIMS_CASES = [[2], [2, 2], [2, 2, 2], [2] * 5, [2] * 6, [3, 2], [2, 3, 2], [3, 3, 2]]


@pytest.mark.parametrize("dims", DIMS_CASES)
def test_apply_vs_tensor(orc, dims):
    rng = np.random.default_rng(sum(dims) * 31 + len(dims))
    n = len(dims)
    D = int(np.prod(dims))
    psi = randket(D, rng)
    rho = randrho(D, rng)
    for k in range(1, min(n, 3) + 1):
        for target in itertools.islice(itertools.permutations(range(n), k), 6):
            DA = int(np.prod([dims[t] for t in target]))
            U = haar(DA, rng)
            assert close(orc.apply(psi, U, list(target), dims), t_apply(psi, U, target, dims))
            assert close(orc.apply(rho, U, list(target), dims), t_apply(rho, U, target, dims))


@pytest.mark.parametrize("dims", [[2, 2], [2] * 4, [2] * 5, [3, 3, 3], [2, 3, 3, 2]])
def test_applyctrl_vs_tensor(orc, dims):
    rng = np.random.default_rng(11 + len(dims))
    n = len(dims)
    D = int(np.prod(dims))
    psi = randket(D, rng)
    rho = randrho(D, rng)
    cases = []
    for c in range(n):
        for t in range(n):
            if c != t and dims[c] == dims[t]:
                cases.append(([c], [t]))
    if n >= 4:
        cases += [([0, 2], [1]), ([3, 1], [0, 2])]
    for ctrl, target in cases:
        if len({dims[c] for c in ctrl}) != 1:
            continue
        d = dims[ctrl[0]]
        DA = int(np.prod([dims[t] for t in target]))
        U = haar(DA, rng)
        for shift in (None, [int(rng.integers(d)) for _ in ctrl]):
            M = t_ctrl_op(U, ctrl, target, dims, shift)
            assert close(orc.applyCTRL(psi, U, ctrl, target, dims, shift), M @ psi)
            assert close(orc.applyCTRL(rho, U, ctrl, target, dims, shift), M @ rho @ M.conj().T)


@pytest.mark.parametrize("dims", [[2] * 4, [2] * 5, [3, 3, 3]])
def test_applyctrl_fan_vs_tensor(orc, dims):
    rng = np.random.default_rng(5)
    n = len(dims)
    D = int(np.prod(dims))
    psi = randket(D, rng)
    rho = randrho(D, rng)
    d = dims[0]
    ctrl, target = [1], [0, n - 1]
    U = haar(d, rng)
    M = np.eye(D, dtype=complex)
    for t in target:
        M = t_ctrl_op(U, ctrl, [t], dims) @ M
    assert close(orc.applyCTRL_fan(psi, U, ctrl, target, dims), M @ psi)
    assert close(orc.applyCTRL_fan(rho, U, ctrl, target, dims), M @ rho @ M.conj().T)


def depolarizing_kraus(p):
    return [np.sqrt(1 - p) * I2, np.sqrt(p / 3) * X, np.sqrt(p / 3) * Y, np.sqrt(p / 3) * Z]


@pytest.mark.parametrize("n,q", [(1, 0), (3, 0), (3, 2), (4, 1)])
def test_kraus_depolarizing_closed_form(orc, n, q):
    """sum_K K rho K^dagger for the Kraus set of noise.hpp:396-400 equals
    (1 - 4p/3) rho + (2p/3) I_q (x) tr_q(rho) (Pauli twirl identity)."""
    rng = np.random.default_rng(n * 10 + q)
    dims = [2] * n
    D = 2 ** n
    rho = randrho(D, rng)
    p = 0.137
    out = orc.apply_kraus(rho, depolarizing_kraus(p), [q], dims)
    red = t_ptrace(rho, [q], dims)  # over the other qubits
    # embed I_q (x) red at position q
    T = np.einsum("ab,...->ab...", I2, red.reshape([2] * (2 * (n - 1))) if n > 1 else red)
    full = np.zeros((D, D), dtype=complex)
    for i in range(D):
        for j in range(D):
            bi = [(i >> (n - 1 - k)) & 1 for k in range(n)]
            bj = [(j >> (n - 1 - k)) & 1 for k in range(n)]
            if bi[q] != bj[q]:
                continue
            ri = [b for k, b in enumerate(bi) if k != q]
            rj = [b for k, b in enumerate(bj) if k != q]
            ii = int("".join(map(str, ri)), 2) if ri else 0
            jj = int("".join(map(str, rj)), 2) if rj else 0
            full[i, j] = red[ii, jj]
    del T
    assert close(out, (1 - 4 * p / 3) * rho + (2 * p / 3) * full)


@pytest.mark.parametrize("dims", [[2] * 3, [2] * 4, [3, 2], [2, 3, 2]])
def test_kraus_vs_tensor(orc, dims):
    rng = np.random.default_rng(19)
    D = int(np.prod(dims))
    rho = randrho(D, rng)
    n = len(dims)
    for target in ([0], [n - 1], [1, 0]):
        DK = int(np.prod([dims[t] for t in target]))
        Ks = [haar(DK, rng) * np.sqrt(w) for w in (0.5, 0.3, 0.2)]
        ref = sum(t_apply(rho, K, target, dims) for K in Ks)
        assert close(orc.apply_kraus(rho, Ks, target, dims), ref)


@pytest.mark.parametrize("dims", [[2] * 3, [2] * 5, [3, 2, 2], [2, 3, 2, 3]])
def test_ptrace_vs_tensor(orc, dims):
    rng = np.random.default_rng(23)
    n = len(dims)
    D = int(np.prod(dims))
    psi = randket(D, rng)
    rho = randrho(D, rng)
    targets = [[], [0], [n - 1], [n - 1, 0], list(range(n)), [1, 0][: n]]
    for t in targets:
        assert close(orc.ptrace(rho, t, dims), t_ptrace(rho, t, dims))
        assert close(orc.ptrace(psi, t, dims), t_ptrace(psi, t, dims))


@pytest.mark.parametrize("dims", [[2] * 3, [2] * 4, [3, 2], [2, 3, 2]])
def test_ptranspose_documented_vs_tensor(orc, dims):
    rng = np.random.default_rng(29)
    n = len(dims)
    D = int(np.prod(dims))
    psi = randket(D, rng)
    rho = randrho(D, rng)
    for t in ([], [0], [n - 1], [1, 0], list(range(n))):
        assert close(orc.ptranspose(rho, t, dims, 0), t_ptranspose(rho, t, dims))
        assert close(orc.ptranspose(psi, t, dims, 0), t_ptranspose(psi, t, dims))


@pytest.mark.parametrize("dims", [[2] * 3, [2] * 4, [3, 2, 2], [2, 3, 4]])
def test_syspermute_documented_vs_tensor(orc, dims):
    rng = np.random.default_rng(31)
    n = len(dims)
    D = int(np.prod(dims))
    psi = randket(D, rng)
    rho = randrho(D, rng)
    for perm in itertools.islice(itertools.permutations(range(n)), 8):
        perm = list(perm)
        assert close(orc.syspermute(psi, perm, dims, 0), t_syspermute(psi, perm, dims))
        assert close(orc.syspermute(rho, perm, dims, 0), t_syspermute(rho, perm, dims))


# ---------------------------------------------------------------------------
# as-built qubit-kernel semantics (ptranspose.hpp:101-119, 189-205;
# syspermute.hpp:70-93, 131-154): closed-form cases
# ---------------------------------------------------------------------------
def basis_rho(x, y, n):
    D = 2 ** n
    r = np.zeros((D, D), dtype=complex)
    r[int(x, 2), int(y, 2)] = 1
    return r


def test_ptranspose_as_built_closed_form(orc):
    # |010><000|, target {1}: documented swaps digit 1 -> |000><010|;
    # as built the kernels transpose complement({n-1-1}) = {0, 2}, where x
    # and y agree, so the element stays at |010><000|
    r = basis_rho("010", "000", 3)
    assert close(orc.ptranspose(r, [1], 2, 0), basis_rho("000", "010", 3))
    assert close(orc.ptranspose(r, [1], 2, 1), r)
    # |100><000|, target {0}: documented -> |000><100|; as built transposes
    # complement({2}) = {0, 1} -> |000><100| as well (digit 0 differs)
    r = basis_rho("100", "000", 3)
    assert close(orc.ptranspose(r, [0], 2, 0), basis_rho("000", "100", 3))
    assert close(orc.ptranspose(r, [0], 2, 1), basis_rho("000", "100", 3))
    # |001><000|, target {0}: documented unchanged; as built transposes
    # {0, 1} -> unchanged; target {2}: as built transposes complement({0}) =
    # {1, 2} -> |000><001|
    r = basis_rho("001", "000", 3)
    assert close(orc.ptranspose(r, [0], 2, 1), r)
    assert close(orc.ptranspose(r, [2], 2, 1), basis_rho("000", "001", 3))
    # trivial targets keep the documented meaning in both modes
    rng = np.random.default_rng(1)
    rho = randrho(8, rng)
    for sem in (0, 1):
        assert close(orc.ptranspose(rho, [], 2, sem), rho)
        assert close(orc.ptranspose(rho, [0, 1, 2], 2, sem), rho.T)


def test_ptranspose_as_built_equals_documented_on_complement_of_reversed(orc):
    rng = np.random.default_rng(2)
    for n in (2, 3, 4):
        rho = randrho(2 ** n, rng)
        psi = randket(2 ** n, rng)
        for k in range(1, n):
            for t in itertools.combinations(range(n), k):
                eff = [q for q in range(n) if q not in {n - 1 - x for x in t}]
                for st in (rho, psi):
                    assert close(orc.ptranspose(st, list(t), 2, 1), orc.ptranspose(st, eff, 2, 0))


def test_syspermute_as_built_closed_form(orc):
    # n = 3, perm {1, 0, 2}: documented |a0 a1 a2> -> |a1 a0 a2>; as built
    # |a0 a1 a2> -> |a0 a2 a1>
    for bits in itertools.product("01", repeat=3):
        s = "".join(bits)
        a0, a1, a2 = s
        assert close(orc.syspermute(kets(s), [1, 0, 2], 2, 0), kets(a1 + a0 + a2))
        assert close(orc.syspermute(kets(s), [1, 0, 2], 2, 1), kets(a0 + a2 + a1))
        # both agree on {1, 2, 0} (|a1 a2 a0>) and {2, 1, 0}
        assert close(orc.syspermute(kets(s), [1, 2, 0], 2, 1), kets(a1 + a2 + a0))
        assert close(orc.syspermute(kets(s), [1, 2, 0], 2, 0), kets(a1 + a2 + a0))
        assert close(orc.syspermute(kets(s), [2, 1, 0], 2, 1), kets(a2 + a1 + a0))


def test_syspermute_as_built_is_reversed_inverse(orc):
    rng = np.random.default_rng(4)
    for n in (2, 3, 4, 5):
        rho = randrho(2 ** n, rng)
        psi = randket(2 ** n, rng)
        for perm in itertools.islice(itertools.permutations(range(n)), 12):
            inv = np.argsort(perm)
            eff = [n - 1 - int(inv[n - 1 - k]) for k in range(n)]
            for st in (rho, psi):
                assert close(orc.syspermute(st, list(perm), 2, 1), orc.syspermute(st, eff, 2, 0))


# ---------------------------------------------------------------------------
# sampling
# ---------------------------------------------------------------------------
def test_probs_are_std_norm(orc):
    rng = np.random.default_rng(8)
    psi = randket(64, rng)
    assert np.array_equal(orc.probs(psi), psi.real * psi.real + psi.imag * psi.imag)
    rho = randrho(16, rng)
    assert np.array_equal(orc.probs(rho), np.diag(rho).real)


def test_sample_index_and_edges(orc):
    # exact table: cp = {0.25, 0.5, 0.75, 1}; lower_bound picks the first
    # cp >= u, so u exactly on an edge selects that edge's index
    p = np.array([0.25, 0.25, 0.25, 0.25])
    assert [orc.discrete_draw(p, u) for u in (0.0, 0.25, 0.2500001, 0.5, 0.74, 0.75, 0.99)] == \
        [0, 0, 1, 1, 2, 2, 3]
    # the last entry is forced to 1 even when the running sum falls short,
    # and zero-weight tails are reachable only through that forced entry
    p = np.array([0.1] * 10)
    cp = orc.discrete_table(p)
    assert cp[-1] == 1.0
    assert orc.discrete_draw([1.0, 0.0, 0.0], 0.999999) == 0
    # a single weight: no table, index 0 (random.tcc:2660-2664, 2700-2701)
    assert orc.discrete_draw([0.7], 0.9) == 0
    # sample() reports pbs[index] (instruments.hpp:808) and target digits
    psi = (kets("01") + kets("10")) / np.sqrt(2)
    idx, dig, pr = orc.sample(psi, [1, 0], 2, 0.3)
    assert idx == 1 and dig == [1, 0] and abs(pr - 0.5) < 1e-15


def test_discrete_draw_matches_libstdcxx(orc):
    """The oracle's discrete_distribution restatement against libstdc++'s own
    outputs (tests/golden/libstdcxx_draws.json, written by
    tests/golden/make_libstdcxx_draws.py with g++ 11.4): fed the oracle's
    generate_canonical draws it reproduces std::discrete_distribution's index
    stream; its cumulative table is the sequential partial sum of libstdc++'s
    normalised probabilities() with the last entry 1.0; and both the oracle's
    and the product package's Rng give libstdc++'s first uniforms."""
    import json
    import os
    import qpp_amd
    with open(os.path.join(os.path.dirname(__file__), "golden", "libstdcxx_draws.json")) as f:
        doc = json.load(f)
    for case in doc["cases"]:
        p = np.array([float.fromhex(h) for h in case["weights_hex"]])
        seed, ref = case["seed"], case["draws"]
        r = orc.Rng(seed)
        assert [orc.discrete_draw(p, r.canonical()) for _ in ref] == ref
        q = [float.fromhex(h) for h in case["probabilities_hex"]]
        cp = list(itertools.accumulate(q))
        cp[-1] = 1.0
        assert list(orc.discrete_table(p)) == cp
        ucan = [float.fromhex(h) for h in case["canonical_hex"]]
        r2 = orc.Rng(seed)
        assert [r2.canonical() for _ in range(4)] == ucan
        py = qpp_amd.Rng(seed)
        assert [py.canonical() for _ in range(4)] == ucan


def test_measure_seq_bell(orc):
    psi = (kets("00") + kets("11")) / np.sqrt(2)
    for u, m in ((0.2, 0), (0.7, 1)):
        ms, ps, st = orc.measure_seq(psi, [0], 2, True, u)
        assert ms == [m] and abs(ps[0] - 0.5) < 1e-15
        assert close(st, kets(str(m)))
        ms, ps, st = orc.measure_seq(psi, [1, 0], 2, False, u)
        assert ms == [m, m] and abs(ps[0] - 0.5) < 1e-15 and abs(ps[1] - 1) < 1e-15
        assert close(st, kets(str(m) * 2))

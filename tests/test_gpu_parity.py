"""GPU engine vs the oracle, through the C ABI, on seeded inputs.

Tolerances: complex128 amplitudes 1e-12 max-abs (the parity bar); index
maps (ptranspose, syspermute, probabilities, sampled indices) bit-exact;
complex64 states 2e-5 max-abs.
"""
import itertools

import numpy as np
import pytest

from helpers import (CNOT, H, I2, TOF, X, Xd, Y, Z, golden_check, golden_known_answers, haar,
                     kets, mket, prj, randket, randrho)

pytestmark = pytest.mark.gpu

TOL = 1e-12


def maxabs(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


_GOLDEN = golden_known_answers()


@pytest.mark.parametrize("case", _GOLDEN, ids=[f"{c['fn']}-{i}" for i, c in enumerate(_GOLDEN)])
def test_known_answers_gpu(qa, case):
    """qpp's own unit-test known answers (tests/golden/known_answers.json),
    computed by the GPU engine through the C ABI."""
    golden_check(qa, case)


def test_library_is_in_tree(qa):
    import os
    assert os.path.dirname(qa.LIB_PATH) == os.path.dirname(os.path.abspath(qa.__file__))
    lib = qa.lib()
    assert lib._name == qa.LIB_PATH


QUBIT_NS = [1, 2, 3, 5, 6, 7, 9, 12]


@pytest.mark.parametrize("n", QUBIT_NS)
def test_apply_ket_qubits(qa, orc, n):
    rng = np.random.default_rng(100 + n)
    psi = randket(2 ** n, rng)
    for k in range(1, min(n, 4) + 1):
        for _ in range(4):
            target = list(rng.permutation(n)[:k])
            U = haar(2 ** k, rng)
            got = qa.apply(psi, U, target, 2)
            ref = orc.apply(psi, U, target, 2)
            assert maxabs(got, ref) <= TOL, (n, target)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 6])
def test_apply_rho_qubits(qa, orc, n):
    rng = np.random.default_rng(200 + n)
    rho = randrho(2 ** n, rng)
    for k in range(1, min(n, 3) + 1):
        for _ in range(3):
            target = list(rng.permutation(n)[:k])
            U = haar(2 ** k, rng)
            assert maxabs(qa.apply(rho, U, target, 2), orc.apply(rho, U, target, 2)) <= TOL


@pytest.mark.parametrize("n", [2, 3, 5, 7, 10])
def test_applyctrl_qubits(qa, orc, n):
    rng = np.random.default_rng(300 + n)
    psi = randket(2 ** n, rng)
    rho = randrho(2 ** min(n, 5), rng)
    nr = min(n, 5)
    for _ in range(6):
        perm = list(rng.permutation(n))
        nc = int(rng.integers(1, max(2, n - 1)))
        nc = min(nc, n - 1)
        ctrl, target = perm[:nc], perm[nc:nc + 1 + int(rng.integers(0, min(2, n - nc)))]
        U = haar(2 ** len(target), rng)
        shift = [int(v) for v in rng.integers(0, 2, len(ctrl))] if rng.random() < 0.5 else None
        assert maxabs(qa.applyCTRL(psi, U, ctrl, target, 2, shift),
                      orc.applyCTRL(psi, U, ctrl, target, 2, shift)) <= TOL
        permr = list(rng.permutation(nr))
        c2, t2 = permr[:1], permr[1:2]
        assert maxabs(qa.applyCTRL(rho, U[:2, :2] if len(target) > 1 else U, c2, t2, 2, None),
                      orc.applyCTRL(rho, U[:2, :2] if len(target) > 1 else U, c2, t2, 2, None)) <= TOL


def test_applyctrl_rho_shift(qa, orc):
    rng = np.random.default_rng(7)
    rho = randrho(2 ** 5, rng)
    U = haar(4, rng)
    for ctrl, target, shift in (([0, 3], [1, 4], [1, 0]), ([4], [0, 2], [1]), ([2, 1], [3, 0], None)):
        assert maxabs(qa.applyCTRL(rho, U, ctrl, target, 2, shift),
                      orc.applyCTRL(rho, U, ctrl, target, 2, shift)) <= TOL


def test_applyctrl_fan(qa, orc):
    rng = np.random.default_rng(8)
    n = 6
    psi = randket(2 ** n, rng)
    rho = randrho(2 ** n, rng)
    U = haar(2, rng)
    for ctrl, target, shift in (([0, 1, 2], [3, 4, 5], None), ([5], [0, 2, 4], [1]), ([1, 3], [5, 0], [1, 0])):
        assert maxabs(qa.applyCTRL_fan(psi, U, ctrl, target, 2, shift),
                      orc.applyCTRL_fan(psi, U, ctrl, target, 2, shift)) <= TOL
        assert maxabs(qa.applyCTRL_fan(rho, U, ctrl, target, 2, shift),
                      orc.applyCTRL_fan(rho, U, ctrl, target, 2, shift)) <= TOL
    # unit_tests/tests/operations.cpp:237-397 known answers on the GPU
    out = qa.applyCTRL_fan(mket([0, 1, 0, 0, 0, 0], 2), X, [0, 1, 2], [3, 4, 5], 2, [1, 0, 1])
    assert maxabs(out, mket([0, 1, 0, 1, 1, 1], 2)) <= TOL


def depol(p):
    return [np.sqrt(1 - p) * I2, np.sqrt(p / 3) * X, np.sqrt(p / 3) * Y, np.sqrt(p / 3) * Z]


@pytest.mark.parametrize("n", [1, 3, 6])
def test_kraus(qa, orc, n):
    rng = np.random.default_rng(400 + n)
    rho = randrho(2 ** n, rng)
    for q in range(n):
        assert maxabs(qa.apply(rho, depol(1e-3), [q], 2), orc.apply_kraus(rho, depol(1e-3), [q], 2)) <= TOL
    if n >= 2:
        Ks = [haar(4, rng) * np.sqrt(w) for w in (0.6, 0.4)]
        assert maxabs(qa.apply(rho, Ks, [n - 1, 0], 2), orc.apply_kraus(rho, Ks, [n - 1, 0], 2)) <= TOL


@pytest.mark.parametrize("n", [2, 4, 7])
def test_ptrace(qa, orc, n):
    rng = np.random.default_rng(500 + n)
    psi = randket(2 ** n, rng)
    rho = randrho(2 ** min(n, 6), rng)
    nr = min(n, 6)
    for t in ([0], [n - 1], list(range(n // 2)), [n - 1, 0], list(range(n)), []):
        assert maxabs(qa.ptrace(psi, t, 2), orc.ptrace(psi, t, 2)) <= TOL
    for t in ([0], sorted({nr - 1, 1}), list(range(nr // 2, nr)), list(range(nr)), []):
        assert maxabs(qa.ptrace(rho, t, 2), orc.ptrace(rho, t, 2)) <= TOL


@pytest.mark.parametrize("sem", [0, 1])
@pytest.mark.parametrize("n", [2, 3, 5, 7])
def test_ptranspose_bit_exact(qa, orc, n, sem):
    rng = np.random.default_rng(600 + n)
    rho = randrho(2 ** n, rng)
    psi = randket(2 ** min(n, 5), rng)
    for k in range(0, n + 1):
        t = list(rng.permutation(n)[:k])
        assert np.array_equal(qa.ptranspose(rho, t, 2, sem), orc.ptranspose(rho, t, 2, sem)), (t, sem)
    m = min(n, 5)
    for k in range(0, m + 1):
        t = list(rng.permutation(m)[:k])
        assert maxabs(qa.ptranspose(psi, t, 2, sem), orc.ptranspose(psi, t, 2, sem)) <= TOL


@pytest.mark.parametrize("sem", [0, 1])
@pytest.mark.parametrize("n", [1, 3, 6, 9])
def test_syspermute_bit_exact(qa, orc, n, sem):
    rng = np.random.default_rng(700 + n)
    psi = randket(2 ** n, rng)
    rho = randrho(2 ** min(n, 6), rng)
    for _ in range(5):
        perm = list(rng.permutation(n))
        assert np.array_equal(qa.syspermute(psi, perm, 2, sem), orc.syspermute(psi, perm, 2, sem))
        pr = list(rng.permutation(min(n, 6)))
        assert np.array_equal(qa.syspermute(rho, pr, 2, sem), orc.syspermute(rho, pr, 2, sem))


# ---------------------------------------------------------------------------
# qudits
# ---------------------------------------------------------------------------
QUDIT_DIMS = [[3, 2], [2, 3, 2], [3, 3, 3], [4, 2, 3]]


@pytest.mark.parametrize("dims", QUDIT_DIMS)
def test_qudit_ops(qa, orc, dims):
    rng = np.random.default_rng(sum(dims))
    n = len(dims)
    D = int(np.prod(dims))
    psi = randket(D, rng)
    rho = randrho(D, rng)
    for k in (1, 2):
        target = list(rng.permutation(n)[:k])
        DA = int(np.prod([dims[t] for t in target]))
        U = haar(DA, rng)
        assert maxabs(qa.apply(psi, U, target, dims), orc.apply(psi, U, target, dims)) <= TOL
        assert maxabs(qa.apply(rho, U, target, dims), orc.apply(rho, U, target, dims)) <= TOL
    for t in ([0], [n - 1, 0], []):
        assert maxabs(qa.ptrace(rho, t, dims), orc.ptrace(rho, t, dims)) <= TOL
        assert maxabs(qa.ptrace(psi, t, dims), orc.ptrace(psi, t, dims)) <= TOL
        assert np.array_equal(qa.ptranspose(rho, t, dims), orc.ptranspose(rho, t, dims))
    for perm in itertools.islice(itertools.permutations(range(n)), 4):
        assert np.array_equal(qa.syspermute(psi, list(perm), dims), orc.syspermute(psi, list(perm), dims))
        assert np.array_equal(qa.syspermute(rho, list(perm), dims), orc.syspermute(rho, list(perm), dims))


def test_qudit_ctrl_known_answers(qa):
    # unit_tests/tests/operations.cpp:101-230
    d, dims = 3, [3] * 4
    U = np.kron(Xd(3), Xd(3))
    assert maxabs(qa.applyCTRL(mket([2, 1, 2, 1], d), U, [0, 2], [1, 3], dims), mket([2, 0, 2, 0], d)) <= TOL
    assert maxabs(qa.applyCTRL(mket([2, 1, 1, 1], d), U, [0, 2], [1, 3], dims, [1, 2]),
                  mket([2, 1, 1, 1], d)) <= TOL
    assert maxabs(qa.applyCTRL(prj(mket([1, 0, 0, 0], d)), U, [0, 2], [1, 3], dims, [1, 2]),
                  prj(mket([1, 2, 0, 2], d))) <= TOL
    assert maxabs(qa.applyCTRL_fan(mket([2, 1, 2, 1], d), Xd(3), [0, 2], [1, 3], dims),
                  mket([2, 0, 2, 0], d)) <= TOL


def test_qudit_ctrl_vs_oracle(qa, orc):
    rng = np.random.default_rng(44)
    dims = [3, 3, 3]
    psi = randket(27, rng)
    rho = randrho(27, rng)
    U = haar(3, rng)
    for ctrl, target, shift in (([0], [2], None), ([2, 0], [1], [1, 2])):
        assert maxabs(qa.applyCTRL(psi, U, ctrl, target, dims, shift),
                      orc.applyCTRL(psi, U, ctrl, target, dims, shift)) <= TOL
        assert maxabs(qa.applyCTRL(rho, U, ctrl, target, dims, shift),
                      orc.applyCTRL(rho, U, ctrl, target, dims, shift)) <= TOL
        assert maxabs(qa.applyCTRL_fan(psi, U, ctrl, [t for t in range(3) if t not in ctrl], dims, shift),
                      orc.applyCTRL_fan(psi, U, ctrl, [t for t in range(3) if t not in ctrl], dims, shift)) <= TOL
    Ks = [haar(3, rng) * np.sqrt(w) for w in (0.7, 0.3)]
    assert maxabs(qa.apply(rho, Ks, [1], dims), orc.apply_kraus(rho, Ks, [1], dims)) <= TOL


# ---------------------------------------------------------------------------
# measurement
# ---------------------------------------------------------------------------
def test_probs_bit_exact(qa, orc):
    rng = np.random.default_rng(9)
    for n in (1, 4, 10, 14):
        psi = randket(2 ** n, rng)
        st = qa.State.from_array(psi, [2] * n)
        assert np.array_equal(st.probs(), orc.probs(psi))
    rho = randrho(64, rng)
    st = qa.State.from_array(rho, [2] * 6)
    assert np.array_equal(st.probs(), orc.probs(rho))


@pytest.mark.parametrize("n", [1, 3, 8, 13, 16])
def test_sample_bit_exact(qa, orc, n):
    rng = np.random.default_rng(1000 + n)
    psi = randket(2 ** n, rng)
    if n >= 8:
        psi[rng.integers(2 ** n, size=2 ** (n - 2))] = 0  # zero runs
        psi /= np.linalg.norm(psi)
    u = list(rng.random(256))
    # uniforms on and next to the table's edges
    cp = orc.discrete_table(orc.probs(psi))
    for k in rng.integers(0, 2 ** n, size=16):
        c = float(cp[k])
        u += [c, np.nextafter(c, 0), np.nextafter(c, 1)]
    u = [min(x, np.nextafter(1.0, 0)) for x in u]
    st = qa.State.from_array(psi, [2] * n)
    target = list(range(n))
    idx, dig, pr = st.sample(target, u)
    ref = [orc.discrete_draw(orc.probs(psi), x) for x in u]
    assert [int(i) for i in idx] == ref
    p = orc.probs(psi)
    assert np.array_equal(pr, p[ref])


def test_sample_exact_table_edges(qa, orc):
    # four equal probabilities: cp = 0.25, 0.5, 0.75, 1 exactly
    psi = np.full(4, 0.5, dtype=complex)
    st = qa.State.from_array(psi, [2, 2])
    u = [1e-300, 0.25, np.nextafter(0.25, 1), 0.5, 0.74, 0.75, 0.999999999]
    idx, dig, pr = st.sample([0, 1], u)
    assert [int(i) for i in idx] == [orc.discrete_draw([0.25] * 4, x) for x in u] == [0, 0, 1, 1, 2, 2, 3]
    assert dig.tolist() == [[0, 0], [0, 0], [0, 1], [0, 1], [1, 0], [1, 0], [1, 1]]


def test_sample_rho(qa, orc):
    rng = np.random.default_rng(5)
    rho = randrho(32, rng)
    st = qa.State.from_array(rho, [2] * 5)
    u = list(rng.random(64))
    idx, dig, pr = st.sample([4, 0], u)
    for i, x in enumerate(u):
        ri, rd, rp = orc.sample(rho, [4, 0], 2, x)
        assert int(idx[i]) == ri and dig[i].tolist() == rd and pr[i] == rp


@pytest.mark.parametrize("destructive", [True, False])
def test_measure_seq(qa, orc, destructive):
    rng = np.random.default_rng(11)
    n = 7
    psi = randket(2 ** n, rng)
    for target in ([0], [6, 2], [1, 5, 3], list(range(n))):
        u = float(rng.random())
        r1, p1, s1 = qa.measure_seq(psi, target, 2, destructive, u)
        r2, p2, s2 = orc.measure_seq(psi, target, 2, destructive, u)
        assert r1 == r2
        assert np.max(np.abs(np.array(p1) - np.array(p2))) <= TOL
        assert maxabs(np.ravel(s1), np.ravel(s2)) <= TOL
    rho = randrho(2 ** 5, rng)
    for target in ([0], [4, 1]):
        u = float(rng.random())
        r1, p1, s1 = qa.measure_seq(rho, target, 2, destructive, u)
        r2, p2, s2 = orc.measure_seq(rho, target, 2, destructive, u)
        assert r1 == r2
        assert np.max(np.abs(np.array(p1) - np.array(p2))) <= TOL
        assert maxabs(s1, s2) <= TOL


def test_measure_seq_qudits_known_answer(qa):
    # unit_tests/tests/instruments.cpp:102-160
    dims = [2, 2, 3, 2, 3, 2]
    mask = [0, 1, 2, 1, 2, 0]
    psi = mket(mask, dims)
    ms, ps, st = qa.measure_seq(psi, [4, 2, 3, 0], dims, True, 0.4)
    assert ms == [2, 2, 1, 0] and ps == [1, 1, 1, 1]
    assert maxabs(st, mket([1, 0], [2, 2])) <= 1e-12
    ms, ps, st = qa.measure_seq(prj(psi), [3, 0, 4, 2], dims, False, 0.4)
    assert ms == [1, 0, 2, 2] and ps == [1, 1, 1, 1] and maxabs(st, prj(psi)) <= 1e-12


def test_normalize(qa):
    rng = np.random.default_rng(3)
    v = randket(256, rng) * 3.5
    st = qa.State.from_array(v, [2] * 8)
    assert abs(st.norm2() - 12.25) < 1e-12
    st.normalize()
    assert maxabs(st.download(), v / np.linalg.norm(v)) <= TOL
    z = qa.State([2, 2])
    z.upload(np.zeros(4, dtype=complex))
    with pytest.raises(qa.OverflowError_):
        z.normalize()


# ---------------------------------------------------------------------------
# complex64 states
# ---------------------------------------------------------------------------
def test_c64_apply(qa, orc):
    rng = np.random.default_rng(13)
    n = 10
    psi = randket(2 ** n, rng)
    st = qa.State.from_array(psi, [2] * n, dtype=qa.C64)
    ref = psi.copy()
    for q in (0, 3, 9):
        U = haar(2, rng)
        st.apply(U, [q])
        ref = orc.apply(ref, U, [q], 2)
    st.apply_ctrl(X, [2], [7])
    ref = orc.applyCTRL(ref, X, [2], [7], 2)
    assert maxabs(st.download(), ref) <= 2e-6


# ---------------------------------------------------------------------------
# validation: qpp's exception classes in qpp's order
# ---------------------------------------------------------------------------
def test_errors(qa):
    psi = kets("000")
    with pytest.raises(qa.ZeroSize):
        qa.apply(psi, X, [], 2)
    with pytest.raises(qa.SubsysMismatchDims):
        qa.apply(psi, X, [3], 2)
    with pytest.raises(qa.SubsysMismatchDims):
        qa.apply(psi, CNOT, [1, 1], 2)
    with pytest.raises(qa.MatrixMismatchSubsys):
        qa.apply(psi, CNOT, [1], 2)
    with pytest.raises(qa.ZeroSize):
        qa.applyCTRL(psi, X, [], [0], 2)
    with pytest.raises(qa.OutOfRange):
        qa.applyCTRL(psi, X, [0], [0], 2)
    with pytest.raises(qa.OutOfRange):
        qa.applyCTRL(psi, X, [0], [1], 2, [2])
    with pytest.raises(qa.DimsNotEqual):
        qa.applyCTRL(mket([0, 0, 0], [2, 3, 2]), X, [0, 1], [2], [2, 3, 2])
    with pytest.raises(qa.MatrixMismatchSubsys):
        qa.applyCTRL_fan(psi, CNOT, [0], [1, 2], 2)
    with pytest.raises(qa.PermInvalid):
        qa.syspermute(psi, [0, 0, 1], 2)
    with pytest.raises(qa.DimsMismatchCvector):
        qa.apply(psi, X, [0], [2, 2])
    try:
        qa.apply(psi, X, [5], 2)
    except qa.SubsysMismatchDims as e:
        assert str(e) == "[WHERE: qpp::apply()] Subsystems mismatch dimensions! [CONTEXT: dims/target]"

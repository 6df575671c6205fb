"""Complex64 path: FP64 comparison kernel vs numpy, c64 circuits vs the c128
oracle (infidelity and max-abs stated per test), FP64 probabilities and
exact sampling on c64 amplitudes, FP64 renormalisation."""
import numpy as np
import pytest

from helpers import X, haar, randket, randrho

import qpp_amd as qa
from qpp_amd.workloads import random_layers

pytestmark = pytest.mark.gpu


def np_compare(a, b):
    a = np.asarray(a, dtype=np.complex128).ravel(order="F")
    b = np.asarray(b, dtype=np.complex128).ravel(order="F")
    ov = np.vdot(a, b)
    return ov, np.vdot(a, a).real, np.vdot(b, b).real, np.max(np.abs(a - b))


@pytest.mark.parametrize("n", [1, 7, 16])
def test_compare_matches_numpy(n):
    rng = np.random.default_rng(n)
    a, b = randket(2 ** n, rng), randket(2 ** n, rng)
    A = qa.State.from_array(a, [2] * n)
    for dt in (qa.C128, qa.C64):
        B = qa.State.from_array(b, [2] * n, dtype=dt)
        r = A.compare(B)
        bb = b.astype(np.complex64) if dt == qa.C64 else b
        ov, na, nb, mx = np_compare(a, bb)
        assert abs(r["overlap"] - ov) < 1e-12 and abs(r["norm2_a"] - na) < 1e-12
        assert abs(r["norm2_b"] - nb) < 1e-12 and abs(r["max_abs"] - mx) < 1e-15
        assert abs(r["infidelity"] - (1 - abs(ov) ** 2 / (na * nb))) < 1e-12


def test_compare_rho_hilbert_schmidt():
    rng = np.random.default_rng(3)
    r1, r2 = randrho(32, rng), randrho(32, rng)
    A = qa.State.from_array(r1, [2] * 5)
    B = qa.State.from_array(r2, [2] * 5)
    r = A.compare(B)
    assert abs(r["overlap"] - np.trace(r1.conj().T @ r2)) < 1e-12
    assert abs(r["hs_distance"] - np.linalg.norm(r1 - r2)) < 1e-12


def test_convert_roundtrip():
    rng = np.random.default_rng(4)
    a = randket(2 ** 10, rng)
    A = qa.State.from_array(a, [2] * 10)
    B = A.astype(qa.C64)
    assert np.array_equal(B.download(), a.astype(np.complex64))
    C = B.astype(qa.C128)
    assert np.array_equal(C.download(), a.astype(np.complex64).astype(np.complex128))


@pytest.mark.parametrize("n,depth", [(12, 20), (18, 10), (22, 6)])
def test_c64_circuit_infidelity(orc, n, depth):
    """c64 tile-fused circuit vs the c128 oracle: infidelity <= 1e-9 and
    max-abs <= 1e-6 at these sizes (FP32 rounding grows ~ sqrt(#gates))."""
    gates = random_layers(n, depth, seed=100 + n)
    qc = qa.QCircuit.from_gates(n, gates)
    ref = np.zeros(2 ** n, dtype=complex)
    ref[0] = 1
    if n <= 18:
        for kind, tg, ctrl, U in gates:
            ref = orc.applyCTRL(ref, U, ctrl, tg, 2) if ctrl else orc.apply(ref, U, tg, 2)
    else:  # c128 device path (itself parity-tested at <= 1e-12) stands in
        st = qa.State([2] * n)
        qa.Program(qc).run(st)
        ref = st.download()
    s32 = qa.State([2] * n, dtype=qa.C64)
    qa.Program(qc, dtype=qa.C64).run(s32)
    R = qa.State.from_array(ref, [2] * n)
    r = R.compare(s32)
    print(f"n={n} depth={depth} infidelity={r['infidelity']:.3e} max_abs={r['max_abs']:.3e}")
    assert r["infidelity"] <= 1e-9
    assert r["max_abs"] <= 1e-6
    assert abs(r["norm2_b"] - 1) <= 1e-5


def test_c64_probabilities_and_sampling_are_fp64(orc):
    """p = re*re + im*im in FP64 on the FP32 amplitudes: identical to the
    oracle on the upcast vector, so sampled indices are bit-exact too."""
    rng = np.random.default_rng(5)
    n = 14
    a = randket(2 ** n, rng).astype(np.complex64)
    S = qa.State.from_array(a, [2] * n, dtype=qa.C64)
    a64 = a.astype(np.complex128)
    assert np.array_equal(S.probs(), orc.probs(a64))
    u = rng.random(257)
    idx, dig, pr = S.sample([0, 5, 13], u)
    for k in range(len(u)):
        i2, d2, p2 = orc.sample(a64, [0, 5, 13], [2] * n, u[k])
        assert idx[k] == i2 and list(dig[k]) == list(d2) and pr[k] == p2


def test_c64_normalize_in_fp64():
    rng = np.random.default_rng(6)
    n = 12
    a = (3.0 * randket(2 ** n, rng)).astype(np.complex64)
    S = qa.State.from_array(a, [2] * n, dtype=qa.C64)
    S.normalize()
    a64 = a.astype(np.complex128)
    ref = (a64 / np.sqrt(np.vdot(a64, a64).real)).astype(np.complex64)
    got = S.download()
    assert np.max(np.abs(got - ref)) <= 2 * np.finfo(np.float32).eps
    assert abs(S.norm2() - 1) < 1e-6


def _dense_cases(n, rng):
    """(target, control) sets for 4- and 5-qubit gates: low, high and mixed
    target bits, with and without controls."""
    return [([0, 1, 2, 3], []), ([n - 1, n - 2, n - 3, n - 4, n - 5], []),
            ([2, n - 1, 5, 0], [n - 3]), ([7, 3, n - 2, 1, 9], [4, n - 1]),
            (list(rng.permutation(n)[:5]), [])]


@pytest.mark.parametrize("valu", [False, True])
def test_c64_dense_gates_mfma_vs_oracle(orc, monkeypatch, valu):
    """4- and 5-qubit dense gates on complex64 kets and density matrices
    (FP32 MFMA kernel; valu: the VALU kernel forced by QPP_AMD_GATEK_VALU)
    against the complex128 oracle: max-abs <= 2e-6 (FP32 sums of up to 32
    terms of a unit column)."""
    if valu:
        monkeypatch.setenv("QPP_AMD_GATEK_VALU", "1")
    rng = np.random.default_rng(41)
    n = 12
    psi = randket(2 ** n, rng)
    for tg, ct in _dense_cases(n, rng):
        U = haar(2 ** len(tg), rng)
        print(f"ket n={n} target={tg} ctrl={ct} valu={valu}", flush=True)
        st = qa.State.from_array(psi, [2] * n, dtype=qa.C64)
        if ct:
            st.apply_ctrl(U, ct, tg)
            ref = orc.applyCTRL(psi, U, ct, tg, [2] * n)
        else:
            st.apply(U, tg)
            ref = orc.apply(psi, U, tg, 2)
        assert np.max(np.abs(st.download() - ref)) <= 2e-6
    nr = 6
    rho = randrho(2 ** nr, rng)
    for tg in ([0, 1, 2, 3], [5, 1, 3, 0, 2]):
        U = haar(2 ** len(tg), rng)
        print(f"rho n={nr} target={tg} valu={valu}", flush=True)
        st = qa.State.from_array(rho, [2] * nr, dtype=qa.C64)
        st.apply(U, tg)
        assert np.max(np.abs(st.download() - orc.apply(rho, U, tg, 2))) <= 2e-6


def test_c64_dense_gate_mfma_equals_valu_at_size(monkeypatch):
    """A 5-qubit dense gate on a 24-qubit complex64 ket: the MFMA and VALU
    kernels agree to FP32 rounding (max-abs <= 1e-6 on a normalised state),
    and unitarity keeps the norm within 1e-6."""
    rng = np.random.default_rng(42)
    n = 24
    U = haar(32, rng)
    a = qa.State([2] * n, dtype=qa.C64)
    a.apply(haar(2, rng), [0])
    for q in range(n):
        a.apply(haar(2, rng), [q])
    b = a.astype(qa.C64)
    a.apply(U, [3, 20, 9, 0, 15])
    monkeypatch.setenv("QPP_AMD_GATEK_VALU", "1")
    b.apply(U, [3, 20, 9, 0, 15])
    r = a.compare(b)
    print(r, flush=True)
    assert r["max_abs"] <= 1e-6 and abs(r["norm2_a"] - 1) <= 1e-6

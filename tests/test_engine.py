"""QEngine / QNoisyEngine: the oracle pinned on the CPU, then the device
engine against the oracle with identical seeds (statistics, classical
dits, probabilities and final state)."""
import numpy as np
import pytest

from helpers import CNOT, CZ, H, X, Z, haar, kets

import qpp_amd as qa


def u3(theta, phi, lam):
    """OpenQASM 2 u3 (qasmtools/qasm/generic qelib1 definition)."""
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return np.array([[c, -np.exp(1j * lam) * s],
                     [np.exp(1j * phi) * s, np.exp(1j * (phi + lam)) * c]])


def teleport_circuit():
    """unit_tests/tests/qasm/circuits/units/teleportation.qasm with the
    classically controlled corrections deferred (controlled X / Z before
    the measurements); the last qubit ends in |0> (qasm.cpp:83-93)."""
    qc = qa.QCircuit(3, 2)
    qc.gate(H, 1).CTRL(X, 1, 2)
    qc.gate(u3(0.3, 0.2, 0.1), 0)
    qc.CTRL(X, 0, 1).gate(H, 0)
    qc.CTRL(X, 1, 2).CTRL(Z, 0, 2)
    qc.gate(u3(-0.3, -0.1, -0.2), 2)
    qc.measure(0, 0).measure(1, 1)
    return qc


def random_circuit(n, depth, seed, measure=True, nc=None):
    rng = np.random.default_rng(seed)
    qc = qa.QCircuit(n, nc if nc is not None else n)
    for _ in range(depth):
        for q in range(n):
            qc.gate(haar(2, rng), q)
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            qc.CTRL(X, int(perm[i]), int(perm[i + 1]))
    if measure:
        qc.measure(list(range(n)), 0)
    return qc


# ---------------------------------------------------------------------------
# oracle (CPU)
# ---------------------------------------------------------------------------
def test_oracle_engine_deterministic(orc):
    qc = qa.QCircuit(3, 3)
    qc.gate(X, 0).CTRL(X, 0, 1)
    qc.measure([0, 1, 2], 0)
    stats, dits, probs, st = orc.engine_run(qc, reps=10)
    assert stats == {(1, 1, 0): 10} and dits == [1, 1, 0] and probs == [1, 1, 1]


def test_oracle_engine_teleport(orc):
    stats, dits, probs, st = orc.engine_run(teleport_circuit(), reps=1, seed=7)
    assert st.shape == (2,)
    assert abs(abs(st[0]) - 1) < 1e-12 and abs(st[1]) < 1e-12


def test_oracle_engine_rng_order(orc):
    """H on q0 then measure q0, reps R (sample-able): rep 0 draws one
    sample and discards it, then measure_seq draws for the recorded
    outcome; reps 1.. draw once each (qengine.hpp:414-473)."""
    qc = qa.QCircuit(1, 1)
    qc.gate(H, 0).measure(0, 0)
    R = 64
    for seed in (1, 2, 3):
        stats, dits, _, _ = orc.engine_run(qc, reps=R, seed=seed)
        r = orc.Rng(seed)
        psi = np.array([1, 1]) / np.sqrt(2)
        p = orc.probs(psi)
        r.canonical()  # discarded sample of rep 0
        outs = [orc.discrete_draw(p, r.canonical()) for _ in range(R)]
        exp = {}
        for o in outs:
            exp[(o,)] = exp.get((o,), 0) + 1
        assert stats == exp
        assert dits == [outs[0]]


def test_oracle_noisy_engine_rng_order(orc):
    """QNoisyEngine: before every step, one draw per live qubit
    (qnoisy_engine.hpp:108-116); p = 0 keeps the state but consumes the
    draws, so the measurement draw is the one after them."""
    qc = qa.QCircuit(2, 2)
    qc.gate(H, 0).measure([0, 1], 0)
    stats, dits, _, _ = orc.engine_run(qc, reps=1, noise_p=0.0, seed=11)
    r = orc.Rng(11)
    for _ in range(2 + 2):  # step 0: 2 live qubits, step 1: 2 live qubits
        r.canonical()
    u = r.canonical()
    p = orc.probs(np.kron(np.array([1, 1]) / np.sqrt(2), [1, 0]))
    idx = orc.discrete_draw(p, u)
    assert dits == [idx >> 1, idx & 1]


# ---------------------------------------------------------------------------
# device engine vs oracle engine
# ---------------------------------------------------------------------------
gpu = pytest.mark.gpu


@gpu
def test_engine_teleport():
    e = qa.QEngine(teleport_circuit(), seed=7).execute(1)
    st = e.get_state().download()
    assert st.shape == (2,)
    assert abs(abs(st[0]) - 1) < 1e-12 and abs(st[1]) < 1e-12


@gpu
@pytest.mark.parametrize("n,depth,reps", [(3, 2, 1), (6, 4, 500), (9, 3, 200), (13, 3, 300), (14, 2, 1)])
def test_engine_vs_oracle(orc, n, depth, reps):
    qc = random_circuit(n, depth, seed=n * 7 + depth)
    e = qa.QEngine(qc, seed=1234 + n).execute(reps)
    stats, dits, probs, st = orc.engine_run(qc, reps=reps, seed=1234 + n)
    assert e.get_stats() == stats
    assert e.get_dits() == dits
    assert np.max(np.abs(np.array(e.get_probs()) - np.array(probs))) <= 1e-12
    got = e.get_state().download()
    assert got.shape == st.shape and np.max(np.abs(got - st)) <= 1e-12


@gpu
def test_engine_midcircuit_measurement(orc):
    """Non-sample-able circuit: gates after a measurement force per-rep
    re-execution (qengine.hpp:316-345); non-destructive and destructive."""
    rng = np.random.default_rng(5)
    qc = qa.QCircuit(5, 5)
    for q in range(5):
        qc.gate(haar(2, rng), q)
    qc.CTRL(X, 0, 3)
    qc.measure(1, 1, destructive=False)
    qc.gate(haar(2, rng), 1).CTRL(X, 1, 4)
    qc.measure(2, 2)
    qc.gate(haar(4, rng), 3, 0)
    qc.measure([0, 3, 4], 0, destructive=False)
    for reps in (1, 40):
        e = qa.QEngine(qc, seed=99).execute(reps)
        stats, dits, probs, st = orc.engine_run(qc, reps=reps, seed=99)
        assert e.get_stats() == stats and e.get_dits() == dits
        assert np.max(np.abs(np.array(e.get_probs()) - np.array(probs))) <= 1e-12
        assert np.max(np.abs(e.get_state().download() - st)) <= 1e-12


@gpu
@pytest.mark.parametrize("n,reps", [(4, 30), (12, 20)])
def test_noisy_engine_vs_oracle(orc, n, reps):
    qc = random_circuit(n, 3, seed=41 + n)
    for p in (0.02, 0.2):
        e = qa.QEngine(qc, noise_p=p, seed=2024).execute(reps)
        stats, dits, probs, st = orc.engine_run(qc, reps=reps, noise_p=p, seed=2024)
        assert e.get_stats() == stats
        assert e.get_dits() == dits
        assert np.max(np.abs(e.get_state().download() - st)) <= 1e-12


@gpu
def test_rho_engine_vs_oracle(orc):
    qc = random_circuit(4, 2, seed=3)
    e = qa.QEngine(qc, kind=qa.RHO, seed=8).execute(50)
    stats, dits, probs, st = orc.engine_run(qc, reps=50, seed=8, is_rho=True)
    assert e.get_stats() == stats and e.get_dits() == dits
    assert np.max(np.abs(e.get_state().download() - st)) <= 1e-12
    qc2 = random_circuit(6, 2, seed=4)  # 12 index bits: fused passes on rho
    e = qa.QEngine(qc2, kind=qa.RHO, seed=9).execute(10)
    stats, dits, probs, st = orc.engine_run(qc2, reps=10, seed=9, is_rho=True)
    assert e.get_stats() == stats and e.get_dits() == dits

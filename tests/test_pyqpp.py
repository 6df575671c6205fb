"""qpp_amd.pyqpp: the pybind11 module with pyqpp's names and signatures
(pyqpp/include/pyqpp/operations_bind.hpp:37-293,
classes/qengine_bind.hpp:52-160), checked against the oracle on the device,
and pyqpp's own teleportation example (pyqpp/examples/teleport.py)."""
import os

import numpy as np
import pytest

from helpers import CNOT, H, X, haar, randket, randrho

import qpp_amd as qa
from qpp_amd import pyqpp

HERE = os.path.dirname(os.path.abspath(__file__))
gpu = pytest.mark.gpu


def col(v):
    return np.asarray(v).reshape(-1, 1)


# ---------------------------------------------------------------------------
# CPU: the module's surface
# ---------------------------------------------------------------------------
def test_module_surface():
    for name in ("apply", "applyCTRL", "applyCTRL_fan", "ptrace", "ptranspose", "syspermute",
                 "sample", "measure_seq", "reset", "QCircuit", "QEngine", "QKetEngine",
                 "QDensityEngine", "QNoisyEngine", "QKetNoisyEngine", "QubitDepolarizingNoise",
                 "gates", "states", "qasm", "seed"):
        assert hasattr(pyqpp, name), name
    assert np.array_equal(pyqpp.gates.CNOT, CNOT) and np.allclose(pyqpp.gates.H, H)
    assert np.array_equal(pyqpp.states.z0, col([1, 0]))


def test_circuits_build_on_cpu():
    qc = pyqpp.QCircuit(3, 2)
    qc.gate(pyqpp.gates.H, 1).CTRL(pyqpp.gates.X, 1, 2).measure([0, 1]).cCTRL(pyqpp.gates.X, 1, 2)
    qc.post_select(0, 1, 0, False).reset([2]).cond_if([0], [1]).gate(X, 2).cond_end().nop()
    assert (qc.get_nq(), qc.get_nc()) == (3, 2)
    q = pyqpp.qasm.read_from_file(os.path.join(HERE, "golden", "qasm", "teleportation.qasm"))
    assert (q.get_nq(), q.get_nc()) == (3, 2)


# ---------------------------------------------------------------------------
# device
# ---------------------------------------------------------------------------
@gpu
def test_operations_vs_oracle(orc):
    rng = np.random.default_rng(1)
    psi, rho = randket(2 ** 6, rng), randrho(2 ** 4, rng)
    U2 = haar(4, rng)
    dims = [2] * 6
    assert np.max(np.abs(pyqpp.apply(psi, U2, [4, 1], dims) - col(orc.apply(psi, U2, [4, 1], dims)))) <= 1e-12
    assert np.max(np.abs(pyqpp.apply(psi, U2, [4, 1]) - col(orc.apply(psi, U2, [4, 1], dims)))) <= 1e-12
    assert np.max(np.abs(pyqpp.apply(rho, U2, [0, 3], d=2) - orc.apply(rho, U2, [0, 3], [2] * 4))) <= 1e-12
    g = 0.3
    Ks = [np.diag([1, np.sqrt(1 - g)]), np.array([[0, np.sqrt(g)], [0, 0]])]
    assert np.max(np.abs(pyqpp.apply(rho, Ks, [2], [2] * 4) - orc.apply_kraus(rho, Ks, [2], [2] * 4))) <= 1e-12
    # apply(A, Ks) on the whole system, rectangular operators: <i| (x) I traces
    # out the first qubit
    Kt = [np.kron(np.eye(2)[i:i + 1, :], np.eye(8)) for i in range(2)]
    assert np.max(np.abs(pyqpp.apply(rho, Kt) - orc.apply_kraus_full(rho, Kt))) <= 1e-12
    U1 = haar(2, rng)
    got = pyqpp.applyCTRL(psi, U1, [0, 5], [2], dims, shift=[1, 0])
    assert np.max(np.abs(got - col(orc.applyCTRL(psi, U1, [0, 5], [2], dims, [1, 0])))) <= 1e-12
    got = pyqpp.applyCTRL_fan(psi, U1, [0], [1, 3])
    assert np.max(np.abs(got - col(orc.applyCTRL_fan(psi, U1, [0], [1, 3], dims)))) <= 1e-12
    assert np.max(np.abs(pyqpp.ptrace(rho, [1, 2]) - orc.ptrace(rho, [1, 2], [2] * 4))) <= 1e-12
    # qubits: qpp's qubit kernels (as built); qudits: the generic code
    assert np.array_equal(pyqpp.ptranspose(rho, [1, 2]), orc.ptranspose(rho, [1, 2], [2] * 4, qa.AS_BUILT))
    assert np.array_equal(pyqpp.syspermute(rho, [2, 0, 3, 1]),
                          orc.syspermute(rho, [2, 0, 3, 1], [2] * 4, qa.AS_BUILT))
    r9 = randrho(9, rng)
    assert np.array_equal(pyqpp.ptranspose(r9, [0], [3, 3]), orc.ptranspose(r9, [0], [3, 3], qa.DOCUMENTED))


@gpu
def test_instruments_follow_the_module_generator(orc):
    rng = np.random.default_rng(2)
    psi = randket(2 ** 5, rng)
    pyqpp.seed(77)
    r = orc.Rng(77)
    for _ in range(3):
        res, p = pyqpp.sample(psi, [3, 0])
        _, dig, pr = orc.sample(psi, [3, 0], [2] * 5, r.canonical())
        assert res == dig and abs(p - pr) <= 1e-12
    res, probs, st = pyqpp.measure_seq(psi, [1, 4])
    eres, eprobs, est = orc.measure_seq(psi, [1, 4], [2] * 5, True, r.canonical())
    assert res == eres and np.max(np.abs(np.array(probs) - eprobs)) <= 1e-12
    assert np.max(np.abs(st - col(est))) <= 1e-12
    got = pyqpp.reset(psi, [2, 3])
    assert np.max(np.abs(got - col(orc.reset(psi, [2, 3], [2] * 5, r.canonical())))) <= 1e-12


@gpu
def test_teleport_example():
    """pyqpp/examples/teleport.py: the output state equals the input."""
    gt, st = pyqpp.gates, pyqpp.states
    pyqpp.seed(3)
    for _ in range(5):
        U = haar(2, np.random.default_rng())
        qc = pyqpp.QCircuit(3, 2)
        qc.gate(U, 0).gate(gt.H, 1).CTRL(gt.X, 1, 2).CTRL(gt.X, 0, 1).gate(gt.H, 0)
        qc.measure([0, 1]).cCTRL(gt.X, 1, 2).cCTRL(gt.Z, 0, 2)
        qe = pyqpp.QEngine(qc)
        qe.execute()
        assert np.linalg.norm(qe.get_state() - U @ st.z0) < 1e-12


@gpu
def test_engines_vs_oracle(orc):
    rng = np.random.default_rng(4)
    Us = [haar(2, rng) for _ in range(6)]

    def build(mod):
        qc = mod.QCircuit(4, 4)
        for q in range(4):
            qc.gate(Us[q], q)
        qc.CTRL(X, 0, 2).CTRL(X, 1, 3).gate(Us[4], 2)
        qc.measure(1, 1, False).cCTRL(Us[5], [1], 3)
        qc.measure([0, 2, 3], 0)
        return qc

    for seed, reps in ((5, 1), (6, 50)):
        e = pyqpp.QEngine(build(pyqpp), seed=seed).execute(reps)
        stats, dits, probs, st = orc.engine_run(build(qa), reps=reps, seed=seed)
        assert e.get_stats() == stats and e.get_dits() == dits
        assert e.get_dit(1) == dits[1] and np.max(np.abs(np.array(e.get_probs()) - probs)) <= 1e-12
        assert np.max(np.abs(e.get_state() - col(st))) <= 1e-12
        e = pyqpp.QDensityEngine(build(pyqpp), seed=seed).execute(reps)
        stats, dits, _, st = orc.engine_run(build(qa), reps=reps, seed=seed, is_rho=True)
        assert e.get_stats() == stats and np.max(np.abs(e.get_state() - st)) <= 1e-12
        e = pyqpp.QNoisyEngine(build(pyqpp), pyqpp.QubitDepolarizingNoise(0.1), seed=seed).execute(reps)
        stats, dits, _, st = orc.engine_run(build(qa), reps=reps, seed=seed, noise_p=0.1)
        assert e.get_stats() == stats and np.max(np.abs(e.get_state() - col(st))) <= 1e-12
        assert e.get_noise_results() == orc.last_noise_results(len(build(qa).steps))
    # post-selection controls (a rep fails after 4 missed draws) and set_state
    def bell_ps(mod):
        qc = mod.QCircuit(2, 2)
        return qc.gate(H, 0).CTRL(X, 0, 1).post_select(0, 1, 0).measure(1, 1)

    e = pyqpp.QEngine(bell_ps(pyqpp), seed=9).set_ensure_post_selection(True)
    e.set_max_post_selection_reps(4)
    assert e.get_ensure_post_selection() and e.get_max_post_selection_reps() == 4
    e.execute(30)
    stats, dits, _, _, ok = orc.engine_run(bell_ps(qa), reps=30, seed=9, ensure_post_selection=True,
                                           max_post_selection_reps=4, with_ps_ok=True)
    assert e.post_select_ok() == ok and e.get_stats() == stats and set(stats) == {(1, 1)}
    e.reset().set_state(np.kron([0, 1], [1, 0]))
    assert np.max(np.abs(e.get_state() - col([0, 0, 1, 0]))) == 0


@gpu
def test_qasm_through_the_module(orc):
    path = os.path.join(HERE, "golden", "qasm", "qec.qasm")
    e = pyqpp.QEngine(pyqpp.qasm.read_from_file(path), seed=21).execute(40)
    stats, dits, _, st = orc.engine_run(qa.qasm.read_from_file(path), reps=40, seed=21)
    assert e.get_stats() == stats and e.get_dits() == dits
    assert np.max(np.abs(e.get_state() - col(st))) <= 1e-12


@gpu
def test_errors_are_qpp_texts():
    psi = randket(8, np.random.default_rng(5))
    with pytest.raises(RuntimeError, match=r"qpp::apply\(\)"):
        pyqpp.apply(psi, np.eye(4), [0, 0])
    with pytest.raises(RuntimeError, match="Invalid conditional"):
        pyqpp.QEngine(pyqpp.QCircuit(1, 1).cond_if([0], [1]))

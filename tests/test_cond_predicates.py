"""Runtime conditions given as predicates over classical dits (qpp's
cond_if(pred) / cond_while(pred), qcircuit.hpp:968-1026, evaluated in
visit_runtime_step_, qengine.hpp:837-880): QCircuit.cond_if(dits, pred)
stores the predicate as a truth table over the dits' values (the
COND_IF_TABLE / COND_WHILE_TABLE steps).

CPU: the oracle engine with a predicate equals the equality form where they
agree, and a parity predicate and a qutrit sum predicate give the outcomes
the circuits force.  GPU: the device engine against the oracle on the same
seeds (a run starts from zero dits and measurements write values in
0..d-1, so the tables cover every value a condition can read)."""
import numpy as np
import pytest

import qpp_amd as qa
from helpers import H, X

gpu = pytest.mark.gpu


def ry(t):
    return np.array([[np.cos(t / 2), -np.sin(t / 2)], [np.sin(t / 2), np.cos(t / 2)]])


def until_one(pred):
    """while c0 == 0: H q0; measure q0 -> c0 (non-destructive); Ry q1."""
    qc = qa.QCircuit(2, 1)
    cond = (lambda v: v[0] == 0) if pred else [0]
    qc.cond_while([0], cond).gate(H, 0).measure(0, 0, destructive=False).gate(ry(0.37), 1)
    return qc.cond_end()


def parity():
    """c2 = c0 XOR c1 by an X on q2 under the predicate "c0 + c1 odd"."""
    qc = qa.QCircuit(3, 3)
    qc.gate(H, 0).gate(H, 1).measure(0, 0, destructive=False).measure(1, 1, destructive=False)
    qc.cond_if([0, 1], lambda v: (v[0] ^ v[1]) == 1).gate(X, 2).cond_else().gate(np.eye(2), 2).cond_end()
    return qc.measure(2, 2, destructive=False)


def qutrit_sum():
    """Two qutrits in uniform superposition, measured; a third gets X3^1
    when (c0 + c1) mod 3 == 2 (an "if" with no "else")."""
    F = np.array([[np.exp(2j * np.pi * r * c / 3) for c in range(3)] for r in range(3)]) / np.sqrt(3)
    X3 = np.roll(np.eye(3), 1, axis=0)
    qc = qa.QCircuit(3, 3, 3)
    qc.gate(F, 0).gate(F, 1).measure(0, 0, destructive=False).measure(1, 1, destructive=False)
    qc.cond_if([0, 1], lambda v: (v[0] + v[1]) % 3 == 2).gate(X3, 2).cond_end()
    return qc.measure(2, 2, destructive=False)


def test_predicate_table_layout():
    qc = qa.QCircuit(1, 2, 3).cond_if([1, 0], lambda v: v == [2, 1]).cond_end()
    table = qc.ps_vals[0]
    assert len(table) == 9 and sum(table) == 1 and table[2 * 3 + 1] == 1
    with pytest.raises(ValueError):
        qa.QCircuit(1, 25).cond_if(list(range(21)), lambda v: True)


def test_oracle_predicate_while_equals_equality_form(orc):
    for seed in range(8):
        a = orc.engine_run(until_one(True), reps=1, seed=seed)
        b = orc.engine_run(until_one(False), reps=1, seed=seed)
        assert a[0] == b[0] and a[1] == b[1] == [1]
        assert np.max(np.abs(a[3] - b[3])) == 0


def test_oracle_parity_and_qutrit_predicates(orc):
    stats, _, _, _ = orc.engine_run(parity(), reps=400, seed=7)
    assert sum(stats.values()) == 400 and len(stats) == 4
    assert all(c2 == c0 ^ c1 for c0, c1, c2 in stats)
    stats, _, _, _ = orc.engine_run(qutrit_sum(), reps=400, seed=8)
    assert sum(stats.values()) == 400 and len(stats) == 9
    assert all(c2 == (1 if (c0 + c1) % 3 == 2 else 0) for c0, c1, c2 in stats)


@gpu
@pytest.mark.parametrize("make", [parity, qutrit_sum, lambda: until_one(True)])
def test_device_engine_predicates_vs_oracle(orc, make):
    qc = make()
    for seed in (3, 11):
        ostats, odits, _, _ = orc.engine_run(qc, reps=64, seed=seed)
        e = qa.QEngine(qc, seed=seed)
        e.execute(64)
        print(f"predicate circuit {qc.nq}q d={qc.d} seed={seed}: {len(ostats)} outcomes", flush=True)
        assert e.get_stats() == ostats and e.get_dits() == odits


@gpu
def test_pyqpp_predicate_form_vs_oracle(orc):
    """pyqpp's cond_if(pred): pred over the list of every classical dit; the
    pyqpp engine against the oracle running the two-dit table form."""
    from qpp_amd import pyqpp
    qc = pyqpp.QCircuit(3, 3)
    qc.gate(H, 0).gate(H, 1).measure(0, 0, False).measure(1, 1, False)
    qc.cond_if(lambda dits: (dits[0] ^ dits[1]) == 1).gate(X, 2).cond_else().gate(np.eye(2), 2).cond_end()
    qc.measure(2, 2, False)
    for seed in (4, 9):
        e = pyqpp.QEngine(qc, seed=seed).execute(50)
        stats, dits, _, _ = orc.engine_run(parity(), reps=50, seed=seed)
        assert e.get_stats() == stats and e.get_dits() == dits
        assert all(c2 == c0 ^ c1 for c0, c1, c2 in stats)

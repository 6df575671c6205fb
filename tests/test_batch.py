"""Batched noisy trajectories (qs_batch_*).

CPU: the host plan (Pauli frame absorbed into 1-qubit gates, conjugated
through Clifford gates) reproduces an explicit simulation of the same
trajectory with the errors applied where QNoisyEngine applies them, drawn
from the same std::mt19937 stream (oracle Rng).  GPU: every trajectory's
outcome equals the oracle QNoisyEngine's for the same seed, in the
persistent launch and on the wide path (QPP_AMD_TRAJ_WIDE_BITS sets the
qubit count from which trajectories run pass by pass; default 18)."""
import numpy as np
import pytest

from helpers import CNOT, CZ, H, I2, X, Y, Z, haar, t_apply

import qpp_amd as qa

PAULI = [I2, X, Y, Z]


def noisy_circuit(n, depth, seed, cz=False, extra=False, measure=True):
    """Haar 1-qubit gates on every qubit, then a CNOT ladder (i, i+1)
    (optionally CZ / X / Z / H Clifford gates too), ending in one
    measurement of every qubit (unless measure is False)."""
    rng = np.random.default_rng(seed)
    qc = qa.QCircuit(n, n)
    for _ in range(depth):
        for q in range(n):
            qc.gate(haar(2, rng), q)
        for i in range(n - 1):
            qc.CTRL(Z if cz and i % 3 == 2 else X, i, i + 1)
        if extra:
            qc.gate(X, 0).gate(Z, n - 1).gate(H, n // 2)
            qc.CTRL(X, n - 1, 0, shift=[1])
            qc.gate(CZ, 1, 3)
    if measure:
        qc.measure(list(range(n)), 0)
    return qc


def explicit_trajectory(orc, qc, seed, p):
    """Statevector of one QNoisyEngine trajectory with explicit Pauli errors
    before every step (qnoisy_engine.hpp:104-121) and the measurement
    uniform; numpy tensor contractions, no frame algebra."""
    n = qc.nq
    psi = np.zeros(2 ** n, dtype=complex)
    psi[0] = 1
    r = orc.Rng(seed)
    w = [1 - p, p / 3, p / 3, p / 3]
    for tp, g, ctrl, tg, shift, creg, dest in qc.steps:
        for q in range(n):
            k = orc.discrete_draw(w, r.canonical())
            if k:
                psi = t_apply(psi, PAULI[k], [q], [2] * n)
        if tp == qa.STEP_MEASURE_MANY:
            return psi, r.canonical()
        U = qc.mats[g]
        if tp == qa.STEP_GATE:
            psi = t_apply(psi, U, tg, [2] * n)
        elif tp == qa.STEP_CTRL:
            psi = orc.applyCTRL(psi, U, ctrl, tg, [2] * n, shift=shift)
    raise AssertionError("no measurement")


def framed_trajectory(orc, qc, plan):
    """The same trajectory from the plan: overridden 1-qubit matrices, then
    the final frame i^phase X^x Z^z."""
    n = qc.nq
    ovr = dict(zip(plan["steps"], plan["mats"]))
    psi = np.zeros(2 ** n, dtype=complex)
    psi[0] = 1
    for s, (tp, g, ctrl, tg, shift, creg, dest) in enumerate(qc.steps):
        if tp == qa.STEP_MEASURE_MANY:
            break
        U = ovr.get(s, qc.mats[g])
        if tp == qa.STEP_GATE:
            psi = t_apply(psi, U, tg, [2] * n)
        elif tp == qa.STEP_CTRL:
            psi = orc.applyCTRL(psi, U, ctrl, tg, [2] * n, shift=shift)
    idx = np.arange(2 ** n)
    zsign = np.where(np.array([bin(int(i) & plan["zmask"]).count("1") % 2 for i in idx]) == 1, -1, 1)
    psi = psi * zsign
    psi = psi[idx ^ plan["xmask"]]
    return psi * (1j ** plan["phase"])


@pytest.mark.parametrize("extra", [False, True])
def test_plan_matches_explicit_trajectory(orc, extra):
    n = 12
    qc = noisy_circuit(n, 2, seed=5, cz=True, extra=extra)
    b = qa.Batch([qc], dtype=qa.C128)
    nerr = 0
    for seed in range(6):
        p = 0.06
        plan = b.plan(0, seed, p)
        ref, u = explicit_trajectory(orc, qc, seed, p)
        got = framed_trajectory(orc, qc, plan)
        assert plan["u"] == u
        assert np.max(np.abs(got - ref)) < 1e-12
        nerr += len(plan["steps"])
    assert nerr > 10  # the noise actually fired


def test_plan_noiseless_draws_only_the_measurement(orc):
    qc = noisy_circuit(12, 1, seed=1)
    b = qa.Batch([qc])
    plan = b.plan(0, 77, -1.0)
    assert plan["steps"] == [] and plan["xmask"] == 0 and plan["phase"] == 0
    assert plan["u"] == orc.Rng(77).canonical()


def test_batch_rejects_unsupported_circuits():
    rng = np.random.default_rng(0)
    qc = qa.QCircuit(12, 12)
    qc.gate(haar(4, rng), 0, 1)  # dense non-Clifford 2-qubit gate
    qc.measure(list(range(12)), 0)
    with pytest.raises(qa.NotImplemented_):
        qa.Batch([qc])
    qc2 = qa.QCircuit(12, 12)
    qc2.gate(H, 0).measure([0, 1], 0)  # not every qubit measured
    with pytest.raises(qa.NotImplemented_):
        qa.Batch([qc2])
    small = qa.QCircuit(6, 6)
    small.gate(H, 0).measure(list(range(6)), 0)
    with pytest.raises(qa.NotImplemented_):
        qa.Batch([small])


# ---------------------------------------------------------------------------
# device
# ---------------------------------------------------------------------------
gpu = pytest.mark.gpu


def oracle_outcome(orc, qc, seed, p):
    stats, dits, probs, st = orc.engine_run(qc, reps=1, noise_p=p, seed=seed)
    return tuple(dits)


@gpu
@pytest.mark.parametrize("p", [0.01, 1e-3, -1.0])
def test_batch_vs_oracle_engine_mixed_sizes(orc, p):
    circs = [noisy_circuit(12, 3, seed=11, cz=True), noisy_circuit(13, 2, seed=12, extra=True)]
    b = qa.Batch(circs, dtype=qa.C128)
    rng = np.random.default_rng(3)
    ntr = 40
    cof = np.array([i % 2 for i in range(ntr)], dtype=np.uint32)
    seeds = rng.integers(0, 2 ** 32, size=ntr, dtype=np.uint64)
    idx, prob = b.run(cof, seeds, p)
    print(b.info(0), b.info(1))
    for i in range(ntr):
        c = int(cof[i])
        got = b.dits(c, idx[i])
        assert got == oracle_outcome(orc, circs[c], int(seeds[i]), p), (i, c, int(seeds[i]))
        assert 0 < prob[i] <= 1


@gpu
@pytest.mark.parametrize("dt", [qa.C128, qa.C64])
@pytest.mark.parametrize("reg_bits", [3, 4])
def test_wide_trajectories_equal_single_launch(monkeypatch, dt, reg_bits):
    """Trajectories at or above the wide threshold run pass by pass over
    all their tiles (many workgroups per state) and are then measured in
    the persistent launch: outcomes and probabilities are bit-identical to
    running every trajectory inside the persistent launch, with wide and
    narrow circuits mixed in one batch."""
    circs = [noisy_circuit(12, 3, seed=71), noisy_circuit(13, 3, seed=72, extra=True),
             noisy_circuit(14, 2, seed=73, cz=True)]
    rng = np.random.default_rng(7)
    ntr = 60
    cof = rng.integers(0, 3, size=ntr).astype(np.uint32)
    seeds = rng.integers(0, 2 ** 32, size=ntr, dtype=np.uint64)
    monkeypatch.setenv("QPP_AMD_TRAJ_WIDE_BITS", "64")
    i0, p0 = qa.Batch(circs, dtype=dt, reg_bits=reg_bits).run(cof, seeds, 0.01)
    monkeypatch.setenv("QPP_AMD_TRAJ_WIDE_BITS", "13")
    b = qa.Batch(circs, dtype=dt, reg_bits=reg_bits)
    i1, p1 = b.run(cof, seeds, 0.01)
    assert np.array_equal(i0, i1) and np.array_equal(p0, p1)
    # the same batch object again (buffers reused), all of it wide
    monkeypatch.setenv("QPP_AMD_TRAJ_WIDE_BITS", "12")
    i2, p2 = b.run(cof, seeds, 0.01)
    assert np.array_equal(i0, i2) and np.array_equal(p0, p2)


@gpu
def test_wide_trajectories_more_than_one_launch_row(monkeypatch):
    """More wide trajectories of one circuit than one launch's second grid
    dimension holds (65535): the passes go out in row chunks; every
    trajectory still equals its single-launch result."""
    monkeypatch.setenv("QPP_AMD_TRAJ_WIDE_BITS", "64")
    circs = [noisy_circuit(12, 2, seed=95)]
    ntr = 70000
    cof = np.zeros(ntr, dtype=np.uint32)
    seeds = np.arange(ntr, dtype=np.uint64) + 12345
    i0, p0 = qa.Batch(circs, dtype=qa.C64).run(cof, seeds, 0.01)
    monkeypatch.setenv("QPP_AMD_TRAJ_WIDE_BITS", "12")
    i1, p1 = qa.Batch(circs, dtype=qa.C64).run(cof, seeds, 0.01)
    assert np.array_equal(i0, i1) and np.array_equal(p0, p1)


@gpu
def test_wide_trajectories_vs_oracle_engine(orc, monkeypatch):
    monkeypatch.setenv("QPP_AMD_TRAJ_WIDE_BITS", "12")
    circs = [noisy_circuit(12, 3, seed=81, cz=True), noisy_circuit(13, 2, seed=82, extra=True)]
    b = qa.Batch(circs, dtype=qa.C128)
    cof = np.array([i % 2 for i in range(24)], dtype=np.uint32)
    seeds = np.arange(900, 924, dtype=np.uint64)
    idx, prob = b.run(cof, seeds, 0.01)
    for i in range(24):
        c = int(cof[i])
        assert b.dits(c, idx[i]) == oracle_outcome(orc, circs[c], int(seeds[i]), 0.01), i
        assert 0 < prob[i] <= 1


@gpu
def test_batch_c64_agrees_with_c128():
    circs = [noisy_circuit(12, 4, seed=21), noisy_circuit(16, 3, seed=22), noisy_circuit(20, 2, seed=23)]
    rng = np.random.default_rng(4)
    ntr = 300
    cof = rng.integers(0, 3, size=ntr).astype(np.uint32)
    seeds = rng.integers(0, 2 ** 32, size=ntr, dtype=np.uint64)
    i64, _ = qa.Batch(circs, dtype=qa.C128).run(cof, seeds, 1e-3)
    b32 = qa.Batch(circs, dtype=qa.C64)
    i32, _ = b32.run(cof, seeds, 1e-3)
    same = np.mean(i64 == i32)
    print("c64 == c128 outcomes:", same, b32.info(2))
    assert same >= 0.97


@gpu
def test_batch_rerun_is_deterministic_and_order_free():
    circs = [noisy_circuit(12, 2, seed=31), noisy_circuit(14, 2, seed=32)]
    b = qa.Batch(circs, dtype=qa.C128)
    cof = np.array([0, 1, 1, 0, 1, 0, 0, 1], dtype=np.uint32)
    seeds = np.arange(100, 108, dtype=np.uint64)
    i1, p1 = b.run(cof, seeds, 0.02)
    i2, p2 = b.run(cof[::-1].copy(), seeds[::-1].copy(), 0.02)
    assert np.array_equal(i1, i2[::-1]) and np.array_equal(p1, p2[::-1])
    # one trajectory alone gives the same outcome as inside the batch
    for k in (0, 5):
        ia, pa = b.run(cof[k:k + 1], seeds[k:k + 1], 0.02)
        assert ia[0] == i1[k] and pa[0] == p1[k]


@gpu
@pytest.mark.parametrize("wide", [False, True])
def test_batch_exact_replay_path_vs_oracle(orc, monkeypatch, wide):
    """Every trajectory's outcome decided by the in-kernel exact sequential
    replay (forced) still equals the oracle's, also on states the wide
    path's pass launches left."""
    monkeypatch.setenv("QPP_AMD_FORCE_EXACT", "1")
    monkeypatch.setenv("QPP_AMD_TRAJ_WIDE_BITS", "12" if wide else "64")
    circs = [noisy_circuit(12, 2, seed=61), noisy_circuit(13, 2, seed=62)]
    cof = np.array([i % 2 for i in range(16)], dtype=np.uint32)
    seeds = np.arange(500, 516, dtype=np.uint64)
    b = qa.Batch(circs, dtype=qa.C128)
    idx, _ = b.run(cof, seeds, 0.01)
    assert b.info(0)["exact_replays"] == 16
    for i in range(16):
        c = int(cof[i])
        assert b.dits(c, idx[i]) == oracle_outcome(orc, circs[c], int(seeds[i]), 0.01)
    b64 = qa.Batch(circs, dtype=qa.C64)
    b64.run(cof, seeds, 0.01)
    assert b64.info(0)["exact_replays"] == 16
    # the bounded parallel search decides the same outcomes
    monkeypatch.delenv("QPP_AMD_FORCE_EXACT")
    i2, _ = qa.Batch(circs, dtype=qa.C128).run(cof, seeds, 0.01)
    assert np.array_equal(i2, idx)


@gpu
@pytest.mark.parametrize("dt", [qa.C128, qa.C64])
def test_run_states_vs_oracle(orc, dt):
    """qs_batch_run_states: every entry's final state (the gate steps only,
    no measurement) equals the oracle engine's state for the same circuit
    without its measurement; entries of one circuit are bit-identical."""
    specs = [(12, 3, 71, dict(cz=True)), (13, 2, 72, dict(extra=True)), (14, 2, 73, {})]
    circs = [noisy_circuit(n, d, s, **kw) for n, d, s, kw in specs]
    bare = [noisy_circuit(n, d, s, measure=False, **kw) for n, d, s, kw in specs]
    b = qa.Batch(circs, dtype=dt)
    cof = [0, 1, 2, 1, 0]
    states = b.run_states(cof)
    tol = 1e-12 if dt == qa.C128 else 2e-6
    for k, st in zip(cof, states):
        print(f"run_states: circuit {k} ({circs[k].nq} qubits), dtype {dt}", flush=True)
        _, _, _, ref = orc.engine_run(bare[k], reps=1)
        got = st.download()
        assert np.max(np.abs(got - ref)) <= tol
    assert np.array_equal(states[1].download(), states[3].download())
    assert np.array_equal(states[0].download(), states[4].download())


@gpu
@pytest.mark.parametrize("dt", [qa.C128, qa.C64])
def test_run_states_wide_equal_single_launch(monkeypatch, dt):
    """run_states with the wide path (passes of the 13- and 14-qubit entries
    run on the callers' states ahead of the persistent launch) leaves
    bit-identical states to running every entry in the persistent launch."""
    circs = [noisy_circuit(12, 3, 91, cz=True), noisy_circuit(13, 2, 92, extra=True),
             noisy_circuit(14, 2, 93)]
    cof = [0, 1, 2, 1, 2, 0]
    monkeypatch.setenv("QPP_AMD_TRAJ_WIDE_BITS", "64")
    ref = [st.download() for st in qa.Batch(circs, dtype=dt).run_states(cof)]
    monkeypatch.setenv("QPP_AMD_TRAJ_WIDE_BITS", "13")
    got = [st.download() for st in qa.Batch(circs, dtype=dt).run_states(cof)]
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


@gpu
def test_run_states_errors():
    circs = [noisy_circuit(12, 1, 81)]
    b = qa.Batch(circs, dtype=qa.C128)
    L = qa.lib()
    import ctypes
    L.qs_batch_run_states.argtypes = [ctypes.c_void_p, ctypes.c_uint64,
                                      ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_void_p)]
    cof = np.zeros(1, dtype=np.uint32)

    def run(states, c=cof):
        hs = (ctypes.c_void_p * len(states))(*[st.handle.value for st in states])
        return L.qs_batch_run_states(b._h, len(states), c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), hs)

    with pytest.raises(qa.TypeMismatch):
        qa._check(run([qa.State([2] * 12, dtype=qa.C64)]))
    with pytest.raises(qa.DimsMismatchCvector):
        qa._check(run([qa.State([2] * 11)]))
    st = qa.State([2] * 12)
    with pytest.raises(qa.OutOfRange):
        qa._check(run([st, st], np.zeros(2, dtype=np.uint32)))
    with pytest.raises(qa.OutOfRange):
        qa._check(run([qa.State([2] * 12)], np.ones(1, dtype=np.uint32)))

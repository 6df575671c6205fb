"""Parity of the benchmark workloads where they run at scale:

* batched noisy trajectories of the trajectory family (depth 30, CNOT
  ladder, depolarizing p = 1e-3 before every step): 12- and 16-qubit
  trajectories against the oracle's QNoisyEngine, 20- and 22-qubit ones
  against the single-state device engine, same seeds, c128 and c64;
* the first two layers of the 12-qubit density-matrix pipeline (84 gates and
  channels) through the fused passes, against the oracle op by op;
* 30 qubits (16 GiB): a GHZ ladder gives exactly two amplitudes equal to H's
  entry, and two Haar layers on top give amplitudes with a closed form
  (1/sqrt2)(prod_q V_q[b_q, 0] + prod_q V_q[b_q, 1]) -- checked on 4096
  sampled indices, so that a wrong qubit-to-bit map above 2^29 shows;
* served requests of 20-24 qubits with 1024 shots against the same QEngine
  run alone (and the oracle engine for the 20-qubit complex128 one)."""
import numpy as np
import pytest

import qpp_amd as qa
from helpers import H, X, haar
from qpp_amd import workloads

pytestmark = pytest.mark.gpu

P_DEP = 1e-3


def traj_circuits():
    return {n: workloads.trajectory_circuit(qa, n, 30, 300 + n) for n in (12, 16, 20, 22)}


@pytest.mark.parametrize("dt", [qa.C128, qa.C64])
def test_trajectories_depth30_vs_oracle_and_engine(orc, dt):
    circs = traj_circuits()
    sizes = [12, 16, 20, 22]
    counts = {12: 8, 16: 3, 20: 3, 22: 3}
    cof = np.concatenate([np.full(counts[n], i, dtype=np.uint32) for i, n in enumerate(sizes)])
    seeds = np.arange(len(cof), dtype=np.uint64) * 7919 + 123
    b = qa.Batch([circs[n] for n in sizes], dtype=dt)
    idx, prob = b.run(cof, seeds, P_DEP)
    print("trajectory batch:", b.info(0), flush=True)
    for j in range(len(cof)):
        c = int(cof[j])
        n = sizes[c]
        got = b.dits(c, idx[j])
        if n <= 16:
            ref = tuple(orc.engine_run(circs[n], 1, P_DEP, int(seeds[j]))[1])
        else:
            e = qa.QEngine(circs[n], dtype=dt, noise_p=P_DEP, seed=int(seeds[j]))
            e.execute(1)
            ref = tuple(e.get_dits())
        print(f"trajectory {j}: {n} qubits, seed {int(seeds[j])}: {got == ref}", flush=True)
        assert got == ref, (j, n, int(seeds[j]))
        assert 0 < prob[j] <= 1


def test_rho12_first_two_layers_vs_oracle(orc):
    n = 12
    rho0, ops, _, _, _ = workloads.dm_pipeline(n, 10, seed=1)
    per_layer = n * 2 + (n // 2) * 3
    ops = ops[: 2 * per_layer]
    assert len(ops) == 84
    st = qa.State.from_array(rho0, [2] * n)
    workloads.dm_program(qa, n, ops).compile().run(st)
    got = st.download()
    ref = rho0
    dims = [2] * n
    for kind, q, M in ops:
        if kind == "u":
            ref = orc.apply(ref, M, q, dims)
        elif kind == "cz":
            ref = orc.applyCTRL(ref, M, [q[0]], [q[1]], dims)
        else:
            ref = orc.apply_kraus(ref, M, q, dims)
    err = float(np.max(np.abs(got - ref)))
    print(f"12-qubit rho, 84 ops fused vs oracle: max-abs {err:.3e}", flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("native", [False, True])
def test_thirty_qubit_closed_forms(native):
    """The general tile kernel and the specialised pass kernels (the
    benchmark's path) against closed forms at 30 qubits, where a wrong
    qubit-to-bit or tile-base mapping shows."""
    n = 30
    D = 2 ** n
    st = qa.State([2] * n)
    st.set_basis(0)

    def program(gates):
        p = qa.Program(n, gates, reg_bits=4, low_bits=3, max_ops=48)
        return p.compile_native() if native else p

    ghz = [("u", [0], [], H)] + [("cx", [q + 1], [q], X) for q in range(n - 1)]
    program(ghz).run(st)
    a = st.download()
    nz = np.flatnonzero(a)
    assert list(nz) == [0, D - 1], nz[:8]
    assert a[0] == H[0, 0] and a[D - 1] == H[1, 0]
    del a
    # two Haar layers on every qubit: V_q = U2_q U1_q
    rng = np.random.default_rng(3030)
    U1 = [haar(2, rng) for _ in range(n)]
    U2 = [haar(2, rng) for _ in range(n)]
    layers = [("u", [q], [], U1[q]) for q in range(n)] + [("u", [q], [], U2[q]) for q in range(n)]
    program(layers).run(st)
    assert abs(st.norm2() - 1) < 1e-12
    a = st.download()
    V = np.array([U2[q] @ U1[q] for q in range(n)])  # (n, 2, 2)
    idx = np.concatenate([rng.integers(0, D, size=4090, dtype=np.int64),
                          [0, D - 1, 1, 1 << (n - 1), 0x15555555, 0x2AAAAAAA]])
    bits = (idx[:, None] >> (n - 1 - np.arange(n))[None, :]) & 1  # qubit q = bit n-1-q
    col0 = V[np.arange(n)[None, :], bits, 0]
    col1 = V[np.arange(n)[None, :], bits, 1]
    expect = (np.prod(col0, axis=1) * H[0, 0] + np.prod(col1, axis=1) * H[1, 0])
    err = float(np.max(np.abs(a[idx] - expect)))
    print(f"30-qubit GHZ + 2 Haar layers (native={native}): max-abs over {len(idx)} "
          f"amplitudes {err:.3e}", flush=True)
    assert err <= 1e-12


def test_served_large_requests_vs_engine(orc):
    reqs = [(20, 14, qa.C128, 2020), (22, 10, qa.C64, 2222), (23, 8, qa.C128, 2323),
            (24, 6, qa.C64, 2424)]
    circs = [workloads.measured_circuit(qa, n, d, sd) for n, d, _, sd in reqs]
    srv = qa.Server(lanes=2)
    ts = [srv.submit(qc, dt, shots=1024, seed=sd) for qc, (_, _, dt, sd) in zip(circs, reqs)]
    served = []
    for t in ts:
        srv.wait(t)
        served.append(srv.stats(t))
    srv.close()
    for qc, (n, d, dt, sd), (stats, last) in zip(circs, reqs, served):
        e = qa.QEngine(qc, dtype=dt, seed=sd)
        e.execute(1024)
        print(f"served {n} qubits depth {d}: {len(stats)} outcomes", flush=True)
        assert e.get_stats() == stats and e.get_dits() == last, (n, d)
        assert sum(stats.values()) == 1024
        if n == 20 and dt == qa.C128:
            ost, odits, _, _ = orc.engine_run(qc, reps=1024, seed=sd)
            assert ost == stats and odits == last

"""Request server (qs_server_*): every request's statistics equal those of
the same QEngine run alone (and the oracle's, for small kets), whatever the
mix it ran in."""
import numpy as np
import pytest

import qpp_amd as qa
from qpp_amd import workloads

pytestmark = pytest.mark.gpu


def mix(seed, count, nmin=10, nmax=18, dmax=12):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        depth = int(rng.integers(2, dmax + 1))
        dt = qa.C64 if rng.random() < 0.5 else qa.C128
        qc = workloads.measured_circuit(qa, n, depth, seed * 1000 + i)
        reqs.append((qc, dt, int(rng.integers(1, 2 ** 31))))
    return reqs


def served_alone(qc, dt, sd, shots):
    """The request on a server of its own: (stats, last dits), batch size."""
    srv = qa.Server(lanes=1)
    t = srv.submit(qc, dt, shots=shots, seed=sd)
    srv.wait(t)
    res, b = srv.stats(t), srv.batch_size(t)
    srv.close()
    return res, b


def tv_distance(a, b, shots):
    keys = set(a) | set(b)
    return sum(abs(a.get(k, 0) - b.get(k, 0)) for k in keys) / (2 * shots)


def test_results_alone_equal_results_in_the_mix(orc):
    """Every request's result equals the same request served alone, bit for
    bit; complex128 results also equal QEngine run alone (and the oracle's
    for small kets).  A batched complex64 request (12-14 qubits) runs its
    gates in another order of operations than QEngine's own passes, so its
    statistics may differ from QEngine's by shots whose uniform falls within
    FP32 rounding of a cumulative-probability edge: at most 5% total
    variation here (expected: a shot or less)."""
    reqs = mix(7, 14)
    srv = qa.Server(lanes=4)
    tickets = [srv.submit(qc, dt, shots=256, seed=sd) for qc, dt, sd in reqs]
    mixed, sizes = [], []
    for t in tickets:
        srv.wait(t)
        mixed.append(srv.stats(t))
        sizes.append(srv.batch_size(t))
    info = srv.info()
    srv.close()
    assert info["completed"] == len(reqs)
    print("batch sizes in the mix:", sizes, flush=True)
    for (qc, dt, sd), (st, last), bs in zip(reqs, mixed, sizes):
        print(f"request n={qc.nq} dtype={dt} batch={bs}", flush=True)
        assert sum(st.values()) == 256
        assert served_alone(qc, dt, sd, 256)[0] == (st, last)
        e = qa.QEngine(qc, dtype=dt, seed=sd).execute(256)
        if dt == qa.C128 or not 12 <= qc.nq <= 14:
            assert e.get_stats() == st and e.get_dits() == last
        else:
            assert tv_distance(e.get_stats(), st, 256) <= 0.05
        if dt == qa.C128 and qc.nq <= 12:
            ost, odits, _, _ = orc.engine_run(qc, reps=256, seed=sd)
            assert ost == st and odits == last


def test_small_requests_run_batched(orc):
    """Small requests queued together run their gate steps in one launch
    (batch size > 1); each result is still the one it has alone, and
    complex128 ones equal QEngine's and the oracle's."""
    reqs = mix(11, 12, nmin=12, nmax=14, dmax=8)
    srv = qa.Server(lanes=1)
    tickets = [srv.submit(qc, dt, shots=512, seed=sd) for qc, dt, sd in reqs]
    res = []
    for t in tickets:
        srv.wait(t)
        res.append((srv.stats(t), srv.batch_size(t)))
    srv.close()
    sizes = [b for _, b in res]
    print("batch sizes:", sizes, flush=True)
    assert max(sizes) >= 2
    for (qc, dt, sd), ((st, last), bs) in zip(reqs, res):
        alone, b1 = served_alone(qc, dt, sd, 512)
        assert b1 == 1 and alone == (st, last)
        if dt == qa.C128:
            e = qa.QEngine(qc, dtype=dt, seed=sd).execute(512)
            assert e.get_stats() == st and e.get_dits() == last
            if qc.nq == 12:
                ost, odits, _, _ = orc.engine_run(qc, reps=512, seed=sd)
                assert ost == st and odits == last


def test_admission_respects_the_budget():
    """A budget of exactly one 18-qubit c128 request's footprint still runs
    every request, one at a time; the peak held never exceeds the budget."""
    reqs = [(workloads.measured_circuit(qa, 18, 3, 50 + i), qa.C128, 11 + i) for i in range(4)]
    budget = qa.Server.footprint(reqs[0][0], qa.C128, 64)
    assert budget >= 3.5 * 16 * 2 ** 18 + 16 * 2 ** 18
    srv = qa.Server(lanes=4, hbm_budget=budget)
    ts = [srv.submit(qc, dt, shots=64, seed=sd) for qc, dt, sd in reqs]
    res = [srv.wait(t) for t in ts]
    info = srv.info()
    srv.close()
    assert info["peak_running"] == 1 and info["peak_bytes"] <= budget
    assert all(r[2] >= r[1] for r in res)


def test_request_above_budget_is_refused():
    srv = qa.Server(lanes=1, hbm_budget=1 << 20)
    with pytest.raises(qa.OutOfMemory):
        srv.submit(workloads.measured_circuit(qa, 16, 1, 3), qa.C128)
    srv.close()

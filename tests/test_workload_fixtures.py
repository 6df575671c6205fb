"""The seeded benchmark workloads against the committed fixtures
(tests/golden/workloads.json, circuit_20q_depth20_seed20.npz, written by
tests/golden/make_workload_fixtures.py).  numpy does not freeze Generator
streams across versions: a drift in the generators fails here first, and
the full-size GPU test runs the 20-qubit circuit from the .npz itself."""
import json
import os

import numpy as np
import pytest

from qpp_amd import workloads

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def fix():
    with open(os.path.join(GOLD, "workloads.json")) as f:
        return json.load(f)


def _mat(m):
    return np.array([[complex(re, im) for re, im in row] for row in m])


def _call(call):
    fn, rest = call.split("(", 1)
    return fn, tuple(int(x) for x in rest.rstrip(")").split(","))


def test_circuits_match_fixtures(fix):
    for name, f in fix["circuits"].items():
        fn, a = _call(f["call"])
        g = getattr(workloads, fn)(*a)
        assert len(g) == f["gates"], name
        assert workloads.gates_digest(g) == f["sha256"], name
        for gate, ref in ((g[0], f["first"]), (g[-1], f["last"])):
            assert (gate[0], gate[1], gate[2]) == (ref["kind"], ref["targets"], ref["controls"])
            assert np.array_equal(gate[3], _mat(ref["matrix"])), name


def test_dm_pipelines_match_fixtures(fix):
    for name, f in fix["dm_pipelines"].items():
        _, a = _call(f["call"])
        rho, ops, tr, ptt, perm = workloads.dm_pipeline(*a)
        assert len(ops) == f["ops"]
        assert workloads.dm_digest(rho, ops, tr, ptt, perm) == f["sha256"], name
        assert rho[0, 0] == complex(*f["rho_00"])
        assert np.array_equal(ops[0][2], _mat(f["first_u"]))
        assert (tr, ptt, perm) == (f["trace_out"], f["ptranspose"], f["perm"])


def test_serving_mix_matches_fixtures(fix):
    for name, f in fix["serving"].items():
        _, (nreq, seed, first) = _call(f["call"])
        reqs = workloads.serve_requests(nreq, seed, first)
        assert workloads.requests_digest(reqs) == f["sha256"], name
        assert reqs[: len(f["first_requests"])] == f["first_requests"]
        for q in reqs:
            assert 10 <= q["n"] <= 26 and 10 <= q["depth"] <= 100


def test_npz_circuit_is_the_generated_one(fix):
    g = workloads.load_gates(os.path.join(GOLD, "circuit_20q_depth20_seed20.npz"))
    assert workloads.gates_digest(g) == fix["circuits"]["random_20q_depth20_seed20"]["sha256"]
    ref = workloads.random_layers(20, 20, 20)
    assert len(g) == len(ref)
    for a, b in zip(g, ref):
        assert a[:3] == b[:3] and np.array_equal(a[3], b[3])


def test_npz_circuit_structure():
    """What the construction fixes, independent of the generator: every
    layer is a 2x2 unitary on each of the 20 qubits in order, then X on
    a perfect pairing of the qubits (10 CNOTs, each qubit in one)."""
    g = workloads.load_gates(os.path.join(GOLD, "circuit_20q_depth20_seed20.npz"))
    n = 20
    per = n + n // 2
    assert len(g) == 20 * per
    for L in range(20):
        layer = g[L * per:(L + 1) * per]
        for q, (kind, t, c, U) in enumerate(layer[:n]):
            assert (kind, t, c) == ("u", [q], [])
            assert np.max(np.abs(U.conj().T @ U - np.eye(2))) <= 1e-14
        seen = []
        for kind, t, c, U in layer[n:]:
            assert kind == "cx" and np.array_equal(U, workloads.X)
            seen += t + c
        assert sorted(seen) == list(range(n))


def test_digest_sees_single_changes():
    g = workloads.random_layers(6, 2, 5)
    d = workloads.gates_digest(g)
    h = [list(x) for x in g]
    U = h[3][3].copy()
    U.view(np.uint64)[0, 0] ^= 1  # one ulp of one real part
    h[3][3] = U
    assert workloads.gates_digest([tuple(x) for x in h]) != d
    k = next(i for i, x in enumerate(g) if x[0] == "cx")
    h = list(g)
    h[k] = ("cx", g[k][2], g[k][1], g[k][3])  # control and target swapped
    assert workloads.gates_digest(h) != d
    assert workloads.gates_digest(g[:-1]) != d

"""Write the seeded workload fixtures:

* workloads.json: for every benchmark workload, the generator call, its
  gate count and its SHA-256 (qpp_amd.workloads.gates_digest / dm_digest /
  requests_digest), the first and last gate's matrix, and the serving
  mix's first request parameters;
* circuit_20q_depth20_seed20.npz: the whole 20-qubit, 20-layer random
  circuit that the full-size GPU test runs against the oracle, as plain
  arrays (qpp_amd.workloads.save_gates).

numpy does not freeze Generator streams across versions; these files pin
the circuits that the device path, the oracle and the CPU baseline share,
and test_workload_fixtures.py fails if the generators drift from them.

Run from the repository root: python tests/golden/make_workload_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from qpp_amd import workloads  # noqa: E402

# (name, generator, args): the bench's and the tests' seeds
CIRCUITS = [
    ("random_20q_depth20_seed30", "random_layers", (20, 20, 30)),
    ("random_20q_depth20_seed20", "random_layers", (20, 20, 20)),
    ("random_30q_depth40_seed30", "random_layers", (30, 40, 30)),
    ("ladder_12q_depth30_seed312", "ladder_layers", (12, 30, 312)),
    ("ladder_16q_depth30_seed316", "ladder_layers", (16, 30, 316)),
    ("ladder_20q_depth30_seed320", "ladder_layers", (20, 30, 320)),
    ("ladder_22q_depth30_seed322", "ladder_layers", (22, 30, 322)),
    ("dm_12q_depth10_seed1", "dm_layers", (12, 10, 1)),
]
PIPELINES = [("dm_pipeline_12q_depth10_seed30", (12, 10, 30)),
             ("dm_pipeline_12q_depth10_seed1", (12, 10, 1))]
SERVING = [("serve_60_seed30", (60, 30, 0)), ("serve_12_seed31_cal", (12, 31, 10000))]


def mat(U):
    return [[[float(z.real), float(z.imag)] for z in row] for row in np.asarray(U)]


def build():
    out = {"circuits": {}, "dm_pipelines": {}, "serving": {}}
    for name, fn, a in CIRCUITS:
        g = getattr(workloads, fn)(*a)
        out["circuits"][name] = {
            "call": f"{fn}{a}", "gates": len(g), "sha256": workloads.gates_digest(g),
            "first": {"kind": g[0][0], "targets": g[0][1], "controls": g[0][2], "matrix": mat(g[0][3])},
            "last": {"kind": g[-1][0], "targets": g[-1][1], "controls": g[-1][2],
                     "matrix": mat(g[-1][3])}}
    for name, a in PIPELINES:
        rho, ops, tr, ptt, perm = workloads.dm_pipeline(*a)
        out["dm_pipelines"][name] = {
            "call": f"dm_pipeline{a}", "ops": len(ops),
            "sha256": workloads.dm_digest(rho, ops, tr, ptt, perm),
            "rho_00": [float(rho[0, 0].real), float(rho[0, 0].imag)],
            "first_u": mat(ops[0][2]), "trace_out": tr, "ptranspose": ptt, "perm": perm}
    for name, (nreq, seed, first) in SERVING:
        reqs = workloads.serve_requests(nreq, seed, first)
        out["serving"][name] = {"call": f"serve_requests{(nreq, seed, first)}",
                                "sha256": workloads.requests_digest(reqs),
                                "first_requests": reqs[:4]}
    return out


def main():
    with open(os.path.join(HERE, "workloads.json"), "w") as f:
        json.dump(build(), f, indent=1)
        f.write("\n")
    workloads.save_gates(os.path.join(HERE, "circuit_20q_depth20_seed20.npz"),
                         workloads.random_layers(20, 20, 20))


if __name__ == "__main__":
    main()

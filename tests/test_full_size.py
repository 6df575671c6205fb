"""The benchmark workloads at their real sizes, where the oracle still runs
in seconds or where size-independent properties decide:

* the 20-qubit complex128 random circuit of 20 layers (read from the
  committed fixture tests/golden/circuit_20q_depth20_seed20.npz), gate by
  gate and as fused passes, against the oracle engine on the same gates;
* the 12-qubit density-matrix pipeline (10 layers, a depolarizing channel
  after every gate, then ptrace / ptranspose / syspermute): fused passes
  against gate-by-gate application, unit trace and hermiticity of the
  result, and the three maps against the oracle on the device's own output
  (ptrace <= 1e-12, ptranspose and syspermute bit-exact).
The 30-qubit circuit is covered by test_gpu_fused.py's forward/inverse test."""
import os

import numpy as np
import pytest

import qpp_amd as qa
from qpp_amd import workloads

pytestmark = pytest.mark.gpu


def test_twenty_qubit_circuit(orc):
    n, depth = 20, 20
    gates = workloads.load_gates(os.path.join(os.path.dirname(__file__), "golden",
                                              "circuit_20q_depth20_seed20.npz"))
    assert len(gates) == depth * (n + n // 2)
    qc = qa.QCircuit.from_gates(n, gates)
    _, _, _, ref = orc.engine_run(qc, reps=1, seed=1)
    st = qa.State([2] * n)
    st.set_basis(0)
    workloads.apply_gates(st, gates)
    assert np.max(np.abs(st.download() - ref)) <= 1e-12
    st.set_basis(0)
    qa.Program(qc, reg_bits=4, low_bits=3, max_ops=48).run(st)
    assert np.max(np.abs(st.download() - ref)) <= 1e-12


def test_twelve_qubit_density_matrix_pipeline(orc):
    n, depth = 12, 10
    rho0, ops, tr, ptt, perm = workloads.dm_pipeline(n, depth, seed=1)
    st = qa.State.from_array(rho0, [2] * n)
    workloads.dm_program(qa, n, ops).compile().run(st)
    fused = st.download()
    s2 = qa.State.from_array(rho0, [2] * n)
    for kind, q, M in ops:
        if kind == "u":
            s2.apply(M, q)
        elif kind == "cz":
            s2.apply_ctrl(M, [q[0]], [q[1]])
        else:
            s2.apply_kraus(M, q)
    assert np.max(np.abs(fused - s2.download())) <= 1e-12
    assert abs(np.trace(fused) - 1) <= 1e-12
    assert np.max(np.abs(fused - fused.conj().T)) <= 1e-14
    dims = [2] * n
    assert np.max(np.abs(st.ptrace(tr).download() - orc.ptrace(fused, tr, dims))) <= 1e-12
    assert np.array_equal(st.ptranspose(ptt, qa.AS_BUILT).download(),
                          orc.ptranspose(fused, ptt, dims, semantics=qa.AS_BUILT))
    assert np.array_equal(st.syspermute(perm, qa.AS_BUILT).download(),
                          orc.syspermute(fused, perm, dims, semantics=qa.AS_BUILT))

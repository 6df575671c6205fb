"""Density-matrix pipeline (gates, depolarizing channels, ptrace,
ptranspose, syspermute) on the device vs the oracle applied op by op."""
import numpy as np
import pytest

import qpp_amd as qa
from qpp_amd import workloads

pytestmark = pytest.mark.gpu


def oracle_pipeline(orc, rho, ops, n):
    for kind, q, M in ops:
        if kind == "u":
            rho = orc.apply(rho, M, q, [2] * n)
        elif kind == "cz":
            rho = orc.applyCTRL(rho, M, [q[0]], [q[1]], [2] * n)
        else:
            rho = orc.apply_kraus(rho, M, q, [2] * n)
    return rho


@pytest.mark.parametrize("n,depth", [(6, 3), (7, 2)])
def test_fused_dm_pipeline_vs_oracle(orc, n, depth):
    rho0, ops, tr, ptt, perm = workloads.dm_pipeline(n, depth, seed=40 + n, p=0.02)
    ref = oracle_pipeline(orc, rho0, ops, n)
    bp = workloads.dm_program(qa, n, ops).compile()
    st = qa.State.from_array(rho0, [2] * n)
    bp.run(st)
    got = st.download()
    assert np.max(np.abs(got - ref)) <= 1e-12
    assert abs(np.trace(got) - 1) < 1e-12
    # gate-by-gate device path agrees too
    s2 = qa.State.from_array(rho0, [2] * n)
    for kind, q, M in ops:
        if kind == "u":
            s2.apply(M, q)
        elif kind == "cz":
            s2.apply_ctrl(M, [q[0]], [q[1]])
        else:
            s2.apply_kraus(M, q)
    assert np.max(np.abs(s2.download() - ref)) <= 1e-12
    # maps on the device's rho: exact against the oracle on the same input
    assert np.max(np.abs(st.ptrace(tr).download() - orc.ptrace(got, tr, [2] * n))) <= 1e-13
    for sem in (qa.DOCUMENTED, qa.AS_BUILT):
        assert np.array_equal(st.ptranspose(ptt, sem).download(),
                              orc.ptranspose(got, ptt, [2] * n, semantics=sem))
        assert np.array_equal(st.syspermute(perm, sem).download(),
                              orc.syspermute(got, perm, [2] * n, semantics=sem))


def test_bitprogram_ket_controls_and_two_bit_gates(orc):
    """Generic bit program on a ket: controlled 1-bit and dense 2-bit ops."""
    rng = np.random.default_rng(3)
    n = 13
    psi = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    psi /= np.linalg.norm(psi)
    bp = qa.BitProgram(n)
    ref = psi.copy()
    for k in range(30):
        q = [int(x) for x in rng.permutation(n)[:3]]
        if k % 3 == 0:
            U = workloads.haar(4, rng)
            bp.gate(U, [n - 1 - q[0], n - 1 - q[1]])
            ref = orc.apply(ref, U, q[:2], [2] * n)
        else:
            U = workloads.haar(2, rng)
            bp.gate(U, [n - 1 - q[0]], cmask=1 << (n - 1 - q[1]), cval=0)
            ref = orc.applyCTRL(ref, U, [q[1]], [q[0]], [2] * n, shift=[1])
    bp.compile()
    st = qa.State.from_array(psi, [2] * n)
    bp.run(st)
    assert np.max(np.abs(st.download() - ref)) <= 1e-12

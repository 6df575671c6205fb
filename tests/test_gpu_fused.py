"""Tile-fused programs vs the oracle applied gate by gate."""
import numpy as np
import pytest

from helpers import CNOT, CZ, H, X, Y, Z, haar, randket, randrho

pytestmark = pytest.mark.gpu

TOL = 1e-12


def maxabs(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


def oracle_run(orc, state, qc):
    """Apply a QCircuit's gate steps with the oracle, step by step."""
    s = state
    for tp, g, ctrl, tg, shift, creg, dest in qc.steps:
        U = qc.mats[g]
        if tp == 1:  # GATE
            s = orc.apply(s, U, tg, 2)
        elif tp == 2:  # FAN
            for t in tg:
                s = orc.apply(s, U, [t], 2)
        elif tp == 3:
            s = orc.applyCTRL(s, U, ctrl, tg, 2, shift)
        elif tp == 4:
            s = orc.applyCTRL_fan(s, U, ctrl, tg, 2, shift)
    return s


def mixed_circuit(qa, n, depth, seed):
    rng = np.random.default_rng(seed)
    qc = qa.QCircuit(n)
    T = np.diag([1, np.exp(1j * np.pi / 4)])
    for _ in range(depth):
        for q in range(n):
            r = rng.random()
            if r < 0.6:
                qc.gate(haar(2, rng), q)
            elif r < 0.7:
                qc.gate(T, q)
            elif r < 0.8:
                qc.gate(H, q)
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            a, b = int(perm[i]), int(perm[i + 1])
            r = rng.random()
            if r < 0.4:
                qc.CTRL(X, a, b)
            elif r < 0.55:
                qc.CTRL(Z, a, b)
            elif r < 0.7:
                qc.gate(haar(4, rng), a, b)
            elif r < 0.8:
                qc.gate(CZ, a, b)
            elif r < 0.9:
                c2 = int(perm[(i + 2) % n])
                if c2 not in (a, b):
                    qc.CTRL(haar(2, rng), [a, c2], b, shift=[int(rng.integers(2)), int(rng.integers(2))])
            else:
                qc.CTRL_fan(haar(2, rng), a, [b] + ([int(perm[(i + 3) % n])] if int(perm[(i + 3) % n]) not in (a, b) else []))
        if rng.random() < 0.3:
            qc.gate_fan(haar(2, rng), [int(x) for x in rng.permutation(n)[:3]])
    return qc


@pytest.mark.parametrize("n,reg", [(12, 4), (12, 3), (14, 4), (17, 4), (20, 3)])
def test_fused_random_layers(qa, orc, n, reg):
    from qpp_amd import workloads
    gates = workloads.random_layers(n, 4, seed=n)
    qc = qa.QCircuit.from_gates(n, gates)
    rng = np.random.default_rng(n)
    psi = randket(2 ** n, rng)
    st = qa.State.from_array(psi, [2] * n)
    prog = qa.Program(qc, reg_bits=reg)
    info = prog.info()
    assert info["gates"] == len(gates)
    prog.run(st)
    ref = oracle_run(orc, psi, qc)
    err = maxabs(st.download(), ref)
    print(f"n={n} reg={reg} {info} err={err:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("n", [12, 13, 16])
def test_fused_mixed_gates(qa, orc, n):
    qc = mixed_circuit(qa, n, 3, seed=100 + n)
    rng = np.random.default_rng(n + 1)
    psi = randket(2 ** n, rng)
    st = qa.State.from_array(psi, [2] * n)
    qa.Program(qc).run(st)
    assert maxabs(st.download(), oracle_run(orc, psi, qc)) <= TOL


@pytest.mark.parametrize("n", [6, 7, 8])
def test_fused_rho(qa, orc, n):
    qc = mixed_circuit(qa, n, 2, seed=200 + n)
    rng = np.random.default_rng(n + 2)
    rho = randrho(2 ** n, rng)
    st = qa.State.from_array(rho, [2] * n)
    qa.Program(qc, kind=qa.RHO).run(st)
    assert maxabs(st.download(), oracle_run(orc, rho, qc)) <= TOL


def test_fused_low_bits_variants(qa, orc):
    n = 15
    qc = mixed_circuit(qa, n, 3, seed=7)
    rng = np.random.default_rng(3)
    psi = randket(2 ** n, rng)
    ref = oracle_run(orc, psi, qc)
    for low, maxops in ((0, 64), (3, 8), (6, 1), (5, 200)):
        st = qa.State.from_array(psi, [2] * n)
        qa.Program(qc, low_bits=low, max_ops=maxops).run(st)
        assert maxabs(st.download(), ref) <= TOL, (low, maxops)


@pytest.mark.parametrize("reg", [4, 3])
def test_fused_many_tiles_per_workgroup(qa, orc, reg):
    """Few persistent workgroups each walking many tiles (grid_cap), so the
    prefetch of the next tile, the LDS remaps and the stores of consecutive
    tiles interleave; several option sets vary the phase layouts, and every
    run must match the oracle (a race would show as a sporadic wrong
    amplitude)."""
    from qpp_amd import workloads
    n = 18
    gates = workloads.random_layers(n, 6, seed=11)
    qc = qa.QCircuit.from_gates(n, gates)
    rng = np.random.default_rng(12)
    psi = randket(2 ** n, rng)
    ref = oracle_run(orc, psi, qc)
    for low, maxops, grid in ((6, 48, 4), (4, 32, 8), (5, 24, 16), (2, 64, 3)):
        for rep in range(2):
            st = qa.State.from_array(psi, [2] * n)
            qa.Program(qc, reg_bits=reg, low_bits=low, max_ops=maxops, grid_cap=grid).run(st)
            assert maxabs(st.download(), ref) <= TOL, (reg, low, maxops, grid, rep)


def test_fused_c64(qa, orc):
    from qpp_amd import workloads
    n = 16
    gates = workloads.random_layers(n, 6, seed=5)
    qc = qa.QCircuit.from_gates(n, gates)
    psi = np.zeros(2 ** n, dtype=complex)
    psi[0] = 1
    st = qa.State.from_array(psi, [2] * n, dtype=qa.C64)
    qa.Program(qc, dtype=qa.C64).run(st)
    ref = oracle_run(orc, psi, qc)
    out = st.download().astype(np.complex128)
    # FP32 rounding over ~100 gates: amplitude error ~1e-7, infidelity ~1e-8
    # (norm drift included); the complex64 path's bar is 1e-5
    infid = 1 - abs(np.vdot(ref, out)) ** 2
    assert maxabs(out, ref) <= 5e-6 and infid <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("c64", [False, True])
def test_full_size_circuit_and_inverse(qa, c64):
    """The benchmark workload at its real size: a 30-qubit complex128 state
    (16 GiB) through the depth-40 random circuit (1200 Haar gates, 600 CNOTs)
    as tile-fused passes, then through the inverse circuit (adjoints in
    reverse order).  The state must come back to |0...0>, with every fused
    pass keeping the norm: size-independent properties that a wrong index,
    lane or phase anywhere in the 2^30 amplitudes would break.  The
    complex64 path keeps its norm and overlap in FP64; its bar is the
    precision path's (infidelity 1e-5)."""
    from qpp_amd import workloads
    n, depth = 30, 40
    gates = workloads.random_layers(n, depth, seed=30)
    inv = [(k, t, c, np.asarray(U).conj().T) for k, t, c, U in reversed(gates)]
    dt = qa.C64 if c64 else qa.C128
    tol_norm, tol_inf, tol_abs = (1e-5, 1e-5, 1e-3) if c64 else (1e-12, 1e-12, 1e-10)
    st = qa.State([2] * n, dtype=dt)
    st.set_basis(0)
    qa.Program(n, gates, dtype=dt, reg_bits=4, low_bits=3, max_ops=48).run(st)
    assert abs(st.norm2() - 1) < tol_norm
    qa.Program(n, inv, dtype=dt, reg_bits=4, low_bits=3, max_ops=48).run(st)
    ref = qa.State([2] * n)
    ref.set_basis(0)
    r = st.compare(ref)
    assert abs(r["norm2_a"] - 1) < tol_norm
    assert r["infidelity"] < tol_inf and r["max_abs"] < tol_abs, r

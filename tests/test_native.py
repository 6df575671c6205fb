"""Specialised pass kernels (qs_program_compile_native): every pass of a
tile program generated as its own kernel and compiled with hipRTC.  Checked
against the oracle applied gate by gate (complex128 <= 1e-12) on circuits
with every op kind the passes hold -- dense and diagonal 1-qubit gates,
dense 2-qubit gates, controls on register, thread and out-of-tile bits,
shifted controls, fans -- on kets and density matrices, with few
persistent workgroups walking many tiles, and at the benchmark's full size
(forward then inverse returns to |0...0>)."""
import numpy as np
import pytest

from helpers import randket, randrho
from test_gpu_fused import maxabs, mixed_circuit, oracle_run

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("n", [12, 13, 16])
def test_native_mixed_gates_vs_oracle(qa, orc, n):
    qc = mixed_circuit(qa, n, 3, seed=300 + n)
    rng = np.random.default_rng(n + 5)
    psi = randket(2 ** n, rng)
    st = qa.State.from_array(psi, [2] * n)
    p = qa.Program(qc).compile_native()
    print(f"native n={n} {p.info()} kernel {p.kernel_name()}", flush=True)
    p.run(st)
    assert maxabs(st.download(), oracle_run(orc, psi, qc)) <= TOL


@pytest.mark.parametrize("n", [6, 7])
def test_native_rho_vs_oracle(qa, orc, n):
    qc = mixed_circuit(qa, n, 2, seed=400 + n)
    rng = np.random.default_rng(n + 6)
    rho = randrho(2 ** n, rng)
    st = qa.State.from_array(rho, [2] * n)
    qa.Program(qc, kind=qa.RHO).compile_native().run(st)
    assert maxabs(st.download(), oracle_run(orc, rho, qc)) <= TOL


def test_native_layouts_and_many_tiles(qa, orc):
    """Layout options that change which phases load, remap and store
    directly, with few persistent workgroups per pass (each walks many
    tiles, so prefetch, LDS layout changes and stores of consecutive tiles
    interleave)."""
    n = 17
    qc = mixed_circuit(qa, n, 3, seed=17)
    rng = np.random.default_rng(18)
    psi = randket(2 ** n, rng)
    ref = oracle_run(orc, psi, qc)
    for low, maxops, grid in ((0, 64, 5), (3, 8, 3), (6, 48, 64), (5, 200, 0)):
        print(f"native layout low={low} max_ops={maxops} grid={grid}", flush=True)
        st = qa.State.from_array(psi, [2] * n)
        qa.Program(qc, low_bits=low, max_ops=maxops, grid_cap=grid).compile_native().run(st)
        assert maxabs(st.download(), ref) <= TOL, (low, maxops, grid)


def test_native_c64_and_equal_to_tile_kernel(qa, orc):
    from qpp_amd import workloads
    n = 16
    gates = workloads.random_layers(n, 6, seed=9)
    qc = qa.QCircuit.from_gates(n, gates)
    psi = np.zeros(2 ** n, dtype=complex)
    psi[0] = 1
    ref = oracle_run(orc, psi, qc)
    a = qa.State.from_array(psi, [2] * n, dtype=qa.C64)
    qa.Program(qc, dtype=qa.C64).compile_native().run(a)
    out = a.download().astype(np.complex128)
    assert maxabs(out, ref) <= 5e-6 and 1 - abs(np.vdot(ref, out)) ** 2 <= 1e-6
    # the same complex products in the same order as the tile kernel
    b = qa.State.from_array(psi, [2] * n, dtype=qa.C64)
    qa.Program(qc, dtype=qa.C64).run(b)
    assert maxabs(out, b.download()) <= 1e-6


def test_native_rho_channels_bit_program(qa):
    """The density-matrix pipeline's gates and Kraus superoperators as one
    bit-level program: specialised kernels equal the tile kernel."""
    from qpp_amd import workloads
    n = 6
    rho, ops, _, _, _ = workloads.dm_pipeline(n, 3, seed=5)
    a = qa.State.from_array(rho, [2] * n)
    b = qa.State.from_array(rho, [2] * n)
    workloads.dm_program(qa, n, ops).compile().compile_native().run(a)
    workloads.dm_program(qa, n, ops).compile().run(b)
    assert maxabs(a.download(), b.download()) <= TOL


def test_native_needs_reg_bits_4(qa):
    from qpp_amd import workloads
    p = qa.Program(12, workloads.random_layers(12, 2, 1), reg_bits=3)
    with pytest.raises(qa.NotImplemented_):
        p.compile_native()


def test_native_full_size_circuit_and_inverse(qa):
    """The 30-qubit depth-40 benchmark circuit and its inverse through
    specialised kernels (complex128): back to |0...0> within 1e-12."""
    from qpp_amd import workloads
    n = 30
    gates = workloads.random_layers(n, 40, seed=30)
    inv = [(k, t, c, np.asarray(U).conj().T) for k, t, c, U in reversed(gates)]
    st = qa.State([2] * n)
    st.set_basis(0)
    qa.Program(n, gates, reg_bits=4, low_bits=3, max_ops=48).compile_native().run(st)
    assert abs(st.norm2() - 1) < 1e-12
    qa.Program(n, inv, reg_bits=4, low_bits=3, max_ops=48).compile_native().run(st)
    ref = qa.State([2] * n)
    ref.set_basis(0)
    r = st.compare(ref)
    assert r["infidelity"] < 1e-12 and r["max_abs"] < 1e-10, r


def test_native_split_layout_changes_vs_oracle(qa, orc):
    """Layout changes split in halves on a wave bit, a lane bit and a
    register bit (complex128; the passes whose changes all split run at half
    the LDS) and moved whole (complex64 keeps whole changes): circuits where
    every form occurs, against the oracle."""
    from qpp_amd import workloads
    seen = dict.fromkeys(("whole", "wave_split", "lane_split", "register_split", "half_lds_passes"), 0)
    for n, seed, low, maxops, dt in ((13, 1, 3, 16, qa.C128), (14, 2, 3, 48, qa.C128),
                                     (16, 1, 5, 48, qa.C128), (14, 3, 3, 48, qa.C64)):
        qc = qa.QCircuit.from_gates(n, workloads.random_layers(n, 12, seed))
        psi = randket(2 ** n, np.random.default_rng(seed + 50))
        st = qa.State.from_array(psi, [2] * n, dtype=dt)
        p = qa.Program(qc, dtype=dt, low_bits=low, max_ops=maxops).compile_native()
        stats = p.native_stats()
        print(f"native splits n={n} seed={seed} low={low} max_ops={maxops} dtype={dt}: {stats}", flush=True)
        p.run(st)
        tol = TOL if dt == qa.C128 else 5e-6
        assert maxabs(st.download().astype(np.complex128), oracle_run(orc, psi, qc)) <= tol, (n, seed, stats)
        for k in seen:
            seen[k] += stats[k]
    assert all(v > 0 for v in seen.values()), seen


def test_native_sign_diagonals_vs_oracle(qa, orc):
    """Diagonal entries of +-1 (Z, CZ, multi-controlled Z) become sign flips:
    on register bits, under lane and tile conditions, and as per-lane and
    per-tile factors (targets on thread bits and outside the tile), for
    complex128 and complex64, against the oracle."""
    from qpp_amd import workloads
    Z = np.diag([1, -1]).astype(complex)
    for n, seed, dt in ((14, 5, qa.C128), (16, 6, qa.C128), (15, 7, qa.C64)):
        rng = np.random.default_rng(seed)
        gates = []
        for kind, t, c, U in workloads.random_layers(n, 8, seed):
            if c:
                gates.append(("cz", t, c, Z))
            elif rng.random() < 0.3:
                gates.append(("z", t, [], Z))
            else:
                gates.append((kind, t, c, U))
        for _ in range(n):  # doubly controlled Z on random qubits
            q = [int(v) for v in rng.choice(n, 3, replace=False)]
            gates.append(("ccz", [q[0]], q[1:], Z))
        qc = qa.QCircuit.from_gates(n, gates)
        psi = randket(2 ** n, rng)
        st = qa.State.from_array(psi, [2] * n, dtype=dt)
        p = qa.Program(qc, dtype=dt).compile_native()
        print(f"native sign diagonals n={n} dtype={dt}: {p.info()}", flush=True)
        p.run(st)
        tol = TOL if dt == qa.C128 else 5e-6
        assert maxabs(st.download().astype(np.complex128), oracle_run(orc, psi, qc)) <= tol, (n, seed)

"""Kraus channels: qpp's whole-system apply(A, Ks) with rectangular operators
(operations.hpp:422-464) and targeted channels wider than the fused
superoperator pass (operations.hpp:476-560).

CPU: the oracle's apply(A, Ks) against closed forms -- the operators
K_i = <i| (x) I trace out the first qubit (so the channel is ptrace), square
operators on every subsystem equal the targeted channel on all targets, and
an isometry split into operators preserves the trace.
GPU: the device against the oracle, c128 and c64, and qpp's errors in
qpp's order (a ket is MatrixNotSquare)."""
import numpy as np
import pytest

import qpp_amd as qa
from helpers import randket

gpu = pytest.mark.gpu


def rand_rho(D, rng, rank=3):
    vs = [randket(D, rng) for _ in range(rank)]
    w = rng.random(rank)
    w /= w.sum()
    return sum(wi * np.outer(v, v.conj()) for wi, v in zip(w, vs))


def ptrace_first(rho, d0, drest):
    r = rho.reshape(d0, drest, d0, drest)
    return np.einsum("iaib->ab", r)


def trace_out_first_ops(d0, drest):
    """K_i = <i| (x) I: drest x (d0 drest)"""
    Ks = []
    for i in range(d0):
        K = np.zeros((drest, d0 * drest), dtype=complex)
        K[:, i * drest:(i + 1) * drest] = np.eye(drest)
        Ks.append(K)
    return Ks


def isometry_ops(Din, Dout, nK, rng):
    """nK operators Dout x Din with sum K^dagger K = I (a random isometry
    Din -> nK Dout, cut into blocks)"""
    Z = rng.standard_normal((nK * Dout, Din)) + 1j * rng.standard_normal((nK * Dout, Din))
    Q, _ = np.linalg.qr(Z)
    return [Q[k * Dout:(k + 1) * Dout, :] for k in range(nK)]


def test_oracle_full_channel_rectangular_is_ptrace(orc):
    rng = np.random.default_rng(1)
    rho = rand_rho(8, rng)
    got = orc.apply_kraus_full(rho, trace_out_first_ops(2, 4))
    assert np.max(np.abs(got - ptrace_first(rho, 2, 4))) < 1e-14


def test_oracle_full_channel_square_equals_targeted_on_all(orc):
    rng = np.random.default_rng(2)
    rho = rand_rho(4, rng)
    Ks = isometry_ops(4, 4, 3, rng)
    full = orc.apply_kraus_full(rho, Ks)
    tgt = orc.apply_kraus(rho, Ks, [0, 1], [2, 2])
    ref = sum(K @ rho @ K.conj().T for K in Ks)
    assert np.max(np.abs(full - ref)) < 1e-14
    assert np.max(np.abs(tgt - ref)) < 1e-14


def test_oracle_full_channel_trace_preserving(orc):
    rng = np.random.default_rng(3)
    rho = rand_rho(16, rng)
    out = orc.apply_kraus_full(rho, isometry_ops(16, 4, 5, rng))
    assert out.shape == (4, 4)
    assert abs(np.trace(out) - 1) < 1e-13
    assert np.max(np.abs(out - out.conj().T)) < 1e-14


@gpu
@pytest.mark.parametrize("dt", [qa.C128, qa.C64])
def test_full_channel_vs_oracle(orc, dt):
    rng = np.random.default_rng(10)
    tol = 1e-12 if dt == qa.C128 else 2e-6
    for Din, Dout, nK in [(8, 4, 2), (16, 8, 3), (12, 12, 4), (64, 16, 5)]:
        print(f"full channel {Din} -> {Dout}, {nK} operators, dtype {dt}", flush=True)
        rho = rand_rho(Din, rng)
        Ks = isometry_ops(Din, Dout, nK, rng) if (Din, Dout) != (8, 4) else trace_out_first_ops(2, 4)
        st = qa.State.from_array(rho, [Din], dtype=dt)
        out = st.apply_kraus_full(Ks).download()
        ref = orc.apply_kraus_full(rho, Ks)
        assert out.shape == (Dout, Dout)
        assert np.max(np.abs(out - ref)) <= tol
        # the module-level form, qpp's apply(A, Ks)
        if dt == qa.C128:
            assert np.max(np.abs(qa.apply(rho, Ks) - ref)) <= tol


@gpu
def test_full_channel_errors():
    rng = np.random.default_rng(11)
    ket = qa.State.from_array(randket(4, rng), [2, 2])
    with pytest.raises(qa.MatrixNotSquare):
        ket.apply_kraus_full([np.eye(4)])
    rho = qa.State.from_array(rand_rho(4, rng), [2, 2])
    with pytest.raises(qa.DimsMismatchMatrix):
        rho.apply_kraus_full([np.eye(3)])
    with pytest.raises(qa.DimsNotEqual):
        rho.apply_kraus_full([np.eye(4), np.eye(2, 4)])
    with pytest.raises(qa.MatrixNotSquare):
        qa.apply(randket(4, rng), [np.eye(4)])


@gpu
@pytest.mark.parametrize("dt", [qa.C128, qa.C64])
def test_wide_targeted_channel_vs_oracle(orc, dt):
    """Targets spanning more than 32 basis states: one term per operator."""
    rng = np.random.default_rng(12)
    tol = 1e-12 if dt == qa.C128 else 2e-6
    for dims, target, nK in [([2] * 7, [1, 2, 3, 4, 5, 6], 3), ([3, 3, 4, 2], [0, 2, 1], 2)]:
        print(f"wide channel dims={dims} target={target}", flush=True)
        D = int(np.prod(dims))
        DK = int(np.prod([dims[t] for t in target]))
        rho = rand_rho(D, rng)
        Ks = isometry_ops(DK, DK, nK, rng)
        st = qa.State.from_array(rho, dims, dtype=dt)
        st.apply_kraus(Ks, target)
        ref = orc.apply_kraus(rho, Ks, target, dims)
        assert np.max(np.abs(st.download() - ref)) <= tol


@gpu
def test_targeted_channel_on_a_ket_is_matrix_not_square():
    rng = np.random.default_rng(13)
    st = qa.State.from_array(randket(8, rng), [2, 2, 2])
    with pytest.raises(qa.MatrixNotSquare):
        st.apply_kraus([np.eye(2)], [0])

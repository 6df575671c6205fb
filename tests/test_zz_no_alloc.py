"""qs_sample_noalloc (the allocation-free variant of qs_sample) against
qs_sample, both through the C ABI: two calls in a row, the second on a
smaller state with fewer shots, give the same indices, digits and
probabilities; the digits-only call (no index buffer) agrees too; sizes above
the fixed scratch are refused before anything runs."""
import ctypes
from ctypes import POINTER, c_double, c_uint32, c_uint64, c_void_p

import numpy as np
import pytest

import qpp_amd as qa
from helpers import randket

pytestmark = pytest.mark.gpu

QS_ERR_OUT_OF_MEMORY = 101


def _f(name):
    f = getattr(qa.lib(), name)
    f.restype = ctypes.c_int
    f.argtypes = [c_void_p, c_uint32, POINTER(c_uint64), c_uint64, POINTER(c_double),
                  POINTER(c_uint64), POINTER(c_uint64), POINTER(c_double)]
    return f


def call(name, st, target, u, want_index=True):
    t = np.asarray(target, dtype=np.uint64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    n = len(u)
    idx = np.full(n, 7, dtype=np.uint64)
    dig = np.full(max(1, n * len(t)), 7, dtype=np.uint64)
    pr = np.full(n, -1.0)
    rc = _f(name)(st.handle, len(t), t.ctypes.data_as(POINTER(c_uint64)), n,
                  u.ctypes.data_as(POINTER(c_double)),
                  idx.ctypes.data_as(POINTER(c_uint64)) if want_index else None,
                  dig.ctypes.data_as(POINTER(c_uint64)), pr.ctypes.data_as(POINTER(c_double)))
    return rc, idx, dig[: n * len(t)].reshape(n, len(t)), pr


def uniforms(st, rng, k):
    """random uniforms plus values on and beside cumulative-table edges"""
    p = st.probs()
    cp = np.cumsum(p / p.sum())
    u = list(rng.random(k))
    for j in rng.integers(0, len(cp), size=16):
        c = float(cp[j])
        u += [c, np.nextafter(c, 0), np.nextafter(c, 1)]
    return np.array([min(x, np.nextafter(1.0, 0)) for x in u] + [0.0])


def test_noalloc_equals_sample_two_calls():
    rng = np.random.default_rng(515)
    cases = [(20, qa.C128, [3, 0, 19, 7], 1000), (12, qa.C64, [11, 2], 60)]
    for n, dt, target, k in cases:  # the second call is smaller
        print(f"noalloc vs sample: n={n} dtype={dt} shots={k}", flush=True)
        psi = randket(2 ** n, rng)
        psi[rng.integers(2 ** n, size=2 ** (n - 3))] = 0
        st = qa.State.from_array(psi / np.linalg.norm(psi), [2] * n, dtype=dt)
        u = uniforms(st, rng, k)
        r1 = call("qs_sample", st, target, u)
        r2 = call("qs_sample_noalloc", st, target, u)
        assert r1[0] == 0 and r2[0] == 0, qa.lib().qs_last_error()
        for a, b in zip(r1[1:], r2[1:]):
            assert np.array_equal(a, b)
        r3 = call("qs_sample_noalloc", st, target, u, want_index=False)
        assert r3[0] == 0 and np.array_equal(r3[2], r1[2]) and np.array_equal(r3[3], r1[3])


def test_noalloc_qudits_and_density_matrix():
    rng = np.random.default_rng(516)
    dims = [3, 5, 2, 4]
    D = int(np.prod(dims))
    st = qa.State.from_array(randket(D, rng), dims)
    u = uniforms(st, rng, 50)
    r1, r2 = call("qs_sample", st, [2, 0], u), call("qs_sample_noalloc", st, [2, 0], u)
    assert r1[0] == r2[0] == 0
    assert all(np.array_equal(a, b) for a, b in zip(r1[1:], r2[1:]))
    v = randket(2 ** 6, rng)
    rho = np.outer(v, v.conj())
    sr = qa.State.from_array(rho, [2] * 6)
    u = rng.random(40)
    r1, r2 = call("qs_sample", sr, [5, 1], u), call("qs_sample_noalloc", sr, [5, 1], u)
    assert r1[0] == r2[0] == 0
    assert all(np.array_equal(a, b) for a, b in zip(r1[1:], r2[1:]))


def test_noalloc_refuses_sizes_above_its_scratch():
    big = qa.State([2] * 21)
    rc, *_ = call("qs_sample_noalloc", big, [0], np.array([0.5]))
    assert rc == QS_ERR_OUT_OF_MEMORY
    assert b"2^20" in qa.lib().qs_last_error()
    small = qa.State([2] * 4)
    small.set_basis(3)
    rc, *_ = call("qs_sample_noalloc", small, [0], np.full(16385, 0.5))
    assert rc == QS_ERR_OUT_OF_MEMORY
    rc, idx, _, _ = call("qs_sample_noalloc", small, [0], np.full(16384, 0.5))
    assert rc == 0 and np.all(idx == 3)

"""syspermute and ptranspose of qubit states at sizes where each persistent
workgroup of the bit-permutation kernel visits several tiles (a 24-qubit ket,
an 11-qubit density matrix), in complex128 and complex64.  The maps only move
values, so both precisions must equal the oracle bit for bit
(qpp/operations.hpp: syspermute, ptranspose; qubit kernels as built)."""
import numpy as np
import pytest

from helpers import randket

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_large_bit_permutations(qa, orc, dtype):
    rng = np.random.default_rng(24)
    dt = qa.C128 if dtype == "c128" else qa.C64
    # values exactly representable in the state's precision
    cast = (lambda a: a.astype(np.complex64).astype(np.complex128)) if dtype == "c64" else (lambda a: a)
    n = 24
    psi = cast(randket(2 ** n, rng))
    st = qa.State.from_array(psi, [2] * n, dtype=dt)
    perms = [list(range(12)) + [int(x) for x in rng.permutation(np.arange(12, 24))],
             [int(x) for x in rng.permutation(n)]]
    for perm in perms:
        print(f"{dtype} ket n={n} syspermute {perm}", flush=True)
        got = st.syspermute(perm, qa.AS_BUILT).download().astype(np.complex128)
        assert np.array_equal(got, orc.syspermute(psi, perm, [2] * n, semantics=qa.AS_BUILT)), perm
    nr = 11
    D = 2 ** nr
    rho = cast((rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D))) / D)
    sr = qa.State.from_array(rho, [2] * nr, dtype=dt)
    for t in ([0, 1, 2], [int(x) for x in rng.permutation(nr)[:5]], [nr - 1]):
        print(f"{dtype} rho n={nr} ptranspose {t}", flush=True)
        got = sr.ptranspose(t, qa.AS_BUILT).download().astype(np.complex128)
        assert np.array_equal(got, orc.ptranspose(rho, t, [2] * nr, semantics=qa.AS_BUILT)), t
    perm = [int(x) for x in rng.permutation(nr)]
    got = sr.syspermute(perm, qa.AS_BUILT).download().astype(np.complex128)
    assert np.array_equal(got, orc.syspermute(rho, perm, [2] * nr, semantics=qa.AS_BUILT)), perm

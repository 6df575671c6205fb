"""The exact sampling table on the device (every shot forced through it)
against the oracle's libstdc++ restatement, including uniforms on and next to
the cumulative table's edges, zero runs, one dominant weight, amplitudes
spanning hundreds of binades and exact powers of two, up to 22 qubits."""
import numpy as np
import pytest

from helpers import randket

pytestmark = pytest.mark.gpu


def make_state(mode, n, rng):
    D = 2 ** n
    if mode == "random":
        psi = randket(D, rng)
        if n >= 9:
            psi[rng.integers(D, size=D // 4)] = 0
    elif mode == "dominant":
        psi = randket(D, rng) * 1e-9
        psi[D // 3] = np.sqrt(0.75)
    elif mode == "tiny":  # |a| from 1e-150 to 1: p across ~1000 binades
        psi = 10.0 ** (-150.0 * rng.random(D)) * np.exp(2j * np.pi * rng.random(D))
        psi[0] = 1e-160
    elif mode == "powers":  # p = 2^-k exactly (left unnormalised: the
        # table divides by the sum)
        return np.ldexp(1.0, -rng.integers(5, 20, size=D)).astype(complex)
    else:
        raise ValueError(mode)
    return psi / np.linalg.norm(psi)


def uniforms(cp, rng, nedge=20):
    u = list(rng.random(200))
    for k in rng.integers(0, len(cp), size=nedge):
        c = float(cp[k])
        u += [c, np.nextafter(c, 0), np.nextafter(c, 1)]
    return [min(x, np.nextafter(1.0, 0)) for x in u] + [np.nextafter(1.0, 0), 1e-300, 0.0]


CASES = [("random", n) for n in (2, 9, 14, 17, 20, 22)] + \
        [(m, n) for m in ("dominant", "tiny", "powers") for n in (12, 20)]


@pytest.mark.parametrize("mode,n", CASES)
def test_forced_exact_table_matches_oracle(qa, orc, mode, n, monkeypatch):
    print(f"exact table: mode={mode} n={n}", flush=True)
    monkeypatch.setenv("QPP_AMD_FORCE_EXACT", "1")
    rng = np.random.default_rng(77 + n)
    psi = make_state(mode, n, rng)
    p = orc.probs(psi)
    cp = orc.discrete_table(p)
    u = uniforms(cp, rng)
    st = qa.State.from_array(psi, [2] * n)
    idx, dig, pr = st.sample(list(range(n)), u)
    # lower_bound in the oracle's sequential table (cp is non-decreasing)
    ref = [int(np.searchsorted(cp, x, side="left")) for x in u]
    assert [int(i) for i in idx] == ref
    if n <= 14:  # the oracle's own draw routine agrees with the table
        assert ref == [orc.discrete_draw(p, x) for x in u]
    assert np.array_equal(pr, p[ref])
    # and the same as the error-bounded parallel search
    monkeypatch.delenv("QPP_AMD_FORCE_EXACT")
    idx2, _, _ = st.sample(list(range(n)), u)
    assert np.array_equal(idx, idx2)

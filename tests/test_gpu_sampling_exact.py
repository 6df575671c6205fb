"""The exact sequential replay of qpp's sampler on the device (every shot
forced through it) against the oracle's libstdc++ restatement, including
uniforms on and next to the cumulative table's edges and zero runs."""
import numpy as np
import pytest

from helpers import randket

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [2, 9, 14, 17])
def test_forced_exact_replay_matches_oracle(qa, orc, n, monkeypatch):
    monkeypatch.setenv("QPP_AMD_FORCE_EXACT", "1")
    rng = np.random.default_rng(77 + n)
    psi = randket(2 ** n, rng)
    if n >= 9:
        psi[rng.integers(2 ** n, size=2 ** (n - 2))] = 0
        psi /= np.linalg.norm(psi)
    p = orc.probs(psi)
    cp = orc.discrete_table(p)
    u = list(rng.random(200))
    for k in rng.integers(0, 2 ** n, size=20):
        c = float(cp[k])
        u += [c, np.nextafter(c, 0), np.nextafter(c, 1)]
    u = [min(x, np.nextafter(1.0, 0)) for x in u] + [np.nextafter(1.0, 0), 1e-300]
    st = qa.State.from_array(psi, [2] * n)
    idx, dig, pr = st.sample(list(range(n)), u)
    ref = [orc.discrete_draw(p, x) for x in u]
    assert [int(i) for i in idx] == ref
    assert np.array_equal(pr, p[ref])
    # and the same as the bounded parallel search
    monkeypatch.delenv("QPP_AMD_FORCE_EXACT")
    idx2, _, _ = st.sample(list(range(n)), u)
    assert np.array_equal(idx, idx2)

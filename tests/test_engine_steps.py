"""reset, discard and post-selection: qpp::reset (instruments.hpp:1148-1196),
the engine's RESET / DISCARD / POST_SELECT steps (qengine.hpp:650-821) and
its post-selection retries (qengine.hpp:240-481, 1177-1227).

The oracle is pinned on the CPU by closed forms and by draw sequences worked
out by hand from qpp's loops for small circuits; the device engine is then
compared with the oracle on the same seeds (statistics, dits, probabilities,
post_select_ok and final state)."""
import numpy as np
import pytest

from helpers import H, X, Xd, haar, randket, randrho

import qpp_amd as qa

INF = (1 << 64) - 1


def bell_ps_circuit(ps_val=1, tail_gate=False):
    """H(0) CNOT(0,1), post-select q0 = ps_val into c0, measure q1 into c1.
    With tail_gate the post-selection is non-destructive and X(0) follows
    it: the canonical form cannot move the post-selection past a gate on its
    own qubit, so the circuit is re-run per rep instead of sampled."""
    qc = qa.QCircuit(2, 2)
    qc.gate(H, 0).CTRL(X, 0, 1).post_select(0, ps_val, 0, destructive=not tail_gate)
    if tail_gate:
        qc.gate(X, 0)
    return qc.measure(1, 1)


# ---------------------------------------------------------------------------
# qpp::reset in the oracle (closed forms)
# ---------------------------------------------------------------------------
def test_oracle_reset_single_qubit(orc):
    """a|0> + b|1>: outcome 0 when u < |a|^2 (keeps a/|a| |0>), else 1,
    whose correction X maps b/|b| |1> to b/|b| |0>."""
    a, b = 0.6, 0.8j
    psi = np.array([a, b])
    for u, exp in ((0.2, np.array([1, 0])), (0.5, np.array([1j, 0]))):
        assert np.max(np.abs(orc.reset(psi, [0], [2], u) - exp)) < 1e-15


def test_oracle_reset_bell_and_qutrits(orc):
    bell = np.array([1, 0, 0, 1]) / np.sqrt(2)
    # measuring q0 of a Bell pair gives r with u < 1/2 <=> r = 0; state |r r>,
    # then X^r on q0: |0 r>
    assert np.max(np.abs(orc.reset(bell, [0], [2, 2], 0.3) - [1, 0, 0, 0])) < 1e-15
    assert np.max(np.abs(orc.reset(bell, [0], [2, 2], 0.7) - [0, 1, 0, 0])) < 1e-15
    # qutrits |1, 2>: reset q1 measures 2, Xd^(3-2) maps |2> -> |0>
    k = np.zeros(9)
    k[1 * 3 + 2] = 1
    exp = np.zeros(9)
    exp[1 * 3 + 0] = 1
    assert np.max(np.abs(orc.reset(k, [1], [3, 3], 0.5) - exp)) < 1e-15
    # both qutrits: Xd^(3-1) on q0 and Xd^(3-2) on q1 -> |0, 0>
    exp0 = np.zeros(9)
    exp0[0] = 1
    assert np.max(np.abs(orc.reset(k, [0, 1], [3, 3], 0.5) - exp0)) < 1e-15


def test_oracle_reset_density_matrix(orc):
    """rho -> X^r P_r rho P_r X^r / tr(P_r rho) on the target, r from the
    sample at u (itself pinned against libstdc++ draws)."""
    rng = np.random.default_rng(3)
    rho = randrho(8, rng)
    for t, u in ((0, 0.1), (1, 0.6), (2, 0.95)):
        _, dig, _ = orc.sample(rho, [t], [2, 2, 2], u)
        r = dig[0]
        P = np.diag([1.0 - r, float(r)])
        ops = [np.eye(2)] * 3
        ops[t] = P
        Pf = np.kron(np.kron(ops[0], ops[1]), ops[2])
        ops[t] = X if r else np.eye(2)
        Xf = np.kron(np.kron(ops[0], ops[1]), ops[2])
        m = Pf @ rho @ Pf
        exp = Xf @ (m / np.trace(m)) @ Xf
        got = orc.reset(rho, [t], [2, 2, 2], u)
        assert np.max(np.abs(got - exp)) < 1e-14
        # the target is in |0>: its reduced state is |0><0|
        red = orc.ptrace(got, [q for q in range(3) if q != t], [2, 2, 2])
        assert np.max(np.abs(red - np.diag([1, 0]))) < 1e-14


# ---------------------------------------------------------------------------
# oracle engine: reset / discard / post-selection
# ---------------------------------------------------------------------------
def test_oracle_engine_reset_and_discard(orc):
    qc = qa.QCircuit(2, 2)
    qc.gate(X, 0).CTRL(X, 0, 1).reset(0).measure([0, 1], 0)
    stats, dits, _, _ = orc.engine_run(qc, reps=20, seed=5)
    assert stats == {(0, 1): 20} and dits == [0, 1]
    # discard q0 of a Bell pair, then measure q1: the same value as the
    # discarded one, either way about half of the time
    qc = qa.QCircuit(2, 1)
    qc.gate(H, 0).CTRL(X, 0, 1).discard(0).measure(1, 0)
    stats, dits, _, st = orc.engine_run(qc, reps=400, seed=6)
    assert set(stats) <= {(0,), (1,)} and 150 < stats.get((0,), 0) < 250
    assert st.shape == (1,) and abs(abs(st[0]) - 1) < 1e-12


def test_oracle_canonical_form_moves_measurements_past_reset(orc):
    """A reset step has no targets for can_swap (qcircuit.hpp:7371-7449), so
    for reps > 1 the canonical form moves measure(0) past reset(0): X, then
    the measurement sees the reset qubit.  reps = 1 keeps program order."""
    qc = qa.QCircuit(1, 1)
    qc.gate(X, 0).measure(0, 0, destructive=False).reset(0)
    stats, dits, _, _ = orc.engine_run(qc, reps=1, seed=1)
    assert stats == {(1,): 1} and dits == [1]
    stats, dits, _, _ = orc.engine_run(qc, reps=5, seed=1)
    assert stats == {(0,): 5} and dits == [0]


def _bell_sampling_expected(orc, seed, reps, ensure, max_ps):
    """The sampling tail (qengine.hpp:414-473) on bell_ps_circuit(1), worked
    out by hand: every attempt draws once and samples (q0, q1) from the Bell
    state (index 0 or 3); the first success runs the steps once -- the
    post-selection drawing until q0 = 1 (ensure on, no limit), then one draw
    for measure q1 -- and adds (1, 1); later successes add (1, 1)."""
    r = orc.Rng(seed)
    p = [0.5, 0, 0, 0.5]
    found, count, rep, att = False, 0, 0, 0
    while rep < reps:
        att += 1
        ok = orc.discrete_draw(p, r.canonical()) == 3
        if not ok:
            if ensure and att < max_ps:
                continue
            rep, att = rep + 1, 0
            continue
        if not found:
            found = True
            while orc.discrete_draw(p, r.canonical()) != 3:
                pass
            r.canonical()
        count += 1
        rep, att = rep + 1, 0
    return ({(1, 1): count} if count else {}), found


@pytest.mark.parametrize("ensure,max_ps", [(False, INF), (True, 2), (True, INF)])
def test_oracle_post_select_sampling_draw_order(orc, ensure, max_ps):
    qc = bell_ps_circuit(1)
    for seed in (1, 2, 3, 4):
        stats, dits, probs, st, ok = orc.engine_run(
            qc, reps=64, seed=seed, ensure_post_selection=ensure,
            max_post_selection_reps=max_ps, with_ps_ok=True)
        exp, found = _bell_sampling_expected(orc, seed, 64, ensure, max_ps)
        assert stats == exp and ok == found
        if found:
            assert dits == [1, 1] and np.allclose(probs, [0.5, 1.0], rtol=0, atol=1e-12)


def test_oracle_post_select_retries_draw_order(orc):
    """Re-run per rep (qengine.hpp:316-345): one draw for the
    post-selection, then -- on success only -- one for the measurement of
    q1, which is 1 (the state is |11>, then X(0): |01>)."""
    qc = bell_ps_circuit(1, tail_gate=True)
    p = [0.5, 0, 0, 0.5]
    for seed in (7, 8):
        stats, dits, _, st, ok = orc.engine_run(qc, reps=40, seed=seed, with_ps_ok=True)
        r = orc.Rng(seed)
        n_ok = 0
        for _ in range(40):
            if orc.discrete_draw(p, r.canonical()) == 3:
                n_ok += 1
                r.canonical()
        assert stats == ({(1, 1): n_ok} if n_ok else {}) and ok == (n_ok > 0)
        assert dits == [1, 1] and st.shape == (2,)
        assert np.max(np.abs(st - [1, 0])) < 1e-12


def test_oracle_post_select_impossible(orc):
    """q0 stays |0>; post-selecting 1 fails every time.  Without ensure the
    failed result is kept (dit 0, state collapsed); with ensure and a limit
    the step gives up and leaves state and dits untouched."""
    qc = qa.QCircuit(2, 2)
    qc.gate(H, 1).post_select(0, 1, 0).measure(1, 1)
    stats, dits, probs, st, ok = orc.engine_run(qc, reps=1, seed=3, with_ps_ok=True)
    assert stats == {} and not ok and dits == [0, 0] and abs(probs[0] - 1) < 1e-12
    assert st.shape == (2,)
    stats, dits, probs, st, ok = orc.engine_run(qc, reps=1, seed=3, ensure_post_selection=True,
                                                max_post_selection_reps=5, with_ps_ok=True)
    assert stats == {} and not ok and probs == [0.0, 0.0] and st.shape == (4,)
    with pytest.raises(RuntimeError):
        orc.engine_run(qc, reps=1, seed=3, ensure_post_selection=True)


# ---------------------------------------------------------------------------
# device: qs_reset and the engine steps against the oracle
# ---------------------------------------------------------------------------
gpu = pytest.mark.gpu


@gpu
def test_reset_vs_oracle(orc):
    rng = np.random.default_rng(11)
    cases = [([2] * 5, [3]), ([2] * 5, [0, 4, 2]), ([2] * 12, [5, 11]), ([3, 3, 3], [2, 0]),
             ([2] * 13, list(range(13)))]
    for dims, target in cases:
        D = int(np.prod(dims))
        for A in (randket(D, rng), randrho(D, rng) if D <= 256 else None):
            if A is None:
                continue
            for u in (0.05, 0.5, 0.93):
                got = qa.reset(A, target, dims, u)
                exp = orc.reset(A, target, dims, u)
                assert np.max(np.abs(got - exp)) <= 1e-12, (dims, target, u)


@gpu
def test_reset_errors():
    psi = randket(6, np.random.default_rng(1))
    with pytest.raises(qa.SubsysMismatchDims):
        qa.reset(psi, [2], [2, 3])
    # the correction is sized by dims[i] (here dims[0] = 2) and qpp's apply
    # rejects it on a qutrit target
    with pytest.raises(qa.MatrixMismatchSubsys):
        qa.reset(psi, [1], [2, 3])
    assert np.max(np.abs(qa.reset(psi, [0], [2, 3], 0.4) -
                         qa.reset(psi, [0], [2, 3], 0.4))) == 0


def random_step_circuit(n, nc, seed, d=2, n_ops=14, tail=True, cctrl=False):
    """Gates, controlled gates, measurements (destructive or not),
    post-selections with random values, resets and discards on live
    qudits, c_regs in range; with cctrl also classically controlled gates
    (plain or fan, with or without shifts)."""
    rng = np.random.default_rng(seed)
    qc = qa.QCircuit(n, nc, d)
    live = list(range(n))
    for _ in range(n_ops):
        kind = rng.choice(["g", "g", "c", "m", "ps", "reset", "discard"] + (["cc", "cc"] if cctrl else []))
        if kind == "cc":
            k = int(rng.integers(1, min(2, nc) + 1))
            dits = [int(x) for x in rng.choice(nc, k, replace=False)]
            shift = [int(rng.integers(0, d)) for _ in dits] if rng.integers(0, 2) else None
            if rng.integers(0, 2) and len(live) > 1:
                t = [int(x) for x in rng.choice(live, 2, replace=False)]
                qc.cCTRL_fan(haar(d, rng), dits, t, shift)
            else:
                qc.cCTRL(haar(d, rng), dits, int(rng.choice(live)), shift)
            continue
        if kind == "c" and len(live) < 2:
            kind = "g"
        if kind == "g":
            qc.gate(haar(d, rng), int(rng.choice(live)))
        elif kind == "c":
            a, b = (int(x) for x in rng.choice(live, 2, replace=False))
            qc.CTRL(Xd(d), a, b)
        elif kind in ("m", "ps"):
            k = int(rng.integers(1, min(2, len(live)) + 1))
            t = [int(x) for x in rng.choice(live, k, replace=False)]
            c = int(rng.integers(0, nc - k + 1))
            dest = bool(rng.integers(0, 2)) and len(live) > k
            if kind == "m":
                qc.measure(t if k > 1 else t[0], c, dest)
            else:
                qc.post_select(t, [int(rng.integers(0, d)) for _ in t], c, dest)
            if dest:
                live = [q for q in live if q not in t]
        elif kind == "reset":
            k = int(rng.integers(1, min(2, len(live)) + 1))
            qc.reset([int(x) for x in rng.choice(live, k, replace=False)])
        elif len(live) > 1:
            q = int(rng.choice(live))
            qc.discard(q)
            live.remove(q)
    if tail:
        qc.measure(live, 0, destructive=False)
    return qc


def _compare(orc, qc, reps, seed, kind=qa.KET, noise_p=-1.0, ensure=False, max_ps=INF):
    e = qa.QEngine(qc, kind=kind, noise_p=noise_p, seed=seed)
    e.set_ensure_post_selection(ensure).set_max_post_selection_reps(max_ps)
    try:
        stats, dits, probs, st, ok = orc.engine_run(
            qc, reps=reps, seed=seed, noise_p=noise_p, is_rho=kind == qa.RHO,
            ensure_post_selection=ensure, max_post_selection_reps=max_ps, with_ps_ok=True)
    except RuntimeError as err:
        # the oracle stopped with qpp's error (QuditAlreadyMeasured for a step
        # the canonical form moved after a discard of its qudit, OutOfRange
        # for a post-selection retried without limit that never succeeds)
        with pytest.raises(qa._BY_CODE[err.code]):
            e.execute(reps)
        return
    noise = orc.last_noise_results(len(qc.steps))
    e.execute(reps)
    assert e.get_stats() == stats
    assert e.get_dits() == dits
    assert e.post_select_ok() == ok
    assert e.get_noise_results() == noise
    assert np.max(np.abs(np.array(e.get_probs()) - np.array(probs))) <= 1e-12
    got = e.get_state().download()
    assert got.shape == st.shape and np.max(np.abs(got - st)) <= 1e-12


@gpu
@pytest.mark.parametrize("ensure,max_ps", [(False, INF), (True, 2), (True, INF)])
def test_engine_bell_post_selection_vs_oracle(orc, ensure, max_ps):
    for tail in (False, True):
        for seed in (1, 2, 3):
            _compare(orc, bell_ps_circuit(1, tail), 64, seed, ensure=ensure, max_ps=max_ps)


@gpu
@pytest.mark.parametrize("seed", range(12))
def test_engine_random_steps_vs_oracle(orc, seed):
    n = 3 + seed % 4
    qc = random_step_circuit(n, n, seed)
    for reps in (1, 30):
        for ensure, max_ps in ((False, INF), (True, 3)):
            _compare(orc, qc, reps, 100 + seed, ensure=ensure, max_ps=max_ps)


@gpu
def test_engine_steps_large_and_rho_and_noise(orc):
    # 13 qubits: gate runs between the steps go through fused tile programs
    qc = random_step_circuit(13, 13, 77, n_ops=18)
    _compare(orc, qc, 1, 5)
    _compare(orc, qc, 6, 6, ensure=True, max_ps=2)
    qc = random_step_circuit(4, 4, 78)
    _compare(orc, qc, 10, 7, kind=qa.RHO)
    _compare(orc, qc, 5, 8, noise_p=0.1)
    _compare(orc, qc, 5, 9, noise_p=0.1, ensure=True, max_ps=2)
    # qutrits: resets shift by Xd powers
    qc = random_step_circuit(3, 3, 79, d=3)
    _compare(orc, qc, 1, 10)
    _compare(orc, qc, 12, 11)


@gpu
def test_engine_post_select_impossible(orc):
    qc = qa.QCircuit(2, 2)
    qc.gate(H, 1).post_select(0, 1, 0).measure(1, 1)
    _compare(orc, qc, 1, 3)
    _compare(orc, qc, 1, 3, ensure=True, max_ps=5)
    e = qa.QEngine(qc, seed=3).set_ensure_post_selection(True)
    with pytest.raises(qa.OutOfRange):
        e.execute(1)
    # the sampling tail (reps > 1, only measurements after the gates) gives
    # up the same way instead of drawing for ever
    qs = qa.QCircuit(2, 2)
    qs.gate(H, 1).post_select(0, 1, 0).measure(1, 1)
    with pytest.raises(RuntimeError):
        orc.engine_run(qs, reps=4, seed=3, ensure_post_selection=True)
    e = qa.QEngine(qs, seed=3).set_ensure_post_selection(True)
    with pytest.raises(qa.OutOfRange):
        e.execute(4)


# ---------------------------------------------------------------------------
# classically controlled gates (qengine.hpp:581-630)
# ---------------------------------------------------------------------------
def teleport_cctrl(U):
    """q0 = U|0> teleported to q2 with cCTRL corrections X^{m1} then Z^{m0}."""
    Z = np.diag([1, -1])
    qc = qa.QCircuit(3, 2)
    qc.gate(U, 0).gate(H, 1).CTRL(X, 1, 2).CTRL(X, 0, 1).gate(H, 0)
    qc.measure(0, 0).measure(1, 1)
    return qc.cCTRL(X, 1, 2).cCTRL(Z, 0, 2)


def test_oracle_cctrl_teleportation(orc):
    rng = np.random.default_rng(9)
    U = haar(2, rng)
    psi = U @ [1, 0]
    for seed in range(6):
        for reps in (1, 5):
            stats, dits, _, st = orc.engine_run(teleport_cctrl(U), reps=reps, seed=seed)
            assert st.shape == (2,) and abs(abs(np.vdot(psi, st)) - 1) < 1e-12
            assert sum(stats.values()) == reps


def test_oracle_cctrl_powers_and_shifts(orc):
    """Qutrits: q0 = |2> measured into c0, then cCTRL(Xd, c0) on q1 applies
    Xd^k with k = (2 + shift) mod 3; two dits that disagree apply nothing."""
    for shift, k in ((None, 2), ([0], 2), ([1], 0), ([2], 1)):
        qc = qa.QCircuit(2, 2, 3)
        qc.gate(Xd(3) @ Xd(3), 0).measure(0, 0, destructive=False)
        qc.cCTRL(Xd(3), [0], 1, shift)
        _, dits, _, st = orc.engine_run(qc, reps=1, seed=1)
        exp = np.zeros(9)
        exp[2 * 3 + k] = 1
        assert dits == [2, 0] and np.max(np.abs(st - exp)) < 1e-12
    qc = qa.QCircuit(2, 2, 3)
    qc.gate(Xd(3), 0).measure(0, 0, destructive=False).cCTRL(Xd(3), [0, 1], 1)
    _, dits, _, st = orc.engine_run(qc, reps=1, seed=1)
    exp = np.zeros(9)
    exp[1 * 3] = 1
    assert dits == [1, 0] and np.max(np.abs(st - exp)) < 1e-12


@gpu
def test_engine_cctrl_vs_oracle(orc):
    rng = np.random.default_rng(9)
    U = haar(2, rng)
    for seed in range(4):
        _compare(orc, teleport_cctrl(U), 1, seed)
        _compare(orc, teleport_cctrl(U), 20, seed)
    for seed in range(10):
        n = 3 + seed % 4
        qc = random_step_circuit(n, n, 500 + seed, cctrl=True)
        _compare(orc, qc, 1, seed)
        _compare(orc, qc, 25, seed, ensure=True, max_ps=2)
    qc = random_step_circuit(13, 13, 600, n_ops=20, cctrl=True)
    _compare(orc, qc, 3, 1)
    qc = random_step_circuit(3, 3, 601, d=3, cctrl=True)
    _compare(orc, qc, 1, 2)
    _compare(orc, qc, 10, 3)
    _compare(orc, random_step_circuit(4, 4, 602, cctrl=True), 6, 4, noise_p=0.1)


# ---------------------------------------------------------------------------
# runtime while loops (qengine.hpp:837-880)
# ---------------------------------------------------------------------------
THETA = 0.37


def ry(t):
    return np.array([[np.cos(t / 2), -np.sin(t / 2)], [np.sin(t / 2), np.cos(t / 2)]])


def until_one(extra_if=False):
    """while c0 == 0: H q0; measure q0 -> c0 (non-destructive); Ry(THETA) q1.
    q1 counts the iterations; with extra_if, X q2 when c0 == 1 inside."""
    qc = qa.QCircuit(3 if extra_if else 2, 1)
    qc.cond_while([0], [0]).gate(H, 0).measure(0, 0, destructive=False).gate(ry(THETA), 1)
    if extra_if:
        qc.cond_if([0], [1]).gate(X, 2).cond_end()
    return qc.cond_end()


def test_oracle_while_loop_iterations(orc):
    """Each iteration measures q0 of H|x> with one draw; q0 = 1 exactly
    when u >= 1/2 (the first two basis indices carry q0 = 0 and probability
    1/2 together).  The loop ends at the first such draw, after k
    iterations: the state is |1> (x) Ry(k THETA)|0> up to a sign."""
    for seed in range(10):
        r = orc.Rng(seed)
        k = 1
        while r.canonical() < 0.5:
            k += 1
        _, dits, _, st = orc.engine_run(until_one(), reps=1, seed=seed)
        exp = np.kron([0, 1], ry(k * THETA) @ [1, 0])
        assert dits == [1] and abs(abs(np.vdot(exp, st)) - 1) < 1e-12


def _loop_runs(orc, seed):
    """Body runs of until_one() for this seed: 1 + the draws below 1/2
    before the first one at or above it."""
    r = orc.Rng(seed)
    k = 1
    while r.canonical() < 0.5:
        k += 1
    return k


def _ps_attempts(orc, seed):
    """Attempts until post-selecting q0 = 1 of H|0> succeeds: the first
    draw at or above 1/2 (one draw per attempt)."""
    r = orc.Rng(seed)
    k = 1
    while r.canonical() < 0.5:
        k += 1
    return k


def post_select_one():
    qc = qa.QCircuit(2, 1)
    return qc.gate(H, 0).post_select(0, 1, 0, destructive=False).gate(ry(THETA), 1)


def test_oracle_give_up_counts_are_exact(orc):
    """The while-loop limit L stops a run whose loop body has run L times
    with its condition still true: a loop that needs k runs passes at
    L = k and stops with OutOfRange at L = k - 1.  An unlimited
    post-selection that succeeds at attempt k has k - 1 failures first: it
    passes with a failure limit P = k and stops at P = k - 1.  0 removes a
    limit."""
    done = 0
    for seed in range(40):
        k = _loop_runs(orc, seed)
        if k < 2:
            continue
        _, dits, _, st = orc.engine_run(until_one(), reps=1, seed=seed, max_loop_reps=k)
        exp = np.kron([0, 1], ry(k * THETA) @ [1, 0])
        assert dits == [1] and abs(abs(np.vdot(exp, st)) - 1) < 1e-12
        orc.engine_run(until_one(), reps=1, seed=seed, max_loop_reps=0)
        with pytest.raises(RuntimeError) as ex:
            orc.engine_run(until_one(), reps=1, seed=seed, max_loop_reps=k - 1)
        assert ex.value.code == qa.OutOfRange.code
        k = _ps_attempts(orc, seed)
        _, dits, _, _, ok = orc.engine_run(post_select_one(), reps=1, seed=seed,
                                           ensure_post_selection=True, max_ps_failures=k,
                                           with_ps_ok=True)
        assert ok and dits == [1]
        orc.engine_run(post_select_one(), reps=1, seed=seed, ensure_post_selection=True,
                       max_ps_failures=0)
        with pytest.raises(RuntimeError) as ex:
            orc.engine_run(post_select_one(), reps=1, seed=seed, ensure_post_selection=True,
                           max_ps_failures=k - 1)
        assert ex.value.code == qa.OutOfRange.code
        done += 1
    assert done >= 10


def test_oracle_while_rejects_destructive_steps(orc):
    bad = qa.QCircuit(2, 1).cond_while([0], [0]).gate(H, 0).measure(0, 0).cond_end()
    with pytest.raises(RuntimeError):
        orc.engine_run(bad, reps=1)
    bad = qa.QCircuit(2, 1).cond_while([0], [0]).discard(0).cond_end()
    with pytest.raises(RuntimeError):
        orc.engine_run(bad, reps=1)


@gpu
def test_engine_while_vs_oracle(orc):
    for seed in range(6):
        _compare(orc, until_one(), 1, seed)
        _compare(orc, until_one(extra_if=True), 12, seed)
        _compare(orc, until_one(extra_if=True), 5, seed, noise_p=0.05)
    bad = qa.QCircuit(2, 1).cond_while([0], [0]).gate(H, 0).measure(0, 0).cond_end()
    with pytest.raises(qa.InvalidConditional):
        qa.QEngine(bad)
    # a loop that can never end (no measurement in its body) stops with an
    # error after 2^24 repetitions instead of running for ever
    forever = qa.QCircuit(1, 1).cond_while([0], [0]).nop().cond_end()
    with pytest.raises(qa.OutOfRange):
        qa.QEngine(forever).execute(1)


@gpu
def test_engine_give_up_counts_vs_oracle(orc):
    """The device engine's settable limits cut off exactly where the
    oracle's do (see test_oracle_give_up_counts_are_exact)."""
    checked = 0
    for seed in range(40):
        k = _loop_runs(orc, seed)
        if k < 2:
            continue
        e = qa.QEngine(until_one(), seed=seed).set_limits(k, 1 << 16).execute(1)
        assert e.get_dits() == [1]
        with pytest.raises(qa.OutOfRange):
            qa.QEngine(until_one(), seed=seed).set_limits(k - 1, 1 << 16).execute(1)
        k = _ps_attempts(orc, seed)
        e = qa.QEngine(post_select_one(), seed=seed).set_ensure_post_selection(True)
        e.set_limits(1 << 24, k).execute(1)
        assert e.get_dits() == [1] and e.post_select_ok()
        e = qa.QEngine(post_select_one(), seed=seed).set_ensure_post_selection(True)
        with pytest.raises(qa.OutOfRange):
            e.set_limits(1 << 24, k - 1).execute(1)
        checked += 1
    assert checked >= 10


# ---------------------------------------------------------------------------
# noise results (qnoisy_engine.hpp:104-121, 174-176)
# ---------------------------------------------------------------------------
def test_oracle_noise_results_follow_the_stream(orc):
    """H(0) then measure both qubits: before each step one draw per live
    qubit (k = 0 I, 1 X, 2 Y, 3 Z from {1-p, p/3, p/3, p/3}), then one draw
    for the measurement; the results accumulate over the reps."""
    qc = qa.QCircuit(2, 2)
    qc.gate(H, 0).measure([0, 1], 0)
    p = 0.6
    w = [1 - p, p / 3, p / 3, p / 3]
    for seed in (1, 2, 3):
        orc.engine_run(qc, reps=3, noise_p=p, seed=seed)
        got = orc.last_noise_results(2)
        r = orc.Rng(seed)
        exp = [[], []]
        for _ in range(3):
            exp[0] += [orc.discrete_draw(w, r.canonical()) for _ in range(2)]
            exp[1] += [orc.discrete_draw(w, r.canonical()) for _ in range(2)]
            r.canonical()
        assert got == exp


@gpu
def test_engine_noise_results_and_reset(orc):
    qc = random_step_circuit(5, 5, 900, cctrl=True)
    _compare(orc, qc, 7, 3, noise_p=0.2)
    e = qa.QEngine(qc, noise_p=0.2, seed=3).execute(2)
    assert sum(len(v) for v in e.get_noise_results()) > 0
    e.reset()
    assert all(v == [] for v in e.get_noise_results())


@gpu
def test_engine_bookkeeping():
    """get_measured_d / get_non_measured_d / was_measured_d, set_dit(s),
    get_dit and reset_stats (qengine.hpp:1035-1164), through the ctypes
    engine and the pyqpp module."""
    from qpp_amd import pyqpp
    for mod in (qa, pyqpp):
        qc = mod.QCircuit(3, 2)
        qc.gate(X, 1).measure(1, 0).measure(2, 1, False)
        e = mod.QEngine(qc, seed=4).execute(5)
        assert e.get_measured_d() == [1] and e.get_non_measured_d() == [0, 2]
        assert e.was_measured_d(1) and not e.was_measured_d(0)
        assert e.get_dits() == [1, 0] and e.get_dit(0) == 1
        e.set_dits([0, 1])
        assert e.get_dits() == [0, 1]
        e.set_dit(0, 1)
        assert e.get_dits() == [1, 1]
        with pytest.raises((qa.SizeMismatch, RuntimeError)):
            e.set_dits([1])
        assert e.get_stats() == {(1, 0): 5}
        e.reset_stats()
        assert e.get_stats() == {}

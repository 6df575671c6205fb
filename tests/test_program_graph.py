"""Programs replayed as HIP graphs (qs_program_capture): the pass launches
of a program captured once for one state, then replayed by qs_program_run
on that state.  Every replay is checked against the same program run
directly on a copy, with the program's own eager work and host
synchronisations between the replays; other states, a new capture and a
recompile fall back to direct runs."""
import numpy as np
import pytest

from helpers import randket
from test_gpu_fused import maxabs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("native", [False, True])
def test_graph_replays_equal_direct_runs(qa, native):
    from qpp_amd import workloads
    n = 16
    qc = qa.QCircuit.from_gates(n, workloads.random_layers(n, 4, seed=21))
    p = qa.Program(qc)
    if native:
        p.compile_native()
    psi = randket(2 ** n, np.random.default_rng(22))
    eager = qa.State.from_array(psi, [2] * n)
    st = qa.State.from_array(psi, [2] * n)
    p.run(eager)
    p.capture(st)  # runs once while capturing
    print(f"graph native={native}: captured", flush=True)
    assert maxabs(st.download(), eager.download()) <= 1e-12
    other = qa.State.from_array(psi, [2] * n)  # not the captured buffer: a direct run
    for rep in range(4):
        p.run(eager)
        p.run(st)  # replay
        p.run(other)
        qa.synchronize()
        ref = eager.download()
        assert maxabs(st.download(), ref) <= 1e-12, rep
        # eager work on the captured state between replays
        st.upload(ref)
        print(f"graph native={native}: replay {rep} ok", flush=True)
    # `other` has seen one run fewer than `eager`
    p.run(other)
    assert maxabs(other.download(), eager.download()) <= 1e-12


def test_graph_dropped_by_recompile_and_recapture(qa):
    from qpp_amd import workloads
    n = 14
    qc = qa.QCircuit.from_gates(n, workloads.random_layers(n, 3, seed=23))
    p = qa.Program(qc)
    psi = randket(2 ** n, np.random.default_rng(24))
    a = qa.State.from_array(psi, [2] * n)
    b = qa.State.from_array(psi, [2] * n)
    ref = qa.State.from_array(psi, [2] * n)
    p.capture(a)
    p.run(ref)
    p.compile_native()  # drops the graph (it held the general kernel's launches)
    p.run(a)
    p.run(ref)
    p.capture(b)  # b: run once while capturing
    p.run(b)      # replay
    qa.synchronize()
    assert maxabs(a.download(), ref.download()) <= 1e-12
    assert maxabs(b.download(), ref.download()) <= 1e-12

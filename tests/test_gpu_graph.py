"""PC steps replayed from captured hipGraphs (HETDDFT_GRAPH=1: both streams
in the graph; 2: the particle chain on the main stream) against the same
steps launched directly (the default): bit-identical state after every
step, over several cycles of the slot states (the rotating psi, H psi and
particle slots return to the same buffers every two steps), with state reads
(and their synchronisation) between the replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _make(hip, wl, precision):
    import inputs
    p = inputs.params_for(wl, precision=precision)
    g = hip.Instance(p)
    g.set_kernels(hip.make_kernels(p))
    psi, r, v = inputs.initial_state(wl)
    g.set_state(psi, r, v)
    g.start()
    return g


@pytest.mark.parametrize("mode", ["1", "2"])
@pytest.mark.parametrize("name", ["t32", "c1"])
@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replay_matches_direct(hip, name, precision, mode, monkeypatch):
    import inputs
    wl = inputs.WORKLOADS[name]
    print(f"graph (mode {mode}) vs direct {name} precision={precision}", flush=True)
    monkeypatch.delenv("HETDDFT_GRAPH", raising=False)
    a = _make(hip, wl, precision)
    monkeypatch.setenv("HETDDFT_GRAPH", mode)
    b = _make(hip, wl, precision)
    for it in range(3 + 14):
        want = it % 3 == 0
        ea = a.iterate(wl.deltat, want_errors=want)
        eb = b.iterate(wl.deltat, want_errors=want)
        if want:
            assert np.array_equal(np.asarray(ea), np.asarray(eb)), it
        for x, y in zip(a.get_state(), b.get_state()):
            assert np.array_equal(x, y), it
    # the replays above really came from graphs, one per slot state
    assert a.graph_info() == (0, False)
    assert b.graph_info() == (2, True)
    for f in ("den", "pot4", "denx", "uimp"):
        assert np.array_equal(a.get_field(f), b.get_field(f)), f
    assert a.energy() == b.energy()
    a.close()
    b.close()

"""Several instances advanced in interleaved rounds (the ensemble mode of the
benchmark: each instance has its own streams, nothing is synchronised between
rounds) end in exactly the state each reaches when it runs alone."""
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


@pytest.mark.parametrize("precision", [0, 1])
def test_interleaved_instances_match_solo(hip, precision):
    import inputs
    wls = [inputs.WORKLOADS["t24"], inputs.WORKLOADS["t32"],
           inputs.Workload("e48", 48, 0.45, 17, 500, 0.5, v0=1.2, vdir=(0.6, 0.0, 0.8), seed=41)]
    nsteps = 9
    print(f"ensemble interleaved vs solo, precision={precision}", flush=True)
    solo = []
    for wl in wls:
        g = _make(hip, wl, precision)
        for _ in range(nsteps):
            g.iterate(wl.deltat, want_errors=False)
        solo.append(g.get_state())
        g.close()
    insts = [_make(hip, wl, precision) for wl in wls]
    for _ in range(nsteps):
        for g, wl in zip(insts, wls):
            g.iterate(wl.deltat, want_errors=False)
    for g, ref in zip(insts, solo):
        got = g.get_state()
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
        g.close()

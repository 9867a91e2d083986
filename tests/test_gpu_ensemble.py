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


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_executor_matches_solo(hip, precision):
    """hetddft_batch_iterate (the library's host threads drive the instances)
    ends in the solo states, with the solo PC error estimates."""
    import inputs
    wls = [inputs.WORKLOADS["t24"], inputs.WORKLOADS["t32"],
           inputs.Workload("e48", 48, 0.45, 17, 500, 0.5, v0=1.2, vdir=(0.6, 0.0, 0.8), seed=41)]
    print(f"batch executor vs solo, precision={precision}", flush=True)
    solo, solo_err = [], []
    for wl in wls:
        g = _make(hip, wl, precision)
        for _ in range(9):
            g.iterate(wl.deltat, want_errors=False)
        solo.append(g.get_state())
        solo_err.append(g.last_errors())
        g.close()
    insts = [_make(hip, wl, precision) for wl in wls]
    b = hip.Batch(insts, [wl.deltat for wl in wls], nthreads=2)
    e = b.iterate(2, want_errors=True)
    assert np.array_equal(e, np.zeros((3, 3)))          # still in the RK start-up
    b.iterate(4)
    err = b.iterate(3, want_errors=True)
    for g, ref, ref_err, got_err in zip(insts, solo, solo_err, err):
        for x, y in zip(g.get_state(), ref):
            assert np.array_equal(x, y)
        assert np.array_equal(np.asarray(ref_err), got_err)
    b.close()
    for g in insts:
        g.close()


def test_batch_executor_reports_the_failing_instance(hip):
    """An instance that was never started fails its first iteration
    (HETDDFT_E_STATE, raised on the host); the batch names it."""
    import inputs
    wls = [inputs.WORKLOADS["t24"], inputs.WORKLOADS["t32"], inputs.WORKLOADS["t24"]]
    insts = [_make(hip, wls[0], 0)]
    p = inputs.params_for(wls[1], precision=0)
    idle = hip.Instance(p)
    idle.set_kernels(hip.make_kernels(p))
    insts.append(idle)
    insts.append(_make(hip, wls[2], 0))
    b = hip.Batch(insts, [wl.deltat for wl in wls], nthreads=3)
    with pytest.raises(hip.HetddftError) as ei:
        b.iterate(3)
    assert ei.value.args[1] == -4 and ei.value.args[2] == 1
    assert "instance 1" in ei.value.args[0]
    b.close()
    for g in insts:
        g.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_shares_launches_across_instances_of_one_grid(hip, precision):
    """Instances of one grid and parameters advance in shared launches (one
    launch per kernel for all of them); each ends bit-identical to its solo
    run, also when their particle-chain buffer sets are out of phase (one
    instance re-ran its potential evaluation) and their droplets differ."""
    import inputs
    base = inputs.WORKLOADS["t32"]
    wls = [inputs.Workload(f"g{i}", base.n, base.h, 12 + 4 * i, base.n_par, 0.5, v0=0.5 * i,
                           vdir=(0.0, 0.6, 0.8), phase_noise=0.01, seed=60 + i) for i in range(3)]
    print(f"shared launches across 3 instances of 32^3, precision={precision}", flush=True)
    solo = []
    for i, wl in enumerate(wls):
        g = _make(hip, wl, precision)
        if i == 1:
            g.potential()
        for _ in range(9):
            g.iterate(wl.deltat, want_errors=False)
        solo.append((g.get_state(), g.last_errors(), g.energy()))
        g.close()
    insts = [_make(hip, wl, precision) for wl in wls]
    insts[1].potential()
    b = hip.Batch(insts, [wl.deltat for wl in wls], nthreads=1)
    b.iterate(2)
    err = b.iterate(7, want_errors=True)
    for g, (ref, ref_err, ref_e), got_err in zip(insts, solo, err):
        for x, y in zip(g.get_state(), ref):
            assert np.array_equal(x, y)
        assert np.array_equal(np.asarray(ref_err), got_err)
        assert g.energy() == ref_e
    b.close()
    for g in insts:
        g.close()

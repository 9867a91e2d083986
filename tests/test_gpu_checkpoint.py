"""Full-history checkpoint: a context restored from an image continues
bit-identically to the context that wrote it, from inside the RK start-up
and from the predictor-corrector phase, in both precisions."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _make(hip, wl, precision):
    import inputs
    p = inputs.params_for(wl, precision=precision)
    g = hip.Instance(p)
    g.set_kernels(hip.make_kernels(p))
    return g


@pytest.mark.parametrize("save_after", [2, 6])
@pytest.mark.parametrize("precision", [0, 1])
def test_resume_is_bit_exact(hip, precision, save_after):
    import inputs
    wl = inputs.WORKLOADS["t32"]
    print(f"checkpoint resume precision={precision} after {save_after}", flush=True)
    a = _make(hip, wl, precision)
    psi, r, v = inputs.initial_state(wl)
    a.set_state(psi, r, v)
    a.start()
    for _ in range(save_after):
        a.iterate(wl.deltat, want_errors=False)
    image = a.checkpoint()
    b = _make(hip, wl, precision)
    b.restore(image)
    for f in ("den", "pot4", "denx", "uimp"):
        assert np.array_equal(a.get_field(f), b.get_field(f)), f
    for it in range(7):
        ea = a.iterate(wl.deltat)
        eb = b.iterate(wl.deltat)
        assert np.array_equal(np.asarray(ea), np.asarray(eb)), it
        for x, y in zip(a.get_state(), b.get_state()):
            assert np.array_equal(x, y), it
    assert a.energy() == b.energy()
    a.close()
    b.close()


def test_restore_rejects_other_shapes(hip):
    import inputs
    wl = inputs.WORKLOADS["t24"]
    a = _make(hip, wl, 0)
    psi, r, v = inputs.initial_state(wl)
    a.set_state(psi, r, v)
    a.start()
    image = a.checkpoint()
    b = _make(hip, inputs.WORKLOADS["t32"], 0)
    with pytest.raises(hip.HetddftError):
        b.restore(image)
    with pytest.raises(hip.HetddftError):
        _make(hip, wl, 0).restore(b"XXXXXXXX" + image[8:])
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_restart_on_a_used_context_equals_a_new_one(hip, precision):
    """set_state + start on a context that has run PC steps restarts exactly
    like a new context: pc, q, pcr, pcv and the histories start at zero as
    after allocation in the reference (dimen.f90:78-79,131-134)."""
    import inputs
    wl = inputs.WORKLOADS["t24"]
    print(f"restart on a used context, precision {precision}", flush=True)
    psi, r, v = inputs.initial_state(wl)
    a = _make(hip, wl, precision)
    a.set_state(psi, r, v)
    a.start()
    for _ in range(6):
        a.iterate(wl.deltat)
    psi2, r2, v2, _ = a.get_state()
    a.set_state(psi2, r2, v2)
    a.start()
    assert a.last_errors() == (0.0, 0.0, 0.0)
    b = _make(hip, wl, precision)
    b.set_state(psi2, r2, v2)
    b.start()
    for it in range(6):
        ea = a.iterate(wl.deltat)
        eb = b.iterate(wl.deltat)
        assert np.array_equal(np.asarray(ea), np.asarray(eb)), it
        for x, y in zip(a.get_state(), b.get_state()):
            assert np.array_equal(x, y), it
    a.close()
    b.close()


def test_errors_after_restore_belong_to_the_restored_state(hip):
    """A restore drops the error estimates of the steps run before it: right
    after a load (or an autotune, which ends with one) there are none."""
    import inputs
    wl = inputs.WORKLOADS["t24"]
    psi, r, v = inputs.initial_state(wl)
    a = _make(hip, wl, 0)
    a.set_state(psi, r, v)
    a.start()
    for _ in range(5):
        a.iterate(wl.deltat)
    image = a.checkpoint()
    ea = a.iterate(wl.deltat)
    a.iterate(wl.deltat, want_errors=False)
    a.restore(image)
    assert a.last_errors() == (0.0, 0.0, 0.0)
    assert np.array_equal(np.asarray(a.iterate(wl.deltat)), np.asarray(ea))
    a.autotune(wl.deltat, steps=1)
    assert a.last_errors() == (0.0, 0.0, 0.0)
    a.close()

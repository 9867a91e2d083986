"""C-ABI surface checks that need no GPU: symbols, defaults, argument errors, byte model."""
import ctypes

import numpy as np
import pytest


def test_library_exports_every_header_symbol(hip):
    lib = hip.load_library()
    names = hip.header_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n


def test_default_params_are_orsay_trento(hip):
    p = hip.default_params()
    assert (p.cp4, p.cpp4, p.alphas, p.den0s) == (-2.41186e4, 1.85850e6, 54.31, 0.04)
    assert (p.h2o2m4, p.C, p.beta, p.den_m, p.h4) == (6.05969638298, 3.1577504e4, 40.0, 0.37, 2.190323)
    assert (p.eps4, p.sigma4, p.core_otc, p.ironx) == (10.22, 2.556, 1, 3)


@pytest.mark.parametrize("bad", [dict(nx=7), dict(nx=14), dict(ny=257), dict(n_par=-1), dict(n_par=2**27 + 1),
                                 dict(precision=3), dict(mimpur=0.0)])
def test_create_rejects_bad_parameters(hip, bad):
    p = hip.default_params(**bad)
    lib = hip.load_library()
    c = ctypes.c_void_p()
    rc = lib.hetddft_create(ctypes.byref(p), ctypes.byref(c))
    assert rc == -1
    assert c.value is None
    assert lib.hetddft_last_error(None)


def test_null_arguments(hip):
    lib = hip.load_library()
    assert lib.hetddft_create(None, None) == -1
    assert lib.hetddft_energy(None, None) == -1
    assert lib.hetddft_destroy(None) == 0


def test_model_bytes():
    """Stage-model algorithmic bytes per PC step (BASELINE model ceilings)."""
    import bench
    assert abs(bench.model_bytes(128, 10000, 0) / 1e9 - 3.98) < 0.01
    assert abs(bench.model_bytes(128, 10000, 1) / 1e9 - 2.30) < 0.01
    assert abs(bench.model_bytes(64, 1000, 0) / 1e9 - 0.50) < 0.01
    assert abs(bench.model_bytes(64, 1000, 1) / 1e9 - 0.29) < 0.01
    assert abs(bench.model_bytes(256, 100000, 0) / 1e9 - 31.8) < 0.1


def test_inputs_are_normalised_and_seeded():
    import inputs
    wl = inputs.WORKLOADS["t32"]
    psi, r, v = inputs.initial_state(wl)
    norm = np.sum(np.abs(psi) ** 2) * wl.h ** 3
    assert abs(norm - wl.n_he) < 1e-9 * wl.n_he
    psi2, r2, v2 = inputs.initial_state(wl)
    assert np.array_equal(psi, psi2) and np.array_equal(r, r2)
    assert np.allclose(v[2], wl.v0 * inputs.KTOPS) and np.all(v[:2] == 0)


def test_ensemble_workloads():
    """Ensemble of uneven instances: sizes in turn, N in range and clipped so
    the droplet stays 5 A inside its box, speeds 0-300 m/s, deterministic."""
    import inputs
    a = inputs.ensemble_c4()
    b = inputs.ensemble_c4()
    assert len(a) == 32
    assert [w.n for w in a[:4]] == [48, 64, 80, 96]
    for w in a:
        assert w.n_par == 2000 and w.h == 0.45
        assert 2.22 * w.n_he ** (1.0 / 3.0) + 5.0 < w.n * w.h / 2.0
        assert w.n_he <= 300 and (w.n_he >= 50 or w.n == 48)
        assert 0.0 <= w.v0 <= 3.0
        assert abs(sum(x * x for x in w.vdir) - 1.0) < 1e-12
    assert [(w.n_he, w.v0, w.vdir) for w in a] == [(w.n_he, w.v0, w.vdir) for w in b]


def test_batch_arguments(hip):
    lib = hip.load_library()
    b = ctypes.c_void_p()
    assert lib.hetddft_batch_create(None, 0, 1, ctypes.byref(b)) == -1
    assert b.value is None
    assert lib.hetddft_batch_create(None, 1, 1, None) == -1
    failed = ctypes.c_int(7)
    assert lib.hetddft_batch_iterate(None, None, 1, None, ctypes.byref(failed)) == -1
    assert failed.value == -1
    assert lib.hetddft_batch_last_error(None) == b""
    assert lib.hetddft_batch_destroy(None) == 0

"""hetddft_particle_means_dev, the entry point that allocates nothing, gives
the means of hetddft_particle_means bit for bit (two calls in a row, the
second with fewer particles), called through the C-ABI on device arrays."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _instance(hip, name, n_par=None, seed=0):
    import inputs
    wl = inputs.WORKLOADS[name]
    p = inputs.params_for(wl)
    psi, r, v = inputs.initial_state(wl)
    if n_par is not None:
        p.n_par = n_par
        g = np.random.default_rng(seed)
        r = 0.8 * g.standard_normal((3, n_par))
        v = g.standard_normal((3, n_par))
    inst = hip.Instance(p)
    inst.set_kernels(hip.make_kernels(p))
    inst.set_state(psi, r, v)
    inst.start()
    for _ in range(4):
        inst.iterate(wl.deltat, want_errors=False)
    return inst


class _Hip:
    """The HIP runtime the library runs on (the caller's own device buffers)."""

    def __init__(self, lib_path):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def download(self, p, n):
        out = np.zeros(n)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _means_dev(hip, r, v, a):
    lib = hip.load_library()
    h = _Hip(hip.lib_path())
    dr, dv, da = h.upload(r), h.upload(v), h.upload(a)
    out = h.upload(np.full(9, np.nan))
    rc = lib.hetddft_particle_means_dev(dr, dv, da, r.shape[1], out, None)
    got = h.download(out, 9)
    h.free()
    return rc, got


def test_means_without_allocation_equal_the_context_means(hip):
    # 20000 particles fill the 64 partial sums per component; 128 take one
    for name, n_par in (("t24", 20000), ("t24", None)):
        print(f"particle_means_dev: {name}, n_par {n_par}", flush=True)
        g = _instance(hip, name, n_par)
        want = g.particle_means()
        _, r, v, a = g.get_state()
        rc, got = _means_dev(hip, r, v, a)
        assert rc == 0, hip.load_library().hetddft_last_error(None)
        assert np.array_equal(got, want), (got, want)
        g.close()


def test_means_without_allocation_rejects_bad_sizes(hip):
    lib = hip.load_library()
    z = np.zeros(9)
    rc = lib.hetddft_particle_means_dev(None, None, None, -1, z.ctypes.data, None)
    assert rc == hip.E_ARG and b"n_par" in lib.hetddft_last_error(None)
    rc = lib.hetddft_particle_means_dev(None, None, None, (1 << 30) + 1, z.ctypes.data, None)
    assert rc == hip.E_ARG and b"n_par" in lib.hetddft_last_error(None)
    rc = lib.hetddft_particle_means_dev(None, None, None, 5, None, None)
    assert rc == hip.E_ARG

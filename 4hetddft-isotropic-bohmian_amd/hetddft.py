"""Host-side mirror of the reference's per-step interface over the HIP C-ABI.

The reference Fortran driver calls ``steprk(deltat)``, ``steppc(deltat, errHe,
errimp, errvimp)``, ``potenimp(); poten(); forceimp()`` and ``energy()``
(DFT4He3d.f90:632-634,653,849-863,882) on module-global arrays.  ``Instance``
exposes the same operations by the same names on one GPU-resident instance;
every call goes through ``libhetddft.so`` (include/hetddft.h).  There is no
CPU fallback: if the library is missing, loading fails with an error.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_NAME = "libhetddft.so"

FP64 = 0
MIXED = 1

FIELDS = {
    "den": 0, "pot4": 1, "uimp": 2, "dencg": 3, "delj4": 4, "falfs": 5,
    "dxden": 6, "dyden": 7, "dzden": 8, "intxalf": 9, "intyalf": 10, "intzalf": 11,
    "denx": 12, "psix": 13, "hpsix": 14, "fxconv": 15, "fyconv": 16, "fzconv": 17,
    "correl": 18, "ualphas_conv": 19,
}

PROF_CLASSES = {
    "fft_xy_forward": 0, "fft_x_forward": 1, "fft_y_forward": 2, "fft_z_kernel": 3,
    "fft_y_inverse": 4, "fft_x_inverse": 5, "fft_xy_inverse": 6, "pot4_assembly": 7,
    "pc_predict": 8, "pc_correct": 9, "grad_den": 10,
}

ENERGY_KEYS = ("etot", "etot4", "ekin4", "elj4", "ealphas", "esolid", "ecor4", "eimpu",
               "ekinx", "ekinqx", "ehex", "norm")

KTOPS = 7.63822291          # DFT4He3d.f90:110
PSTOK = 1.0 / 7.63822291


E_ARG, E_HIP, E_NOMEM, E_STATE, E_NONFINITE = -1, -2, -3, -4, -5   # HETDDFT_E_*


class HetddftError(RuntimeError):
    """A failed library call: args = (message[, HETDDFT_E_* code[, instance]]);
    ``code`` and ``instance`` name the last two (None when absent)."""

    def __init__(self, msg, *extra):
        super().__init__(msg, *extra)
        self.code = extra[0] if extra else None
        self.instance = extra[1] if len(extra) > 1 else None

    def __str__(self):
        return str(self.args[0])


class Params(ctypes.Structure):
    """Mirror of ``hetddft_params`` (include/hetddft.h)."""
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("xmax", ctypes.c_double), ("ymax", ctypes.c_double), ("zmax", ctypes.c_double),
        ("n_par", ctypes.c_int), ("mimpur", ctypes.c_double), ("ironx", ctypes.c_int),
        ("core_otc", ctypes.c_int),
        ("lambdah", ctypes.c_double), ("tzmean", ctypes.c_double), ("tzsurf", ctypes.c_double),
        ("cp4", ctypes.c_double), ("cpp4", ctypes.c_double), ("alphas", ctypes.c_double),
        ("l", ctypes.c_double), ("den0s", ctypes.c_double), ("h2o2m4", ctypes.c_double),
        ("C", ctypes.c_double), ("beta", ctypes.c_double), ("den_m", ctypes.c_double),
        ("h4", ctypes.c_double), ("eps4", ctypes.c_double), ("sigma4", ctypes.c_double),
        ("precision", ctypes.c_int),
    ]

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def hshape(self):
        return (self.nz, self.ny, self.nx // 2 + 1)


class Morse(ctypes.Structure):
    """Mirror of ``hetddft_morse``."""
    _fields_ = [("D", ctypes.c_double), ("re", ctypes.c_double), ("alpha", ctypes.c_double),
                ("rc", ctypes.c_double), ("umax", ctypes.c_double), ("nq", ctypes.c_int)]


class Droplet(ctypes.Structure):
    """Mirror of ``hetddft_droplet`` (fresh start built on the device)."""
    _fields_ = [("n_he", ctypes.c_double), ("rho0", ctypes.c_double), ("r0", ctypes.c_double),
                ("a", ctypes.c_double), ("cavity", ctypes.c_double), ("cavity_w", ctypes.c_double),
                ("ripple", ctypes.c_double), ("clm", ctypes.c_double * 24), ("centre", ctypes.c_double * 3),
                ("lamda", ctypes.c_double * 3), ("phase_noise", ctypes.c_double), ("sigma", ctypes.c_double),
                ("v0", ctypes.c_double * 3), ("seed", ctypes.c_ulonglong)]


def droplet(n_he, seed, rho0=0.0218, r0=2.22, a=0.7, cavity=0.0, cavity_w=0.3, ripple=0.0, clm=None,
            centre=(0.0, 0.0, 0.0), lamda=(0.0, 0.0, 0.0), phase_noise=0.0, sigma=0.0, v0=(0.0, 0.0, 0.0)) -> Droplet:
    """A ``hetddft_droplet`` with the Woods-Saxon defaults of the workloads."""
    d = Droplet()
    d.n_he, d.rho0, d.r0, d.a = n_he, rho0, r0, a
    d.cavity, d.cavity_w, d.ripple = cavity, cavity_w, ripple
    for k, c in enumerate(clm if clm is not None else ()):
        d.clm[k] = c
    for k in range(3):
        d.centre[k], d.lamda[k], d.v0[k] = centre[k], lamda[k], v0[k]
    d.phase_noise, d.sigma, d.seed = phase_noise, sigma, int(seed)
    return d


TUNE_CLASSES = {"grad_den": 0, "pot4_assembly": 1, "pc_predict": 2, "pc_correct": 3, "rk_stage": 4}
TILES = {0: (64, 4), 1: (32, 8), 2: (32, 16), 3: (64, 8)}


class Tuning(ctypes.Structure):
    """Mirror of ``hetddft_tuning``: launch shape per stencil class and the z-transform tile."""
    _fields_ = [("tile", ctypes.c_int * 5), ("zc", ctypes.c_int * 5), ("prefetch", ctypes.c_int * 5),
                ("fft_z_tile", ctypes.c_int), ("fft_z_wgs", ctypes.c_int), ("ms_per_step", ctypes.c_double)]

    def as_dict(self):
        d = {name: {"tile": "x".join(map(str, TILES[self.tile[k]])), "zc": self.zc[k],
                    "prefetch": self.prefetch[k]}
             for name, k in TUNE_CLASSES.items()}
        d["fft_z_tile"] = self.fft_z_tile
        d["fft_z_wgs"] = self.fft_z_wgs
        d["ms_per_step"] = self.ms_per_step
        return d


_lib = None

_DP = ctypes.POINTER(ctypes.c_double)
_VP = ctypes.c_void_p


def lib_path():
    """libhetddft.so next to this file, or the file named by HETDDFT_LIBRARY."""
    env = os.environ.get("HETDDFT_LIBRARY")
    return Path(env) if env else _HERE / LIB_NAME


def load_library():
    """Load libhetddft.so from the package directory (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not path.exists():
        raise HetddftError(f"{path} not found: build it first (__graft_entry__.build())")
    lib = ctypes.CDLL(str(path))
    P = ctypes.POINTER(Params)
    C = ctypes.c_void_p
    sig = {
        "hetddft_default_params": ([P], ctypes.c_int),
        "hetddft_create": ([P, ctypes.POINTER(C)], ctypes.c_int),
        "hetddft_destroy": ([C], ctypes.c_int),
        "hetddft_last_error": ([C], ctypes.c_char_p),
        "hetddft_make_kernels": ([P, ctypes.POINTER(Morse), _VP, _VP, _VP, _VP], ctypes.c_int),
        "hetddft_set_kernels": ([C, _VP, _VP, _VP, _VP], ctypes.c_int),
        "hetddft_set_state": ([C, _VP, _VP, _VP], ctypes.c_int),
        "hetddft_start": ([C], ctypes.c_int),
        "hetddft_potential": ([C], ctypes.c_int),
        "hetddft_step_rk": ([C, ctypes.c_double], ctypes.c_int),
        "hetddft_step_pc": ([C, ctypes.c_double, _VP], ctypes.c_int),
        "hetddft_last_errors": ([C, _VP], ctypes.c_int),
        "hetddft_iterate": ([C, ctypes.c_double, _VP], ctypes.c_int),
        "hetddft_particle_means": ([C, _VP], ctypes.c_int),
        "hetddft_particle_means_dev": ([_VP, _VP, _VP, ctypes.c_int, _VP, _VP], ctypes.c_int),
        "hetddft_energy": ([C, _VP], ctypes.c_int),
        "hetddft_get_state": ([C, _VP, _VP, _VP, _VP], ctypes.c_int),
        "hetddft_get_field": ([C, ctypes.c_int, _VP], ctypes.c_int),
        "hetddft_convolve": ([C, ctypes.c_int, _VP, _VP, _VP], ctypes.c_int),
        "hetddft_sync": ([C], ctypes.c_int),
        "hetddft_graph_info": ([C, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "hetddft_diagnostics": ([C, _VP], ctypes.c_int),
        "hetddft_checkpoint_size": ([C, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
        "hetddft_sample_particles": ([C, _VP, _VP, _VP, ctypes.c_ulonglong, ctypes.c_longlong, _VP], ctypes.c_int),
        "hetddft_init_kernels": ([C, ctypes.POINTER(Morse)], ctypes.c_int),
        "hetddft_init_droplet": ([C, ctypes.POINTER(Droplet)], ctypes.c_int),
        "hetddft_get_kernels": ([C, _VP, _VP, _VP, _VP], ctypes.c_int),
        "hetddft_checkpoint_save": ([C, _VP], ctypes.c_int),
        "hetddft_checkpoint_load": ([C, _VP], ctypes.c_int),
        "hetddft_current": ([C, _VP, _VP, _VP], ctypes.c_int),
        "hetddft_profile": ([C, ctypes.c_int], ctypes.c_int),
        "hetddft_profile_read": ([C, ctypes.c_int, _VP, _VP, _VP], ctypes.c_int),
        "hetddft_device_bytes": ([C], ctypes.c_longlong),
        "hetddft_autotune": ([C, ctypes.c_double, ctypes.c_int, ctypes.POINTER(Tuning)], ctypes.c_int),
        "hetddft_get_tuning": ([C, ctypes.POINTER(Tuning)], ctypes.c_int),
        "hetddft_set_tuning": ([C, ctypes.POINTER(Tuning)], ctypes.c_int),
        "hetddft_batch_create": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.POINTER(C)], ctypes.c_int),
        "hetddft_batch_iterate": ([C, _VP, ctypes.c_int, _VP, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "hetddft_batch_last_error": ([C], ctypes.c_char_p),
        "hetddft_batch_destroy": ([C], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def default_params(**kw) -> Params:
    lib = load_library()
    p = Params()
    lib.hetddft_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def make_kernels(p: Params, morse: Morse | None = None) -> dict:
    """fvlj4, wcgk, kalfs, vq on the half grid (hetddft_make_kernels)."""
    lib = load_library()
    out = {k: np.zeros(p.hshape) for k in ("fvlj4", "wcgk", "kalfs", "vq")}
    rc = lib.hetddft_make_kernels(ctypes.byref(p), ctypes.byref(morse) if morse else None,
                                  _ptr(out["fvlj4"]), _ptr(out["wcgk"]), _ptr(out["kalfs"]),
                                  _ptr(out["vq"]))
    if rc != 0:
        raise HetddftError(f"hetddft_make_kernels failed ({rc})")
    return out


class Instance:
    """One droplet + impurity instance resident on the GPU."""

    def __init__(self, p: Params):
        self.lib = load_library()
        self.p = p
        self._c = ctypes.c_void_p()
        rc = self.lib.hetddft_create(ctypes.byref(p), ctypes.byref(self._c))
        if rc != 0:
            raise HetddftError(f"hetddft_create failed ({rc}): "
                               f"{self.lib.hetddft_last_error(None).decode()}", rc)

    # -- plumbing --
    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.hetddft_last_error(self._c).decode()
            raise HetddftError(f"{what} failed ({rc}): {msg}", rc)

    def close(self):
        if self._c:
            self.lib.hetddft_destroy(self._c)
            self._c = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def device_bytes(self):
        return int(self.lib.hetddft_device_bytes(self._c))

    # -- state --
    def set_kernels(self, k: dict):
        arrs = [np.ascontiguousarray(k[n], dtype=np.float64) for n in ("fvlj4", "wcgk", "kalfs", "vq")]
        for a in arrs:
            assert a.shape == self.p.hshape, (a.shape, self.p.hshape)
        self._check(self.lib.hetddft_set_kernels(self._c, *[_ptr(a) for a in arrs]), "set_kernels")

    def set_state(self, psi, rimp, vimp):
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        assert psi.shape == self.p.shape
        rimp = np.ascontiguousarray(rimp, dtype=np.float64)
        vimp = np.ascontiguousarray(vimp, dtype=np.float64)
        assert rimp.shape == (3, self.p.n_par) and vimp.shape == (3, self.p.n_par)
        self._check(self.lib.hetddft_set_state(self._c, _ptr(psi), _ptr(rimp), _ptr(vimp)), "set_state")

    def start(self):
        self._check(self.lib.hetddft_start(self._c), "start")

    # -- the reference's per-step operations --
    def potential(self):
        """potenimp(); poten(); forceimp()"""
        self._check(self.lib.hetddft_potential(self._c), "potential")

    def steprk(self, deltat):
        self._check(self.lib.hetddft_step_rk(self._c, float(deltat)), "steprk")

    def steppc(self, deltat, want_errors=True):
        err = np.zeros(3) if want_errors else None
        self._check(self.lib.hetddft_step_pc(self._c, float(deltat), _ptr(err)), "steppc")
        return tuple(err) if want_errors else None

    def iterate(self, deltat, want_errors=True):
        err = np.zeros(3) if want_errors else None
        self._check(self.lib.hetddft_iterate(self._c, float(deltat), _ptr(err)), "iterate")
        return tuple(err) if want_errors else None

    def last_errors(self):
        err = np.zeros(3)
        self._check(self.lib.hetddft_last_errors(self._c, _ptr(err)), "last_errors")
        return tuple(err)

    def particle_means(self):
        out = np.zeros(9)
        self._check(self.lib.hetddft_particle_means(self._c, _ptr(out)), "particle_means")
        return out

    def energy(self):
        out = np.zeros(12)
        self._check(self.lib.hetddft_energy(self._c, _ptr(out)), "energy")
        return dict(zip(ENERGY_KEYS, out))

    def get_state(self):
        psi = np.zeros(self.p.shape, np.complex128)
        n = self.p.n_par
        r, v, a = np.zeros((3, n)), np.zeros((3, n)), np.zeros((3, n))
        self._check(self.lib.hetddft_get_state(self._c, _ptr(psi), _ptr(r), _ptr(v), _ptr(a)),
                    "get_state")
        return psi, r, v, a

    def get_field(self, name):
        out = np.zeros(self.p.shape)
        self._check(self.lib.hetddft_get_field(self._c, FIELDS[name], _ptr(out)), f"get_field({name})")
        return out

    def convolve(self, fields, kern):
        """c2r(r2c(f) * kern) * renor for up to five real fields."""
        ins = [np.ascontiguousarray(f, dtype=np.float64) for f in fields]
        outs = [np.zeros(self.p.shape) for _ in ins]
        kern = np.ascontiguousarray(kern, dtype=np.float64)
        inp = (ctypes.c_void_p * len(ins))(*[a.ctypes.data for a in ins])
        outp = (ctypes.c_void_p * len(outs))(*[a.ctypes.data for a in outs])
        self._check(self.lib.hetddft_convolve(self._c, len(ins), ctypes.cast(inp, ctypes.c_void_p),
                                              _ptr(kern), ctypes.cast(outp, ctypes.c_void_p)),
                    "convolve")
        return outs

    def sync(self):
        self._check(self.lib.hetddft_sync(self._c), "sync")

    def diagnostics(self):
        """Device-side per-step diagnostics (hetddft_diagnostics): He count, the
        centre-of-mass moments sum(den x_c) dxyz, chemical potential, integral of
        the current, impurity mean-square radii, min/max of vimp and aimp."""
        out = np.zeros(23)
        self._check(self.lib.hetddft_diagnostics(self._c, _ptr(out)), "diagnostics")
        return {"norm": out[0], "cm_moments": out[1:4].copy(), "mu": out[4], "jtot": out[5:8].copy(),
                "msx": out[8:11].copy(), "vmin": out[11:17:2].copy(), "vmax": out[12:17:2].copy(),
                "amin": out[17:23:2].copy(), "amax": out[18:23:2].copy()}

    def init_kernels(self, morse=None):
        """Build the k-space kernels on the device (hetddft_init_kernels)."""
        m = ctypes.byref(morse) if morse is not None else None
        self._check(self.lib.hetddft_init_kernels(self._c, m), "init_kernels")

    def get_kernels(self):
        """The context's kernels (dict of (nz, ny, nx//2+1) arrays)."""
        shape = (self.p.nz, self.p.ny, self.p.nx // 2 + 1)
        k = {n: np.zeros(shape) for n in ("fvlj4", "wcgk", "kalfs", "vq")}
        self._check(self.lib.hetddft_get_kernels(self._c, _ptr(k["fvlj4"]), _ptr(k["wcgk"]), _ptr(k["kalfs"]),
                                                 _ptr(k["vq"])), "get_kernels")
        return k

    def sample_particles(self, denx, lo, hi, seed, max_attempts=1 << 24):
        """N_par positions (3, N_par) drawn by rejection from denx (nz, ny, nx), as sorteo."""
        d = np.ascontiguousarray(denx, dtype=np.float64)
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        hi = np.ascontiguousarray(hi, dtype=np.float64)
        out = np.zeros((3, self.p.n_par))
        self._check(self.lib.hetddft_sample_particles(self._c, _ptr(d), _ptr(lo), _ptr(hi), int(seed),
                                                      int(max_attempts), _ptr(out)), "sample_particles")
        return out

    def init_droplet(self, d: Droplet):
        """Fresh start built on the device (hetddft_init_droplet); replaces the
        state like set_state, so start() follows."""
        self._check(self.lib.hetddft_init_droplet(self._c, ctypes.byref(d)), "init_droplet")

    def checkpoint(self):
        """Full-history checkpoint image (bytes) for bit-exact resume."""
        n = ctypes.c_longlong(0)
        self._check(self.lib.hetddft_checkpoint_size(self._c, ctypes.byref(n)), "checkpoint_size")
        buf = np.zeros(n.value, dtype=np.uint8)
        self._check(self.lib.hetddft_checkpoint_save(self._c, _ptr(buf)), "checkpoint_save")
        return buf.tobytes()

    def restore(self, image):
        """Restore a checkpoint image (after set_kernels) and re-run the potential evaluation."""
        buf = np.frombuffer(image, dtype=np.uint8).copy()
        self._check(self.lib.hetddft_checkpoint_load(self._c, _ptr(buf)), "checkpoint_load")

    def current(self):
        """Current density j = 2 h2o2m4 Im(conj(psi) grad psi), three (nz, ny, nx) arrays."""
        shape = (self.p.nz, self.p.ny, self.p.nx)
        j = [np.zeros(shape) for _ in range(3)]
        self._check(self.lib.hetddft_current(self._c, _ptr(j[0]), _ptr(j[1]), _ptr(j[2])), "current")
        return tuple(j)

    def autotune(self, deltat, steps=4):
        """Time the launch shapes on this instance's state and keep the fastest (a Tuning)."""
        t = Tuning()
        self._check(self.lib.hetddft_autotune(self._c, float(deltat), int(steps), ctypes.byref(t)), "autotune")
        return t

    def get_tuning(self):
        t = Tuning()
        self._check(self.lib.hetddft_get_tuning(self._c, ctypes.byref(t)), "get_tuning")
        return t

    def set_tuning(self, t):
        self._check(self.lib.hetddft_set_tuning(self._c, ctypes.byref(t)), "set_tuning")

    def graph_info(self):
        """(captured PC-step graphs, graphs enabled) of this instance."""
        n, on = ctypes.c_int(0), ctypes.c_int(0)
        self._check(self.lib.hetddft_graph_info(self._c, ctypes.byref(n), ctypes.byref(on)), "graph_info")
        return n.value, bool(on.value)

    def profile(self, enable=True):
        """Reset and enable (or disable) per-kernel-class event timing."""
        self._check(self.lib.hetddft_profile(self._c, int(bool(enable))), "profile")

    def profile_read(self):
        """{class: (total_ms, launches, algorithmic_bytes)} for classes that ran."""
        out = {}
        for name, cls in PROF_CLASSES.items():
            ms = ctypes.c_double()
            n = ctypes.c_longlong()
            b = ctypes.c_double()
            self._check(self.lib.hetddft_profile_read(self._c, cls, ctypes.byref(ms), ctypes.byref(n),
                                                      ctypes.byref(b)), "profile_read")
            if n.value:
                out[name] = (ms.value, n.value, b.value)
        return out


class Batch:
    """An ensemble of started instances advanced together by the library's
    host threads (hetddft_batch_*): ``iterate(steps)`` runs ``steps``
    iterations of the reference time loop on every instance."""

    def __init__(self, instances, deltats, nthreads=0):
        self.lib = load_library()
        self.instances = list(instances)
        self._ptrs = (ctypes.c_void_p * len(self.instances))(*[g._c.value for g in self.instances])
        self.deltats = np.ascontiguousarray(deltats, dtype=np.float64)
        if self.deltats.shape != (len(self.instances),):
            raise ValueError("one deltat per instance")
        self._b = ctypes.c_void_p()
        rc = self.lib.hetddft_batch_create(ctypes.cast(self._ptrs, ctypes.c_void_p), len(self.instances),
                                           int(nthreads), ctypes.byref(self._b))
        if rc != 0:
            raise HetddftError(f"batch_create failed ({rc})")

    def iterate(self, steps=1, want_errors=False):
        """Advance every instance `steps` iterations; (n, 3) PC errors of the last one if asked."""
        err = np.zeros((len(self.instances), 3)) if want_errors else None
        failed = ctypes.c_int(-1)
        rc = self.lib.hetddft_batch_iterate(self._b, _ptr(self.deltats), int(steps),
                                            _ptr(err) if err is not None else None, ctypes.byref(failed))
        if rc != 0:
            msg = self.lib.hetddft_batch_last_error(self._b)
            raise HetddftError(f"batch_iterate failed ({rc}) at instance {failed.value}: "
                               f"{msg.decode() if msg else ''}", rc, failed.value)
        return err

    def close(self):
        if self._b:
            self.lib.hetddft_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def header_functions(header=None):
    """Names of the functions declared in include/hetddft.h."""
    import re
    header = Path(header) if header else _HERE.parent / "include" / "hetddft.h"
    text = header.read_text()
    return sorted(set(re.findall(r"\b(hetddft_[a-z0-9_]+)\s*\(", text)))

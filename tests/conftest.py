import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "4hetddft-isotropic-bohmian_amd"
for p in (str(PKG), str(ROOT)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def hip():
    import hetddft
    hetddft.load_library()
    return hetddft


def oracle_params(wl, **kw):
    from oracle import hetddft_oracle as orc
    kw.setdefault("ironx", wl.ironx)
    return orc.Params(nx=wl.n, ny=wl.n, nz=wl.n, xmax=wl.xmax, ymax=wl.xmax, zmax=wl.xmax,
                      n_par=wl.n_par, mimpur=wl.mimpur, **kw)


def rel_max(a, b):
    """max|a-b| / max|b| (scale-free comparison of whole fields)."""
    import numpy as np
    s = float(np.abs(b).max())
    return float(np.abs(a - b).max()) / (s if s > 0 else 1.0)

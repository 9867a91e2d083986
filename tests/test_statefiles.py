"""printoutc / readenc state files (statefiles.py): round trips, the
list-directed forms Fortran compilers write, and the reference's own
printoutc and readenc compiled from its sources (oracle/Makefile)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import statefiles as sf

ROOT = Path(__file__).resolve().parents[1]
HARNESS = ROOT / "oracle" / "_ref" / "ref_harness"
needs_ref = pytest.mark.skipif(not HARNESS.exists(), reason="reference harness not built (oracle/Makefile)")


def _state(nx=6, ny=5, nz=4, n_par=7, seed=3):
    rng = np.random.default_rng(seed)
    psi = rng.standard_normal((nz, ny, nx)) + 1j * rng.standard_normal((nz, ny, nx))
    psi[0, 0, :3] = 0.0                 # repeated values (repeat counts in some writers)
    rimp = rng.standard_normal((3, n_par)) * 3.0
    vimp = rng.standard_normal((3, n_par)) * 1e-3
    box = dict(xmax=nx * 0.4 / 2, ymax=ny * 0.45 / 2, zmax=nz * 0.5 / 2, hx=0.4, hy=0.45, hz=0.5)
    return psi, rimp, vimp, box


def test_text_round_trip_exact(tmp_path):
    psi, rimp, vimp, box = _state()
    sf.write_text(tmp_path / "d", tmp_path / "i", psi, rimp, vimp, iteration=12, deltatps=5e-4, time=0.006, **box)
    p2, r2, v2, hdr = sf.read_text(tmp_path / "d", tmp_path / "i")
    assert np.array_equal(p2, psi) and np.array_equal(r2, rimp) and np.array_equal(v2, vimp)
    assert (hdr["nx"], hdr["ny"], hdr["nz"]) == (6, 5, 4) and hdr["hy"] == 0.45
    head = (tmp_path / "d").read_text().splitlines()[:5]
    assert head[0] == "#  Density after              12 iterations."
    assert head[1] == "#  Actual delatatps....:    5.0000000000E-04"
    assert head[4] == "#  Backup copy."


def test_binary_round_trip_exact(tmp_path):
    psi, rimp, vimp, box = _state()
    sf.write_binary(tmp_path / "d", tmp_path / "i", psi, rimp, vimp, **box)
    raw = (tmp_path / "d").read_bytes()
    assert int.from_bytes(raw[:4], "little") == 60           # 6 real*8 + 3 int*4
    p2, r2, v2, hdr = sf.read_binary(tmp_path / "d", tmp_path / "i")
    assert np.array_equal(p2, psi) and np.array_equal(r2, rimp) and np.array_equal(v2, vimp)
    assert hdr["nz"] == 4 and hdr["xmax"] == box["xmax"]


def test_list_directed_forms(tmp_path):
    """Repeat counts, D exponents, comma separators and line breaks."""
    (tmp_path / "d").write_text("# c1\n#c2\n 1.0D0, 2.0 3.0 0.5 0.5 0.5 2 1 2\n"
                                " 2*(0.0,0.0), (1.5D-01, -2.0E+00)\n (3,4)\n")
    (tmp_path / "i").write_text("#\n 1\n 1.0 2.0\n 3.0\n 3*0.25\n")
    psi, rimp, vimp, hdr = sf.read_text(tmp_path / "d", tmp_path / "i")
    assert hdr["xmax"] == 1.0 and hdr["nx"] == 2 and hdr["nz"] == 2
    assert np.array_equal(psi.reshape(-1), np.array([0, 0, 0.15 - 2j, 3 + 4j]))
    assert np.array_equal(rimp[:, 0], [1.0, 2.0, 3.0]) and np.array_equal(vimp[:, 0], [0.25] * 3)


def test_centre_of_mass_matches_oracle_moments():
    from oracle import hetddft_oracle as orc
    psi, rimp, vimp, box = _state(nx=8, ny=8, nz=8)
    prm = orc.Params(nx=8, ny=8, nz=8, xmax=box["xmax"], ymax=box["ymax"], zmax=box["zmax"], n_par=7)
    d = orc.diagnostics(prm, psi, np.zeros(psi.shape), np.zeros(psi.shape), vimp, vimp)
    cm = sf.centre_of_mass(psi, **{k: box[k] for k in ("xmax", "ymax", "zmax")}, hx=prm.hx, hy=prm.hy, hz=prm.hz)
    assert np.allclose(cm, d[1:4] / d[0], rtol=1e-13, atol=1e-15)


def _harness(cmd, cwd, params, arrays):
    (cwd / "params.txt").write_text(params + "\n")
    for name, a in arrays.items():
        np.asarray(a, dtype=np.float64).tofile(cwd / name)
    res = subprocess.run([str(HARNESS), cmd, "."], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr


def _fortran_arrays(psi, rimp, vimp):
    raw = np.stack([psi.real, psi.imag], axis=-1).reshape(-1)
    return {"psi.bin": raw, "rimp.bin": rimp.reshape(-1), "vimp.bin": vimp.reshape(-1)}


@needs_ref
@pytest.mark.parametrize("iprint", [3, 1, -3])
def test_reference_printoutc_files_are_read(tmp_path, iprint):
    """Files written by the reference's printoutc (text and binary) are read back
    by statefiles: values exact (text to the digits the compiler writes)."""
    psi, rimp, vimp, box = _state()
    nz, ny, nx = psi.shape
    _harness("printoutc", tmp_path,
             f"{nx} {ny} {nz} {rimp.shape[1]} {box['xmax']!r} {box['ymax']!r} {box['zmax']!r} "
             f"{box['hx']!r} {box['hy']!r} {box['hz']!r} 0.0125 {iprint} 25 5.0e-4",
             _fortran_arrays(psi, rimp, vimp))
    read = sf.read_binary if iprint < 0 else sf.read_text
    p2, r2, v2, hdr = read(tmp_path / "den.out", tmp_path / "imp.out")
    assert (hdr["nx"], hdr["ny"], hdr["nz"]) == (nx, ny, nz)
    assert np.array_equal(p2, psi) and np.array_equal(r2, rimp) and np.array_equal(v2, vimp)
    if iprint > 0:
        ref_head = [l for l in (tmp_path / "den.out").read_text().splitlines() if l.startswith("#")]
        sf.write_text(tmp_path / "d", tmp_path / "i", psi, rimp, vimp, iteration=25, deltatps=5e-4, time=0.0125,
                      iprint=iprint, **box)
        own_head = [l for l in (tmp_path / "d").read_text().splitlines() if l.startswith("#")]
        assert own_head == [l.rstrip() for l in ref_head]


@needs_ref
def test_reference_readenc_reads_own_files(tmp_path):
    """The reference's readenc (mode 2) reads the files statefiles writes."""
    psi, rimp, vimp, box = _state()
    nz, ny, nx = psi.shape
    sf.write_text(tmp_path / "den.in", tmp_path / "imp.in", psi, rimp, vimp, **box)
    _harness("readenc", tmp_path, f"{nx} {ny} {nz} {rimp.shape[1]}", {})
    raw = np.fromfile(tmp_path / "psi.bin")
    p2 = (raw[0::2] + 1j * raw[1::2]).reshape(nz, ny, nx)
    assert np.array_equal(p2, psi)
    assert np.array_equal(np.fromfile(tmp_path / "rimp.bin").reshape(3, -1), rimp)
    assert np.array_equal(np.fromfile(tmp_path / "vimp.bin").reshape(3, -1), vimp)

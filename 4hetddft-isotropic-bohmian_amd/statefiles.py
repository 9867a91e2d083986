"""State files of the reference solver, for hosts other than its Fortran driver.

Text form (``printoutc`` with iprint > 0, printoutc.f90:76-116), read back by
``readenc`` mode 2 (readenc.f90:96-114):

* density file: comment lines starting with '#' (iteration and time step,
  total time, helium centre of mass, the kind of copy; ``titols`` skips every
  leading line whose first column is '#', titols.f90:12-41), then list-directed
  ``xmax ymax zmax hx hy hz nx ny nz`` and ``psi(nx,ny,nz)`` as ``(re,im)``
  pairs in Fortran order;
* impurity file: the same comment lines, then ``N_par``, ``rimp(N_par,3)``,
  ``vimp(N_par,3)`` list-directed.

Binary form (iprint < 0, printoutc.f90:66-75): Fortran unformatted sequential
records with 4-byte length markers: ``[xmax ymax zmax hx hy hz (real*8),
nx ny nz (int*4)]``, ``[psi]``; and ``[N_par]``, ``[rimp]``, ``[vimp]``.

Arrays follow this package's conventions: psi has shape (nz, ny, nx) (the
Fortran (nx, ny, nz) layout byte for byte), rimp and vimp shape (3, N_par)
(Fortran (N_par, 3)).  The list-directed reader accepts what Fortran
compilers write: blanks or commas between items, ``(re,im)`` complex items,
``r*value`` repeat counts and D exponents.
"""
from __future__ import annotations

import re
import struct
from pathlib import Path

import numpy as np

_KIND = {1: "#  Last printout of the run.", 2: "#  Change of Paflov parameter.", 3: "#  Backup copy."}


def _fortran_e(x, width, d):
    """Fortran 1P,Ew.d edit descriptor: one digit before the point, d after."""
    s = f"{x:.{d}E}"
    mant, exp = s.split("E")
    e = int(exp)
    return f"{mant}E{'+' if e >= 0 else '-'}{abs(e):02d}".rjust(width)


def _header_lines(iteration, deltatps, time, cm, iprint):
    """The comment block of printoutc.f90:79-91 (formats 1020-1050)."""
    return [
        f"#  Density after {iteration:15d} iterations.",
        f"#  Actual delatatps....:{_fortran_e(deltatps, 20, 10)}",
        f"#  Total evolution time(ps)...:{_fortran_e(time, 20, 12)}",
        "#  xcm,ycm,zcm:" + "".join(_fortran_e(v, 16, 8) for v in cm),
        _KIND.get(iprint, _KIND[3]),
    ]


def centre_of_mass(psi, xmax, ymax, zmax, hx, hy, hz):
    """Helium centre of mass of the comment line (printoutc.f90:35-58)."""
    nz, ny, nx = psi.shape
    den = np.abs(psi) ** 2
    x = -xmax + hx * np.arange(nx)
    y = -ymax + hy * np.arange(ny)
    z = -zmax + hz * np.arange(nz)
    norm = den.sum()
    return ((den * x[None, None, :]).sum() / norm, (den * y[None, :, None]).sum() / norm,
            (den * z[:, None, None]).sum() / norm)


def _num(x):
    return repr(float(x))


def write_text(den_path, imp_path, psi, rimp, vimp, xmax, ymax, zmax, hx, hy, hz,
               iteration=0, deltatps=0.0, time=0.0, iprint=3):
    """printoutc text files (iprint = 1, 2 or 3).  Values carry 17 significant
    digits, so reading them back is exact."""
    nz, ny, nx = psi.shape
    head = _header_lines(iteration, deltatps, time, centre_of_mass(psi, xmax, ymax, zmax, hx, hy, hz), iprint)
    flat = np.ascontiguousarray(psi).reshape(-1)
    with open(den_path, "w") as f:
        f.write("\n".join(head) + "\n")
        f.write(" " + " ".join(_num(v) for v in (xmax, ymax, zmax, hx, hy, hz)) + f" {nx} {ny} {nz}\n")
        lines = [f" ({_num(v.real)},{_num(v.imag)})" for v in flat]
        f.write("\n".join(lines) + "\n")
    n_par = rimp.shape[1]
    with open(imp_path, "w") as f:
        f.write("\n".join(head) + "\n")
        f.write(f" {n_par}\n")
        for a in (rimp, vimp):
            f.write("\n".join(" " + _num(v) for v in np.ascontiguousarray(a).reshape(-1)) + "\n")


_ITEM = re.compile(r"""\s*,?\s*(?:(\d+)\*)?(\([^)]*\)|[^\s,()]+)""")


def _items(text):
    """List-directed items of text, repeat counts expanded."""
    pos, out = 0, []
    text = text.strip()
    while pos < len(text):
        m = _ITEM.match(text, pos)
        if not m:
            raise ValueError(f"cannot parse list-directed input near {text[pos:pos + 40]!r}")
        rep = int(m.group(1)) if m.group(1) else 1
        out.extend([m.group(2)] * rep)
        pos = m.end()
        while pos < len(text) and text[pos] in " \t\r\n,":
            pos += 1
    return out


def _real(tok):
    return float(tok.replace("D", "E").replace("d", "e"))


def _cplx(tok):
    if not (tok.startswith("(") and tok.endswith(")")):
        raise ValueError(f"complex item expected, got {tok!r}")
    re_, im_ = tok[1:-1].split(",")
    return complex(_real(re_), _real(im_))


def _skip_comments(lines):
    """titols.f90:12-41: skip the leading lines whose first column is '#'."""
    k = 0
    while k < len(lines) and lines[k][:1] == "#":
        k += 1
    return lines[k:]


def read_text(den_path, imp_path):
    """readenc mode 2 (readenc.f90:96-114): returns psi, rimp, vimp and the
    header dict (xmax ymax zmax hx hy hz nx ny nz)."""
    body = _skip_comments(Path(den_path).read_text().splitlines())
    items = _items("\n".join(body))
    hdr = dict(zip(("xmax", "ymax", "zmax", "hx", "hy", "hz"), (_real(t) for t in items[:6])))
    nx, ny, nz = (int(t) for t in items[6:9])
    hdr.update(nx=nx, ny=ny, nz=nz)
    n = nx * ny * nz
    vals = items[9:9 + n]
    if len(vals) != n:
        raise ValueError(f"{den_path}: {len(vals)} of {n} psi values")
    psi = np.array([_cplx(t) for t in vals], dtype=np.complex128).reshape(nz, ny, nx)
    body = _skip_comments(Path(imp_path).read_text().splitlines())
    items = _items("\n".join(body))
    n_par = int(items[0])
    nums = np.array([_real(t) for t in items[1:1 + 6 * n_par]])
    if nums.size != 6 * n_par:
        raise ValueError(f"{imp_path}: {nums.size} of {6 * n_par} particle values")
    return psi, nums[:3 * n_par].reshape(3, n_par), nums[3 * n_par:].reshape(3, n_par), hdr


def _record(f, payload: bytes):
    f.write(struct.pack("<i", len(payload)))
    f.write(payload)
    f.write(struct.pack("<i", len(payload)))


def _records(data: bytes):
    pos, out = 0, []
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        body = data[pos + 4:pos + 4 + n]
        (n2,) = struct.unpack_from("<i", data, pos + 4 + n)
        if n2 != n or len(body) != n:
            raise ValueError("bad Fortran record markers")
        out.append(body)
        pos += 8 + n
    return out


def write_binary(den_path, imp_path, psi, rimp, vimp, xmax, ymax, zmax, hx, hy, hz):
    """printoutc binary files (iprint < 0)."""
    nz, ny, nx = psi.shape
    with open(den_path, "wb") as f:
        _record(f, struct.pack("<6d3i", xmax, ymax, zmax, hx, hy, hz, nx, ny, nz))
        _record(f, np.ascontiguousarray(psi, dtype=np.complex128).tobytes())
    with open(imp_path, "wb") as f:
        _record(f, struct.pack("<i", rimp.shape[1]))
        _record(f, np.ascontiguousarray(rimp, dtype=np.float64).tobytes())
        _record(f, np.ascontiguousarray(vimp, dtype=np.float64).tobytes())


def read_binary(den_path, imp_path):
    """Inverse of write_binary: psi, rimp, vimp, header dict."""
    r = _records(Path(den_path).read_bytes())
    vals = struct.unpack("<6d3i", r[0])
    hdr = dict(zip(("xmax", "ymax", "zmax", "hx", "hy", "hz", "nx", "ny", "nz"), vals))
    psi = np.frombuffer(r[1], dtype=np.complex128).reshape(hdr["nz"], hdr["ny"], hdr["nx"]).copy()
    r = _records(Path(imp_path).read_bytes())
    (n_par,) = struct.unpack("<i", r[0])
    rimp = np.frombuffer(r[1], dtype=np.float64).reshape(3, n_par).copy()
    vimp = np.frombuffer(r[2], dtype=np.float64).reshape(3, n_par).copy()
    return psi, rimp, vimp, hdr

"""Extract the 13-point finite-difference weight blocks from the reference table.

Reads the BLOCK DATA table of the reference (``mates.f``, M=13 blocks for K=1
at A(1041..1125) and K=2 at A(1126..1210), see ``mates.f:786-879``) and writes
``stencil_m13.json``: the decimal literals exactly as the table spells them,
so tests can check the oracle's closed-form Fornberg weights against the
reference's own numbers without reading the reference at test time.

Run once in a container where the reference tree is mounted:
    python tests/golden/make_stencil_golden.py /root/reference/mates.f
"""
import json
import re
import sys
from pathlib import Path


def read_table(path):
    text = Path(path).read_text(errors="replace").splitlines()
    vals = {}
    cur = None
    num = re.compile(r"[-+]?\d+\.\d+D[-+]\d+")
    for line in text:
        m = re.match(r"\s+DATA \(A\(I\),I=\s*(\d+),\s*(\d+)\)/", line)
        if m:
            cur = int(m.group(1))
            continue
        if cur is None:
            continue
        if not line.startswith("     &"):
            cur = None
            continue
        for tok in num.findall(line):
            vals[cur] = tok
            cur += 1
    return vals


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/mates.f"
    vals = read_table(src)
    out = {
        "source": "mates.f:786-879 (BLOCK DATA BLKDS, M=13)",
        "layout": "85 leading entries of the 13x13 row-major weight matrix; "
                  "entries 79..85 (1-based) are the centre row up to the centre point",
        "K1": [vals[i] for i in range(1041, 1126)],
        "K2": [vals[i] for i in range(1126, 1211)],
    }
    dst = Path(__file__).with_name("stencil_m13.json")
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", dst)


if __name__ == "__main__":
    main()

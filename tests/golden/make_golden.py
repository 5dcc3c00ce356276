"""Regenerate tests/golden/ref_L8.npz from the reference itself.

Builds oracle/_ref/ref_harness (the reference's own sources compiled in place
at the 8^3 size its macro.h ships, linked with oracle/ref_harness.c), runs it,
and converts its named binary records into a compressed .npz of plain arrays
(no pickles).  Only this container has the reference checkout; the committed
.npz is what the tests read everywhere else.

Usage:  python tests/golden/make_golden.py
"""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
ORACLE = os.path.join(ROOT, "oracle")
OUT = os.path.join(ROOT, "tests", "golden", "ref_L8.npz")


def read_records(path):
    recs = {}
    with open(path, "rb") as f:
        while True:
            hdr = f.read(40)
            if not hdr:
                break
            name = hdr[:32].split(b"\0", 1)[0].decode()
            kind = chr(hdr[32])
            (count,) = struct.unpack("<q", f.read(8))
            dt = {"d": np.float64, "i": np.int32, "b": np.uint8}[kind]
            data = np.frombuffer(f.read(count * np.dtype(dt).itemsize), dtype=dt).copy()
            recs[name] = data
    return recs


def main():
    subprocess.run(["make", "-C", ORACLE, "ref"], check=True)
    with tempfile.TemporaryDirectory() as td:
        binp = os.path.join(td, "golden.bin")
        acc = os.path.join(td, "acc.txt")
        subprocess.run([os.path.join(ORACLE, "_ref", "ref_harness"), binp, acc], check=True)
        recs = read_records(binp)
    np.savez_compressed(OUT, **recs)
    print(f"wrote {OUT}: {len(recs)} records")


if __name__ == "__main__":
    sys.exit(main())

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
import shutil
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
ORACLE = os.path.join(ROOT, "oracle")
OUT = os.path.join(ROOT, "tests", "golden", "ref_L8.npz")
OUT_DRIVER = os.path.join(ROOT, "tests", "golden", "ref_driver_L8.npz")


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
    # every output of the reference's build and runs stays under oracle/_ref/
    run = os.path.join(ORACLE, "_ref", "harness_run")
    shutil.rmtree(run, ignore_errors=True)
    os.makedirs(run)
    binp = os.path.join(run, "golden.bin")
    acc = os.path.join(run, "acc.txt")
    subprocess.run([os.path.join(ORACLE, "_ref", "ref_harness"), binp, acc], check=True)
    recs = read_records(binp)
    np.savez_compressed(OUT, **recs)
    print(f"wrote {OUT}: {len(recs)} records")
    # the reference's own driver flow (oracle/driver_main.c): the files its
    # first measurement block writes
    subprocess.run(["make", "-C", ORACLE, os.path.join("_ref", "ref_driver")], check=True)
    run = os.path.join(ORACLE, "_ref", "driver_run")
    shutil.rmtree(run, ignore_errors=True)
    os.makedirs(run)
    subprocess.run([os.path.join(ORACLE, "_ref", "ref_driver"), run], check=True, stdout=subprocess.DEVNULL)
    files = {kind: driver_file(run, kind) for kind in ("data", "acceptance", "terma")}
    np.savez_compressed(OUT_DRIVER, **{k: np.frombuffer(open(v, "rb").read(), dtype=np.uint8) for k, v in files.items()})
    print(f"wrote {OUT_DRIVER}")


def driver_file(workdir, kind):
    """The one file the driver wrote under dati_simulation3d/<kind>_*/."""
    base = os.path.join(workdir, "dati_simulation3d")
    dirs = [d for d in os.listdir(base) if d.startswith(kind + "_")]
    assert len(dirs) == 1, dirs
    names = os.listdir(os.path.join(base, dirs[0]))
    assert len(names) == 1, names
    return os.path.join(base, dirs[0], names[0])


if __name__ == "__main__":
    sys.exit(main())

"""Writes tests/golden/ref_setup.json from the reference's own setup numerics.

Runs oracle/_ref/ref_setup (built by ``make -C oracle ref`` from the
reference's MainVar3D.f90 and Subroutineso.f90, compiled where they lie) and
stores its printed values as a fixture: every record is the name of the
reference routine, and the list of values it returned, in call order.  The
driver's call sequence is in oracle/ref_setup.f90.

This script only runs where the reference tree is mounted; the fixture it
writes is data and travels with the tests.
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_setup")
    if not os.path.exists(exe):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    out = subprocess.check_output([exe], text=True)
    rec = {}
    for line in out.splitlines():
        parts = line.split()
        if not parts:
            continue
        name, n = parts[0], int(parts[1])
        vals = [float(v) for v in parts[2:2 + n]]
        assert len(vals) == n, line
        rec.setdefault(name, []).append(vals)
    path = os.path.join(HERE, "ref_setup.json")
    with open(path, "w") as f:
        json.dump({"source": "reference src/Subroutineso.f90 + src/MainVar3D.f90 via oracle/ref_setup.f90",
                   "records": rec}, f, indent=0)
    print("wrote", path, {k: len(v) for k, v in rec.items()})


if __name__ == "__main__":
    sys.exit(main())

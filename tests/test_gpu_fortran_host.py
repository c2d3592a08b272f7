"""A Fortran host (tests/fortran/run_case.f90, through the iso_c_binding
module tandemproject_amd/fortran/tandem_hip.f90) runs a case on the GPU and
gets the same final state, bit for bit, as the Python host over the same C
ABI: the drop-in works from the reference's own language."""
import os
import subprocess
import sys

import numpy as np
import pytest

from caseutil import aerofoil_case
from tandemproject_amd import api

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = "/opt/rocm/lib/llvm/bin/flang"


@pytest.mark.skipif(not os.path.exists(FLANG), reason="no Fortran compiler in this image")
@pytest.mark.parametrize("exact", [1, 0], ids=["sequential", "chunked"])
def test_fortran_host_runs_steps(exact, tmp_path):
    print("Fortran host, 5 steps,", "sequential" if exact else "chunked", flush=True)
    lay, ss, p = aerofoil_case(lxibk=(16, 20, 16), letbk=(14, 14))
    nsteps = 5
    arrs, ipar, dpar = api.case_args(lay, p)
    d, b, n, s, j, v = arrs
    case = tmp_path / "case.bin"
    with open(case, "wb") as f:
        np.array([lay.nb, lay.njct], np.int32).tofile(f)
        for a in (d, b, n, s):
            a.astype(np.int32).ravel().tofile(f)
        if lay.njct:
            j.astype(np.int32).ravel().tofile(f)
            v.astype(np.int32).ravel().tofile(f)
        ipar.astype(np.int32).tofile(f)
        dpar.astype(np.float64).tofile(f)
        for blk in range(lay.nb):
            np.asarray(ss[blk], np.float64).ravel().tofile(f)
        np.array([nsteps, exact], np.int32).tofile(f)
    mod = os.path.join(ROOT, "tandemproject_amd", "fortran", "tandem_hip.f90")
    prog = os.path.join(ROOT, "tests", "fortran", "run_case.f90")
    libdir = os.path.join(ROOT, "tandemproject_amd")
    subprocess.check_call([FLANG, "-c", "-module-dir", str(tmp_path), mod, "-o", str(tmp_path / "m.o")])
    exe = str(tmp_path / "run_case")
    subprocess.check_call([FLANG, "-I", str(tmp_path), "-module-dir", str(tmp_path), prog, str(tmp_path / "m.o"),
                           "-L" + libdir, "-ltandem_hip", "-Wl,-rpath," + libdir, "-o", exe])
    out = tmp_path / "qa.bin"
    r = subprocess.run([exe, str(case), str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    got = np.fromfile(out, np.float64)
    # the same case through the Python host
    sv = api.Solver([(lay, p)], exact_lines=bool(exact))
    for blk in range(lay.nb):
        sv.set_grid(0, blk, ss[blk])
    sv.setup_geometry(); sv.initialo(0)
    sv.step(nsteps); sv.sync()
    want = np.concatenate([sv.download(0, blk, "qa") for blk in range(lay.nb)])
    assert got.size == want.size
    np.testing.assert_array_equal(got, want)
    assert sv.scalars(0)["n"] == nsteps

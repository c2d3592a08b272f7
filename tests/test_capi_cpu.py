"""CPU-side checks of the C ABI: the library builds, loads, exports every
entry point declared in include/tandem_hip.h, and refuses to run without a
GPU (there is no CPU fallback)."""
import ctypes
import os
import re

import pytest

from tandemproject_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "tandem_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tp_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = api.load()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(api.EXPORTS) == names


def test_no_cpu_fallback_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = api.load()
    h = ctypes.c_void_p()
    rc = lib.tp_create(ctypes.byref(h))
    assert rc != 0
    assert b"no HIP device" in lib.tp_last_error()


def test_layout_matches_reference_decomposition():
    """3x2 H layout, walls at blocks 1 and 4, spanwise split into 2 sub-blocks
    (src/SingleTurbulence3D.f90:121-164, src/Main3D.f90:132-178)."""
    from tandemproject_amd.layout import h_layout
    L = h_layout([24, 32, 24], [20, 20], [31], npcz=[2])
    assert L.nb == 12
    blk = lambda mb, kp: mb * 2 + kp
    # block 1 (bottom middle): eta end is the viscous wall, block 4 eta start too
    assert tuple(L.nbc[blk(1, 0), 1]) == (10, 25)
    assert tuple(L.nbc[blk(4, 1), 1]) == (25, 10)
    # zeta: periodic at the outer ends, closure interface (40) between halves
    assert tuple(L.nbc[blk(0, 0), 2]) == (45, 40) and tuple(L.nbc[blk(0, 1), 2]) == (40, 45)
    assert L.ncd[blk(0, 0), 2, 0] == blk(0, 1) and L.ncd[blk(0, 0), 2, 1] == blk(0, 1)
    # xi interfaces pair up with the same spanwise half
    assert L.ncd[blk(0, 1), 0, 1] == blk(1, 1) and L.ncd[blk(1, 1), 0, 0] == blk(0, 1)
    # sub-block sizes: 32 points split as 16 + 16
    assert L.dims[blk(2, 0), 2] == 15 and L.dims[blk(2, 1), 2] == 15
    # junctions: leading and trailing edge per spanwise half, velocities from the wall blocks
    assert L.njct == 4
    le = L.jct[0]
    assert [m[0] for m in le] == [blk(0, 0), blk(1, 0), blk(3, 0), blk(4, 0)]
    assert list(L.jvel[0]) == [1, 3] and list(L.jvel[1]) == [0, 2]


FLANG = "/opt/rocm/lib/llvm/bin/flang"


@pytest.mark.skipif(not os.path.exists(FLANG), reason="no Fortran compiler in this image")
def test_fortran_binding_compiles_links_and_runs(tmp_path):
    """The iso_c_binding module a Fortran host uses (tandemproject_amd/fortran)
    declares the same entry points as the C header, compiles, links against
    the library and makes calls that need no GPU."""
    import subprocess
    import torch
    mod = os.path.join(ROOT, "tandemproject_amd", "fortran", "tandem_hip.f90")
    src = open(mod).read().lower()
    bound = set(re.findall(r'bind\(c, name="(tp_[a-z0-9_]+)"\)', src))
    assert bound == set(declared_symbols())
    prog = os.path.join(ROOT, "tests", "fortran", "check_binding.f90")
    libdir = os.path.join(ROOT, "tandemproject_amd")
    subprocess.check_call([FLANG, "-c", "-module-dir", str(tmp_path), mod, "-o", str(tmp_path / "m.o")])
    exe = str(tmp_path / "check_binding")
    subprocess.check_call([FLANG, "-I", str(tmp_path), "-module-dir", str(tmp_path), prog, str(tmp_path / "m.o"),
                           "-L" + libdir, "-ltandem_hip", "-Wl,-rpath," + libdir, "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "binding ok" in out.stdout
    if not torch.cuda.is_available():
        assert "tp_create failed as expected" in out.stdout and "no HIP device" in out.stdout


def test_case_checks_without_gpu():
    """tp_check_case refuses, without a context, what tp_add_case refuses:
    a block of 2^29 points or more (32-bit byte offsets in the line solves),
    periodic lines shorter than the 13 points their closure reads, unknown
    boundary codes; and accepts the reference's 3x2 layout."""
    from caseutil import BASE_PARAMS
    from tandemproject_amd import layout as L
    api.check_case(L.h_layout([24, 32, 24], [20, 20], [31], npcz=[2]), BASE_PARAMS)
    with pytest.raises(api.TPError, match="2\\^29"):
        api.check_case(L.single_block(1023, 1023, 511, [[10, 10], [10, 10], [45, 45]]), BASE_PARAMS)
    api.check_case(L.single_block(1023, 1023, 510, [[10, 10], [10, 10], [45, 45]]), BASE_PARAMS)
    with pytest.raises(api.TPError, match="its faces need >= 13"):
        api.check_case(L.single_block(20, 20, 8, [[10, 10], [10, 10], [45, 45]]), BASE_PARAMS)
    with pytest.raises(api.TPError, match="unknown boundary code"):
        api.check_case(L.single_block(20, 20, 20, [[10, 11], [10, 10], [45, 45]]), BASE_PARAMS)


def test_oracle_refuses_bad_cases():
    """The oracle answers a case it cannot hold with an error, not an abort:
    a 9-point periodic line (its closure reads 13 points) and a selector out
    of range raise OracleError in this process."""
    import pyoracle as po
    from caseutil import BASE_PARAMS
    from tandemproject_amd import layout as L
    with pytest.raises(po.OracleError, match="has 9 points; its faces need >= 13"):
        po.Oracle(L.single_block(20, 20, 8, [[10, 10], [10, 10], [45, 45]]), BASE_PARAMS)
    o = po.Oracle(L.single_block(20, 20, 20, [[10, 10], [10, 10], [45, 45]]), BASE_PARAMS)
    with pytest.raises(po.OracleError, match="out of range"):
        o.deriv(4, 1, 1)
    with pytest.raises(po.OracleError, match="out of range"):
        o.stage_part(9, "begin")


def test_host_buffers_match_the_header_sizes():
    """The Python host's parameter and scalar buffers have the sizes the C
    entry points write (tp_case_params copies all TP_D_NPARAM doubles, a
    shorter buffer would be overrun), and no output buffer is sized from the
    name lists, which name fewer entries than the arrays hold."""
    src = open(os.path.join(ROOT, "include", "tandem_hip.h")).read()
    size = {k: int(v) for k, v in re.findall(r"\b(TP_[ID]_NPARAM|TP_S_NSCALAR) = (\d+)", src)}
    assert api.N_IPARAM == size["TP_I_NPARAM"] and api.N_DPARAM == size["TP_D_NPARAM"]
    assert len(api.I_NAMES) <= api.N_IPARAM and len(api.D_NAMES) <= api.N_DPARAM
    assert len(api.SCALARS) <= size["TP_S_NSCALAR"]
    host = open(os.path.join(ROOT, "tandemproject_amd", "api.py")).read()
    assert not re.search(r"np\.(zeros|empty)\(len\((I_NAMES|D_NAMES|SCALARS)\)", host)
    assert f"np.zeros({size['TP_S_NSCALAR']})" in host     # tp_get_scalars' buffer

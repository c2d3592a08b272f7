"""Python host over the C ABI of libtandem_hip.so.

``Solver`` mirrors the reference's operator interface for the hot path
(src/Subroutines3D.f90 ``mpigo``/``deriv``/``filte`` and the time loop of
src/Main3D.f90:552-965) on device-resident state.  The HIP library is the
only compute path: importing this module on a machine where the library
cannot be loaded, or creating a ``Solver`` without a GPU, raises.
"""
import atexit
import ctypes
import os
import sys
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtandem_hip.so")

I_NAMES = ["nkrk", "nviscous", "nsmf", "les", "nextrabc", "nextgcic", "inflow", "nmodes", "seed", "ntz", "forcing"]
D_NAMES = ["reoo", "tempoo", "amach1", "amach2", "amach3", "wtemp", "cfl", "tmax",
           "timf", "fltk", "fltkbc", "dto", "aoa0", "aoa1", "tps", "tp", "smago1",
           "smago2", "domlen", "szth1", "szth2", "szxt", "szco", "span", "wlew", "wlea", "tgusto",
           "slit", "tu", "lturb", "gustw", "gustk", "cfit", "amfor", "xfor", "yfor", "rfor", "tsfor", "tefor"]
N_IPARAM, N_DPARAM = 16, 48
ARRAYS = ["qa", "qo", "de", "rr", "ss", "xim", "etm", "zem", "p", "yaco", "tv",
          "drva1", "drva2", "drva3", "drvb1", "drvb2", "drvb3", "send1", "send2", "send3",
          "recv1", "recv2", "recv3", "cm1", "cm2", "cm3", "xu", "xl", "yu", "yl", "rpex"]
SCALARS = ["n", "ndt", "dt", "dts", "dte", "timo", "umf1", "umf2", "umf3", "dudt1",
           "dudt2", "dudt3", "dtk", "dtko", "aoa"]
KCLASS = ["deriv", "filter", "exchange", "point", "flux", "viscous", "gcbc", "rk", "misc"]
EXPORTS = ["tp_create", "tp_destroy", "tp_last_error", "tp_add_case", "tp_finalize",
           "tp_set_grid", "tp_setup_geometry", "tp_initialo", "tp_upload", "tp_download",
           "tp_array_count", "tp_get_scalars", "tp_set_scalars", "tp_mpigo", "tp_deriv",
           "tp_filte", "tp_filter_step", "tp_stage", "tp_step", "tp_step_graph", "tp_sync",
           "tp_profile", "tp_profile_classes", "tp_profile_read", "tp_info", "tp_set_option", "tp_autotune",
           "tp_col_config", "tp_inflow_modes", "tp_gust_mode", "tp_case_bytes", "tp_admit", "tp_admit_cases", "tp_remove_cases",
           "tp_remove_case", "tp_case_info", "tp_server_create", "tp_server_destroy",
           "tp_server_context", "tp_server_error", "tp_server_clock", "tp_server_submit",
           "tp_server_iterate", "tp_server_run", "tp_server_result", "tp_server_result_count",
           "tp_server_stats", "tp_server_set_fill", "tp_stats_start", "tp_stats_forces",
           "tp_stats_wall_count", "tp_stats_wall", "tp_set_parents", "tp_get_parents",
           "tp_case_dims", "tp_case_params", "tp_write_restart", "tp_read_restart", "tp_write_grid",
           "tp_read_grid", "tp_write_plot3d", "tp_get_grid", "tp_check_case",
           "tp_primitives", "tp_viscous", "tp_flux_divergence", "tp_sponge", "tp_gcbc", "tp_rk_update",
           "tp_wall", "tp_junction", "tp_interface_avg", "tp_field_deriv", "tp_field_deriv_noalloc", "tp_adopt_cases", "tp_get_option", "tp_server_set_staging",
           "tp_server_stats2", "tp_eddy_params",
           "tp_set_inflow_eddies"]
# the sections of a stage, in the order tp_stage runs them (name, takes nk)
STAGE_PARTS = [("primitives", True), ("viscous", False), ("flux_divergence", True), ("sponge", False),
               ("gcbc", False), ("rk_update", True), ("wall", False), ("junction", False),
               ("interface_avg", False)]

_lib = None


class TPError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise TPError(f"{path} missing: build it with `python -m tandemproject_amd.build` "
                      "(there is no CPU fallback)")
    # One HIP runtime per process.  PyTorch ships its own libamdhip64 and
    # libhsa-runtime64 and asks for them by the names libamdhip64.so and
    # libhsa-runtime64.so; this library asks for the sonames .so.7 / .so.1.
    # Without care a process that loads this library and then imports torch
    # maps two HIP and two HSA runtimes, and the second finds no GPU.  The
    # system runtime (the one this library is built against) is therefore
    # loaded under torch's names first, so that a later `import torch` binds
    # to it.  A process that imported torch before has torch's runtime
    # loaded; this library then binds to that one (same soname).
    if "torch" not in sys.modules:
        for name in ("libhsa-runtime64.so", "libamdhip64.so"):
            try:
                ctypes.CDLL(name, mode=ctypes.RTLD_GLOBAL)
            except OSError:
                pass
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    i32, i64, dbl = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    pi32 = ctypes.POINTER(ctypes.c_int32)
    pd = ctypes.POINTER(ctypes.c_double)
    L.tp_last_error.restype = ctypes.c_char_p
    sig = {
        "tp_create": [ctypes.POINTER(P)], "tp_destroy": [P],
        "tp_add_case": [P, i32, pi32, pi32, pi32, pi32, i32, pi32, pi32, pi32, pd, pi32],
        "tp_finalize": [P], "tp_set_grid": [P, i32, i32, pd], "tp_setup_geometry": [P],
        "tp_initialo": [P, i32], "tp_upload": [P, i32, i32, i32, pd, i64],
        "tp_download": [P, i32, i32, i32, pd, i64], "tp_get_scalars": [P, i32, pd],
        "tp_set_scalars": [P, i32, pd], "tp_mpigo": [P, i32, i32, i32],
        "tp_deriv": [P, i32, i32, i32], "tp_filte": [P, i32, i32], "tp_filter_step": [P],
        "tp_stage": [P, i32], "tp_step": [P, i32],
        "tp_check_case": [i32, pi32, pi32, pi32, i32, pi32, pi32, pi32],
        "tp_primitives": [P, i32], "tp_viscous": [P], "tp_flux_divergence": [P, i32], "tp_sponge": [P],
        "tp_gcbc": [P], "tp_rk_update": [P, i32], "tp_wall": [P], "tp_junction": [P],
        "tp_interface_avg": [P],
        "tp_field_deriv": [i32, i32, i32, pi32, i32, P, P, P],
        "tp_field_deriv_noalloc": [i32, i32, i32, pi32, i32, P, P, P],
        "tp_eddy_params": [i32, pd, pd, pd],
        "tp_set_inflow_eddies": [P, i32, i32, pd],
        "tp_adopt_cases": [P, i32, ctypes.POINTER(P), pi32, pi32],
        "tp_get_option": [P, i32, ctypes.POINTER(i64)],
        "tp_server_set_staging": [P, i32], "tp_server_stats2": [P, pd], "tp_step_graph": [P, i32], "tp_sync": [P],
        "tp_profile": [P, i32], "tp_profile_classes": [P, i32],
        "tp_profile_read": [P, i32, pd, ctypes.POINTER(i64), pd, i32],
        "tp_info": [P, ctypes.POINTER(i64)],
        "tp_set_option": [P, i32, i64],
        "tp_autotune": [P, i32, pi32, pd],
        "tp_col_config": [P, pi32],
        "tp_inflow_modes": [i32, ctypes.c_uint64, dbl, dbl, pd, dbl, dbl, pd],
        "tp_gust_mode": [dbl, dbl, pd, pd, pd],
        "tp_case_bytes": [P, i32, ctypes.POINTER(i64)],
        "tp_admit": [P, i32], "tp_remove_case": [P, i32], "tp_admit_cases": [P, i32, pi32, pi32],
        "tp_remove_cases": [P, i32, pi32],
        "tp_case_info": [P, i32, ctypes.POINTER(i64)],
        "tp_server_create": [i64, i32, i32, ctypes.POINTER(P)],
        "tp_server_destroy": [P],
        "tp_server_submit": [P, i32, pi32, pi32, pi32, pi32, i32, pi32, pi32, pi32, pd,
                             ctypes.POINTER(pd), i32, dbl, pi32],
        "tp_server_iterate": [P, i32, pi32], "tp_server_run": [P, i32],
        "tp_server_result": [P, i32, pd, i64, pd],
        "tp_server_stats": [P, pd], "tp_server_set_fill": [P, i64],
        "tp_stats_start": [P, i32, i32], "tp_stats_forces": [P, i32, pd],
        "tp_stats_wall": [P, i32, i32, pd, pd, pd, pd, i64],
        "tp_set_parents": [P, i32, i32, pi32, pi32, pi32],
        "tp_get_parents": [P, i32, pi32, pi32, pi32, pi32],
        "tp_case_dims": [P, i32, pi32], "tp_case_params": [P, i32, pi32, pd],
        "tp_write_restart": [P, i32, ctypes.c_char_p], "tp_read_restart": [P, i32, ctypes.c_char_p],
        "tp_write_grid": [P, i32, ctypes.c_char_p], "tp_read_grid": [P, i32, ctypes.c_char_p],
        "tp_write_plot3d": [P, i32, ctypes.c_char_p, i32],
        "tp_get_grid": [P, i32, i32, pd, i64],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = i32
    L.tp_array_count.argtypes = [P, i32, i32, i32]
    L.tp_array_count.restype = i64
    L.tp_server_context.argtypes = [P]
    L.tp_server_context.restype = P
    L.tp_server_error.argtypes = [P]
    L.tp_server_error.restype = ctypes.c_char_p
    L.tp_server_clock.argtypes = [P]
    L.tp_server_clock.restype = dbl
    L.tp_server_result_count.argtypes = [P, i32]
    L.tp_server_result_count.restype = i64
    L.tp_stats_wall_count.argtypes = [P, i32, i32]
    L.tp_stats_wall_count.restype = i64
    _lib = L
    return L


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc):
    if rc != 0:
        raise TPError(load().tp_last_error().decode())


def inflow_modes(nmodes, seed, tu, lturb, uoo, slit, span):
    """Seeded random Fourier modes of the inflow turbulence (tp_inflow_modes):
    array (nmodes, 8) of k (3), a (3), phase, 0.  Needs no GPU."""
    out = np.zeros((int(nmodes), 8))
    u = np.ascontiguousarray(uoo, np.float64)
    _check(load().tp_inflow_modes(int(nmodes), int(seed), float(tu), float(lturb), _dp(u), float(slit),
                                  float(span), _dp(out)))
    return out


def gust_mode(tu, kr, uoo):
    """Single Fourier mode of a sinusoidal transverse gust (tp_gust_mode):
    (array (1, 8) in the layout of inflow_modes, x-period slit).  Needs no GPU."""
    out = np.zeros((1, 8))
    u = np.ascontiguousarray(uoo, np.float64)
    slit = ctypes.c_double()
    _check(load().tp_gust_mode(float(tu), float(kr), _dp(u), _dp(out), ctypes.byref(slit)))
    return out, slit.value


def case_args(layout, params):
    """Arrays of tp_add_case for a block layout and a parameter dict: the
    layout arrays, iparams and dparams (unnamed parameters are 0, nkrk 4,
    nviscous 1)."""
    ipar = np.zeros(N_IPARAM, np.int32)
    for i, k in enumerate(I_NAMES):
        ipar[i] = int(params.get(k, {"nkrk": 4, "nviscous": 1}.get(k, 0)))
    dpar = np.zeros(N_DPARAM)
    for i, k in enumerate(D_NAMES):
        dpar[i] = float(params.get(k, 0.0))
    arrs = [np.ascontiguousarray(a, np.int32) for a in
            (layout.dims, layout.nbc, layout.ncd, layout.nsz, layout.jct, layout.jvel)]
    return arrs, ipar, dpar


def _hip():
    """The HIP runtime this library is bound to (by its soname)."""
    load()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    h.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
    h.hipHostFree.argtypes = [ctypes.c_void_p]
    return h


def device_memory():
    """(free, total) device bytes of the current device (hipMemGetInfo)."""
    fr, tot = ctypes.c_size_t(), ctypes.c_size_t()
    if _hip().hipMemGetInfo(ctypes.byref(fr), ctypes.byref(tot)) != 0:
        raise TPError("hipMemGetInfo failed")
    return fr.value, tot.value


def device_synchronize():
    if _hip().hipDeviceSynchronize() != 0:
        raise TPError("hipDeviceSynchronize failed")


class DeviceArray:
    """n float64 values of device memory on the current device, for the
    context-free entry points (tp_field_deriv*) that take caller-owned
    device pointers.  Copies are synchronous (hipMemcpy)."""

    def __init__(self, n):
        self._hip = _hip()
        self.n = int(n)
        self.ptr = ctypes.c_void_p()
        if self._hip.hipMalloc(ctypes.byref(self.ptr), 8 * max(self.n, 1)) != 0:
            self.ptr = None
            raise TPError("hipMalloc failed")

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64).reshape(-1)
        if a.size != self.n:
            raise TPError(f"expected {self.n} values, got {a.size}")
        if self._hip.hipMemcpy(self.ptr, a.ctypes.data, 8 * self.n, 1) != 0:   # host to device
            raise TPError("hipMemcpy to the device failed")
        return self

    def download(self):
        out = np.empty(self.n)
        if self._hip.hipMemcpy(out.ctypes.data, self.ptr, 8 * self.n, 2) != 0:   # device to host
            raise TPError("hipMemcpy to the host failed")
        return out

    @property
    def address(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            self._hip.hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


class PinnedBuffer:
    """n float64 values of page-locked host memory (hipHostMalloc).  Device
    downloads into it run at full link speed; a freshly allocated pageable
    array takes a page fault per page during the copy, several times slower
    for states of hundreds of MB.  `array` is valid until free()."""

    def __init__(self, n):
        self._p = None
        self.array = None
        self._hip = _hip()
        self._p = ctypes.c_void_p()
        m = max(int(n), 1)
        if self._hip.hipHostMalloc(ctypes.byref(self._p), ctypes.c_size_t(8 * m), ctypes.c_uint(0)) != 0:
            self._p = None
            raise RuntimeError("hipHostMalloc failed")
        self.array = np.ctypeslib.as_array((ctypes.c_double * m).from_address(self._p.value))[:int(n)]

    def free(self):
        if self._p:
            self.array = None
            self._hip.hipHostFree(self._p)
            self._p = None

    def __del__(self):
        self.free()


def eddy_params(nits, randnum, deck):
    """tp_eddy_params: the reference's synthetic eddies, (nits, 10) array of
    centre (3), 1/size^2 (3), amplitude (3), type.  deck: dict with span,
    slit, gaus, cutlb, domlen, szth1, tl0, tl1, tlw (2 each), tam (2 x 3)."""
    r = np.ascontiguousarray(randnum[:8 * nits], np.float64)
    d = np.array([deck["span"], deck["slit"], deck["gaus"], deck["cutlb"], deck["domlen"], deck["szth1"],
                  deck["tl0"][0], deck["tl1"][0], deck["tlw"][0], deck["tl0"][1], deck["tl1"][1], deck["tlw"][1],
                  *deck["tam"][0], *deck["tam"][1]], np.float64)
    out = np.zeros((nits, 10))
    _check(load().tp_eddy_params(int(nits), _dp(r), _dp(d), _dp(out)))
    return out


def field_deriv(dims3, nbc2, nn, src_ptr, dst_ptr, stream=None):
    """tp_field_deriv on device addresses (e.g. torch tensors' data_ptr())."""
    nbc = np.ascontiguousarray(nbc2, np.int32)
    _check(load().tp_field_deriv(int(dims3[0]), int(dims3[1]), int(dims3[2]), _ip(nbc), int(nn),
                                 ctypes.c_void_p(src_ptr), ctypes.c_void_p(dst_ptr), ctypes.c_void_p(stream)))


def field_deriv_noalloc(dims3, nbc2, nn, src_ptr, dst_ptr, stream=None):
    """tp_field_deriv_noalloc on device addresses: no allocation, work area
    in the library's per-GPU array; enqueued only (synchronise the stream)."""
    nbc = np.ascontiguousarray(nbc2, np.int32)
    _check(load().tp_field_deriv_noalloc(int(dims3[0]), int(dims3[1]), int(dims3[2]), _ip(nbc), int(nn),
                                         ctypes.c_void_p(src_ptr), ctypes.c_void_p(dst_ptr),
                                         ctypes.c_void_p(stream)))


def check_case(layout, params):
    """tp_check_case: the checks tp_add_case makes, without a context (raises TPError)."""
    arrs, ipar, _ = case_args(layout, params)
    d, b, n, _, j, v = arrs
    _check(load().tp_check_case(layout.nb, _ip(d), _ip(b), _ip(n), layout.njct, _ip(j), _ip(v), _ip(ipar)))


class Solver:
    """Device-resident state of one or more cases (each a list of sub-blocks)."""

    def __init__(self, cases, chunks=0, exact_lines=False):
        """cases: list of (layout, params) pairs.  chunks: fixed number of
        chunks per line of the chunked line sweeps (1 to 16; 0: chosen
        per direction); exact_lines: sequential sweeps, one lane per line."""
        L = load()
        self.h = ctypes.c_void_p()
        _check(L.tp_create(ctypes.byref(self.h)))
        self.layouts = {}
        self._keep = []
        for layout, params in cases:
            self.add_case(layout, params)
        _check(L.tp_set_option(self.h, 1, int(chunks)))
        _check(L.tp_finalize(self.h))
        if exact_lines:
            self.set_exact_lines(True)

    def add_case(self, layout, params):
        """Add a case; returns its handle.  After construction the case is
        pending until admit()."""
        arrs, ipar, dpar = case_args(layout, params)
        self._keep.append((arrs, ipar, dpar))
        cid = ctypes.c_int32()
        d, b, n, s, j, v = arrs
        _check(load().tp_add_case(self.h, layout.nb, _ip(d), _ip(b), _ip(n), _ip(s), layout.njct,
                                  _ip(j), _ip(v), _ip(ipar), _dp(dpar), ctypes.byref(cid)))
        self.layouts[cid.value] = layout
        return cid.value

    def set_parents(self, case, layout):
        """Parent blocks of the reference's file formats from a layout with
        parent/start/blockdims (layout.h_layout fills them)."""
        par = np.ascontiguousarray(layout.parent, np.int32)
        st = np.ascontiguousarray(layout.start, np.int32).ravel()
        pd_ = np.ascontiguousarray(layout.blockdims, np.int32).ravel()
        _check(load().tp_set_parents(self.h, case, int(len(layout.blockdims)), _ip(par), _ip(st), _ip(pd_)))

    def write_restart(self, case, rundir):
        _check(load().tp_write_restart(self.h, case, str(rundir).encode()))

    def read_restart(self, case, rundir):
        _check(load().tp_read_restart(self.h, case, str(rundir).encode()))

    def write_grid(self, case, rundir):
        _check(load().tp_write_grid(self.h, case, str(rundir).encode()))

    def read_grid(self, case, rundir):
        _check(load().tp_read_grid(self.h, case, str(rundir).encode()))

    def write_plot3d(self, case, rundir, grid=True, solution=True):
        _check(load().tp_write_plot3d(self.h, case, str(rundir).encode(), (1 if grid else 0) | (2 if solution else 0)))

    def case_bytes(self, case):
        """Device bytes the case holds once admitted (needs its grids)."""
        n = ctypes.c_int64()
        _check(load().tp_case_bytes(self.h, case, ctypes.byref(n)))
        return n.value

    def admit(self, case):
        """Put a pending case on the device (allocation, geometry); its state
        is zero until initialo() or upload()."""
        _check(load().tp_admit(self.h, case))

    def admit_cases(self, cases):
        """Put several pending cases on the device at once (one rebuild of
        the launch tables); returns the per-case status codes."""
        ids = np.ascontiguousarray(cases, np.int32)
        st = np.zeros(len(ids), np.int32)
        rc = load().tp_admit_cases(self.h, len(ids), _ip(ids), _ip(st))
        if rc and not st.any():
            _check(rc)
        return st

    def remove_case(self, case):
        _check(load().tp_remove_case(self.h, case))
        self.layouts.pop(case, None)

    def remove_cases(self, cases):
        ids = np.ascontiguousarray(cases, np.int32)
        _check(load().tp_remove_cases(self.h, len(ids), _ip(ids)))
        for c in cases:
            self.layouts.pop(c, None)

    def case_info(self, case):
        out = (ctypes.c_int64 * 4)()
        _check(load().tp_case_info(self.h, case, out))
        return dict(state=out[0], points=out[1], blocks=out[2], device_bytes=out[3])

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            load().tp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_exact_lines(self, exact=True):
        """Sequential one-lane-per-line sweeps (bit-identical to the reference's
        order) instead of the chunked sweeps."""
        _check(load().tp_set_option(self.h, 0, 1 if exact else 0))

    def set_keep_stresses(self, keep=True):
        """Leave the viscous stresses in tv at every point after every stage
        (default: only at the wall rows the surface forces read)."""
        _check(load().tp_set_option(self.h, 2, 1 if keep else 0))

    def set_chunks(self, chunks):
        """Chunks per line of the chunked sweeps (0: the library's choice)."""
        _check(load().tp_set_option(self.h, 1, int(chunks)))

    def autotune(self, reps=10):
        """Time every line-solver configuration on these blocks and keep the
        fastest per direction and kind.  Returns {(direction, kind): (chunks,
        microseconds per launch)}; chunks 0 is the sequential kernel.
        Overwrites the stage's gradient and flux scratch (fx), so call it
        between steps."""
        ch = np.zeros(6, np.int32); us = np.zeros(6)
        _check(load().tp_autotune(self.h, int(reps), _ip(ch), _dp(us)))
        return {(i // 2, ("deriv", "filter")[i % 2]): (int(ch[i]), float(us[i])) for i in range(6)}

    def col_config(self):
        """Chunks per line in use: {(direction, kind): chunks} (0: sequential)."""
        ch = np.zeros(6, np.int32)
        _check(load().tp_col_config(self.h, _ip(ch)))
        return {(i // 2, ("deriv", "filter")[i % 2]): int(ch[i]) for i in range(6)}

    # ---- data ----
    def set_grid(self, case, block, xyz):
        x = np.ascontiguousarray(xyz, np.float64)
        _check(load().tp_set_grid(self.h, case, block, _dp(x)))

    def setup_geometry(self):
        _check(load().tp_setup_geometry(self.h))

    def initialo(self, case=0):
        _check(load().tp_initialo(self.h, case))

    def count(self, case, block, name):
        return int(load().tp_array_count(self.h, case, block, ARRAYS.index(name)))

    def upload(self, case, block, name, values):
        v = np.ascontiguousarray(values, np.float64).ravel()
        _check(load().tp_upload(self.h, case, block, ARRAYS.index(name), _dp(v), v.size))

    def download(self, case, block, name, out=None):
        """The array's values in the reference's layout.  out: a C-contiguous
        float64 array of count() values to fill (e.g. a PinnedBuffer's);
        default a new array."""
        n = self.count(case, block, name)
        if out is None:
            out = np.empty(max(n, 0), np.float64)
        elif out.dtype != np.float64 or not out.flags.c_contiguous or out.size != n:
            raise ValueError(f"download: out must hold {n} contiguous float64 values")
        _check(load().tp_download(self.h, case, block, ARRAYS.index(name), _dp(out), n))
        return out

    def scalars(self, case=0):
        s = np.zeros(16)
        _check(load().tp_get_scalars(self.h, case, _dp(s)))
        return {k: float(s[i]) for i, k in enumerate(SCALARS)}

    def set_scalars(self, case, values):
        s = np.zeros(16)
        _check(load().tp_get_scalars(self.h, case, _dp(s)))
        for k, v in values.items():
            s[SCALARS.index(k)] = v
        _check(load().tp_set_scalars(self.h, case, _dp(s)))

    # ---- surface statistics ----
    def stats_start(self, case=0, every=1):
        """Sample wall forces and pressure every `every` steps (0: stop)."""
        _check(load().tp_stats_start(self.h, case, int(every)))

    def stats_forces(self, case=0):
        o = np.zeros(8)
        _check(load().tp_stats_forces(self.h, case, _dp(o)))
        return dict(samples=int(o[0]), t0=o[1], t1=o[2], cl=o[3], cd=o[4], cl_mean=o[5], cd_mean=o[6],
                    window=o[7])

    def stats_wall(self, case, block):
        """(mean p, rms p, area, unit normal (3, n)) per wall point, or None."""
        L = load()
        n = int(L.tp_stats_wall_count(self.h, case, block))
        if n < 0:
            raise TPError("bad case/block")
        if n == 0:
            return None
        m = np.zeros(n); r = np.zeros(n); a = np.zeros(n); w = np.zeros(3 * n)
        _check(L.tp_stats_wall(self.h, case, block, _dp(m), _dp(r), _dp(a), _dp(w), n))
        return m, r, a, w.reshape(3, n)

    def stats_cp(self, case, block):
        """(mean Cp, rms Cp) per wall point, or None: Cp = (p - poo) 2 / M^2
        with poo = 1/gamma and M the free-stream Mach number
        (src/rptPost.f90:693-694, src/Main3D.f90:61)."""
        w = self.stats_wall(case, block)
        if w is None:
            return None
        dp = np.zeros(N_DPARAM)   # tp_case_params writes all TP_D_NPARAM values
        _check(load().tp_case_params(self.h, case, None, _dp(dp)))
        m2 = sum(dp[D_NAMES.index(k)] ** 2 for k in ("amach1", "amach2", "amach3"))
        return (w[0] - 1.0 / 1.4) * 2.0 / m2, w[1] * 2.0 / m2

    # ---- reference routines (all sub-blocks) ----
    def mpigo(self, nt, nrt, n45):
        _check(load().tp_mpigo(self.h, nt, nrt, n45))

    def deriv(self, nn, nz, m):
        _check(load().tp_deriv(self.h, nn, nz, m))

    def filte(self, nn, nz):
        _check(load().tp_filte(self.h, nn, nz))

    def filter_step(self):
        _check(load().tp_filter_step(self.h))

    def stage(self, nk):
        _check(load().tp_stage(self.h, nk))

    def adopt(self, others):
        """Move cases set up in other Solvers into this one: others = list of
        (solver, case handle); returns the new handles here."""
        n = len(others)
        srcs = (ctypes.c_void_p * n)(*[o.h.value for o, _ in others])
        ids = np.array([c for _, c in others], np.int32)
        new = np.zeros(n, np.int32)
        _check(load().tp_adopt_cases(self.h, n, srcs, _ip(ids), _ip(new)))
        for (o, c), nc in zip(others, new):
            self.layouts[int(nc)] = o.layouts.pop(c, None) if hasattr(o, "layouts") else None
        return [int(x) for x in new]

    def set_inflow_eddies(self, case, eddies):
        e = np.ascontiguousarray(eddies, np.float64)
        _check(load().tp_set_inflow_eddies(self.h, int(case), e.shape[0], _dp(e)))

    def stage_part(self, nk, part):
        """One section of stage nk (a name of STAGE_PARTS): tp_primitives,
        tp_viscous, ... tp_interface_avg."""
        takes_nk = dict(STAGE_PARTS)[part]
        fn = getattr(load(), "tp_" + part)
        _check(fn(self.h, nk) if takes_nk else fn(self.h))

    def step(self, n=1, graph=False):
        if graph:
            _check(load().tp_step_graph(self.h, n))
        else:
            _check(load().tp_step(self.h, n))

    def sync(self):
        _check(load().tp_sync(self.h))

    # ---- measurement ----
    def profile(self, on=True):
        _check(load().tp_profile(self.h, 1 if on else 0))

    def profile_classes(self, classes):
        """Time only these kernel classes (names from KCLASS)."""
        mask = 0
        for k in classes:
            mask |= 1 << KCLASS.index(k)
        _check(load().tp_profile_classes(self.h, mask))

    def profile_read(self, kclass, reset=False):
        ms = ctypes.c_double(); n = ctypes.c_int64(); by = ctypes.c_double()
        _check(load().tp_profile_read(self.h, KCLASS.index(kclass), ctypes.byref(ms), ctypes.byref(n),
                                      ctypes.byref(by), 1 if reset else 0))
        return ms.value, n.value, by.value

    def info(self):
        out = (ctypes.c_int64 * 4)()
        _check(load().tp_info(self.h, out))
        return dict(points=out[0], blocks=out[1], cases=out[2], device_bytes=out[3])


# servers still open when the interpreter exits are closed before it tears
# its modules down: each joins its staging thread and frees its contexts while
# the HIP runtime is whole
_OPEN_SERVERS = weakref.WeakSet()


@atexit.register
def _close_open_servers():
    for s in list(_OPEN_SERVERS):
        s.close()


class Server:
    """Request loop over one context (tp_server_*): cases queue with an arrival
    time, are admitted while their device bytes fit the HBM budget and a slot
    is free, advance together in shared launches, and leave with their final
    qa copied to the host."""

    STATES = ("queued", "running", "done", "failed")

    def __init__(self, budget_bytes=0, max_active=64, graph=True, fill_points=0):
        L = load()
        self.h = ctypes.c_void_p()
        rc = L.tp_server_create(int(budget_bytes), int(max_active), 1 if graph else 0, ctypes.byref(self.h))
        if rc != 0:
            raise TPError(L.tp_last_error().decode())
        self.ctx = ctypes.c_void_p(L.tp_server_context(self.h))
        self.layouts = {}
        _OPEN_SERVERS.add(self)
        if fill_points:
            self._check(L.tp_server_set_fill(self.h, int(fill_points)))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            load().tp_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise TPError(load().tp_server_error(self.h).decode())

    def set_option(self, option, value):
        _check(load().tp_set_option(self.ctx, int(option), int(value)))

    def clock(self):
        return load().tp_server_clock(self.h)

    def submit(self, layout, params, grids, nsteps, t_arrival=-1.0):
        """Queue a case: grids[b] = (3, N_b) coordinates of sub-block b."""
        arrs, ipar, dpar = case_args(layout, params)
        g = [np.ascontiguousarray(x, np.float64) for x in grids]
        ptrs = (ctypes.POINTER(ctypes.c_double) * len(g))(*[_dp(x) for x in g])
        req = ctypes.c_int32()
        d, b, n, s, j, v = arrs
        self._check(load().tp_server_submit(self.h, layout.nb, _ip(d), _ip(b), _ip(n), _ip(s), layout.njct,
                                            _ip(j), _ip(v), _ip(ipar), _dp(dpar), ptrs, int(nsteps),
                                            float(t_arrival), ctypes.byref(req)))
        self.layouts[req.value] = layout
        return req.value

    def iterate(self, max_steps=10):
        st = np.zeros(6, np.int32)
        self._check(load().tp_server_iterate(self.h, int(max_steps), _ip(st)))
        return dict(zip(("active", "queued", "waiting", "done", "admitted", "steps"), map(int, st)))

    def run(self, max_steps=10):
        self._check(load().tp_server_run(self.h, int(max_steps)))

    def run_stream(self, items, max_steps=10):
        """Serve a stream: a client thread submits items[i] = (layout,
        params, grids, nsteps, t) at t seconds after the call (each arrival
        is the time of its submission) while this thread runs the loop, until
        every request has finished.  Returns the request ids in item order."""
        import threading
        import time
        ids, err = [None] * len(items), []
        t0 = self.clock()

        def client():
            try:
                for i, (lay, params, grids, nsteps, t) in enumerate(items):
                    wait = t0 + t - self.clock()
                    if wait > 0:
                        time.sleep(wait)
                    ids[i] = self.submit(lay, params, grids, nsteps)
            except Exception as e:   # reported by the loop's thread
                err.append(e)

        th = threading.Thread(target=client, daemon=True)
        th.start()
        try:
            while th.is_alive():
                self.run(max_steps)     # returns when nothing submitted is left open
                time.sleep(0.001)
            self.run(max_steps)         # the last submissions
        finally:
            th.join()
        if err:
            raise err[0]
        return ids

    def result(self, req, with_state=True):
        """(info dict, list of per-block qa arrays or None)."""
        info = np.zeros(6)
        L = load()
        self._check(L.tp_server_result(self.h, int(req), None, 0, _dp(info)))
        d = dict(state=self.STATES[int(info[0])], arrival=info[1], admitted=info[2], done=info[3],
                 steps=int(info[4]), device_bytes=int(info[5]))
        if not with_state or d["state"] != "done":
            return d, None
        n = L.tp_server_result_count(self.h, int(req))
        out = np.empty(n)
        self._check(L.tp_server_result(self.h, int(req), _dp(out), n, None))
        lay = self.layouts[req]
        parts, off = [], 0
        for b in range(lay.nb):
            m = 5 * lay.npoints(b)
            parts.append(out[off:off + m])
            off += m
        return d, parts

    def set_staging(self, ahead):
        """Requests set up ahead of admission on the staging thread (>= 1)."""
        self._check(load().tp_server_set_staging(self.h, int(ahead)))

    def stats(self):
        o = np.zeros(16)
        self._check(load().tp_server_stats(self.h, _dp(o)))
        keys = ("finished", "failed", "p50_s", "p99_s", "mean_s", "max_s", "wall_s", "cases_per_hour",
                "peak_bytes", "peak_cases", "joined_running", "stepping_s", "admit_s", "retire_s",
                "capture_s", "captures")
        out = dict(zip(keys, (float(x) for x in o)))
        o2 = np.zeros(6)
        self._check(load().tp_server_stats2(self.h, _dp(o2)))
        out.update(zip(("stage_s", "staged", "stage_failed", "staged_peak", "eager_iterations", "staging_window"),
                       (float(x) for x in o2)))
        return out

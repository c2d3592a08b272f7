"""ctypes wrapper over oracle/liboracle.so (the CPU restatement).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by ``bench.py`` for its ``cpu_baseline`` leg.  Product
code never imports this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB: another build of the oracle (the tests that plant mistakes in a
# copy of it load that copy instead)
LIB_PATH = os.environ.get("ORACLE_LIB", os.path.join(HERE, "liboracle.so"))

I_NKRK, I_NVISCOUS, I_NSMF, I_LES, I_NEXTRABC, I_NEXTGCIC, I_INFLOW, I_NMODES, I_SEED, I_NTZ, I_FORCING = range(11)
N_IPARAM = 16
D_NAMES = ["reoo", "tempoo", "amach1", "amach2", "amach3", "wtemp", "cfl", "tmax",
           "timf", "fltk", "fltkbc", "dto", "aoa0", "aoa1", "tps", "tp", "smago1",
           "smago2", "domlen", "szth1", "szth2", "szxt", "szco", "span", "wlew", "wlea",
           "tgusto", "slit", "tu", "lturb", "gustw", "gustk", "cfit", "amfor", "xfor", "yfor", "rfor",
           "tsfor", "tefor"]
N_DPARAM = 48
A_NAMES = ["qa", "qo", "de", "rr", "ss", "xim", "etm", "zem", "p", "yaco", "tv",
           "drva1", "drva2", "drva3", "drvb1", "drvb2", "drvb3", "send1", "send2", "send3",
           "recv1", "recv2", "recv3", "cm1", "cm2", "cm3", "xu", "xl", "yu", "yl",
           "rpex", "sbcc", "asz", "bsz", "vito", "atz"]
# sections of a stage, in the order orc_stage runs them (tp_oracle.h)
PARTS = ["begin", "viscous", "flux", "sponge", "gcbc", "rk", "wall", "junction", "avg"]
S_NAMES = ["n", "ndt", "dt", "dts", "dte", "timo", "umf1", "umf2", "umf3", "dudt1",
           "dudt2", "dudt3", "dtk", "dtko", "aoa"]

_lib = None


class OracleError(RuntimeError):
    """A case or selector the oracle refused (its message from orc_last_error)."""


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: build with `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int32)
        L.orc_create.restype = P
        L.orc_create.argtypes = [ctypes.c_int, ip, ip, ip, ip, ctypes.c_int, ip, ip, ip, dp]
        L.orc_destroy.argtypes = [P]
        L.orc_npoints.argtypes = [P, ctypes.c_int]
        L.orc_count.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.orc_array.restype = dp
        L.orc_array.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.orc_scalars.restype = dp
        L.orc_scalars.argtypes = [P]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_coefficients.argtypes = [P, dp, dp, dp, dp, dp, dp]
        for f in ("orc_metrics", "orc_initialo", "orc_filter_step"):
            getattr(L, f).argtypes = [P]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_mpigo.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_deriv.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_filte.argtypes = [P, ctypes.c_int, ctypes.c_int]
        for f in ("orc_mpigo", "orc_deriv", "orc_filte"):
            getattr(L, f).restype = ctypes.c_int
        L.orc_stage.argtypes = [P, ctypes.c_int]
        L.orc_stage_part.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.orc_stage_part.restype = ctypes.c_int
        L.orc_step.argtypes = [P, ctypes.c_int]
        L.orc_stats_start.argtypes = [P, ctypes.c_int]
        L.orc_stats_forces.argtypes = [P, dp]
        L.orc_stats_wall.argtypes = [P, ctypes.c_int, dp, dp, dp, dp]
        L.orc_stats_wall.restype = ctypes.c_int
        L.orc_penta.argtypes = [ctypes.c_int, dp, dp, ctypes.c_double, ctypes.c_double, dp, dp]
        L.orc_fcint.argtypes = [ctypes.c_double, ctypes.c_double, dp]
        L.orc_fcbcm.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp, dp, dp]
        L.orc_sbcco.argtypes = [ctypes.c_double, dp, dp, dp]
        L.orc_mtrxi.argtypes = [ctypes.c_int, dp, dp, dp]
        L.orc_xtq2r.argtypes = [dp, ctypes.c_double, dp, dp, dp]
        L.orc_xtr2q.argtypes = [dp, ctypes.c_double, dp, dp, dp]
        L.orc_movef.argtypes = [ctypes.c_int] + [ctypes.c_double] * 4 + [dp] + \
            [ctypes.c_double] * 5 + [dp, dp, dp]
        L.orc_sponge_list.argtypes = [P, ctypes.c_int, ip, ip]
        L.orc_sponge_list.restype = ctypes.c_int
        L.orc_set_inflow_eddies.argtypes = [P, ctypes.c_int, dp]
        L.orc_force_list.argtypes = [P, ctypes.c_int, ip, dp]
        L.orc_force_list.restype = ctypes.c_int
        L.orc_set_inflow_modes.argtypes = [P, ctypes.c_int, dp]
        L.orc_inflow_weights.restype = ctypes.c_int
        L.orc_inflow_weights.argtypes = [P, ctypes.c_double, ctypes.c_double, ip, dp, dp]
        L.orc_pow15.argtypes = [ctypes.c_double]
        L.orc_pow15.restype = ctypes.c_double
        L.orc_sincos.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def _check(rc):
    if rc:
        raise OracleError(lib().orc_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def set_threads(n):
    lib().orc_set_threads(int(n))


class Oracle:
    """One multi-block case held by the CPU restatement."""

    def __init__(self, layout, params):
        L = lib()
        self.layout = layout
        ipar = np.zeros(N_IPARAM, np.int32)
        ipar[I_NKRK] = params.get("nkrk", 4)
        ipar[I_NVISCOUS] = params.get("nviscous", 1)
        ipar[I_NSMF] = params.get("nsmf", 0)
        ipar[I_LES] = params.get("les", 0)
        ipar[I_NEXTRABC] = params.get("nextrabc", 0)
        ipar[I_NEXTGCIC] = params.get("nextgcic", 0)
        ipar[I_INFLOW] = params.get("inflow", 0)
        ipar[I_NMODES] = params.get("nmodes", 0)
        ipar[I_SEED] = params.get("seed", 0)
        ipar[I_NTZ] = params.get("ntz", 0)
        ipar[I_FORCING] = params.get("forcing", 0)
        dpar = np.zeros(N_DPARAM)
        for i, k in enumerate(D_NAMES):
            dpar[i] = params.get(k, 0.0)
        self._keep = [np.ascontiguousarray(a, np.int32) for a in
                      (layout.dims, layout.nbc, layout.ncd, layout.nsz, layout.jct, layout.jvel)]
        d, b, n, s, j, v = self._keep
        self.h = L.orc_create(layout.nb, _ip(d), _ip(b), _ip(n), _ip(s), layout.njct,
                              _ip(j), _ip(v), _ip(ipar), _dp(dpar))
        if not self.h:
            raise OracleError(L.orc_last_error().decode())
        self.nb = layout.nb

    def set_inflow_modes(self, modes):
        """Fourier modes (n, 8) the inflow table is built from (before metrics)."""
        m = np.ascontiguousarray(modes, np.float64)
        self._modes = m
        lib().orc_set_inflow_modes(self.h, m.shape[0], _dp(m))

    def set_inflow_eddies(self, eddies):
        """Synthetic eddies (n, 10) of inflow 3 (before metrics)."""
        e = np.ascontiguousarray(eddies, np.float64)
        self._eddies = e
        lib().orc_set_inflow_eddies(self.h, e.shape[0], _dp(e))

    def force_list(self, b):
        """(point index, xafor (3, n)) of the forcing points of block b"""
        n = lib().orc_force_list(self.h, int(b), None, None)
        l = np.zeros(max(n, 1), np.int32); x = np.zeros(3 * max(n, 1))
        lib().orc_force_list(self.h, int(b), _ip(l), _dp(x))
        return l[:n], x[:3 * n].reshape(3, n)

    def inflow_weights(self, timo, dtk):
        """Table rows, node times and weights of the time interpolation, or None."""
        il = np.zeros(5, np.int32); tl = np.zeros(5); w = np.zeros(5)
        on = lib().orc_inflow_weights(self.h, float(timo), float(dtk), _ip(il), _dp(tl), _dp(w))
        return (il, tl, w) if on else None

    def sponge_list(self, b):
        """(point index, inflow index or -1) of every sponge point of block b"""
        n = lib().orc_sponge_list(self.h, int(b), None, None)
        l = np.zeros(max(n, 1), np.int32); g = np.zeros(max(n, 1), np.int32)
        lib().orc_sponge_list(self.h, int(b), _ip(l), _ip(g))
        return l[:n], g[:n]

    def stats_start(self, every):
        """Sample the surface statistics every `every` steps (0: off)."""
        lib().orc_stats_start(self.h, int(every))

    def stats_forces(self):
        o = np.zeros(8)
        lib().orc_stats_forces(self.h, _dp(o))
        return dict(samples=int(o[0]), t0=o[1], t1=o[2], cl=o[3], cd=o[4], cl_mean=o[5], cd_mean=o[6],
                    window=o[7])

    def stats_wall(self, b):
        """(mean p, rms p, area, unit normal (3, n)) per wall point of block b, or None."""
        n = lib().orc_stats_wall(self.h, int(b), None, None, None, None)
        if n <= 0:
            return None
        m = np.zeros(n); r = np.zeros(n); a = np.zeros(n); w = np.zeros(3 * n)
        lib().orc_stats_wall(self.h, int(b), _dp(m), _dp(r), _dp(a), _dp(w))
        return m, r, a, w.reshape(3, n)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def npoints(self, b):
        return lib().orc_npoints(self.h, b)

    def array(self, b, name):
        """Live numpy view of a block array (Fortran-order flat)."""
        w = A_NAMES.index(name)
        n = lib().orc_count(self.h, b, w)
        if n == 0:
            return np.zeros(0)
        ptr = lib().orc_array(self.h, b, w)
        return np.ctypeslib.as_array(ptr, shape=(n,))

    def scalars(self):
        ptr = lib().orc_scalars(self.h)
        a = np.ctypeslib.as_array(ptr, shape=(16,))
        return {k: float(a[i]) for i, k in enumerate(S_NAMES)}

    def set_scalar(self, name, value):
        ptr = lib().orc_scalars(self.h)
        a = np.ctypeslib.as_array(ptr, shape=(16,))
        a[S_NAMES.index(name)] = value

    def set_grid(self, b, xyz):
        self.array(b, "ss")[:] = np.asarray(xyz, float).reshape(-1)

    def metrics(self):
        lib().orc_metrics(self.h)

    def initialo(self):
        lib().orc_initialo(self.h)

    def mpigo(self, nt, nrt, n45):
        _check(lib().orc_mpigo(self.h, nt, nrt, n45))

    def deriv(self, nn, nz, m):
        _check(lib().orc_deriv(self.h, nn, nz, m))

    def filte(self, nn, nz):
        _check(lib().orc_filte(self.h, nn, nz))

    def filter_step(self):
        lib().orc_filter_step(self.h)

    def stage(self, nk):
        lib().orc_stage(self.h, nk)

    def stage_part(self, nk, part):
        """One section of stage nk (a name of PARTS)."""
        _check(lib().orc_stage_part(self.h, nk, PARTS.index(part)))

    def step(self, n=1):
        lib().orc_step(self.h, n)

    def coefficients(self):
        pbci = np.zeros(48); pbco = np.zeros(48); pbcot = np.zeros(4)
        albef = np.zeros(30); fabc = np.zeros(14); sbci = np.zeros(3)
        lib().orc_coefficients(self.h, _dp(pbci), _dp(pbco), _dp(pbcot), _dp(albef),
                               _dp(fabc), _dp(sbci))
        return dict(pbci=pbci, pbco=pbco, pbcot=pbcot, albef=albef, fabc=fabc, sbci=sbci)


def pow15(x):
    """x**1.5 correctly rounded (the oracle's viscosity law)."""
    return lib().orc_pow15(float(x))


def sincos(x):
    """(sin x, cos x) correctly rounded (the oracle's per-stage scalars)."""
    s, c = ctypes.c_double(), ctypes.c_double()
    lib().orc_sincos(float(x), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def penta(n, albes, albee, alpha, beta):
    xu = np.zeros(3 * n); xl = np.zeros(2 * n)
    a = np.ascontiguousarray(albes, float); e = np.ascontiguousarray(albee, float)
    lib().orc_penta(n, _dp(a), _dp(e), alpha, beta, _dp(xu), _dp(xl))
    return xu, xl


def fcint(fltk, fltr=0.5):
    o = np.zeros(5)
    lib().orc_fcint(fltk, fltr, _dp(o))
    return o


def fcbcm(fltk, fltkbc):
    albef = np.zeros(15); fam = np.zeros(3); fbm = np.zeros(3); fcm = np.zeros(3)
    lib().orc_fcbcm(fltk, fltkbc, _dp(albef), _dp(fam), _dp(fbm), _dp(fcm))
    return albef, fam, fbm, fcm


def sbcco(fltk):
    pbci = np.zeros(48); pbco = np.zeros(48); pbcot = np.zeros(4)
    lib().orc_sbcco(fltk, _dp(pbci), _dp(pbco), _dp(pbcot))
    return pbci.reshape(2, 2, 12), pbco.reshape(2, 2, 12), pbcot.reshape(2, 2)


def mtrxi(a):
    a = np.ascontiguousarray(a, float)
    n = a.shape[0]
    s = np.zeros_like(a); r0 = np.zeros(1)
    lib().orc_mtrxi(n, _dp(a), _dp(s), _dp(r0))
    return s, r0[0]


def xtq2r(q, p, cm, umf):
    o = np.zeros(25)
    q = np.ascontiguousarray(q, float); cm = np.ascontiguousarray(cm, float)
    u = np.ascontiguousarray(umf, float)
    lib().orc_xtq2r(_dp(q), p, _dp(cm), _dp(u), _dp(o))
    return o.reshape(5, 5)


def xtr2q(q, p, cm, umf):
    o = np.zeros(25)
    q = np.ascontiguousarray(q, float); cm = np.ascontiguousarray(cm, float)
    u = np.ascontiguousarray(umf, float)
    lib().orc_xtr2q(_dp(q), p, _dp(cm), _dp(u), _dp(o))
    return o.reshape(5, 5)


def movef(nsmf, timf, timo, dtko, dtk, uoo, tp, tps, aoa0, aoa1, amachoo):
    umf = np.zeros(3); dudt = np.zeros(3); aoa = np.array([aoa0], float)
    u = np.ascontiguousarray(uoo, float)
    lib().orc_movef(nsmf, timf, timo, dtko, dtk, _dp(u), tp, tps, aoa0, aoa1, amachoo,
                    _dp(umf), _dp(dudt), _dp(aoa))
    return umf, dudt, aoa[0]

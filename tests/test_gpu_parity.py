"""Parity of the HIP path with the CPU oracle, through the C ABI.

Per-call drop-ins (mpigo/deriv/filte, src/Subroutines3D.f90) must agree with
the oracle to 1e-13 relative (normwise over each array); metrics to 1e-13;
whole time steps to 1e-10 relative after 100 steps.  The library is compiled
without FP contraction, so most of these agree to the last bit.
"""
import sys

import numpy as np
import pytest

import pyoracle as po
from caseutil import BASE_PARAMS, aerofoil_case, seeded_field
from tandemproject_amd import api
from tandemproject_amd import layout as L

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, float); b = np.asarray(b, float)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def pair(layout, params, exact=False, chunks=0):
    o = po.Oracle(layout, params)
    s = api.Solver([(layout, params)], chunks=chunks)
    if exact:
        s.set_exact_lines(True)
    return o, s


# line-solver configurations of the per-call tests: the sequential sweep, and
# chunked sweeps with fixed chunk counts so that every line is split over
# several waves (the automatic choice may keep short lines in one chunk);
# over the cases below they use every chunk capacity from 8 to 32 (the
# 64-point lines 32, 24 and 12, the 240-point lines 28, 20 and 16)
CALL_MODES = {"sequential": dict(exact=True), "chunks2": dict(chunks=2), "chunks3": dict(chunks=3),
              "chunks4": dict(chunks=4), "chunks7": dict(chunks=7), "chunks9": dict(chunks=9),
              "chunks12": dict(chunks=12), "chunks16": dict(chunks=16), "auto": dict()}


LINE_CASES = {
    "one-sided both ends": L.single_block(20, 17, 14, [[10, 30], [25, 10], [20, 30]]),
    "periodic zeta": L.single_block(18, 16, 15, [[10, 10], [10, 25], [45, 45]]),
    "all periodic": L.single_block(14, 13, 12, [[45, 45], [45, 45], [45, 45]]),
    "closure 35 self": L.single_block(13, 12, 13, [[35, 35], [10, 10], [45, 45]]),
    "3x2 split zeta": L.h_layout([14, 18, 14], [12, 14], [31], npcz=[2]),
    "3x2 split xi": L.h_layout([14, 27, 14], [12, 14], [12], npcx=[1, 2, 1]),
    "long lines": L.single_block(239, 63, 12, [[10, 30], [25, 10], [45, 45]]),
}


def fill_rr(o, s, layout, seed):
    for b in range(layout.nb):
        n = layout.npoints(b)
        shape = tuple((layout.dims[b] + 1)[::-1])
        f = np.concatenate([seeded_field(shape, seed + 17 * b + c) for c in range(3)])
        o.array(b, "rr")[:] = f
        s.upload(0, b, "rr", f)
    return


@pytest.mark.parametrize("mode", list(CALL_MODES))
@pytest.mark.parametrize("name", list(LINE_CASES))
def test_deriv_per_call(name, mode):
    """chunked sweeps: <= 1e-13; sequential sweeps: bit-identical"""
    say("deriv per call:", name, mode)
    lay = LINE_CASES[name]
    exact = mode == "sequential"
    o, s = pair(lay, BASE_PARAMS, **CALL_MODES[mode])
    for nrt in (1, 0):
        fill_rr(o, s, lay, 5 + nrt)
        o.mpigo(0, nrt, 0); s.mpigo(0, nrt, 0)
        for nn, nz, m in ((3, 1 if nrt else 3, 1), (2, 1 if nrt else 2, 2), (1, 1, 3)):
            o.deriv(nn, nz, m); s.deriv(nn, nz, m)
        s.sync()
        worst = 0.0
        for b in range(lay.nb):
            worst = max(worst, rel(s.download(0, b, "rr"), o.array(b, "rr")))
            for d in (1, 2, 3):
                worst = max(worst, rel(s.download(0, b, f"drva{d}"), o.array(b, f"drva{d}")))
                if np.abs(o.array(b, f"recv{d}")).max() > 0:
                    worst = max(worst, rel(s.download(0, b, f"recv{d}"), o.array(b, f"recv{d}")))
        say("  nrt", nrt, "max rel diff", worst)
        assert worst == 0.0 if exact else worst <= 1e-13


@pytest.mark.parametrize("mode", list(CALL_MODES))
@pytest.mark.parametrize("name", list(LINE_CASES))
def test_filte_per_call(name, mode):
    say("filte per call:", name, mode)
    lay = LINE_CASES[name]
    exact = mode == "sequential"
    o, s = pair(lay, BASE_PARAMS, **CALL_MODES[mode])
    fill_rr(o, s, lay, 11)
    o.mpigo(1, 1, 0); s.mpigo(1, 1, 0)
    for nn in (1, 2, 3):
        o.filte(nn, 1); s.filte(nn, 1)
    s.sync()
    worst = max(rel(s.download(0, b, "rr"), o.array(b, "rr")) for b in range(lay.nb))
    say("  max rel diff", worst)
    assert worst == 0.0 if exact else worst <= 1e-13


def test_mpigo_periodic_jump():
    say("mpigo with the periodic jump correction")
    lay = LINE_CASES["3x2 split zeta"]
    o, s = pair(lay, BASE_PARAMS)
    fill_rr(o, s, lay, 23)
    for nt in (0, 1):
        o.mpigo(nt, 1, 1); s.mpigo(nt, 1, 1)
        s.sync()
        for b in range(lay.nb):
            for d in (1, 2, 3):
                ref = o.array(b, f"recv{d}")
                if np.abs(ref).max() == 0:
                    continue
                assert rel(s.download(0, b, f"recv{d}"), ref) <= 1e-13


def geometry_pair(lay, ss, p, exact=False):
    o, s = pair(lay, p, exact)
    for b in range(lay.nb):
        o.set_grid(b, ss[b]); s.set_grid(0, b, ss[b])
    o.metrics(); s.setup_geometry()
    return o, s


def test_metrics_and_normals():
    say("metrics, Jacobian, face normals on the 3x2 aerofoil grid")
    lay, ss, p = aerofoil_case()
    o, s = geometry_pair(lay, ss, p)
    for b in range(lay.nb):
        for name in ("xim", "etm", "zem", "yaco", "cm1", "cm2", "cm3"):
            r = rel(s.download(0, b, name), o.array(b, name))
            assert r <= 1e-13, (b, name, r)


def run_steps(lay, ss, p, nsteps, graph=False, exact=True):
    o, s = geometry_pair(lay, ss, p, exact)
    o.initialo(); s.initialo(0)
    o.step(nsteps)
    s.step(nsteps, graph=graph)
    s.sync()
    return o, s


def compare_state(o, s, lay, tol, case=0, acoustic=False):
    """Max difference of the conserved variables relative to their magnitude:
    rho and E by their own maxima, the momentum components by the maximum of
    the momentum vector (a spanwise-uniform flow has rho*w at round-off level,
    so normalising it by itself would measure noise against noise).
    acoustic: the momentum scale is at least rho_max * a_inf (a_inf = 1 in the
    reference's scaling) -- during the moving-frame ramp the momentum is
    still ~1e-6 of that, and line solves that round differently from the
    sequential sweep differ from it by ~1e-16 of rho*a_inf."""
    qs = [s.download(case, b, "qa").reshape(5, -1) for b in range(lay.nb)]
    rs = [o.array(b, "qa").reshape(5, -1) for b in range(lay.nb)]
    for q in qs:
        assert np.isfinite(q).all()
    scale = [max(np.abs(r[v]).max() for r in rs) for v in range(5)]
    mom = max(scale[1:4])
    if acoustic:
        mom = max(mom, scale[0])
    scale = [scale[0], mom, mom, mom, scale[4]]
    worst = 0.0
    for q, r in zip(qs, rs):
        for v in range(5):
            worst = max(worst, float(np.abs(q[v] - r[v]).max() / scale[v]))
    so, sg = o.scalars(), s.scalars(case)
    for k in ("n", "dt", "timo", "umf1"):
        assert abs(so[k] - sg[k]) <= tol * max(abs(so[k]), 1e-300), (k, so[k], sg[k])
    say("  max rel diff of conserved variables", worst)
    assert worst <= tol
    return worst


LINE_MODES = pytest.mark.parametrize("exact", [True, False], ids=["sequential", "chunked"])


@LINE_MODES
def test_first_steps_tight(exact):
    say("four full steps (dt > 0 from the third), viscous 3x2 aerofoil case,",
        "sequential" if exact else "chunked")
    lay, ss, p = aerofoil_case()
    o, s = run_steps(lay, ss, p, 4, exact=exact)
    assert o.scalars()["dt"] > 0
    compare_state(o, s, lay, 1e-12, acoustic=not exact)


@LINE_MODES
@pytest.mark.parametrize("nviscous", [1, 0])
def test_stage_arrays(nviscous, exact):
    """Every array a stage writes, after one filter sweep and one stage with
    dt > 0.  Sequential line solves: each component against its own maximum,
    to 1e-12.  Chunked line solves round the chunk carries differently, so
    an array that is still round-off (the right-hand side de of the
    near-uniform start holds ~1e-15) is compared against the magnitude of
    the conserved state instead (rho*a_inf scale)."""
    say("stage arrays, nviscous =", nviscous, "sequential" if exact else "chunked")
    lay, ss, p = aerofoil_case(nviscous=nviscous)
    o, s = run_steps(lay, ss, p, 2, exact=exact)
    s.set_keep_stresses(True)   # the stresses everywhere, not only at the wall rows
    o.filter_step(); s.filter_step()
    o.stage(1); s.stage(1)
    s.sync()
    names = ["qa", "qo", "p", "de"] + (["tv"] if nviscous else [])
    worst = 0.0
    for b in range(lay.nb):
        for name in names:
            g = s.download(0, b, name); r = o.array(b, name)
            k = 9 if name == "tv" else (5 if name in ("qa", "qo", "de") else 1)
            g = g.reshape(k, -1); r = r.reshape(k, -1)
            whole = max(np.abs(r).max(), np.abs(o.array(b, "qa")).max())
            for v in range(k):
                d = rel(g[v], r[v]) if exact else float(np.abs(g[v] - r[v]).max() / whole)
                worst = max(worst, d)
                assert d <= 1e-12, (b, name, v, d)
    say("  max rel diff over stage arrays", worst)
    for key in ("dt", "umf1", "dudt1", "dtk"):
        assert abs(o.scalars()[key] - s.scalars(0)[key]) <= 1e-14 * max(1.0, abs(o.scalars()[key]))


def test_stresses_kept_only_where_read():
    """By default the flux pass keeps the viscous stresses in registers and
    leaves them in tv only at the wall rows of wall blocks on the first stage
    (for the surface forces); with TP_OPT_STRESSES it writes them everywhere.
    The stage's results are bit-identical either way, and so are the wall-row
    stresses, on the first stage and across a whole step."""
    say("stress option: default against kept everywhere")
    lay, ss, p = aerofoil_case()
    sols = []
    for keep in (False, True):
        s = api.Solver([(lay, p)], exact_lines=True)
        s.set_keep_stresses(keep)
        for b in range(lay.nb):
            s.set_grid(0, b, ss[b])
        s.setup_geometry(); s.initialo(0)
        s.step(2)
        s.filter_step(); s.stage(1)
        s.sync()
        sols.append(s)
    d, k = sols
    walls = 0
    for b in range(lay.nb):
        for name in ("qa", "qo", "p", "de"):
            np.testing.assert_array_equal(d.download(0, b, name), k.download(0, b, name))
        nbc = lay.nbc[b][1]
        if 25 not in nbc:
            continue
        walls += 1
        nx, ny, nz = lay.dims[b] + 1
        j = 0 if nbc[0] == 25 else ny - 1
        td = d.download(0, b, "tv").reshape(9, nz, ny, nx)[:, :, j, :]
        tk = k.download(0, b, "tv").reshape(9, nz, ny, nx)[:, :, j, :]
        np.testing.assert_array_equal(td, tk)
        assert np.abs(tk).max() > 0
    assert walls == 2
    for s in sols:
        for nk in range(2, 5):
            s.stage(nk)
        s.sync()
    for b in range(lay.nb):
        np.testing.assert_array_equal(d.download(0, b, "qa"), k.download(0, b, "qa"))


@LINE_MODES
@pytest.mark.parametrize("variant", ["viscous", "euler", "les-gcic"])
def test_hundred_steps(variant, exact):
    say("100 steps:", variant, "sequential" if exact else "chunked")
    if variant == "viscous":
        lay, ss, p = aerofoil_case()
    elif variant == "euler":
        lay, ss, p = aerofoil_case(nviscous=0)
    else:
        # left-handed grid (z flipped) so that the Smagorinsky volume term
        # (-1/yaco)**1.5 is real, wall slots near the leading edge exchanged
        # through the aerofoil (nextrabc = nextgcic = 1)
        lay, ss, p = aerofoil_case(lzebk=(31,), npcz=[2])
        ss = [np.concatenate([x[:2 * (x.size // 3)], -x[2 * (x.size // 3):]]) for x in ss]
        p = dict(p); p.update(les=1, nextgcic=1, nextrabc=1, wlea=0.4, wlew=0.1)
    o, s = run_steps(lay, ss, p, 100, exact=exact)
    if variant == "les-gcic":
        assert max(o.array(b, "rpex").max() if o.array(b, "rpex").size else 0 for b in range(lay.nb)) == 1
    # (the Smagorinsky volume term (-1/yaco)**1.5, which the reference keeps
    # in single precision, src/MainVar3D.f90:105, Main3D.f90:664, is formed
    # once on the host with the same pow() as the oracle)
    compare_state(o, s, lay, 1e-10, acoustic=not exact)


def test_graph_replay_matches_eager():
    say("graph replay of a captured step vs eager launches")
    lay, ss, p = aerofoil_case(lxibk=(16, 20, 16), letbk=(14, 14))
    s1 = api.Solver([(lay, p)]); s2 = api.Solver([(lay, p)])
    for s in (s1, s2):
        for b in range(lay.nb):
            s.set_grid(0, b, ss[b])
        s.setup_geometry(); s.initialo(0)
    s1.step(12)
    for _ in range(3):
        s2.step(4, graph=True)
        s2.sync()
        s2.scalars(0)   # host work between replays
    for b in range(lay.nb):
        np.testing.assert_array_equal(s1.download(0, b, "qa"), s2.download(0, b, "qa"))


@pytest.mark.parametrize("mode", ["sequential", "chunks4"])
def test_ensemble_equals_alone(mode):
    """Two cases batched in one context compute exactly what each computes
    alone (same line-solver configuration)."""
    say("two cases in one context == each alone:", mode)
    kw = dict(exact_lines=True) if mode == "sequential" else dict(chunks=4)
    lay1, ss1, p1 = aerofoil_case(lxibk=(16, 20, 16), letbk=(14, 14))
    lay2, ss2, p2 = aerofoil_case(lxibk=(20, 24, 18), letbk=(16, 12), amach1=0.3, reoo=4e4)
    both = api.Solver([(lay1, p1), (lay2, p2)], **kw)
    a1 = api.Solver([(lay1, p1)], **kw); a2 = api.Solver([(lay2, p2)], **kw)
    for s, cs, lay, ss in ((both, 0, lay1, ss1), (both, 1, lay2, ss2), (a1, 0, lay1, ss1), (a2, 0, lay2, ss2)):
        for b in range(lay.nb):
            s.set_grid(cs, b, ss[b])
    for s in (both, a1, a2):
        s.setup_geometry()
    both.initialo(0); both.initialo(1); a1.initialo(0); a2.initialo(0)
    both.step(15); a1.step(15); a2.step(15)
    for b in range(lay1.nb):
        np.testing.assert_array_equal(both.download(0, b, "qa"), a1.download(0, b, "qa"))
    for b in range(lay2.nb):
        np.testing.assert_array_equal(both.download(1, b, "qa"), a2.download(0, b, "qa"))


@LINE_MODES
def test_inflow_turbulence_steps(exact):
    """Turbulent inflow through the sponge target: the device table (built
    from the library's seeded Fourier modes), the per-stage time
    interpolation and the relaxation against the oracle's restatement of
    spongego over 40 steps; and the inflow does change the flow."""
    say("inflow turbulence, 40 steps,", "sequential" if exact else "chunked")
    lay, ss, p = aerofoil_case()
    p = dict(p)
    p.update(inflow=1, nmodes=48, seed=11, ntz=16, tgusto=0.0, slit=2.0, tu=0.05, lturb=0.3, gustw=1.5)
    uoo = np.array([p["amach1"], p["amach2"], p["amach3"]]) * np.sqrt(1.4 * (1 / 1.4) / 1.0)
    modes = api.inflow_modes(p["nmodes"], p["seed"], p["tu"], p["lturb"], uoo, p["slit"], p["span"])
    o = po.Oracle(lay, p)
    o.set_inflow_modes(modes)
    s = api.Solver([(lay, p)], exact_lines=exact)
    plain = api.Solver([(lay, dict(p, inflow=0))], exact_lines=exact)
    for b in range(lay.nb):
        o.set_grid(b, ss[b]); s.set_grid(0, b, ss[b]); plain.set_grid(0, b, ss[b])
    o.metrics(); s.setup_geometry(); plain.setup_geometry()
    o.initialo(); s.initialo(0); plain.initialo(0)
    o.step(40); s.step(40); plain.step(40)
    s.sync()
    compare_state(o, s, lay, 1e-10, acoustic=not exact)
    moved = max(float(np.abs(s.download(0, b, "qa") - plain.download(0, b, "qa")).max()) for b in range(lay.nb))
    say("  max change of the state by the inflow", moved)
    assert moved > 1e-6


def test_inflow_gust_steps():
    """Sinusoidal transverse gust (TP_I_INFLOW = 2) at 5 degrees incidence:
    the library derives the single mode and the table period itself; the
    oracle is given the same mode (tp_gust_mode) and period, 40 steps."""
    say("inflow gust, 40 steps, sequential")
    lay, ss, p = aerofoil_case()
    a = np.radians(5.0)
    p = dict(p)
    p.update(amach1=0.4 * np.cos(a), amach2=0.4 * np.sin(a), inflow=2, ntz=24, tgusto=0.0, tu=0.04,
             gustk=1.2, gustw=1.5)
    uoo = np.array([p["amach1"], p["amach2"], 0.0])
    mode, slit = api.gust_mode(p["tu"], p["gustk"], uoo)
    o = po.Oracle(lay, dict(p, slit=slit, nmodes=1))
    o.set_inflow_modes(mode)
    s = api.Solver([(lay, p)], exact_lines=True)
    for b in range(lay.nb):
        o.set_grid(b, ss[b]); s.set_grid(0, b, ss[b])
    o.metrics(); s.setup_geometry()
    o.initialo(); s.initialo(0)
    o.step(40); s.step(40)
    s.sync()
    compare_state(o, s, lay, 1e-10)


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_surface_statistics(graph):
    """Wall cell areas and normals (src/rpt.f90:774-864), the force
    coefficients of clpost (:1017-1056) and the wall-pressure time averages,
    sampled every 2 steps over 20 steps (inside the captured step for graph
    replays), against the oracle's restatement."""
    say("surface statistics, 20 steps sampled every 2,", "graph" if graph else "eager")
    lay, ss, p = aerofoil_case()
    o, s = geometry_pair(lay, ss, p, exact=True)
    walls = 0
    for b in range(lay.nb):
        ro, rs = o.stats_wall(b), s.stats_wall(0, b)
        assert (ro is None) == (rs is None), b
        if ro is None:
            continue
        walls += 1
        assert rel(rs[2], ro[2]) <= 1e-14 and rel(rs[3], ro[3]) <= 1e-14
    assert walls == 2
    o.initialo(); s.initialo(0)
    o.step(3); s.step(3)
    o.stats_start(2); s.stats_start(0, 2)
    o.step(20); s.step(20, graph=graph)
    s.sync()
    fo, fs = o.stats_forces(), s.stats_forces(0)
    say("  oracle", fo)
    say("  device", fs)
    assert fo["samples"] == fs["samples"] == 10
    for k in ("t0", "t1", "window"):
        assert abs(fo[k] - fs[k]) <= 1e-14 * abs(fo[k]), k
    for k in ("cl", "cd", "cl_mean", "cd_mean"):
        assert abs(fo[k] - fs[k]) <= 1e-12 * max(abs(fo["cl"]), abs(fo["cd"])), (k, fo[k], fs[k])
    for b in range(lay.nb):
        ro = o.stats_wall(b)
        if ro is None:
            continue
        rs = s.stats_wall(0, b)
        assert rel(rs[0], ro[0]) <= 1e-13
        assert np.abs(rs[1] ** 2 - ro[1] ** 2).max() <= 1e-12 * np.abs(ro[0]).max() ** 2
        assert ro[1].max() > 0


def test_surface_statistics_long_run():
    """Mean and rms wall pressure and the mean force coefficients over a long
    gust-response run on the production path (chunked line solves, replayed
    step graph): 600 steps, sampled every 4 steps after 40 steps of start-up,
    against the oracle within 1 % (the bar for long runs; the agreement
    reached is printed and is far tighter)."""
    say("surface statistics, gust response, 600 steps sampled every 4, chunked + graph")
    lay, ss, p = aerofoil_case()
    a = np.radians(4.0)
    p = dict(p)
    p.update(amach1=0.4 * np.cos(a), amach2=0.4 * np.sin(a), inflow=2, ntz=32, tgusto=0.0, tu=0.04,
             gustk=2.0, gustw=1.5)
    uoo = np.array([p["amach1"], p["amach2"], 0.0])
    mode, slit = api.gust_mode(p["tu"], p["gustk"], uoo)
    o = po.Oracle(lay, dict(p, slit=slit, nmodes=1))
    o.set_inflow_modes(mode)
    s = api.Solver([(lay, p)])
    for b in range(lay.nb):
        o.set_grid(b, ss[b]); s.set_grid(0, b, ss[b])
    o.metrics(); s.setup_geometry()
    o.initialo(); s.initialo(0)
    o.step(40); s.step(40)
    o.stats_start(4); s.stats_start(0, 4)
    o.step(600); s.step(600, graph=True)
    s.sync()
    fo, fs = o.stats_forces(), s.stats_forces(0)
    say("  oracle", fo)
    say("  device", fs)
    assert fo["samples"] == fs["samples"] == 150
    worst = 0.0
    for k in ("cl_mean", "cd_mean"):
        e = abs(fs[k] - fo[k]) / max(abs(fo["cl_mean"]), abs(fo["cd_mean"]))
        worst = max(worst, e)
        assert e <= 0.01, (k, fo[k], fs[k])
    for b in range(lay.nb):
        ro = o.stats_wall(b)
        if ro is None:
            continue
        rs = s.stats_wall(0, b)
        em = np.abs(rs[0] - ro[0]).max() / np.abs(ro[0]).max()
        er = np.abs(rs[1] - ro[1]).max() / ro[1].max()
        worst = max(worst, em, er)
        assert em <= 0.01 and er <= 0.01, (b, em, er)
        assert ro[1].max() > 1e-4 * np.abs(ro[0]).max()
        cm, cr = s.stats_cp(0, b)
        m2 = p["amach1"] ** 2 + p["amach2"] ** 2
        np.testing.assert_allclose(cm, (ro[0] - 1 / 1.4) * 2 / m2, rtol=0, atol=0.01 * np.abs(cm).max())
        np.testing.assert_allclose(cr, ro[1] * 2 / m2, rtol=0, atol=0.01 * cr.max())
    say("  worst relative difference of the statistics", worst)


@pytest.mark.parametrize("mode", ["chunked", "sequential"])
def test_profile_classes(mode):
    """tp_profile_classes times only the classes asked for; tp_profile times
    them all; the launches and algorithmic bytes of a step are counted.  The
    line solves carry their timing events in their own dispatches, in the
    chunked launches (one or more per direction) and the sequential ones."""
    say("profiling by kernel class:", mode)
    lay, ss, p = aerofoil_case(lxibk=(16, 20, 16), letbk=(14, 14))
    s = api.Solver([(lay, p)], exact_lines=mode == "sequential")
    for b in range(lay.nb):
        s.set_grid(0, b, ss[b])
    s.setup_geometry(); s.initialo(0)
    s.step(1)
    s.profile_classes(["deriv"])
    s.step(2)
    s.sync()
    s.profile(False)
    got = {k: s.profile_read(k) for k in api.KCLASS}
    s.profile_read("deriv", reset=True)
    assert got["deriv"][1] == 2 * 4 * 6 and got["deriv"][0] > 0 and got["deriv"][2] > 0
    assert all(got[k][1] == 0 for k in api.KCLASS if k != "deriv")
    s.profile(True)
    s.step(1)
    s.sync()
    s.profile(False)
    got = {k: s.profile_read(k) for k in api.KCLASS}
    for k in ("deriv", "filter", "point", "flux", "gcbc", "rk"):
        assert got[k][1] > 0, k
    assert got["deriv"][1] == 4 * 6


def eddy_inputs(p, nits=24):
    """The reference's synthetic eddies from its own random numbers
    (tests/golden/randnum_head.json) on this case's domain, with non-zero
    amplitudes (the shipped deck's are zero), and the band half-width the
    reference derives, tlb (2.5 - cutlb) (src/SingleTurbulence3D.f90:255)."""
    import json
    import os
    ref = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "randnum_head.json")))
    d = dict(ref["deck"])
    d.update(span=p["span"], slit=2.0, domlen=p["domlen"], szth1=p["szth1"],
             tam=[[0.05, 0.05, 0.05], [0.03, 0.03, 0.03]])
    eddies = api.eddy_params(nits, np.array(ref["randnum"]), d)
    tlb = max(max(d["tl0"]), max(d["tl1"]))
    return eddies, dict(inflow=3, ntz=16, tgusto=0.0, slit=2.0, cfit=d["cfit"], gustw=tlb * (2.5 - d["cutlb"]))


@LINE_MODES
def test_inflow_eddies_steps(exact):
    """Inflow from the reference's synthetic eddies (TP_I_INFLOW = 3): the
    device table, interpolation and sponge target against the oracle over 40
    steps, and the eddies do change the flow."""
    say("inflow eddies, 40 steps,", "sequential" if exact else "chunked")
    lay, ss, p = aerofoil_case()
    eddies, over = eddy_inputs(p)
    p = dict(p); p.update(over)
    o = po.Oracle(lay, p)
    o.set_inflow_eddies(eddies)
    s = api.Solver([(lay, p)], exact_lines=exact)
    s.set_inflow_eddies(0, eddies)
    plain = api.Solver([(lay, dict(p, inflow=0))], exact_lines=exact)
    for b in range(lay.nb):
        o.set_grid(b, ss[b]); s.set_grid(0, b, ss[b]); plain.set_grid(0, b, ss[b])
    o.metrics(); s.setup_geometry(); plain.setup_geometry()
    assert sum(o.array(b, "atz").size for b in range(lay.nb)) > 0
    o.initialo(); s.initialo(0); plain.initialo(0)
    o.step(40); s.step(40); plain.step(40)
    s.sync()
    compare_state(o, s, lay, 1e-10, acoustic=not exact)
    moved = max(float(np.abs(s.download(0, b, "qa") - plain.download(0, b, "qa")).max()) for b in range(lay.nb))
    say("  max change of the state by the eddies", moved)
    assert moved > 1e-7


@LINE_MODES
def test_forcing_steps(exact):
    """The reference's body force (forceup/forcego, src/rpt.f90:728-769)
    switched on behind the aerofoil: 40 steps against the oracle."""
    say("forcing, 40 steps,", "sequential" if exact else "chunked")
    lay, ss, p = aerofoil_case()
    p = dict(p)
    p.update(forcing=1, amfor=0.05, xfor=1.6, yfor=0.1, rfor=0.5, tsfor=0.0, tefor=100.0)
    o, s = run_steps(lay, ss, p, 40, exact=exact)
    compare_state(o, s, lay, 1e-10, acoustic=not exact)
    ref = api.Solver([(lay, dict(p, forcing=0))], exact_lines=exact)
    for b in range(lay.nb):
        ref.set_grid(0, b, ss[b])
    ref.setup_geometry(); ref.initialo(0); ref.step(40); ref.sync()
    moved = max(float(np.abs(s.download(0, b, "qa") - ref.download(0, b, "qa")).max()) for b in range(lay.nb))
    say("  max change of the state by the force", moved)
    assert moved > 1e-7


@LINE_MODES
def test_pitching_frame_steps(exact):
    """Pitching aerofoil (movef/pitch, src/Subroutineso.f90:222-259): the
    frame velocity rotates from aoa0 to aoa1 over tp from tps; 100 steps
    across the whole pitch window against the oracle."""
    say("pitching frame, 100 steps,", "sequential" if exact else "chunked")
    lay, ss, p = aerofoil_case()
    p = dict(p)
    p.update(tps=0.01, tp=0.04, aoa0=0.0, aoa1=6.0)
    o, s = run_steps(lay, ss, p, 100, exact=exact)
    so, sg = o.scalars(), s.scalars(0)
    say("  time", so["timo"], "angle of attack", so["aoa"], sg["aoa"])
    assert so["timo"] > p["tps"] + p["tp"]          # the whole window was crossed
    assert abs(sg["aoa"] - 6.0) < 1e-12 and abs(so["aoa"] - sg["aoa"]) < 1e-12
    assert abs(sg["umf2"]) > 0.01                  # the frame turned
    compare_state(o, s, lay, 1e-10, acoustic=not exact)

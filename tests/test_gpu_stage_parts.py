"""The sections of a stage as drop-ins (tp_primitives ... tp_interface_avg),
through the C ABI.

* called in tp_stage's order they compute tp_stage's results bit for bit,
  over whole time steps, with sequential and chunked line solves;
* each section, started from the oracle's state, agrees with the matching
  oracle section (src/Main3D.f90:567-958, src/SingleTurbulence3D.f90:
  332-443; orc_stage_part) to 1e-13 of each output array's magnitude.
"""
import sys

import numpy as np
import pytest

import pyoracle as po
from caseutil import aerofoil_case
from tandemproject_amd import api

pytestmark = pytest.mark.gpu

PARTS = [name for name, _ in api.STAGE_PARTS]
ORACLE_PART = dict(zip(PARTS, po.PARTS))
SCALARS = ["n", "ndt", "dt", "dts", "dte", "timo", "umf1", "umf2", "umf3", "dudt1", "dudt2", "dudt3",
           "dtk", "dtko", "aoa"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gust_case():
    lay, ss, p = aerofoil_case()
    a = np.radians(4.0)
    p = dict(p)
    p.update(amach1=0.4 * np.cos(a), amach2=0.4 * np.sin(a), inflow=2, ntz=24, tgusto=0.0, tu=0.04,
             gustk=1.2, gustw=1.5)
    mode, slit = api.gust_mode(p["tu"], p["gustk"], np.array([p["amach1"], p["amach2"], 0.0]))
    return lay, ss, p, mode, slit


def solver(lay, ss, p, **kw):
    s = api.Solver([(lay, p)], **kw)
    for b in range(lay.nb):
        s.set_grid(0, b, ss[b])
    s.setup_geometry()
    s.initialo(0)
    return s


def run_parts(s, nk):
    for part in PARTS:
        s.stage_part(nk, part)


@pytest.mark.parametrize("mode", ["sequential", "chunked"])
def test_parts_compose_to_the_stage(mode):
    say("stage sections in order == tp_stage, 3 steps,", mode)
    lay, ss, p, mode_, slit = gust_case()
    kw = dict(exact_lines=True) if mode == "sequential" else dict(chunks=3)
    whole = solver(lay, ss, p, **kw)
    parts = solver(lay, ss, p, **kw)
    whole.step(3)
    for _ in range(3):
        parts.filter_step()
        for nk in range(1, 5):
            run_parts(parts, nk)
        sc = parts.scalars(0)
        sc["n"] += 1
        sc["timo"] += sc["dt"]
        parts.set_scalars(0, sc)
    whole.sync(); parts.sync()
    for b in range(lay.nb):
        for name in ("qa", "qo", "de", "p"):
            np.testing.assert_array_equal(whole.download(0, b, name), parts.download(0, b, name), err_msg=name)
    for k in ("n", "dt", "timo", "umf1", "dtk"):
        assert whole.scalars(0)[k] == parts.scalars(0)[k], k


# arrays each section writes (compared) -- the rest of its inputs are uploaded
# from the oracle before the section runs
OUTPUTS = {
    "primitives": ["de", "p", "ss"],
    "viscous": ["tv"],
    "flux_divergence": ["de", "drva1", "drva2", "drva3"],
    "sponge": ["de"],
    "gcbc": ["de"],
    "rk_update": ["qa"],
    "wall": ["qa"],
    "junction": ["qa"],
    "interface_avg": ["qa"],
}
INPUTS = ["qa", "qo", "de", "p", "tv", "drva1", "drva2", "drva3"]


def upload_state(o, s, lay, scalars=True):
    for b in range(lay.nb):
        for name in INPUTS:
            s.upload(0, b, name, o.array(b, name))
        # ss(:,1) is the viscosity during a stage
        s.upload(0, b, "ss", o.array(b, "ss"))
    if scalars:
        sc = s.scalars(0)
        so = o.scalars()
        for k in SCALARS:
            sc[k] = so[k]
        s.set_scalars(0, sc)


def les_gcic_case():
    """Smagorinsky LES with wall slots exchanged through the aerofoil
    (nextrabc = nextgcic = 1) on a left-handed grid (z flipped, so that the
    volume term (-1/yaco)**1.5 is real), spanwise split in two."""
    lay, ss, p = aerofoil_case(lzebk=(31,), npcz=[2])
    ss = [np.concatenate([x[:2 * (x.size // 3)], -x[2 * (x.size // 3):]]) for x in ss]
    p = dict(p); p.update(les=1, nextgcic=1, nextrabc=1, wlea=0.4, wlew=0.1)
    return lay, ss, p


@pytest.mark.parametrize("case", ["gust", "les-gcic"])
def test_parts_against_oracle_sections(case):
    """Lock-step over stages 1 and 2 of a step, with the sequential line
    solves: before each section the device gets the oracle's arrays, then
    both run the section and its outputs are compared.  (The chunked solves
    round the chunk carries differently; their per-call parity is
    test_deriv_per_call's, and a flux divergence is a sum of nearly
    cancelling derivatives, so it is not compared at this bar.)"""
    say("stage sections against the oracle's sections,", case)
    if case == "gust":
        lay, ss, p, gmode, slit = gust_case()
        o = po.Oracle(lay, dict(p, slit=slit, nmodes=1))
        o.set_inflow_modes(gmode)
    else:
        lay, ss, p = les_gcic_case()
        o = po.Oracle(lay, p)
    s = api.Solver([(lay, p)], exact_lines=True)
    for b in range(lay.nb):
        o.set_grid(b, ss[b]); s.set_grid(0, b, ss[b])
    o.metrics(); s.setup_geometry()
    for b in range(lay.nb):
        for name in ("xim", "etm", "zem", "yaco"):
            d = float(np.abs(s.download(0, b, name) - o.array(b, name)).max())
            if d:
                say(f"  metrics differ: block {b} {name} {d:.2e}")
    o.initialo(); s.initialo(0)
    o.step(3)
    upload_state(o, s, lay)
    o.filter_step(); s.filter_step()
    s.sync()
    worst = {}
    for nk in (1, 2):
        for part in PARTS:
            upload_state(o, s, lay)
            o.stage_part(nk, ORACLE_PART[part])
            s.stage_part(nk, part)
            s.sync()
            for name in OUTPUTS[part]:
                for b in range(lay.nb):
                    r = o.array(b, name)
                    g = s.download(0, b, name)
                    scale = max(float(np.abs(r).max()), 1e-300)
                    d = float(np.abs(g - r).max()) / scale
                    worst[(nk, part, name)] = max(worst.get((nk, part, name), 0.0), d)
            if part == "primitives":
                # the stage's scalars (frame velocity and acceleration, time
                # step) use correctly rounded sin, cos on both sides
                so, sg = o.scalars(), s.scalars(0)
                for k in ("dt", "dtk", "dtko", "umf1", "umf2", "umf3", "dudt1", "dudt2", "dudt3"):
                    assert so[k] == sg[k], (nk, k, so[k], sg[k])
    for k, v in sorted(worst.items()):
        say(f"  stage {k[0]} {k[1]:16s} {k[2]:6s} {v:.2e}")
    bad = {k: v for k, v in worst.items() if v > 1e-13}
    assert not bad, bad

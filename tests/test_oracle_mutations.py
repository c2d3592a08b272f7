"""The oracle's pins must fail when the oracle is wrong.

Each case plants one plausible mistake (a sign, a constant, a swapped index,
a wrong weight) in a copy of oracle/tp_oracle.c, builds it into a temporary
directory and runs the pins of that part (tests/test_oracle_stage_pins.py,
tests/test_rounding_cpu.py) against that copy in a child process: at least one
pin must fail.  CPU only.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "tp_oracle.c")

# (what the mistake is, text of the oracle, the mistaken text)
MUTATIONS = {
    "energy flux frame term sign": ("r1 = d5 * s1 - pl * (umf[0] * x1", "r1 = d5 * s1 + pl * (umf[0] * x1"),
    "dilatation factor": ("d5 = two_three * (txx + tyy + tzz);\n          double ntxx = r1 * (2 * txx - d5), ntyy",
                          "d5 = 0.5 * (txx + tyy + tzz);\n          double ntxx = r1 * (2 * txx - d5), ntyy"),
    "Prandtl number": ("double r2 = gamm1prndtli * r1, d5;", "double r2 = 1.0 / (gamm1 * 0.72) * r1, d5;"),
    "stress transposed": ("double ntxy = r1 * (txy + hzz), ntyz = r1 * (tyz + hxx), ntzx = r1 * (tzx + hyy);\n          txx = ntxx; tyy = ntyy; tzz = ntzz; txy = ntxy; tyz = ntyz; tzx = ntzx;\n        }",
                          "double ntxy = r1 * (txy + txy), ntyz = r1 * (tyz + hxx), ntzx = r1 * (tzx + hyy);\n          txx = ntxx; tyy = ntyy; tzz = ntzz; txy = ntxy; tyz = ntyz; tzx = ntzx;\n        }"),
    "viscous time-step limit": ("if (c->nviscous == 1) ra1 = half / fv;", "if (c->nviscous == 1) ra1 = 0.4 / fv;"),
    "ramp period": ("double s = crsin(0.05 * PI * (double)(n - (int)sc[ORC_S_NDT]));",
                    "double s = crsin(0.055 * PI * (double)(n - (int)sc[ORC_S_NDT]));"),
    "incoming wave index": ("if (ra0 * (vv + e.ao) > zero) cha[3] = zero;\n              if (ra0 * (vv - e.ao) > zero) cha[4] = zero;",
                            "if (ra0 * (vv + e.ao) > zero) cha[4] = zero;\n              if (ra0 * (vv - e.ao) > zero) cha[3] = zero;"),
    "wall reflection sign": ("cha[3 + ipp] = cha[4 - ipp] + two * ra0", "cha[3 + ipp] = -cha[4 - ipp] + two * ra0"),
    "junction velocity members": ("for (int m = 1; m <= 3; m++) r[m] = half * (v[va][m] + v[vb][m]);",
                                  "for (int m = 1; m <= 3; m++) r[m] = 0.25 * (v[0][m] + v[1][m] + v[2][m] + v[3][m]);"),
    "isentropic exponent": ("vit[2] * vit[2]), hamm1);", "vit[2] * vit[2]), ham);"),
    "RK stage fraction": ("double dtk = dt / (nkrk - nk + 1);", "double dtk = dt / (nkrk - nk + 2);"),
    "wall temperature factor": ("double ra0 = ham * hamm1 * c->wtemp;", "double ra0 = hamm1 * c->wtemp;"),
    "averaging weight": ("double r2 = one / rr[l];\n              rr[N + l] = r2;\n              for (int v = 0; v < 5; v++)\n                qa[(size_t)v * N + l] = r2 * ((rr[l] - one)",
                         "double r2 = 0.4;\n              rr[N + l] = r2;\n              for (int v = 0; v < 5; v++)\n                qa[(size_t)v * N + l] = r2 * ((rr[l] - one)"),
    "GCBC face weight": ("cnt[l] = cnt[l] + one; wgt[ll] = res; lst[ll] = l;", "cnt[l] = cnt[l] + one; wgt[ll] = 0.5 * res; lst[ll] = l;"),
    "sponge outflow zone without the extension": ("szr[1][0] = 1 / (c->szth1 + c->szxt);", "szr[1][0] = 1 / c->szth1;"),
    "sponge edge literal taken as double": ("double p07 = (double)0.7f;", "double p07 = 0.7;"),
    "sponge upper eta edge": ("szp[1][1] = c->domlen + c->szxt - c->szth2;", "szp[1][1] = c->domlen - c->szth2;"),
    "sponge profile factor": ("(one - 0.125 * (one + cos(PI * r[0]))", "(one - 0.25 * (one + cos(PI * r[0]))"),
    "sponge streamwise weight period": ("double ra1 = 0.25, ra2 = PI / (two * c->domlen);", "double ra1 = 0.25, ra2 = PI / c->domlen;"),
    "sponge energy weight": ("B->bsz[q] = res * hamm1 * sg[l];", "B->bsz[q] = res * ham * sg[l];"),
    "wall slots: one pitch too many": ("int ie = (int)((c->span + sml) / c->wlew) - 1;", "int ie = (int)((c->span + sml) / c->wlew);"),
    "wall slots: quarter-pitch offset dropped": ("double ra1 = -half * c->span + 0.25 * c->wlew - half * slitw, ra2",
                                                 "double ra1 = -half * c->span - half * slitw, ra2"),
    "wall slots: chordwise extent": ("B->xbc_ra0 = -half + c->wlea + slitl;", "B->xbc_ra0 = -half + c->wlea;"),
    "wall slots: width": ("double slitl = c->wlea, slitw = 0.4 * slitl;", "double slitl = c->wlea, slitw = 0.2 * slitl;"),
    "wall slots: switch ignored": ("if (res < zero && c->nextrabc == 1) flag = one;", "if (res < zero) flag = one;"),
    # the order of the sections in a whole stage (orc_stage only; the
    # sections themselves are unchanged)
    "stage: stresses after the fluxes": ("ORC_P_BEGIN, ORC_P_VISCOUS, ORC_P_FLUX,", "ORC_P_BEGIN, ORC_P_FLUX, ORC_P_VISCOUS,"),
    "stage: characteristic faces before the fluxes": ("ORC_P_FLUX, ORC_P_SPONGE, ORC_P_GCBC,\n    ORC_P_RK,",
                                                      "ORC_P_GCBC, ORC_P_FLUX, ORC_P_SPONGE,\n    ORC_P_RK,"),
    "stage: update before the characteristic faces": ("ORC_P_GCBC,\n    ORC_P_RK, ORC_P_WALL,", "ORC_P_RK,\n    ORC_P_GCBC, ORC_P_WALL,"),
    "stage: wall before the update": ("ORC_P_RK, ORC_P_WALL, ORC_P_JUNCTION, ORC_P_AVG};", "ORC_P_WALL, ORC_P_RK, ORC_P_JUNCTION, ORC_P_AVG};"),
    "stage: averaging before the update": ("ORC_P_RK, ORC_P_WALL, ORC_P_JUNCTION, ORC_P_AVG};", "ORC_P_AVG, ORC_P_RK, ORC_P_WALL, ORC_P_JUNCTION};"),
    "stage: sponge after the characteristic faces": ("ORC_P_SPONGE, ORC_P_GCBC,\n    ORC_P_RK,", "ORC_P_GCBC, ORC_P_SPONGE,\n    ORC_P_RK,"),
    "interface: the entering acoustic wave not taken": (
        "              if (ra0 * (vv - e.ao) > zero) cha[4] = dha[4];\n            }\n            xtr2q(&e, cmv, xt);",
        "            }\n            xtr2q(&e, cmv, xt);"),
    "interface: the entering entropy and vorticity waves not taken": (
        "            } else {\n              if (ra0 * vv > zero) { cha[0] = dha[0]; cha[1] = dha[1]; cha[2] = dha[2]; }",
        "            } else {\n              if (ra0 * vv > zero) { cha[0] = dha[0]; }"),
    # the Smagorinsky subgrid stress: pinned by tests/test_oracle_stage_pins.py::test_les_subgrid_stress_form
    "LES: a strain-rate term dropped": ("(hyy + tzx) * (hyy + tzx) + (hxx + tyz) * (hxx + tyz);", "(hyy + tzx) * (hyy + tzx);"),
    "LES: eddy viscosity linear in smago1": ("double r3 = qa[l] * (c->smago1 * c->smago1) * varr * sqrt(2 * d1);",
                                             "double r3 = qa[l] * c->smago1 * varr * sqrt(2 * d1);"),
    "LES: eddy viscosity without the density": ("double r3 = qa[l] * (c->smago1 * c->smago1) * varr * sqrt(2 * d1);",
                                                "double r3 = (c->smago1 * c->smago1) * varr * sqrt(2 * d1);"),
    "LES: strain-rate norm not rooted": ("double r3 = qa[l] * (c->smago1 * c->smago1) * varr * sqrt(2 * d1);",
                                         "double r3 = qa[l] * (c->smago1 * c->smago1) * varr * (2 * d1);"),
    "LES: isotropic part added": ("double ntxx = r1 * (2 * txx - d5) - r3, ntyy = r1 * (2 * tyy - d5) - r3,",
                                  "double ntxx = r1 * (2 * txx - d5) + r3, ntyy = r1 * (2 * tyy - d5) + r3,"),
    "wall slots: the entering acoustic waves not taken": (
        "/* extrabcs, src/SingleTurbulence3D.f90:468-474 */\n                if (ra0 * vv > zero) { cha[0] = dha[0]; cha[1] = dha[1]; cha[2] = dha[2]; }\n                if (ra0 * (vv + e.ao) > zero) cha[3] = dha[3];\n                if (ra0 * (vv - e.ao) > zero) cha[4] = dha[4];",
        "/* extrabcs, src/SingleTurbulence3D.f90:468-474 */\n                if (ra0 * vv > zero) { cha[0] = dha[0]; cha[1] = dha[1]; cha[2] = dha[2]; }"),
    "wall slots: reflected like the wall": ("if (B->rpex[lp] == zero) {", "if (B->rpex[lp] == zero || 1) {"),
    "stage: averaging before the junction": ("ORC_P_RK, ORC_P_WALL, ORC_P_JUNCTION, ORC_P_AVG};", "ORC_P_RK, ORC_P_WALL, ORC_P_AVG, ORC_P_JUNCTION};"),
    "stage: junction before the wall": ("ORC_P_RK, ORC_P_WALL, ORC_P_JUNCTION, ORC_P_AVG};", "ORC_P_RK, ORC_P_JUNCTION, ORC_P_WALL, ORC_P_AVG};"),
}
# the pins a mutation is run against (default: the stage pins)
PINS = {name: "test_rounding_cpu.py" for name in MUTATIONS if name.startswith(("x**1.5", "sin/cos"))}


@pytest.mark.parametrize("name", list(MUTATIONS))
def test_planted_mistake_fails_a_pin(name, tmp_path):
    good, bad = MUTATIONS[name]
    src = open(SRC).read()
    assert src.count(good) >= 1, "the oracle text this case edits has changed"
    mut = tmp_path / "tp_oracle.c"
    mut.write_text(src.replace(good, bad, 1))
    lib = tmp_path / "liboracle.so"
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c11",
                           "-I" + os.path.join(ROOT, "oracle"), "-shared", "-o", str(lib), str(mut), "-lm"])
    env = dict(os.environ, ORACLE_LIB=str(lib), OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", PINS.get(name, "test_oracle_stage_pins.py"))],
                       capture_output=True, text=True, env=env, cwd=ROOT, timeout=300)
    tail = r.stdout.strip().splitlines()[-1] if r.stdout.strip() else r.stderr[-500:]
    assert r.returncode != 0 and "failed" in tail, f"no pin caught '{name}': {tail}"

"""Writes tests/golden/randnum_head.json: the first 1024 numbers of the
reference's input file randnum.dat (one value per line; the synthetic-eddy
inflow reads them in order, src/SingleTurbulence3D.f90:88-93) and its
eddy-related deck values from inputp.dat (slit-gaus-cfit, the two eddy types'
tl0/tl1/tlw, cutlb, span, domlen, szth1).  Reads data files only; run where
the reference tree is mounted."""
import json
import os

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

vals = [float(t) for t in open(os.path.join(REF, "randnum.dat")).read().split()][:1024]
deck = {}
for line in open(os.path.join(REF, "inputp.dat")):
    f = line.split()
    if not f:
        break
    deck[f[0]] = [float(x.replace("d", "e")) for x in f[1:4] if x[0] in "0123456789.-"]
out = {"source": "randnum.dat (first 1024 values), inputp.dat",
       "randnum": vals,
       "deck": {"span": deck["span-wlew-wlea"][0], "slit": deck["slit-gaus-cfit"][0],
                "gaus": deck["slit-gaus-cfit"][1], "cfit": deck["slit-gaus-cfit"][2],
                "tl0": [deck["tlg0-g1-gw"][0], deck["tlm0-m1-mw"][0]],
                "tl1": [deck["tlg0-g1-gw"][1], deck["tlm0-m1-mw"][1]],
                "tlw": [deck["tlg0-g1-gw"][2], deck["tlm0-m1-mw"][2]],
                "cutlb": deck["cutlb"][0], "domlen": deck["smgrid-domlen"][1],
                "szth1": deck["szth1-th2-xt-co"][0]}}
json.dump(out, open(os.path.join(HERE, "randnum_head.json"), "w"), indent=0)
print(len(vals), out["deck"])

"""Throughput of the FP64 compact-scheme time step on one MI355X.

Workload (default): the NACA0012 turbulence-interaction shape at real size on
one GPU -- the reference's 3x2 H-type block layout split into 2 spanwise
sub-blocks (12 sub-blocks), 523 x 402 in-plane points, nz = 96 (20.2 M points),
M = 0.3, Re_c = 4e5, viscous laminar, FP64.  One step = one filter sweep plus
four RK stages over every point (src/Main3D.f90:552-965).

Prints one JSON line: Mpoint-steps/s over the timed steps (state resident in
HBM), the roofline of the dominant kernel (compact derivative line solves,
HIP events around every launch in the timed region), and the CPU oracle timed
on this host's cores on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from tandemproject_amd import api  # noqa: E402
from tandemproject_amd import grid as G  # noqa: E402
from tandemproject_amd import layout as L  # noqa: E402

METRIC = "FP64 Mpoint-RK-steps/s + %HBM peak; ensemble cases/h and p99 latency"
HBM_PEAK_GBS = 8000.0


def workload(scale=1.0):
    lx = [int(round(140 * scale)), int(round(240 * scale)), int(round(140 * scale))]
    ly = [int(round(200 * scale)), int(round(200 * scale))]
    lz = [95 if scale >= 1.0 else 31]
    lay = L.h_layout(lx, ly, lz, npcz=[2], nviscous=1)
    blocks = G.h_grid_blocks(lx, ly, lz, xu=7.0, xd=10.0, yb=7.0, yt=7.0, span=0.2,
                             thick=0.12, wall_spacing=4e-3)
    ss = G.subblock_grids(lay, blocks)
    # inflow: synthetic isotropic turbulence from 4096 seeded random Fourier
    # modes (von Karman spectrum, Tu = 2.5 %, integral length 0.1 chord),
    # frozen in a periodic box of 4 chords convected with u_inf, imposed
    # through the inflow sponge in the band |y| < 2
    params = dict(reoo=4e5 / 0.3, tempoo=273.0, amach1=0.3, amach2=0.0, amach3=0.0, wtemp=1.0,
                  cfl=1.0, tmax=1e9, timf=15.0, fltk=0.85, fltkbc=0.85, dto=-1.0, domlen=10.0,
                  szth1=2.0, szth2=2.0, szxt=4.0, szco=12.0, span=0.2, wlew=0.11, wlea=0.015,
                  nkrk=4, nviscous=1, nsmf=0, les=0, nextrabc=0, nextgcic=0,
                  inflow=1, nmodes=4096, seed=2, ntz=64, tgusto=0.0, slit=4.0, tu=0.025, lturb=0.1,
                  gustw=2.0)
    return lay, ss, params, dict(lxibk=lx, letbk=ly, lzebk=lz)


def cpu_baseline(seconds_target=15.0):
    """The oracle (C restatement, OpenMP) on a bounded sample of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    threads = min(os.cpu_count() or 1, 16)
    po.set_threads(threads)
    lay, ss, params, dims = workload(scale=0.2)
    o = po.Oracle(lay, params)
    if params.get("inflow"):
        # free-stream velocity in units of aoo = sqrt(gam poo / rhooo) = 1
        uoo = np.array([params["amach1"], params["amach2"], params["amach3"]])
        o.set_inflow_modes(api.inflow_modes(params["nmodes"], params["seed"], params["tu"], params["lturb"],
                                            uoo, params["slit"], params["span"]))
    for b in range(lay.nb):
        o.set_grid(b, ss[b])
    o.metrics()
    o.initialo()
    o.step(2)
    t0 = time.time(); steps = 0
    while time.time() - t0 < seconds_target and steps < 200:
        o.step(2); steps += 2
    dt = time.time() - t0
    pts = lay.npoints()
    return {"value": pts * steps / dt / 1e6, "unit": "Mpoint-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle/tp_oracle.c on {lay.nb} sub-blocks, {pts / 1e6:.3f} M points "
                      f"(lxibk={dims['lxibk']}, letbk={dims['letbk']}, nz={dims['lzebk'][0] + 1}), "
                      f"{steps} steps, {dt:.1f} s"}


def measured_traffic():
    """HBM bytes per derivative launch from the rocprofv3 FETCH_SIZE/WRITE_SIZE
    passes summarised under profiles/ (counter values scaled by a copy of
    known size); None when no such summary is present."""
    path = os.path.join(ROOT, "profiles", "line_solve_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


def copy_probe(mib):
    """One device-to-device copy of mib MiB through the HIP runtime (a blit
    kernel): a reference of known traffic for scaling hardware byte counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a); hip.hipFree(b)
    if rc != 0:
        raise RuntimeError("hipMemcpy failed")
    print(f"[bench] copy probe: {n} bytes read, {n} bytes written", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--graph", action="store_true", help="time HIP-graph replays (no per-kernel events)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default line-solver launch shapes")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="after the run, one plain device-to-device copy of MB MiB (a reference kernel "
                         "of known traffic for scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding by block is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    lay, ss, params, dims = workload(args.scale)
    print(f"[bench] building case: {lay.nb} sub-blocks, {lay.npoints() / 1e6:.2f} M points", file=sys.stderr, flush=True)
    s = api.Solver([(lay, params)])
    for b in range(lay.nb):
        s.set_grid(0, b, ss[b])
    s.setup_geometry()
    tuned = None
    if not args.no_tune:
        tuned = s.autotune(10)
        print("[bench] line solves tuned (chunks per line, us per launch): "
              + ", ".join(f"{'xez'[d]}-{k} {c} {t:.0f}" for (d, k), (c, t) in tuned.items()),
              file=sys.stderr, flush=True)
    s.initialo(0)
    print("[bench] geometry ready; warmup", file=sys.stderr, flush=True)
    s.step(args.warmup)
    s.sync()
    npts = lay.npoints()

    s.profile(not args.graph)
    s.profile_read("deriv", reset=True)
    s.sync()
    t0 = time.perf_counter()
    s.step(args.steps, graph=args.graph)
    s.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    s.profile(False)
    prof = {k: s.profile_read(k) for k in api.KCLASS}
    sc = s.scalars(0)
    q = s.download(0, 0, "qa")
    finite = bool(np.isfinite(q).all())
    print(f"[bench] {args.steps} steps in {elapsed:.3f} s; n={sc['n']:.0f} dt={sc['dt']:.3e} finite={finite}",
          file=sys.stderr, flush=True)

    value = npts * args.steps / elapsed / 1e6
    roof = None
    ms, nl, by = prof["deriv"]
    if nl > 0 and ms > 0:
        avg_s = ms / 1e3 / nl
        # 16 B/pt per launch (field in, derivative out), 24 B/pt for the
        # launches that accumulate into their output (the eta and zeta flux
        # derivatives): the library counts these per launch
        achieved = by / (ms / 1e3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                "kernel": "compact derivative line solves (k_col / k_deriv, 3 directions)",
                "avg_launch_us": round(avg_s * 1e6, 2), "launches": nl,
                "bytes_per_launch": round(by / nl)}
    breakdown = {k: round(v[0] / args.steps, 3) for k, v in prof.items() if v[1]}
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_seconds)
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpoint-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: uniform-flow start with the moving-frame ramp on an algebraic 3x2 H-grid, "
                "seeded Fourier-mode turbulence inflow (4096 modes, von Karman, Tu 2.5 %, L 0.1c)",
        "config": {"workload": "naca0012-turbulence-interaction shape: 3x2 H blocks x 2 spanwise "
                               f"sub-blocks, lxibk={dims['lxibk']}, letbk={dims['letbk']}, nz={dims['lzebk'][0] + 1}, "
                               f"{npts / 1e6:.2f} M points, M=0.3, Re_c=4e5, viscous, 4-stage RK, "
                               "turbulent inflow Tu=2.5%, L=0.1c, 4096 modes",
                   "points": npts, "sub_blocks": lay.nb, "graph": bool(args.graph),
                   "line_chunks": {f"{'xez'[d]}-{k}": c for (d, k), c in s.col_config().items()},
                   "ms_per_step_by_class": breakdown},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

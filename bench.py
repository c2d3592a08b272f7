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
from tandemproject_amd import workloads as W  # noqa: E402

METRIC = "FP64 Mpoint-RK-steps/s + %HBM peak; ensemble cases/h and p99 latency"
HBM_PEAK_GBS = 8000.0


def workload(scale=1.0):
    return W.naca0012_20m(scale)


def run_plain(lay, ss, p, nsteps):
    """One case alone on the plain path: its own context, set-up, nsteps
    eager steps, final qa to page-locked host memory.  Returns (seconds, qa
    per block)."""
    t0 = time.perf_counter()
    s = api.Solver([(lay, p)])
    for b in range(lay.nb):
        s.set_grid(0, b, ss[b])
    s.setup_geometry()
    s.initialo(0)
    s.step(nsteps)
    # the final state into page-locked host memory, as the request loop keeps it
    counts = [s.count(0, b, "qa") for b in range(lay.nb)]
    buf = api.PinnedBuffer(sum(counts))
    offs = np.concatenate([[0], np.cumsum(counts)])
    views = [s.download(0, b, "qa", out=buf.array[offs[b]:offs[b + 1]]) for b in range(lay.nb)]
    dt = time.perf_counter() - t0
    qa = [v.copy() for v in views]
    buf.free()
    s.close()
    return dt, qa


def state_diff(qs, rs):
    """Max difference of the conserved variables, each relative to the
    largest of rho, |rho u| and E over the case."""
    qs = [q.reshape(5, -1) for q in qs]
    rs = [r.reshape(5, -1) for r in rs]
    scale = [max(np.abs(r[v]).max() for r in rs) for v in range(5)]
    mom = max(max(scale[1:4]), scale[0])
    scale = [scale[0], mom, mom, mom, scale[4]]
    return max(float(np.abs(q[v] - r[v]).max() / scale[v]) for q, r in zip(qs, rs) for v in range(5))


def run_logged(srv, max_steps, label, every_s=20.0):
    """The request loop in pieces of at most max_steps steps, with a progress
    line every every_s seconds (a long ensemble run prints nothing else)."""
    t0 = last = time.perf_counter()
    while True:
        st = srv.iterate(max_steps)
        if st["active"] == 0 and st["queued"] == 0 and st["waiting"] == 0:
            return
        if st["active"] == 0:
            time.sleep(0.001)
        now = time.perf_counter()
        if now - last > every_s:
            log(f"{label}: {st['done']} done, {st['active']} on the device, {now - t0:.0f} s")
            last = now


def ensemble_bench(ncase, nsteps, nplain):
    """cases/h of the batched ensemble (every case in one request loop,
    advanced in shared launches) against the plain path (one case at a time,
    measured on a seeded sample of the same cases).  Both are timed from the
    cases' grids on the host to their final states on the host; grid
    generation is outside."""
    cases = W.ensemble_cases(ncase)
    pts = [lay.npoints() for lay, _, _ in cases]
    log(f"ensemble: {ncase} cases, {sum(pts) / 1e6:.1f} M points, {nsteps} steps each")
    srv = api.Server(budget_bytes=0, max_active=ncase, graph=True)
    t0 = time.perf_counter()
    reqs = [srv.submit(lay, p, ss, nsteps, 0.0) for lay, ss, p in cases]
    run_logged(srv, min(nsteps, 50), "ensemble")
    wall_b = time.perf_counter() - t0
    st = srv.stats()
    step_s = st["stepping_s"] / nsteps
    setup_b = wall_b - st["stepping_s"]
    sample = np.random.default_rng(41).choice(ncase, size=min(nplain, ncase), replace=False)
    t_plain, diff = [], 0.0
    for i in sample:
        lay, ss, p = cases[i]
        dt, qa = run_plain(lay, ss, p, nsteps)
        t_plain.append(dt)
        _, qb = srv.result(reqs[i])
        diff = max(diff, state_diff(qb, qa))
    srv.close()
    mean_plain = float(np.mean(t_plain))
    # per-case plain time split into set-up and steps for the projection to
    # the configured 2000 steps: the steps part scales with the step count
    setup_p = []
    for i in sample[:2]:
        lay, ss, p = cases[i]
        setup_p.append(run_plain(lay, ss, p, 0)[0])
    setup_plain = float(np.mean(setup_p))
    cph_b = 3600.0 * ncase / wall_b
    cph_p = 3600.0 / mean_plain
    proj_b = 3600.0 * ncase / (setup_b + 2000 * step_s)
    proj_p = 3600.0 / (setup_plain + 2000 * max(mean_plain - setup_plain, 1e-9) / nsteps)
    return {"cases": ncase, "steps_per_case": nsteps, "points_total": int(sum(pts)),
            "points_range": [int(min(pts)), int(max(pts))],
            "cases_per_hour": round(cph_b, 1), "plain_cases_per_hour": round(cph_p, 1),
            "speedup_vs_plain": round(cph_b / cph_p, 2),
            "wall_s": round(wall_b, 3), "setup_s": round(setup_b, 3), "ms_per_step": round(1e3 * step_s, 3),
            "mpoint_steps_per_s": round(sum(pts) / step_s / 1e6, 1),
            "plain_sample": int(len(sample)), "plain_mean_case_s": round(mean_plain, 4),
            "max_rel_diff_vs_plain": diff,
            "projected_2000_steps": {"cases_per_hour": round(proj_b, 1), "plain_cases_per_hour": round(proj_p, 1),
                                     "speedup_vs_plain": round(proj_b / proj_p, 2)},
            "workload": "single-aerofoil sinusoidal-gust cases (AoA 0-8 deg, M 0.2-0.5, amplitude 1-5 %, "
                        "reduced frequency 0.5-8), 0.15-2 M points, nz 13-32, seed 4"}


def tandem_bench(nsteps, warmup, tune):
    """Batching within one case: the tandem configuration's 24 uneven blocks
    (192 x 96 to 512 x 192 in-plane, nz = 64, 47.2 M points) in shared
    launches, line-solve shapes autotuned, steps replayed from the captured
    step graph."""
    lay, ss, params, dims = W.tandem_45m()
    npts = lay.npoints()
    log(f"tandem: {lay.nb} blocks, {npts / 1e6:.2f} M points")
    s = api.Solver([(lay, params)])
    for b in range(lay.nb):
        s.set_grid(0, b, ss[b])
    s.setup_geometry()
    tuned = s.autotune(5) if tune else None
    s.initialo(0)
    s.step(warmup, graph=True)
    s.sync()
    t0 = time.perf_counter()
    s.step(nsteps, graph=True)
    s.sync()
    dt = time.perf_counter() - t0
    finite = all(bool(np.isfinite(s.download(0, b, "qa")).all()) for b in (0, lay.nb - 1))
    s.close()
    out = {"value": round(npts * nsteps / dt / 1e6, 2), "unit": "Mpoint-steps/s", "steps": nsteps,
           "ms_per_step": round(dt / nsteps * 1e3, 3), "points": npts, "blocks": lay.nb, "graph": True,
           "finite": finite,
           "line_chunks": {f"{'xez'[d]}-{k}": c for (d, k), (c, _) in tuned.items()} if tuned else None,
           "workload": "two NACA0012 in tandem 3c apart, 24 uneven blocks, in-plane 192x96 to 512x192 "
                       f"(lxibk={dims['lxibk']}, letbk={dims['letbk']}), nz={dims['lzebk'][0] + 1}, M=0.3, "
                       "Re_c=2e5, viscous, turbulent inflow (2048 modes)"}
    log(f"tandem: {out['value']} Mpoint-steps/s ({out['ms_per_step']} ms per step)")
    return out


def serving_bench(nreq, nsteps, load, fill):
    """Poisson stream of served cases through the request loop: p50/p99 time
    to solution (arrival to final state on the host) and cases/h, against
    the plain path serving the same stream one case at a time (first come,
    first served, each case's plain time measured alone)."""
    cases = W.served_cases(nreq)
    pts = [lay.npoints() for lay, _, _ in cases]
    log(f"serving: {nreq} requests, {sum(pts) / 1e6:.1f} M points, {nsteps} steps each")
    # service time of each request alone on the plain path
    t_plain, qa_plain = [], []
    for lay, ss, p in cases:
        dt, qa = run_plain(lay, ss, p, nsteps)
        t_plain.append(dt)
        qa_plain.append(qa)
    # arrivals: Poisson at the rate that offers `load` times what the plain
    # path can serve
    rate = load * nreq / sum(t_plain)
    arr = np.cumsum(np.random.default_rng(55).exponential(1.0 / rate, nreq))
    arr -= arr[0]
    srv = api.Server(budget_bytes=0, max_active=64, graph=True, fill_points=fill)
    # a client thread submits each request at its arrival into the running
    # loop (the arrival is the submission call)
    reqs = srv.run_stream([(lay, p, ss, nsteps, a) for (lay, ss, p), a in zip(cases, arr)], max_steps=10)
    st = srv.stats()
    diff = 0.0
    for r, qa in zip(reqs, qa_plain):
        _, qb = srv.result(r)
        diff = max(diff, state_diff(qb, qa))
    srv.close()
    # the plain path serving the same arrivals, one at a time in arrival order
    done, tts = 0.0, []
    for a, t in zip(arr, t_plain):
        done = max(done, a) + t
        tts.append(done - a)
    wall_p = done - arr[0]
    tts = np.sort(np.array(tts))
    pct = lambda v, q: float(v[max(int(np.ceil(q * len(v))), 1) - 1])
    return {"requests": nreq, "steps_per_case": nsteps, "arrival_rate_per_s": round(rate, 3),
            "points_range": [int(min(pts)), int(max(pts))],
            "p50_s": round(st["p50_s"], 3), "p99_s": round(st["p99_s"], 3), "mean_s": round(st["mean_s"], 3),
            "cases_per_hour": round(st["cases_per_hour"], 1),
            "plain_p50_s": round(pct(tts, 0.5), 3), "plain_p99_s": round(pct(tts, 0.99), 3),
            "plain_cases_per_hour": round(3600.0 * nreq / wall_p, 1),
            "speedup_vs_plain": round(st["cases_per_hour"] / (3600.0 * nreq / wall_p), 2),
            "peak_cases_on_device": int(st["peak_cases"]), "peak_device_gb": round(st["peak_bytes"] / 1e9, 2),
            "joined_running": int(st["joined_running"]), "max_rel_diff_vs_plain": diff,
            "fill_points": int(fill), "admit_s": round(st["admit_s"], 3), "retire_s": round(st["retire_s"], 3),
            "stepping_s": round(st["stepping_s"], 3),
            "workload": "single-aerofoil turbulence-response cases (1024 Fourier modes, Tu 1-5 %, "
                        "L 0.05-0.2c), 0.2-20 M points, nz 16-96, seed 5; Poisson arrivals offering "
                        f"{load:g}x the plain path's capacity"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(seconds_target=15.0):
    """The oracle (C restatement, OpenMP) on a bounded sample of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    # the host cores this process may run on, and the share of them the
    # process is given (OMP_NUM_THREADS, when the host sets one)
    avail = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(avail, share) if share > 0 else avail
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
    return {"value": pts * steps / dt / 1e6, "unit": "Mpoint-steps/s", "cores": threads,
            "cores_available": avail, "kind": "port",
            "sample": f"oracle/tp_oracle.c on {lay.nb} sub-blocks, {pts / 1e6:.3f} M points "
                      f"(lxibk={dims['lxibk']}, letbk={dims['letbk']}, nz={dims['lzebk'][0] + 1}), "
                      f"{steps} steps, {dt:.1f} s"}


def measured_traffic_ratio():
    """HBM bytes moved by the derivative line solves of one time step over
    their algorithmic bytes, from the rocprofv3 FETCH_SIZE/WRITE_SIZE passes
    summarised under profiles/ (counter values scaled by a copy of known
    size); None when no such summary is present."""
    path = os.path.join(ROOT, "profiles", "line_solve_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["traffic_over_algorithmic"]
    except (OSError, ValueError, KeyError):
        return None


# Bytes per point and time step of a minimal viscous schedule (FP64, O(face)
# and sponge-zone traffic left out), in doubles per stage: three gradient
# passes reading Q (5) and writing 4 derivatives, 3 x 9 = 27; stresses and
# heat fluxes reading 12 gradients, 9 metrics, Q and the Jacobian and writing
# 9, 36; three flux passes reading Q, one metric row (3) and the 9 stresses
# and heat fluxes, plus the residual accumulated (25), 3 x 17 + 25 = 76; the
# update reading Q0, the residual and the Jacobian and writing Q, 16.  155
# doubles per stage, four stages, plus the filter sweep's 3 x 10 doubles:
# 1240 x 4 + 240 = 5200 B.
STEP_BYTES_PER_POINT = 5200


def measured_step_bytes():
    """HBM bytes per point moved by every kernel of one time step of the
    default workload (FETCH_SIZE/WRITE_SIZE passes summarised under
    profiles/); None when no such summary is present."""
    path = os.path.join(ROOT, "profiles", "step_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["bytes_per_point_step_measured"]
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
    ap.add_argument("--graph-steps", type=int, default=20,
                    help="steps of the extra timed pass of graph replays (0: none)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default line-solver launch shapes")
    ap.add_argument("--ensemble-cases", type=int, default=64, help="0: skip the ensemble measurement")
    ap.add_argument("--ensemble-steps", type=int, default=40)
    ap.add_argument("--ensemble-plain", type=int, default=8, help="cases timed on the plain path")
    ap.add_argument("--serve-requests", type=int, default=24, help="0: skip the serving measurement")
    ap.add_argument("--serve-steps", type=int, default=40)
    ap.add_argument("--serve-load", type=float, default=0.9,
                    help="offered load as a multiple of the plain path's capacity")
    ap.add_argument("--serve-fill", type=float, default=24e6,
                    help="admission stops while the cases on the device hold this many points")
    ap.add_argument("--tandem-steps", type=int, default=10,
                    help="steps of the tandem (24 uneven blocks, 47 M points) measurement (0: skip)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="after the run, one plain device-to-device copy of MB MiB (a reference kernel "
                         "of known traffic for scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU only; --gpus {args.gpus} is not supported "
              "(sharding the sub-blocks over GPUs is a design, not built)", file=sys.stderr)
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

    # HIP events around the launches of the dominant kernel class only (each
    # event pair costs the stream a few microseconds; around every launch
    # they would slow the step by ~3 %)
    if not args.graph:
        s.profile_classes(["deriv"])
    for k in api.KCLASS:
        s.profile_read(k, reset=True)
    s.sync()
    t0 = time.perf_counter()
    s.step(args.steps, graph=args.graph)
    s.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    s.profile(False)
    prof = {k: s.profile_read(k) for k in api.KCLASS}
    s.profile_read("deriv", reset=True)   # clears every class
    # the per-class breakdown from a short pass with every class timed
    breakdown_steps = max(1, min(5, args.steps))
    s.profile(True)
    s.step(breakdown_steps)
    s.sync()
    s.profile(False)
    prof_all = {k: s.profile_read(k) for k in api.KCLASS}
    s.profile_read("deriv", reset=True)
    sc = s.scalars(0)
    q = s.download(0, 0, "qa")
    line_chunks = {f"{'xez'[d]}-{k}": c for (d, k), c in s.col_config().items()}
    finite = bool(np.isfinite(q).all())
    print(f"[bench] {args.steps} steps in {elapsed:.3f} s; n={sc['n']:.0f} dt={sc['dt']:.3e} finite={finite}",
          file=sys.stderr, flush=True)
    # the same number of steps again as replays of the captured step (one
    # graph launch per step, no per-kernel events): the rate a host that
    # drives the solver through tp_step_graph sees
    graph_replay = None
    if not args.graph and args.graph_steps > 0:
        s.step(1, graph=True)   # capture
        s.sync()
        tg0 = time.perf_counter()
        s.step(args.graph_steps, graph=True)
        s.sync()
        tg = time.perf_counter() - tg0
        graph_replay = {"value": round(npts * args.graph_steps / tg / 1e6, 2), "steps": args.graph_steps,
                        "ms_per_step": round(tg / args.graph_steps * 1e3, 3)}
        log(f"graph replay: {graph_replay['value']} Mpoint-steps/s")

    value = npts * args.steps / elapsed / 1e6
    roof = None
    ms, nl, by = prof["deriv"]
    if nl > 0 and ms > 0:
        avg_s = ms / 1e3 / nl
        # 16 B/pt per field (field in, derivative out), 24 B/pt for the
        # launches that accumulate into their output (the eta and zeta flux
        # derivatives), times the fields of the launch (4 gradients, 5
        # fluxes): the library counts these per launch
        achieved = by / (ms / 1e3) / 1e9
        ratio = measured_traffic_ratio()
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": round(ratio * by / nl) if ratio else None,
                "kernel": "compact derivative line solves (k_col / k_deriv, 3 directions)",
                "avg_launch_us": round(avg_s * 1e6, 2), "launches": nl,
                "bytes_per_launch": round(by / nl)}
    # the whole step against the same peak: the minimal schedule's bytes per
    # point-step at the measured throughput, and the bytes the kernels were
    # measured to move
    step_roof = {"bound": "hbm", "bytes_per_point_step": STEP_BYTES_PER_POINT,
                 "achieved": round(value * 1e6 * STEP_BYTES_PER_POINT / 1e9, 1), "peak": HBM_PEAK_GBS,
                 "unit": "GB/s", "frac": round(value * 1e6 * STEP_BYTES_PER_POINT / 1e9 / HBM_PEAK_GBS, 4),
                 "measured_bytes_per_point_step": measured_step_bytes()}
    breakdown = {k: round(v[0] / breakdown_steps, 3) for k, v in prof_all.items() if v[1]}
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
    s.close()
    tandem = tandem_bench(args.tandem_steps, 3, not args.no_tune) if args.tandem_steps > 0 else None
    ens = ensemble_bench(args.ensemble_cases, args.ensemble_steps, args.ensemble_plain) \
        if args.ensemble_cases > 0 else None
    if ens:
        log(f"ensemble: {ens['cases_per_hour']} cases/h vs plain {ens['plain_cases_per_hour']} "
            f"({ens['speedup_vs_plain']}x)")
    srv = serving_bench(args.serve_requests, args.serve_steps, args.serve_load, args.serve_fill) \
        if args.serve_requests > 0 else None
    if srv:
        log(f"serving: p50 {srv['p50_s']} s p99 {srv['p99_s']} s, {srv['cases_per_hour']} cases/h "
            f"vs plain {srv['plain_cases_per_hour']} ({srv['speedup_vs_plain']}x)")
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
                   "line_chunks": line_chunks,
                   "ms_per_step_by_class": breakdown},
        "roofline": roof, "step_roofline": step_roof, "cpu_baseline": cpu, "graph_replay": graph_replay, "tandem": tandem, "ensemble": ens,
        "serving": srv,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

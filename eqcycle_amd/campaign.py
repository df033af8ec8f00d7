"""Earthquake-cycle campaign: the plain FP64 solver and the learned start with
the FP32 V-cycle over the same stretch of simulated time, with throughput, CG
iterations and the event catalogues side by side.

    python -m eqcycle_amd.campaign [--config cfg5] [--window-years 2] [--cycles 0] [--chunk-steps 500]

Learning phase: the plain FP64 solver runs from the seeded initial state and
records one (slip increment, displacement) snapshot per quasi-static step until
it holds --snapshots of them (the config's count, 3000 for cfg5); the POD basis
is fitted from them.  Windows (--windows): "learned" starts where the learning
phase ended, on simulated time the basis has not seen; "start" starts from the
seeded initial state, over the stretch the snapshots came from.  In each
window every run restarts from the window's first state and advances in chunks
of --chunk-steps (one run: the event bookkeeping carries over between chunks; a
progress line after each)
until --window-years have passed, or --cycles earthquake cycles have completed
when --cycles > 0 (a cycle: a cluster of event starts separated from the
previous one by more than --gap-s, ended by its last event end), or a run has
taken --max-wall-s, or the rate-state search has failed (FBC!,
src/NRsearch.jl:15-54: the run stops there, as the reference's does, and the
line records where).  Catalogue per event: eqcycle_amd.catalogue
(post/earthquake_cycles.jl).  One JSON line per run and a summary line
comparing throughput, iterations, event onsets and magnitudes.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

YR = 365 * 24 * 60 * 60


def cycle_onsets(tStart, gap_s):
    out, last = [], None
    for t in tStart:
        if last is None or t - last > gap_s:
            out.append(float(t))
        last = t
    return out


def run_cycles(ctx, a, label, t_end):
    """Advance to t_end (or until a.cycles cycles are complete, when > 0);
    returns stats and events."""
    t_wall = 0.0
    stopped = None
    t0 = ctx.state_get()["t"]
    keys = ("tStart", "tEnd", "hypo", "taubefore", "tauafter", "delfafter")
    ev = {k: [] for k in keys}
    n_qs = n_dyn = iters = 0
    wall_qs = wall_dyn = 0.0
    while True:
        w0 = time.perf_counter()
        R = ctx.run(total_time=t_end, max_steps=a.chunk_steps, snapshots=False, events=True,
                    cap_steps=a.chunk_steps + 1, cap_events=4000, keep_on_nr_failure=True)
        t_wall += time.perf_counter() - w0
        for k in keys:
            if len(R[k]):
                ev[k].append(R[k])
        n_qs += int(R["n_qs"])
        n_dyn += int(R["n_dyn"])
        iters += int(R["total_pcg_iters"])
        wall_qs += R["wall_qs"]
        wall_dyn += R["wall_dyn"]
        s = ctx.state_get()
        tS = np.concatenate(ev["tStart"]) if ev["tStart"] else np.zeros(0)
        tE = np.concatenate(ev["tEnd"]) if ev["tEnd"] else np.zeros(0)
        onsets = cycle_onsets(tS, a.gap_s)
        # complete: the (cycles + 1)-th onset has been reached, or the last event ended
        # after the cycles-th onset and the regime is quasi-static again
        done = s["t"] >= t_end or (a.cycles > 0 and (
            len(onsets) > a.cycles or (len(onsets) == a.cycles and len(tE) == len(tS) and s["isolver"] == 1
                                       and s["t"] - tE[-1] > a.gap_s)))
        nq = max(1, int(R["n_qs"]))
        print(f"[campaign] {label}: t = {s['t'] / YR:.2f} yr, {len(onsets)} cycles, {len(tS)} event starts, "
              f"wall {t_wall:.1f} s; chunk: {int(R['n_qs'])} qs / {int(R['n_dyn'])} dyn steps, "
              f"{R['total_pcg_iters'] / (2 * nq):.1f} iters/solve, {R['wall_qs'] * 1e3 / nq:.2f} ms/qs step", flush=True)
        if R["error"]:
            stopped = {"error": R["error"], "t_yr": s["t"] / YR}
            print(f"[campaign] {label}: stopped at t = {s['t'] / YR:.6f} yr: {R['error']}", flush=True)
            break
        if done or t_wall > a.max_wall_s:
            break
    out = {k: (np.concatenate(v) if v else np.zeros((0,))) for k, v in ev.items()}
    years = (ctx.state_get()["t"] - t0) / YR
    stats = {"run": label, "t_start_yr": t0 / YR, "t_end_yr": t0 / YR + years, "sim_years": years, "wall_s": t_wall, "sim_years_per_hour": years / (t_wall / 3600),
             "qs_steps": n_qs, "dyn_steps": n_dyn, "pcg_iters_per_solve": iters / max(1, 2 * n_qs),
             "wall_qs_s": wall_qs, "wall_dyn_s": wall_dyn, "events": int(len(out["tStart"])),
             "cycle_onsets_yr": [t / YR for t in cycle_onsets(out["tStart"], a.gap_s)], "stopped": stopped}
    return stats, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="cfg5")
    ap.add_argument("--window-years", type=float, default=2.0)
    ap.add_argument("--cycles", type=int, default=0, help="> 0: stop a run once this many cycles are complete")
    ap.add_argument("--chunk-steps", type=int, default=500)
    ap.add_argument("--max-wall-s", type=float, default=480.0, help="stop a run after this much wall time")
    ap.add_argument("--gap-s", type=float, default=1e7, help="event starts further apart open a new cycle")
    ap.add_argument("--snapshots", type=int, default=None, help="default: the config's (3000 for cfg5)")
    ap.add_argument("--mu", type=float, default=None, help="rigidity for Mw (default rho vs^2 of the host rock)")
    ap.add_argument("--runs", default="plain,galerkin")
    ap.add_argument("--windows", default="learned",
                    help="learned: from the state where the learning phase ended (time the basis has not seen); "
                         "start: from the seeded initial state (the stretch the snapshots came from)")
    ap.add_argument("--graphs", type=int, default=1, help="replay CG iteration batches as HIP graphs")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    from eqcycle_amd import learned
    from eqcycle_amd.api import EqContext
    from eqcycle_amd.catalogue import catalogue
    from eqcycle_amd.configs import CONFIGS
    from eqcycle_amd.params import set_parameters

    cfg = CONFIGS[a.config]
    P = set_parameters(cfg)
    tol = cfg.pcg_tol
    mu = a.mu or cfg.rho * cfg.vs ** 2
    w0 = time.perf_counter()
    ctx = EqContext(P, amg={"hierarchy": "pmg", "degree": 5})
    print(f"[campaign] {a.config}: {P.nglob} nodes, context and preconditioner set up in "
          f"{time.perf_counter() - w0:.1f} s", flush=True)

    # learning phase: plain FP64 steps from the seeded state, one snapshot per quasi-static step
    nsnap = a.snapshots or cfg.n_snapshots or 2000
    w0 = time.perf_counter()
    ctx.set_solver(tol=tol, pc="amg", prec="fp64", guess="previous", graphs=bool(a.graphs))
    ctx.state_init()
    ctx.learn_init(capacity=nsnap, nbasis=64)
    ctx.learn_record(True)
    while ctx.learn_info()["nsnap"] < nsnap:
        ctx.run(total_time=1e30, max_steps=a.chunk_steps, snapshots=False)
        print(f"[campaign] learning phase: {ctx.learn_info()['nsnap']} snapshots, "
              f"t = {ctx.state_get()['t'] / YR:.3f} yr", flush=True)
    ctx.learn_record(False)
    S = ctx.state_get()
    fit = None
    if any(r != "plain" for r in a.runs.split(",")):
        fit = learned.fit_basis(ctx, tol=tol)
    learning = {"snapshots": nsnap, "basis_used": None if fit is None else fit["used"],
                "learning_s": time.perf_counter() - w0, "window_start_yr": S["t"] / YR}
    print(json.dumps(learning), flush=True)
    lines = []
    for window in a.windows.split(","):
        t_start = S["t"] if window == "learned" else 0.0
        t_end = t_start + a.window_years * YR
        wl = []
        for name in a.runs.split(","):
            if name == "plain":
                ctx.set_solver(tol=tol, pc="amg", prec="fp64", guess="previous", graphs=bool(a.graphs))
            else:
                ctx.set_solver(tol=tol, pc="amg", prec=cfg.precision, guess="galerkin", graphs=bool(a.graphs),
                               basis_fp16=True)
            if window == "learned":
                ctx.state_set(d=S["d"], v=S["v"], a=S["a"], psi=S["psi"], tau=S["tau"], t=S["t"], dt=S["dt"],
                              isolver=S["isolver"])
            elif window == "start":
                ctx.state_init()
            else:
                raise ValueError(f"unknown window {window!r}")
            stats, ev = run_cycles(ctx, a, f"{window}/{name}", t_end)
            stats.update(window=window, prec="fp64" if name == "plain" else cfg.precision)
            cat = catalogue(ev, P.FltX, mu) if len(ev["delfafter"]) else []
            stats["Mw"] = [round(e["Mw"], 3) for e in cat]
            print(json.dumps(stats), flush=True)
            wl.append(stats)
        summary = {"config": a.config, "nodes": int(P.nglob), "tol": tol, "window": window,
                   "window_years": a.window_years, "cycles_requested": a.cycles, **learning}
        if len(wl) == 2:
            # the plain run and the learned one, in whichever order --runs gave them
            p, g = sorted(wl, key=lambda r: 0 if r["run"].endswith("/plain") else 1)
            n = min(len(p["cycle_onsets_yr"]), len(g["cycle_onsets_yr"]))
            summary.update({
                "t_end_yr": [p["t_end_yr"], g["t_end_yr"]],
                "t_end_rel_diff": abs(g["t_end_yr"] - p["t_end_yr"]) / max(p["t_end_yr"], 1e-30),
                "speedup": g["sim_years_per_hour"] / p["sim_years_per_hour"],
                "iters_ratio": g["pcg_iters_per_solve"] / max(p["pcg_iters_per_solve"], 1e-30),
                "cycles_compared": n,
                "cycle_onset_rel_diff_max": max((abs(g["cycle_onsets_yr"][k] - p["cycle_onsets_yr"][k]) /
                                                 p["cycle_onsets_yr"][k] for k in range(n)), default=None),
                "Mw_abs_diff_max": max((abs(x - y) for x, y in zip(p["Mw"], g["Mw"])), default=None),
            })
        print(json.dumps(summary), flush=True)
        lines += wl + [summary]
    if a.out:
        with open(a.out, "w") as f:
            for line in lines:
                f.write(json.dumps(line) + "\n")


if __name__ == "__main__":
    main()

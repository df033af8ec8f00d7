"""Earthquake-cycle campaign: the plain FP64 solver and the learned start with
the FP32 V-cycle over the same number of earthquake cycles, with throughput,
CG iterations and the event catalogues side by side.

    python -m eqcycle_amd.campaign [--config cfg5] [--cycles 20] [--chunk-years 100]

Both runs start from the seeded initial state and advance in chunks of
--chunk-years (one run: the event bookkeeping carries over between chunks)
until --cycles earthquake cycles have completed (a cycle: a cluster of event
starts separated from the previous one by more than --gap-s, ended by its last
event end) or --max-years.  The learned run first records --snapshots plain
quasi-static solves from the seeded state (not timed, reported) and builds the
POD basis; its start is the Galerkin start on that basis.  Catalogue per event:
eqcycle_amd.catalogue (post/earthquake_cycles.jl).  One JSON line per run and a
summary line comparing cycle onsets and magnitudes.
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


def run_cycles(ctx, a, label):
    """Advance until a.cycles cycles are complete; returns stats and events."""
    t_wall = 0.0
    keys = ("tStart", "tEnd", "hypo", "taubefore", "tauafter", "delfafter")
    ev = {k: [] for k in keys}
    n_qs = n_dyn = iters = 0
    wall_qs = wall_dyn = 0.0
    while True:
        t0 = ctx.state_get()["t"]
        w0 = time.perf_counter()
        R = ctx.run(total_time=t0 + a.chunk_years * YR, snapshots=False, events=True, cap_steps=4_000_000,
                    cap_events=4000)
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
        done = len(onsets) > a.cycles or (len(onsets) == a.cycles and len(tE) == len(tS) and s["isolver"] == 1
                                          and s["t"] - tE[-1] > a.gap_s)
        print(f"[campaign] {label}: t = {s['t'] / YR:.1f} yr, {len(onsets)} cycles, {len(tS)} event starts, "
              f"wall {t_wall:.1f} s", flush=True)
        if done or s["t"] / YR >= a.max_years or t_wall > a.max_wall_s:
            break
    out = {k: (np.concatenate(v) if v else np.zeros((0,))) for k, v in ev.items()}
    years = ctx.state_get()["t"] / YR
    stats = {"run": label, "sim_years": years, "wall_s": t_wall, "sim_years_per_hour": years / (t_wall / 3600),
             "qs_steps": n_qs, "dyn_steps": n_dyn, "pcg_iters_per_solve": iters / max(1, 2 * n_qs),
             "wall_qs_s": wall_qs, "wall_dyn_s": wall_dyn, "events": int(len(out["tStart"])),
             "cycle_onsets_yr": [t / YR for t in cycle_onsets(out["tStart"], a.gap_s)]}
    return stats, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="cfg5")
    ap.add_argument("--cycles", type=int, default=20)
    ap.add_argument("--chunk-years", type=float, default=100.0)
    ap.add_argument("--max-years", type=float, default=5000.0)
    ap.add_argument("--max-wall-s", type=float, default=480.0, help="stop a run after this much wall time")
    ap.add_argument("--gap-s", type=float, default=1e7, help="event starts further apart open a new cycle")
    ap.add_argument("--snapshots", type=int, default=None, help="default: the config's (3000 for cfg5)")
    ap.add_argument("--mu", type=float, default=None, help="rigidity for Mw (default rho vs^2 of the host rock)")
    ap.add_argument("--runs", default="plain,galerkin")
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
    ctx = EqContext(P, amg={"hierarchy": "pmg", "degree": 5})
    lines, cats = [], {}
    for name in a.runs.split(","):
        info = {}
        if name == "plain":
            ctx.set_solver(tol=tol, pc="amg", prec="fp64", guess="previous")
            ctx.state_init()
        else:
            nsnap = a.snapshots or cfg.n_snapshots or 2000
            w0 = time.perf_counter()
            ctx.set_solver(tol=tol, pc="amg", prec="fp64", guess="previous")
            ctx.state_init()
            ctx.learn_init(capacity=nsnap, nbasis=64)
            ctx.learn_record(True)
            while ctx.learn_info()["nsnap"] < nsnap:
                ctx.run(total_time=1e30, max_steps=400, snapshots=False)
            ctx.learn_record(False)
            fit = learned.fit_basis(ctx, tol=tol)
            info = {"snapshots": nsnap, "used": fit["used"], "learning_s": time.perf_counter() - w0}
            ctx.set_solver(tol=tol, pc="amg", prec=cfg.precision, guess="galerkin")
            ctx.state_init()
        stats, ev = run_cycles(ctx, a, name)
        stats.update(info)
        cats[name] = catalogue(ev, P.FltX, mu) if len(ev["delfafter"]) else []
        stats["Mw"] = [round(e["Mw"], 3) for e in cats[name]]
        print(json.dumps(stats), flush=True)
        lines.append(stats)
    summary = {"config": a.config, "nodes": int(P.nglob), "tol": tol, "cycles_requested": a.cycles}
    if len(lines) == 2:
        p, g = lines
        n = min(len(p["cycle_onsets_yr"]), len(g["cycle_onsets_yr"]))
        summary.update({
            "speedup": g["sim_years_per_hour"] / p["sim_years_per_hour"],
            "iters_ratio": g["pcg_iters_per_solve"] / max(p["pcg_iters_per_solve"], 1e-30),
            "cycles_compared": n,
            "cycle_onset_rel_diff_max": max((abs(g["cycle_onsets_yr"][k] - p["cycle_onsets_yr"][k]) /
                                             p["cycle_onsets_yr"][k] for k in range(n)), default=None),
        })
    print(json.dumps(summary), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            for line in lines + [summary]:
                f.write(json.dumps(line) + "\n")


if __name__ == "__main__":
    main()

"""Error of the mixed-precision solvers against the FP64 path on one fault model.

    python -m eqcycle_amd.precision_report [--config cfg3] [--years 300] [--tol 1e-8]

Runs main(P) on the GPU from the seeded initial state with

  fp64      FP64 CG, FP64 V-cycle, tolerance --tol
  fp64_ref  FP64 CG at --tol / 100 (the reference trajectory)
  mixed     FP32 inner CG inside FP64 iterative refinement, tolerance --tol
  f32pc     FP64 CG with the FP32 V-cycle, tolerance --tol

and reports, for each run against fp64_ref: the relative difference of the
final displacement field, of the final fault slip, the largest difference of
log10(Vfmax) on a common time grid (quasi-static history, interpolated), and
the relative difference of each event-cluster onset time.  The fp64 run at
the same tolerance as the mixed runs shows how far two correct FP64 solves
already drift apart; a mixed-precision error of that size is the solver
tolerance, not the precision.  One JSON line per run, then a summary line.
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


def clusters(ts, gap_s):
    """Onset of each group of event starts separated by more than gap_s."""
    out = []
    for t in ts:
        if not out or t - last > gap_s:
            out.append(float(t))
        last = t
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="cfg3")
    ap.add_argument("--years", type=float, default=300.0)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--runs", default="fp64_ref,fp64,mixed,f32pc")
    ap.add_argument("--out", default=None, help="write the JSON lines to this file too")
    a = ap.parse_args()
    from eqcycle_amd.api import EqContext
    from eqcycle_amd.configs import CONFIGS
    from eqcycle_amd.params import set_parameters

    cfg = CONFIGS[a.config]
    P = set_parameters(cfg)
    ctx = EqContext(P, amg={"hierarchy": "pmg", "degree": 5})
    iflt = P.iFlt - 1
    res = {}
    lines = []
    for name in a.runs.split(","):
        prec = {"fp64_ref": "fp64", "fp64": "fp64", "mixed": "mixed", "f32pc": "f32pc"}[name]
        tol = a.tol / 100 if name == "fp64_ref" else a.tol
        ctx.set_solver(tol=tol, pc="amg", prec=prec)
        ctx.state_init()
        w0 = time.perf_counter()
        R = ctx.run(total_time=a.years * YR, snapshots=False, cap_steps=4_000_000)
        wall = time.perf_counter() - w0
        s = ctx.state_get()
        res[name] = dict(R=R, d=s["d"], t=s["t"])
        line = {"run": name, "prec": prec, "tol": tol, "wall_s": wall, "sim_years": s["t"] / YR,
                "sim_years_per_hour": s["t"] / YR / (wall / 3600), "steps": int(R["n_steps"]),
                "qs_steps": int(R["n_qs"]), "dyn_steps": int(R["n_dyn"]),
                "pcg_iters_per_solve": R["total_pcg_iters"] / max(1, 2 * R["n_qs"]),
                "events": int(len(R["tStart"])), "cluster_onsets_yr": [t / YR for t in clusters(R["tStart"], 1e7)]}
        print(json.dumps(line), flush=True)
        lines.append(line)
    ref = res.get("fp64_ref")
    summary = {"config": a.config, "years": a.years, "tol": a.tol, "nodes": int(P.nglob), "vs_fp64_ref": {}}
    if ref is not None:
        tr = ref["R"]["time_"]
        grid = np.linspace(tr[0], min(x["t"] for x in res.values()), 2000)
        lr = np.interp(grid, tr, np.log10(ref["R"]["Vfmax"]))
        c_ref = clusters(ref["R"]["tStart"], 1e7)
        for name, x in res.items():
            if name == "fp64_ref":
                continue
            lx = np.interp(grid, x["R"]["time_"], np.log10(x["R"]["Vfmax"]))
            c_x = clusters(x["R"]["tStart"], 1e7)
            n = min(len(c_ref), len(c_x))
            summary["vs_fp64_ref"][name] = {
                "d_rel": float(np.max(np.abs(x["d"] - ref["d"])) / np.max(np.abs(ref["d"]))),
                "fault_slip_rel": float(np.max(np.abs(x["d"][iflt] - ref["d"][iflt])) /
                                        np.max(np.abs(ref["d"][iflt]))),
                "log10_Vfmax_max_abs_diff": float(np.max(np.abs(lx - lr))),
                "event_clusters": [len(c_x), len(c_ref)],
                "cluster_onset_rel_diff": [abs(c_x[k] - c_ref[k]) / c_ref[k] for k in range(n)],
            }
    print(json.dumps(summary), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            for line in lines + [summary]:
                f.write(json.dumps(line) + "\n")


if __name__ == "__main__":
    main()

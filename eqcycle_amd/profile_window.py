"""One solver variant of the time loop over a window of steps, for profilers.

    python -m eqcycle_amd.profile_window [--config cfg4] [--variant galerkin] [--snapshots 400] [--steps 300]

Builds the learned POD basis from --snapshots plain quasi-static solves (the
same recording and POD as bench.py; the network is trained only for the
`network` variant), restarts from the seeded initial state, advances
--warmup-years with the plain solver, then runs --steps steps of the chosen
variant and prints the wall time per step type.  Right before the window the
script runs the 512 MiB calibration copy once (`convert_kernel<double,
double>`, through eq_time_kernel): in a `rocprofv3 --kernel-trace` of this
program the window is every dispatch after that copy.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

YR = 365 * 24 * 60 * 60


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="cfg4")
    ap.add_argument("--variant", default="galerkin", help="plain | network | galerkin")
    ap.add_argument("--prec", default=None, help="default: the config's precision for learned variants, fp64 plain")
    ap.add_argument("--snapshots", type=int, default=400)
    ap.add_argument("--epochs", type=int, default=100)
    ap.add_argument("--warmup-years", type=float, default=20.0)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--graphs", type=int, default=1,
                    help="replay CG batches as HIP graphs (0 under rocprofv3 --kernel-trace, which crashed in "
                         "graph capture here)")
    a = ap.parse_args()
    from eqcycle_amd import learned
    from eqcycle_amd.api import EqContext
    from eqcycle_amd.configs import CONFIGS
    from eqcycle_amd.params import set_parameters

    cfg = CONFIGS[a.config]
    P = set_parameters(cfg)
    ctx = EqContext(P, amg={"hierarchy": "pmg", "degree": 5})
    tol = cfg.pcg_tol
    guess = {"plain": "previous", "network": "learned", "galerkin": "galerkin"}[a.variant]
    prec = a.prec or ("fp64" if a.variant == "plain" else cfg.precision)
    if a.variant != "plain":
        ctx.set_solver(tol=tol, pc="amg", prec="fp64", guess="previous", graphs=bool(a.graphs))
        ctx.state_init()
        ctx.learn_init(capacity=a.snapshots, nbasis=64)
        ctx.learn_record(True)
        while ctx.learn_info()["nsnap"] < a.snapshots:
            ctx.run(total_time=1e30, max_steps=200, snapshots=False)
        ctx.learn_record(False)
        learned.fit_basis(ctx, tol=tol)
        if a.variant == "network":
            learned.train(ctx, epochs=a.epochs, batch=64, lr=1e-3)
    print("[profile_window] learning done; window follows", flush=True)
    ctx.set_solver(tol=tol, pc="amg", prec="fp64", guess="previous", graphs=bool(a.graphs))
    ctx.state_init()
    ctx.run(total_time=a.warmup_years * YR, snapshots=False)
    ctx.set_solver(tol=tol, pc="amg", prec=prec, guess=guess, graphs=bool(a.graphs))
    ctx.time_kernel("copy", "fp64", reps=1)   # trace marker: the window starts after this copy
    t0 = time.perf_counter()
    R = ctx.run(total_time=1e30, max_steps=a.steps, snapshots=False)
    wall = time.perf_counter() - t0
    out = {"config": a.config, "variant": a.variant, "prec": prec, "steps": int(R["n_steps"]),
           "qs_steps": int(R["n_qs"]), "dyn_steps": int(R["n_dyn"]), "wall_s": wall,
           "ms_per_qs_step": R["wall_qs"] * 1e3 / max(1, R["n_qs"]),
           "us_per_dyn_step": R["wall_dyn"] * 1e6 / max(1, R["n_dyn"]),
           "pcg_iters_per_solve": R["total_pcg_iters"] / max(1, 2 * R["n_qs"])}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""One pass of a solver variant over the bench's earthquake cycle, for profilers.

    python -m eqcycle_amd.profile_window [--config cfg4] [--variant galerkin] [--graphs 1]

Builds the bench's checkpoint (the plain solver from the seeded state to year
110, recording the training snapshots on the way; POD basis, network and
linear predictor fitted as bench.py does), runs one untimed pass of the
variant, then the 512 MiB calibration copy once (`convert_kernel<double,
double>`, through eq_time_kernel) and one more pass: in a `rocprofv3
--kernel-trace` of this program the pass is every dispatch after that copy
(eqcycle_amd.trace_window).  CG batches launch eagerly unless --graphs 1;
keep them eager under kernel tracing: the tracer's interception of HIP stream
capture has crashed the host process on these machines.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="cfg4")
    ap.add_argument("--variant", default="galerkin", help="plain | network | linear | galerkin")
    ap.add_argument("--degree", type=int, default=None,
                    help="fine-level Chebyshev degree of the measured pass (default as bench.py: 5 plain, 7 else)")
    ap.add_argument("--degree-plain", type=int, default=5, help="degree of the checkpoint run")
    ap.add_argument("--graphs", type=int, default=0)
    ap.add_argument("--extrapolate", type=int, default=1, help="extrapolated Galerkin start, as bench.py")
    a = ap.parse_args()
    import bench
    from eqcycle_amd.api import EqContext
    from eqcycle_amd.configs import CONFIGS
    from eqcycle_amd.params import set_parameters

    cfg = CONFIGS[a.config]
    P = set_parameters(cfg)
    ctx = EqContext(P, amg={"hierarchy": "pmg", "degree": a.degree_plain})
    tol = cfg.pcg_tol
    ck, _ = bench.checkpoint(ctx, cfg, tol, cfg.learned, 300, a.graphs)
    degree = a.degree if a.degree is not None else (a.degree_plain if a.variant == "plain" else 7)
    if degree != a.degree_plain:
        ctx.amg_setup(hierarchy="pmg", degree=degree)
    guess = {"plain": "previous", "network": "learned", "linear": "linear", "galerkin": "galerkin"}[a.variant]
    prec = "fp64" if a.variant == "plain" else cfg.precision
    ctx.set_solver(tol=tol, pc="amg", prec=prec, guess=guess, graphs=bool(a.graphs), basis_fp16=guess == "galerkin",
                   extrapolate=guess == "galerkin" and bool(a.extrapolate))
    t_end = ck["t"] + bench.SEGMENT_YEARS * bench.YR
    bench.one_pass(ctx, ck, t_end)
    print("[profile_window] checkpoint and warm-up pass done; the measured pass follows", flush=True)
    ctx.time_kernel("copy", "fp64", reps=1)   # trace marker: the pass starts after this copy
    wall, R = bench.one_pass(ctx, ck, t_end)
    out = {"config": a.config, "variant": a.variant, "prec": prec, "degree": degree, "graphs": bool(a.graphs),
           "wall_s": wall,
           "qs_steps": int(R["n_qs"]), "dyn_steps": int(R["n_dyn"]),
           "pcg_iters_per_solve": R["total_pcg_iters"] / max(1, 2 * R["n_qs"]),
           "wall_qs_s": R["wall_qs"], "wall_dyn_s": R["wall_dyn"], "events": int(len(R["tStart"]))}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark: simulated years per hour of the earthquake-cycle time loop on one MI355X.

Workload: the config-2 fault model (160 x 160 elements, NGLL 5, 410,881 nodes,
trapezoidal damage zone, FP64, AMG-preconditioned CG to 1e-8 relative residual)
run by the device-resident time loop (main(P)) from its seeded initial state.
One bench step is one simulated year: --warmup W years are simulated untimed,
then exactly --steps K years are timed (the loop stops at the first time step
that reaches the target time, so the timed window is K years plus at most one
step; the JSON reports the simulated years actually covered).

Output: one JSON line (metric, value, unit, ..., roofline, cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

YR = 365 * 24 * 60 * 60
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed simulated years")
    ap.add_argument("--warmup", type=int, default=1, help="untimed simulated years before the timed window")
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--prec", default=None, help="fp64 | mixed (default: the config's)")
    ap.add_argument("--pc", default="amg")
    ap.add_argument("--hierarchy", default="pmg", help="pmg (default) | rs")
    ap.add_argument("--degree", type=int, default=3, help="Chebyshev degree on the fine level (pmg)")
    ap.add_argument("--tol", type=float, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    return ap.parse_args()


def cpu_baseline(P, cfg, tol, stats, window_years, budget_s):
    """Time the oracle (C FP64 restatement, one thread) on a bounded sample of the
    same workload and convert to simulated years per hour with the GPU run's step mix."""
    import numpy as np

    from oracle import oracle as O
    t_setup = time.perf_counter()
    sim = O.Simulation(P, tol=tol, solver="amg")
    t_setup = time.perf_counter() - t_setup
    # (i) AMG-PCG iterations on a quasi-static system: time per iteration
    n = sim.kni.n
    rng = np.random.default_rng(0)
    b = rng.standard_normal(n)
    t0 = time.perf_counter()
    O.cg(sim.kni, np.zeros(n), b, sim.ml, 1e-30, maxiter=2)
    per_it = (time.perf_counter() - t0) / 2
    m = max(2, int(0.6 * budget_s / per_it))
    t0 = time.perf_counter()
    O.cg(sim.kni, np.zeros(n), b, sim.ml, 1e-30, maxiter=m)
    per_it = (time.perf_counter() - t0) / m
    # SpMVs and slrFunc! of one QS pass
    t0 = time.perf_counter()
    for _ in range(3):
        sim.K.matvec(np.zeros(P.nglob))
    spmv = (time.perf_counter() - t0) / 3
    # (ii) Newmark steps
    sim.isolver = 2
    sim.dt = sim.dtmin
    k = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 0.3 * budget_s and k < 200:
        sim.dyn_step()
        k += 1
    per_dyn = (time.perf_counter() - t0) / k
    iters_per_solve = stats["iters_per_solve"]
    per_qs = 2 * (iters_per_solve * per_it + 2 * spmv)
    cpu_s = stats["n_qs"] * per_qs + stats["n_dyn"] * per_dyn
    value = window_years / (cpu_s / 3600.0)
    return {
        "value": value, "unit": "sim-years/hour", "cores": 1, "kind": "port",
        "sample": (f"oracle (C FP64 restatement: assembled CSR K, RS-AMG + symmetric Gauss-Seidel V-cycle, "
                   f"cg!, serial FBC!/NRsearch!) on the same mesh: {m} AMG-PCG iterations "
                   f"({per_it * 1e3:.1f} ms each) and {k} Newmark steps ({per_dyn * 1e3:.1f} ms each), "
                   f"scaled by the GPU window's {stats['n_qs']} QS steps x 2 solves x "
                   f"{iters_per_solve:.1f} iterations (the GPU's count, favourable to the CPU) and "
                   f"{stats['n_dyn']} dynamic steps; set-up {t_setup:.1f} s not counted"),
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); every GPU box here has one GPU and the "
              "multi-GPU decomposition is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)
    from eqcycle_amd.api import EqContext
    from eqcycle_amd.configs import CONFIGS
    from eqcycle_amd.params import set_parameters

    cfg = CONFIGS[a.config]
    prec = a.prec or cfg.precision
    tol = a.tol or cfg.pcg_tol
    P = set_parameters(cfg)
    amg = {"hierarchy": a.hierarchy, "degree": a.degree} if a.pc == "amg" else False
    ctx = EqContext(P, amg=amg)
    ctx.set_solver(tol=tol, pc=a.pc, prec=prec)
    ctx.state_init()
    # warm-up window (untimed)
    if a.warmup > 0:
        ctx.run(total_time=a.warmup * YR, snapshots=False)
    s0 = ctx.state_get()
    t_start = s0["t"]
    # timed window: K simulated years; every step ends with a device sync
    w0 = time.perf_counter()
    R = ctx.run(total_time=t_start + a.steps * YR, snapshots=False)
    w1 = time.perf_counter()
    t_end = ctx.state_get()["t"]
    wall = w1 - w0
    years = (t_end - t_start) / YR
    n_solves = 2 * R["n_qs"]
    stats = dict(n_qs=R["n_qs"], n_dyn=R["n_dyn"],
                 iters_per_solve=R["total_pcg_iters"] / max(1, n_solves))
    value = years / (wall / 3600.0)

    # roofline of the dominant kernel (largest share of device time in the
    # rocprof summary): the fused Chebyshev smoothing step on the matrix-free
    # fine level.  Per node it reads the window vector d, mu, dinv, r, x and
    # writes r, d', x: 8 scalars of s bytes (the 4-node LDS halo is not counted).
    dt_ms = ctx.time_kernel("cheb", "fp64" if prec == "fp64" else "fp32", reps=200)
    s = 8 if prec == "fp64" else 4
    traffic_alg = 8 * s * P.nglob
    achieved = traffic_alg / (dt_ms * 1e-3) / 1e9
    roof = {"kernel": "tile_kernel<OpCheb> (fine-level Chebyshev step: K_II d fused with the r, d, x updates)",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "avg_launch_us": dt_ms * 1e3,
            "alg_bytes_per_launch": traffic_alg}

    out = {
        "metric": "sim-years/hour", "value": value, "unit": "sim-years/hour", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": wall * 1e3 / max(a.steps, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == "fp64" else
        "f32+f64", "data": "synthetic (seeded fault model, config 2)",
        "config": {"workload": f"{a.config}: {cfg.nelx}x{cfg.nely} elements NGLL 5 ({P.nglob} nodes), "
                               f"{cfg.material} damage zone, AMG-PCG tol {tol:g}, {prec}",
                   "sim_years_timed": years, "qs_steps": R["n_qs"], "dyn_steps": R["n_dyn"],
                   "pcg_iters_per_solve": stats["iters_per_solve"], "wall_qs_s": R["wall_qs"],
                   "wall_dyn_s": R["wall_dyn"], "pc": a.pc, "hierarchy": a.hierarchy,
                   "amg_levels": ctx.amg_info() if a.pc == "amg" else None},
        "roofline": roof,
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(P, cfg, tol, stats, years, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

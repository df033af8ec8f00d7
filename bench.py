"""Benchmark: simulated years per hour and CG iterations per solve of the
earthquake-cycle time loop on one MI355X, plain FP64 against learned initial
guess + FP32 V-cycle at the same relative residual.

Default workload (config 4): 160 x 160 elements, NGLL 5 (410,881 nodes),
trapezoidal damage zone, p-multigrid preconditioned CG to 1e-6 relative
residual, the whole time loop (quasi-static steps, Newmark steps, rate-state
Newton searches, adaptive dt, event bookkeeping) resident on the GPU.

One bench step is one simulated year.  Each variant starts from the same state:
the seeded initial state advanced --warmup W years with the plain solver
(untimed); then exactly --steps K years are timed (the loop stops at the first
time step that reaches the target time, so a window is K years plus at most one
step; the JSON reports the simulated years actually covered and uses them).

  plain   : FP64 CG, FP64 V-cycle, previous-step initial guess (src/main.jl:241)
  galerkin: FP64 CG, FP32 V-cycle, start x + Phi c on the POD basis Phi learned
            from the plain solver's own quasi-static snapshots, c the Galerkin
            (K_II-norm optimal) coefficients, both quasi-static passes
  network : FP64 CG, FP32 V-cycle, pass-1 start from the conv1d + MLP network's
            predicted POD coefficients with an energy-optimal step
  (snapshot recording, POD and network training run beforehand, untimed and
  reported)

`value` is the galerkin variant's sim-years/hour (the plain one's for configs
without a learned basis); all variants, their iteration counts and their
agreement with the plain run are in `config`.  Output: one JSON line with `roofline` (dominant
kernel, HIP-event timed) and `cpu_baseline` (the oracle, one core, bounded
sample).
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

# HBM bytes per launch from two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE)
# calibrated by a 512 MiB copy, on the 410,881-node mesh
# (profiles/newmark_cheb_pair_cfg4_hbm_traffic.md).  Keyed by nodes, and
# precision for the Chebyshev pair.
NEWMARK_TRAFFIC = {410881: 27.19e6}
CHEB_PAIR_TRAFFIC = {(410881, "fp64"): 28.07e6, (410881, "fp32"): 14.29e6}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed simulated years per variant")
    ap.add_argument("--warmup", type=int, default=1, help="untimed simulated years before each timed window")
    ap.add_argument("--config", default="cfg4")
    ap.add_argument("--tol", type=float, default=None)
    ap.add_argument("--hierarchy", default="pmg", help="pmg (default) | rs (Ruge-Stueben on K_II)")
    ap.add_argument("--degree", type=int, default=5, help="Chebyshev degree on the fine level (pmg)")
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain-only", action="store_true")
    return ap.parse_args()


def cpu_baseline(P, tol, stats, window_years, budget_s):
    """Time the oracle (C FP64 restatement, one thread) on a bounded sample of the
    same workload and convert to simulated years per hour with the GPU plain
    run's step mix and iteration counts."""
    import numpy as np

    from oracle import oracle as O
    t_setup = time.perf_counter()
    sim = O.Simulation(P, tol=tol, solver="amg")
    t_setup = time.perf_counter() - t_setup
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
    t0 = time.perf_counter()
    for _ in range(3):
        sim.K.matvec(np.zeros(P.nglob))
    spmv = (time.perf_counter() - t0) / 3
    sim.isolver = 2
    sim.dt = sim.dtmin
    k = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 0.3 * budget_s and k < 200:
        sim.dyn_step()
        k += 1
    per_dyn = (time.perf_counter() - t0) / k
    per_qs = 2 * (stats["iters_per_solve"] * per_it + 2 * spmv)
    cpu_s = stats["n_qs"] * per_qs + stats["n_dyn"] * per_dyn
    return {
        "value": window_years / (cpu_s / 3600.0), "unit": "sim-years/hour", "cores": 1, "kind": "port",
        "sample": (f"oracle (C FP64 restatement: assembled CSR K, RS-AMG + symmetric Gauss-Seidel V-cycle, "
                   f"cg!, serial FBC!/NRsearch!) on the same mesh: {m} AMG-PCG iterations "
                   f"({per_it * 1e3:.1f} ms each) and {k} Newmark steps ({per_dyn * 1e3:.1f} ms each), "
                   f"scaled to the GPU plain window's {stats['n_qs']} QS steps x 2 solves x "
                   f"{stats['iters_per_solve']:.2f} iterations (the GPU's count, favourable to the CPU) and "
                   f"{stats['n_dyn']} dynamic steps; set-up {t_setup:.1f} s not counted"),
    }


def window(ctx, cfg, tol, prec, guess, warmup, steps):
    """State init, `warmup` plain years (untimed), then `steps` timed years."""
    import numpy as np
    ctx.set_solver(tol=tol, pc="amg", prec="fp64", guess="previous")
    ctx.state_init()
    if warmup > 0:
        ctx.run(total_time=warmup * YR, snapshots=False)
    ctx.set_solver(tol=tol, pc="amg", prec=prec, guess=guess)
    t_start = ctx.state_get()["t"]
    w0 = time.perf_counter()
    R = ctx.run(total_time=t_start + steps * YR, snapshots=False)
    wall = time.perf_counter() - w0
    s = ctx.state_get()
    years = (s["t"] - t_start) / YR
    it = R["pcg_iters"]
    qs = R["isolver"] == 1
    st = {
        "sim_years_per_hour": years / (wall / 3600.0), "sim_years": years, "wall_s": wall,
        "qs_steps": int(R["n_qs"]), "dyn_steps": int(R["n_dyn"]),
        "pcg_iters_per_solve": R["total_pcg_iters"] / max(1, 2 * R["n_qs"]),
        "pcg_iters_pass1": float(np.mean(it[qs, 0])) if qs.any() else 0.0,
        "pcg_iters_pass2": float(np.mean(it[qs, 1])) if qs.any() else 0.0,
        "wall_qs_s": R["wall_qs"], "wall_dyn_s": R["wall_dyn"], "events": int(len(R["tStart"])),
        "prec": prec, "guess": guess,
    }
    return st, s, R


def train_learned(ctx, cfg, tol, epochs):
    """Record cfg.n_snapshots pass-1 snapshots with the plain solver from the
    seeded initial state, fit the POD basis, train the network."""
    from eqcycle_amd import learned
    t0 = time.perf_counter()
    ctx.set_solver(tol=tol, pc="amg", prec="fp64", guess="previous")
    ctx.state_init()
    ctx.learn_init(capacity=cfg.n_snapshots, nbasis=64)
    ctx.learn_record(True)
    while ctx.learn_info()["nsnap"] < cfg.n_snapshots:
        ctx.run(total_time=1e30, max_steps=400, snapshots=False)
    ctx.learn_record(False)
    t_rec = time.perf_counter() - t0
    years = ctx.state_get()["t"] / YR
    t0 = time.perf_counter()
    fit = learned.fit_basis(ctx, tol=tol)
    t_pod = time.perf_counter() - t0
    t0 = time.perf_counter()
    hist = learned.train(ctx, epochs=epochs, batch=64, lr=1e-3)
    t_train = time.perf_counter() - t0
    return {"snapshots": cfg.n_snapshots, "used": fit["used"], "record_s": t_rec, "sim_years_recorded": years,
            "pod_s": t_pod, "energy_captured": fit["energy_captured"], "train_s": t_train, "epochs": epochs,
            "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
            "nparams": ctx.learn_info()["nparams"]}


def roofline(ctx, P, prec):
    """The Newmark step, the kernel with the largest share of device time in
    this window (one launch per dynamic step, ~20,000 steps against ~10,000
    Chebyshev pair launches): per node it reads the predicted displacement
    (the step's new d), v, a, M, mu and writes v, a and the next step's
    predictor -- 8 FP64 values (the LDS halo and the fault-row stencil reads
    are not counted).  Timed with HIP events on the library's stream."""
    dt_ms = ctx.time_kernel("newmark", "fp64", reps=200)
    alg = 8 * 8 * P.nglob
    achieved = alg / (dt_ms * 1e-3) / 1e9
    return {"kernel": "newmark_kernel (whole Newmark step from the carried predictor: -K d, absorbing edges, "
                      "FBC! on the fault row in blocks of their own, corrector, next predictor)",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": NEWMARK_TRAFFIC.get(P.nglob), "avg_launch_us": dt_ms * 1e3,
            "alg_bytes_per_launch": alg}


def roofline_cheb_pair(ctx, P, prec):
    """Two fine-level Chebyshev steps in one launch (the quasi-static solve's
    largest kernel), in the V-cycle precision of the value's variant.  The
    algorithmic bytes are those of one step, 8 values of s bytes per node
    (window vector, mu, D^-1, r, x in; r, d, x out): the first step's outputs
    never leave the chip."""
    kp = "fp32" if prec in ("f32pc", "mixed", "fp32") else "fp64"
    s = 4 if kp == "fp32" else 8
    dt_ms = ctx.time_kernel("cheb2", kp, reps=200)
    alg = 8 * s * P.nglob
    achieved = alg / (dt_ms * 1e-3) / 1e9
    return {"kernel": f"cheb_pair_kernel<{'float' if s == 4 else 'double'}>", "bound": "hbm",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": CHEB_PAIR_TRAFFIC.get((P.nglob, kp)), "avg_launch_us": dt_ms * 1e3,
            "alg_bytes_per_launch": alg}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}): the library drives one device, and the "
              "decomposition over several GPUs is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    from eqcycle_amd.api import EqContext
    from eqcycle_amd.configs import CONFIGS
    from eqcycle_amd.params import set_parameters

    cfg = CONFIGS[a.config]
    tol = a.tol or cfg.pcg_tol
    P = set_parameters(cfg)
    ctx = EqContext(P, amg={"hierarchy": a.hierarchy, "degree": a.degree})
    learned = cfg.learned and not a.plain_only

    plain, s_plain, R_plain = window(ctx, cfg, tol, "fp64", "previous", a.warmup, a.steps)
    extra = {"plain_fp64": plain}
    main_st, prec = plain, "fp64"
    if learned:
        extra["training"] = train_learned(ctx, cfg, tol, a.epochs)
        lrn, s_lrn, R_lrn = window(ctx, cfg, tol, cfg.precision, "learned", a.warmup, a.steps)
        gal, s_gal, R_gal = window(ctx, cfg, tol, cfg.precision, "galerkin", a.warmup, a.steps)
        extra["network_f32"] = lrn
        extra["galerkin_f32"] = gal
        main_st, prec = gal, cfg.precision

        def versus(st, s_x, R_x):
            dref = np.max(np.abs(s_plain["d"]))
            nt = min(len(R_x["Vfmax"]), len(R_plain["Vfmax"]))
            return {
                "speedup": st["sim_years_per_hour"] / plain["sim_years_per_hour"],
                "iters_ratio": st["pcg_iters_per_solve"] / max(plain["pcg_iters_per_solve"], 1e-30),
                "d_max_rel_diff": float(np.max(np.abs(s_x["d"] - s_plain["d"])) / dref),
                "t_end_rel_diff": abs(s_x["t"] - s_plain["t"]) / max(s_plain["t"], 1e-30),
                "event_starts": [float(x) for x in R_x["tStart"]],
                "log10_Vfmax_max_abs_diff_common_steps": float(np.max(np.abs(
                    np.log10(R_x["Vfmax"][:nt]) - np.log10(R_plain["Vfmax"][:nt])))) if nt else None,
            }
        extra["plain_event_starts"] = [float(x) for x in R_plain["tStart"]]
        extra["galerkin_vs_plain"] = versus(gal, s_gal, R_gal)
        extra["network_vs_plain"] = versus(lrn, s_lrn, R_lrn)
    years = main_st["sim_years"]
    out = {
        "metric": "sim-years/hour", "value": main_st["sim_years_per_hour"], "unit": "sim-years/hour",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": main_st["wall_s"] * 1e3 / max(years, 1e-30),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == "fp64" else "f64+f32",
        "data": f"synthetic (seeded fault model, {a.config}; network trained on the plain solver's own snapshots)"
                if learned else f"synthetic (seeded fault model, {a.config})",
        "config": {"workload": f"{a.config}: {cfg.nelx}x{cfg.nely} elements NGLL 5 ({P.nglob} nodes), "
                               f"{cfg.material} damage zone, CG tol {tol:g}, "
                               + ("x0 from the learned POD basis (Galerkin coefficients) + FP32 V-cycle; "
                                  "plain FP64 and network-predicted x0 alongside" if learned else "plain FP64"),
                   "pcg_iters_per_solve": main_st["pcg_iters_per_solve"], "hierarchy": a.hierarchy,
                   "amg_levels": ctx.amg_info(), **extra},
        "roofline": roofline(ctx, P, prec),
        "roofline_cheb_pair": roofline_cheb_pair(ctx, P, prec),
    }
    if not a.no_cpu_baseline:
        stats = dict(n_qs=plain["qs_steps"], n_dyn=plain["dyn_steps"], iters_per_solve=plain["pcg_iters_per_solve"])
        out["cpu_baseline"] = cpu_baseline(P, tol, stats, plain["sim_years"], a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

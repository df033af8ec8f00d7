"""Benchmark: simulated years per hour and CG iterations per solve of the
earthquake-cycle time loop on one MI355X, plain FP64 against learned initial
guess + FP32 V-cycle at the same relative residual.

Default workload (config 4): 160 x 160 elements, NGLL 5 (410,881 nodes),
trapezoidal damage zone, p-multigrid preconditioned CG to 1e-6 relative
residual, the whole time loop (quasi-static steps, Newmark steps, rate-state
Newton searches, adaptive dt, event bookkeeping) resident on the GPU.

One bench step is one full earthquake cycle.  The plain solver runs the seeded
model from t = 0 (recording the training snapshots of the learned start on
the way) to year 110, past the second event cluster (4.15 and 101.87 years on
config 4); its state there is the checkpoint.  A step restores the checkpoint
(eq_state_set, untimed) and runs the loop over the next 100 simulated years:
interseismic loading, nucleation and the third event cluster.  Every step
repeats the same work bit for bit.  --warmup W untimed steps, then exactly
--steps K timed steps per variant; each step's loop returns after its last
host synchronisation, and the timed region is the sum of the K loop runs.

  plain   : FP64 CG, FP64 V-cycle, previous-step initial guess (src/main.jl:241)
  galerkin: FP64 CG, FP32 V-cycle, start x + Phi c on the POD basis Phi learned
            from the plain solver's own quasi-static snapshots of the first
            cycle, c the Galerkin (K_II-norm optimal) coefficients, both passes
  network : FP64 CG, FP32 V-cycle, pass-1 start from the conv1d + MLP network's
            predicted POD coefficients with an energy-optimal step
  linear  : as network, with the coefficients from the linear least-squares
            map of the same input (the bar the network has to beat)
  (snapshot recording, POD and network training run beforehand, untimed and
  reported)

`value` is the galerkin variant's sim-years/hour (the plain one's for configs
without a learned basis).  `roofline` is the kernel with the largest device
time in one further pass of the value's variant with every launch of the
fine-level kernels bracketed by HIP events on the library's stream
(eq_launch_timing); the shares of all of them are under config.kernel_time.
`cpu_baseline` times the oracle (C FP64 restatement of the reference: CSR K,
Ruge-Stueben AMG-PCG, serial Newton searches) on a bounded sample of the same
cycle's quasi-static and dynamic steps, on all host cores and on one.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

YR = 365 * 24 * 60 * 60
HBM_PEAK_GBS = 8000.0
CHECKPOINT_YEARS = 110.0
SEGMENT_YEARS = 100.0
# the Galerkin start's first pass from the last increment extrapolated (eq_solver_opts.extrapolate)
EXTRAPOLATE = True
# its expansion and start residual in one launch (eq_solver_opts.galerkin_split = 0)
GALERKIN_FUSED = True

# HBM bytes per launch from two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE)
# calibrated by a 512 MiB copy on the 410,881-node mesh, keyed by (kernel, nodes)
# (the tile sweeps: the mean of the pre- and post-smoothing sweep, as they alternate; FP32 post in
# the F32PC form)
TRAFFIC = {("newmark_kernel", 410881): 27.19e6,
           ("cheb_pair_kernel<double>", 410881): 28.07e6,
           ("cheb_pair_kernel<float>", 410881): 14.29e6,
           ("cheb_sweep_kernel<float>", 410881): 10.80e6,
           ("cheb_sweep_kernel<double>", 410881): 16.14e6}


def alg_bytes(kernel, nglob, prec, smoother="tiles"):
    """Algorithmic HBM bytes of one launch: the values per node that the
    kernel must read and write once (halo re-reads are not counted)."""
    s = 4 if "float" in kernel else 8
    # single Chebyshev steps: with tile sweeps only the residual after pre-smoothing
    # (x, mu, b in; r out), otherwise a whole step
    step = 4 * s if smoother == "tiles" else 8 * s
    per_node = {
        "newmark_kernel": 64,                    # predictor, v, a, M, mu in; v, a, next predictor out
        "cheb_pair_kernel<float>": 8 * s,        # window vector, mu, D^-1, r, x in; r, d, x out (one step's)
        "cheb_pair_kernel<double>": 8 * s,
        "tile_kernel<float, OpCheb>": step,
        "tile_kernel<double, OpCheb>": step,
        "tile_kernel<T, OpCgDir>": 5 * 8,        # c, p, mu in; p, q out (FP64 CG vectors)
        "cg_update_kernel<T>": 6 * 8 + (4 if prec == "f32pc" else 0),   # x, r, p, q in; x, r out (+ FP32 r)
        # a whole smoothing sweep per launch, pre and post taking turns (mean of the two):
        #   pre: b, mu, D^-1 in; x out.  post: x, b, mu, D^-1 in; c out; and with the FP32 V-cycle
        #   of f32pc the FP64 residual in and c widened to FP64 out (4 s + 8 + 8).
        "cheb_sweep_kernel<float>": (4 * s + (4 * s + 16 if prec == "f32pc" else 5 * s)) // 2,
        "cheb_sweep_kernel<double>": (4 * s + 5 * s) // 2,
    }[kernel]
    return per_node * nglob


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed earthquake cycles per variant")
    ap.add_argument("--warmup", type=int, default=1, help="untimed cycles before the timed ones")
    ap.add_argument("--config", default="cfg4")
    ap.add_argument("--tol", type=float, default=None)
    ap.add_argument("--hierarchy", default="pmg", help="pmg (default) | rs (Ruge-Stueben on K_II)")
    ap.add_argument("--degree", type=int, default=7,
                    help="fine-level Chebyshev degree of the learned-start variants (pmg)")
    ap.add_argument("--degree-plain", type=int, default=5,
                    help="fine-level Chebyshev degree of the plain FP64 variant and the checkpoint run (pmg)")
    ap.add_argument("--smoother", default="tiles",
                    help="fine-level smoother: tiles (one launch per Chebyshev sweep) | chebyshev | jacobi")
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--graphs", type=int, default=0,
                    help="replay CG batches as HIP graphs (1) or launch them eagerly (0, faster here)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU budget per thread setting")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain-only", action="store_true")
    ap.add_argument("--nbasis", type=int, default=16, help="POD basis vectors of the learned starts")
    ap.add_argument("--max-coarse", type=int, default=None,
                    help="largest coarsest level of the hierarchy (default: the library's, 2000 for pmg)")
    ap.add_argument("--galerkin-fused", type=int, default=1,
                    help="Galerkin expansion and start residual in one launch (1) or two (0)")
    ap.add_argument("--extrapolate", type=int, default=1,
                    help="Galerkin start from the last increment extrapolated (1) or from the last solution (0)")
    return ap.parse_args()


# ---------------------------------------------------------------------------
# GPU side
# ---------------------------------------------------------------------------
def checkpoint(ctx, cfg, tol, learned, epochs, graphs, nbasis=16):
    """The plain solver from the seeded state to CHECKPOINT_YEARS; with the
    learned start, its first cfg.n_snapshots pass-1 solves are the training
    snapshots (POD basis and network fitted afterwards)."""
    from eqcycle_amd import learned as L
    ctx.set_solver(tol=tol, pc="amg", prec="fp64", guess="previous", graphs=bool(graphs))
    ctx.state_init()
    info = {}
    t0 = time.perf_counter()
    if learned:
        ctx.learn_init(capacity=cfg.n_snapshots, nbasis=nbasis)
        ctx.learn_record(True)
        while ctx.learn_info()["nsnap"] < cfg.n_snapshots and ctx.state_get()["t"] < CHECKPOINT_YEARS * YR:
            ctx.run(total_time=CHECKPOINT_YEARS * YR, max_steps=400, snapshots=False)
        ctx.learn_record(False)
        info["record_s"] = time.perf_counter() - t0
        info["sim_years_recorded"] = ctx.state_get()["t"] / YR
    ctx.run(total_time=CHECKPOINT_YEARS * YR, snapshots=False)
    s = ctx.state_get()
    if learned:
        t1 = time.perf_counter()
        fit = L.fit_basis(ctx, tol=tol)
        t_pod = time.perf_counter() - t1
        t1 = time.perf_counter()
        hist = L.train(ctx, epochs=epochs, batch=64, lr=1e-3)
        info.update(snapshots=cfg.n_snapshots, used=fit["used"], pod_s=t_pod,
                    energy_captured=fit["energy_captured"], train_s=time.perf_counter() - t1, epochs=epochs,
                    loss_first=float(hist[0]), loss_last=float(hist[-1]), nparams=ctx.learn_info()["nparams"])
        t1 = time.perf_counter()
        lin = L.fit_linear(ctx, fit["use"])
        info.update(linear_fit_s=time.perf_counter() - t1, linear_train_rel_residual=lin["train_rel_residual"])
    return s, info


def one_pass(ctx, ck, t_end):
    ctx.state_set(d=ck["d"], v=ck["v"], a=ck["a"], psi=ck["psi"], tau=ck["tau"], t=ck["t"], dt=ck["dt"],
                  isolver=ck["isolver"])
    w0 = time.perf_counter()
    R = ctx.run(total_time=t_end, snapshots=False)
    return time.perf_counter() - w0, R


def variant(ctx, ck, tol, prec, guess, warmup, steps, graphs):
    """warmup + steps passes over the cycle; statistics of the timed ones."""
    import numpy as np
    ctx.set_solver(tol=tol, pc="amg", prec=prec, guess=guess, graphs=bool(graphs), basis_fp16=guess == "galerkin",
                   extrapolate=guess == "galerkin" and EXTRAPOLATE,
                   galerkin_fused=GALERKIN_FUSED)
    t_end = ck["t"] + SEGMENT_YEARS * YR
    for _ in range(warmup):
        one_pass(ctx, ck, t_end)
    walls, sig = [], None
    for _ in range(steps):
        w, R = one_pass(ctx, ck, t_end)
        walls.append(w)
        s = ctx.state_get()
        this = (R["n_qs"], R["n_dyn"], R["total_pcg_iters"], s["t"])
        if sig is not None and this != sig:
            raise RuntimeError(f"passes over the same cycle differ: {this} vs {sig}")
        sig = this
    years = (s["t"] - ck["t"]) / YR
    wall = float(sum(walls))
    it = R["pcg_iters"]
    qs = R["isolver"] == 1
    st = {
        "sim_years_per_hour": years * steps / (wall / 3600.0), "sim_years_per_pass": years,
        "wall_s": wall, "wall_per_pass_s": wall / steps, "wall_pass_min_max_s": [min(walls), max(walls)],
        "qs_steps": int(R["n_qs"]), "dyn_steps": int(R["n_dyn"]),
        "pcg_iters_per_solve": R["total_pcg_iters"] / max(1, 2 * R["n_qs"]),
        "pcg_iters_pass1": float(np.mean(it[qs, 0])) if qs.any() else 0.0,
        "pcg_iters_pass2": float(np.mean(it[qs, 1])) if qs.any() else 0.0,
        "wall_qs_s": R["wall_qs"], "wall_dyn_s": R["wall_dyn"], "events": int(len(R["tStart"])),
        "event_starts_yr": [float(x) / YR for x in R["tStart"]], "prec": prec, "guess": guess,
    }
    if st["dyn_steps"] == 0 or st["events"] == 0:
        raise RuntimeError(f"the timed cycle of {guess}/{prec} holds no event: {st}")
    return st, s, R


def kernel_time(ctx, ck, tol, prec, guess, P, wall_pass, smoother):
    """One more pass of the variant with every fine-level launch bracketed by
    HIP events (CG batches launch eagerly: same kernels); device time per
    kernel, the dominant one, and its roofline."""
    ctx.set_solver(tol=tol, pc="amg", prec=prec, guess=guess, graphs=False, basis_fp16=guess == "galerkin",
                   extrapolate=guess == "galerkin" and EXTRAPOLATE,
                   galerkin_fused=GALERKIN_FUSED)
    ctx.launch_timing(True)
    w, R = one_pass(ctx, ck, ck["t"] + SEGMENT_YEARS * YR)
    kt = ctx.launch_timing_get()
    ctx.launch_timing(False)
    table = {}
    for name, (n, ms) in kt.items():
        if n == 0:
            continue
        avg_us = ms * 1e3 / n
        alg = alg_bytes(name, P.nglob, prec, smoother)
        table[name] = {"launches": n, "total_ms": ms, "avg_us": avg_us, "share_of_pass_wall": ms / 1e3 / wall_pass,
                       "alg_bytes_per_launch": alg, "achieved_GBs": alg / (avg_us * 1e-6) / 1e9}
    dom = max(table, key=lambda k: table[k]["total_ms"])
    d = table[dom]
    roof = {"kernel": dom, "bound": "hbm", "achieved": d["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": d["achieved_GBs"] / HBM_PEAK_GBS, "traffic": TRAFFIC.get((dom, P.nglob)),
            "avg_launch_us": d["avg_us"], "alg_bytes_per_launch": d["alg_bytes_per_launch"],
            "launches_per_cycle": d["launches"],
            "measured": "HIP events around each launch on the library's stream, one pass over the bench cycle"}
    return roof, table, w


# ---------------------------------------------------------------------------
# CPU baseline: the oracle on a bounded sample of the same cycle
# ---------------------------------------------------------------------------
def host_info():
    cores, model = os.cpu_count(), ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except (OSError, subprocess.SubprocessError, ValueError):
        nproc = cores
    return {"nproc": nproc, "os_cpu_count": cores, "lscpu_model": model}


def cpu_baseline(P, tol, ck, dyn_state, mix, budget_s):
    """The oracle (C FP64 restatement of the reference path, its own
    Ruge-Stueben AMG-PCG, previous-step guess, the bench tolerance) on a
    bounded sample of the same cycle:
      * one whole quasi-static step from the checkpoint on all host threads:
        the oracle's own iteration counts of its two solves, and its time;
      * per setting (all threads, one thread): the time of one of those CG
        iterations (a few timed on the checkpoint's first system) and of one
        Newmark step from a state inside the cycle's event;
    then sim-years/hour = years / (n_qs t_qs + n_dyn t_dyn), with the GPU plain
    pass's step counts and t_qs = (its_1 + its_2) t_iter + 4 SpMV."""
    import numpy as np

    from oracle import oracle as O
    hi = host_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or hi["nproc"] or 1
    t0 = time.perf_counter()
    sim = O.Simulation(P, tol=tol, solver="amg")
    t_setup = time.perf_counter() - t0

    def load(st, isolver):
        sim.d, sim.v, sim.a = st["d"].copy(), st["v"].copy(), st["a"].copy()
        sim.psi = st["psi"].copy()
        sim.t, sim.isolver = st["t"], isolver
        if isolver == 1:
            sim.tau, sim.dt = st["tau"].copy(), st["dt"]
        else:
            sim.dt = sim.dtmin

    O.set_threads(threads)
    load(ck, 1)
    t0 = time.perf_counter()
    its = sim.qs_step()
    t_qs_whole = time.perf_counter() - t0
    # the checkpoint's first quasi-static system (src/main.jl:237-256)
    F = np.zeros(P.nglob)
    F[sim.iFlt] = ck["d"][sim.iFlt] + ck["v"][sim.iFlt] * ck["dt"]
    rhs = sim.K.matvec(F)[sim.FltNI]
    x0 = ck["d"][sim.FltNI].copy()
    legs = {}
    for th in (threads, 1):
        O.set_threads(th)
        m = 3
        t0 = time.perf_counter()
        O.cg(sim.kni, x0.copy(), rhs, sim.ml, 1e-300, maxiter=m)
        t_iter = (time.perf_counter() - t0) / m
        m = max(3, min(40, int(0.3 * budget_s / t_iter)))
        t0 = time.perf_counter()
        O.cg(sim.kni, x0.copy(), rhs, sim.ml, 1e-300, maxiter=m)
        t_iter = (time.perf_counter() - t0) / m
        t0 = time.perf_counter()
        for _ in range(10):
            sim.K.matvec(F)
        t_spmv = (time.perf_counter() - t0) / 10
        load(dyn_state, 2)
        nd = 0
        t0 = time.perf_counter()
        while nd < 20 or (time.perf_counter() - t0 < 0.3 * budget_s and nd < 1000):
            sim.dyn_step()
            nd += 1
        t_dyn = (time.perf_counter() - t0) / nd
        t_qs = (its[0] + its[1]) * t_iter + 4 * t_spmv
        cpu_s = mix["n_qs"] * t_qs + mix["n_dyn"] * t_dyn
        legs[th] = {"value": mix["years"] / (cpu_s / 3600.0), "threads": th, "cg_iterations_timed": m,
                    "per_cg_iteration_s": t_iter, "per_qs_step_s": t_qs, "dyn_steps_timed": nd,
                    "per_dyn_step_s": t_dyn}
    O.set_threads(1)
    multi, single = legs[threads], legs[1]
    return {
        "value": multi["value"], "unit": "sim-years/hour", "cores": threads, "kind": "port",
        "sample": (f"oracle (C FP64 restatement of the reference: assembled CSR K, Ruge-Stueben AMG + symmetric "
                   f"Gauss-Seidel V-cycle, cg!, serial FBC!/NRsearch!; OpenMP over SpMV and CG vector "
                   f"operations, Gauss-Seidel sequential as in the reference) at the bench checkpoint: one whole "
                   f"quasi-static step ({its[0]} + {its[1]} of its own AMG-PCG iterations, {t_qs_whole:.1f} s), "
                   f"{multi['cg_iterations_timed']} CG iterations ({multi['per_cg_iteration_s'] * 1e3:.1f} ms "
                   f"each) and {multi['dyn_steps_timed']} Newmark steps from inside the cycle's event "
                   f"({multi['per_dyn_step_s'] * 1e3:.2f} ms each), scaled to the GPU plain pass's {mix['n_qs']} "
                   f"quasi-static and {mix['n_dyn']} dynamic steps over {mix['years']:.1f} years; oracle set-up "
                   f"{t_setup:.1f} s not counted"),
        "oracle_pcg_iters_qs_step": list(its), "qs_step_whole_s": t_qs_whole,
        "single_thread": single, "multi_thread": multi, "host": hi,
    }


def main():
    a = parse()
    global EXTRAPOLATE, GALERKIN_FUSED
    EXTRAPOLATE = bool(a.extrapolate)
    GALERKIN_FUSED = bool(a.galerkin_fused)
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
    amg_plain = {"hierarchy": a.hierarchy, "degree": a.degree_plain, "smoother": a.smoother,
                 "max_coarse": a.max_coarse}
    ctx = EqContext(P, amg=amg_plain)
    learned = cfg.learned and not a.plain_only

    ck, training = checkpoint(ctx, cfg, tol, learned, a.epochs, a.graphs, a.nbasis)
    print(f"checkpoint at {ck['t'] / YR:.3f} years", file=sys.stderr, flush=True)
    plain, s_plain, R_plain = variant(ctx, ck, tol, "fp64", "previous", a.warmup, a.steps, a.graphs)
    print(f"plain: {plain['sim_years_per_hour']:.0f} sim-years/hour", file=sys.stderr, flush=True)
    extra = {"plain_fp64": plain}
    main_st, prec, guess = plain, "fp64", "previous"
    if learned:
        extra["training"] = training
        # each variant with its best fine-level degree (set-up outside the timed passes)
        ctx.amg_setup(hierarchy=a.hierarchy, degree=a.degree, smoother=a.smoother, max_coarse=a.max_coarse)
        lrn, s_lrn, R_lrn = variant(ctx, ck, tol, cfg.precision, "learned", a.warmup, a.steps, a.graphs)
        print(f"network: {lrn['sim_years_per_hour']:.0f}", file=sys.stderr, flush=True)
        lin, s_lin, R_lin = variant(ctx, ck, tol, cfg.precision, "linear", a.warmup, a.steps, a.graphs)
        print(f"linear: {lin['sim_years_per_hour']:.0f}", file=sys.stderr, flush=True)
        gal, s_gal, R_gal = variant(ctx, ck, tol, cfg.precision, "galerkin", a.warmup, a.steps, a.graphs)
        print(f"galerkin: {gal['sim_years_per_hour']:.0f}", file=sys.stderr, flush=True)
        extra["network_f32"] = lrn
        extra["linear_f32"] = lin
        extra["galerkin_f32"] = gal
        main_st, prec, guess = gal, cfg.precision, "galerkin"

        def versus(st, s_x, R_x):
            dref = np.max(np.abs(s_plain["d"]))
            ev = [abs(x - y) / y for x, y in zip(R_x["tStart"], R_plain["tStart"])]
            return {
                "speedup": st["sim_years_per_hour"] / plain["sim_years_per_hour"],
                "iters_ratio": st["pcg_iters_per_solve"] / max(plain["pcg_iters_per_solve"], 1e-30),
                "d_max_rel_diff_end": float(np.max(np.abs(s_x["d"] - s_plain["d"])) / dref),
                "event_start_max_rel_diff": float(max(ev)) if ev else None,
                "event_count": [len(R_x["tStart"]), len(R_plain["tStart"])],
            }
        extra["galerkin_vs_plain"] = versus(gal, s_gal, R_gal)
        extra["network_vs_plain"] = versus(lrn, s_lrn, R_lrn)
        extra["linear_vs_plain"] = versus(lin, s_lin, R_lin)
    roof, ktable, w_acc = kernel_time(ctx, ck, tol, prec, guess, P, main_st["wall_per_pass_s"], a.smoother)
    roof["newmark_isolated_us"] = ctx.time_kernel("newmark", "fp64", reps=200) * 1e3
    print(f"dominant kernel {roof['kernel']}: {roof['frac']:.3f} of HBM peak", file=sys.stderr, flush=True)
    out = {
        "metric": "sim-years/hour", "value": main_st["sim_years_per_hour"], "unit": "sim-years/hour",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": main_st["wall_per_pass_s"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == "fp64" else "f64+f32",
        "data": (f"synthetic (seeded fault model, {a.config}; POD basis and network learned from the plain "
                 f"solver's own snapshots of the first cycle)" if learned
                 else f"synthetic (seeded fault model, {a.config})"),
        "config": {"workload": (f"{a.config}: {cfg.nelx}x{cfg.nely} elements NGLL 5 ({P.nglob} nodes), "
                                f"{cfg.material} damage zone, CG tol {tol:g}; one step = one earthquake cycle "
                                f"(years {ck['t'] / YR:.1f} to {ck['t'] / YR + SEGMENT_YEARS:.1f} from the plain "
                                f"solver's checkpoint, third event cluster inside); "
                                + ("value: x0 from the learned POD basis (Galerkin coefficients, expansion from "
                                   "its FP16 copy) + FP32 V-cycle, "
                                   "plain FP64, network-predicted and linear-predicted x0 alongside" if learned
                                   else "plain FP64")),
                   "pcg_iters_per_solve": main_st["pcg_iters_per_solve"], "hierarchy": a.hierarchy, "smoother": a.smoother,
                   "chebyshev_degree": {"plain": a.degree_plain, "learned_starts": a.degree if learned else None},
                   "graphs": bool(a.graphs), "amg_levels": ctx.amg_info(), **extra,
                   "kernel_time": {"variant": f"{guess}/{prec}", "pass_wall_s_with_events": w_acc,
                                   "kernels": ktable}},
        "roofline": roof,
    }
    if not a.no_cpu_baseline:
        # a state inside the cycle's event for the oracle's Newmark steps
        i_dyn = int(np.argmax(R_plain["isolver"] == 2))
        if learned:
            ctx.amg_setup(**amg_plain)
        ctx.set_solver(tol=tol, pc="amg", prec="fp64", guess="previous", graphs=bool(a.graphs))
        ctx.state_set(d=ck["d"], v=ck["v"], a=ck["a"], psi=ck["psi"], tau=ck["tau"], t=ck["t"], dt=ck["dt"],
                      isolver=ck["isolver"])
        ctx.run(total_time=ck["t"] + SEGMENT_YEARS * YR, max_steps=i_dyn + 2000, snapshots=False)
        dyn_state = ctx.state_get()
        mix = dict(n_qs=plain["qs_steps"], n_dyn=plain["dyn_steps"], years=plain["sim_years_per_pass"])
        out["cpu_baseline"] = cpu_baseline(P, tol, ck, dyn_state, mix, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Record a configuration's training snapshots, then time the POD set-up, for profilers.

    python -m eqcycle_amd.profile_pod [--config cfg4] [--reps 3]

Runs the plain FP64 solver from the seeded state until cfg.n_snapshots pass-1
solves are recorded (config 4: 2,000 snapshots of 410,240 interior nodes),
then --reps times the snapshot Gram S^T S (gram_kernel on the matrix cores)
and once the whole POD fit (Gram, host eigendecomposition, basis and its
Gram for the Cholesky-QR pass), then (--epochs > 0) trains the network that
many epochs (batches of 64), printing wall times.  CG batches launch one
kernel at a time (no HIP stream capture), so the program can run under
rocprofv3 --kernel-trace.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="cfg4")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--epochs", type=int, default=0, help="> 0: then train the network this many epochs")
    a = ap.parse_args()
    from eqcycle_amd import learned
    from eqcycle_amd.api import EqContext
    from eqcycle_amd.configs import CONFIGS
    from eqcycle_amd.params import set_parameters

    cfg = CONFIGS[a.config]
    P = set_parameters(cfg)
    ctx = EqContext(P, amg={"hierarchy": "pmg", "degree": 5})
    ctx.set_solver(tol=cfg.pcg_tol, pc="amg", prec="fp64", guess="previous", graphs=False)
    ctx.state_init()
    ctx.learn_init(capacity=cfg.n_snapshots, nbasis=64)
    ctx.learn_record(True)
    while ctx.learn_info()["nsnap"] < cfg.n_snapshots:
        ctx.run(total_time=1e30, max_steps=400, snapshots=False)
    ctx.learn_record(False)
    m = ctx.learn_info()["nsnap"]
    print(f"[profile_pod] {m} snapshots of {ctx.nI} interior nodes", flush=True)
    walls = []
    for _ in range(a.reps):
        t = time.perf_counter()
        ctx.learn_gram()
        walls.append(time.perf_counter() - t)
    t = time.perf_counter()
    fit = learned.fit_basis(ctx, tol=cfg.pcg_tol)
    out = {"config": a.config, "snapshots": m, "nI": ctx.nI, "gram_wall_s": walls,
           "fit_basis_wall_s": time.perf_counter() - t, "energy_captured": fit["energy_captured"]}
    if a.epochs > 0:
        t = time.perf_counter()
        hist = learned.train(ctx, epochs=a.epochs, batch=64, lr=1e-3)
        out.update(train_epochs=a.epochs, train_wall_s=time.perf_counter() - t, loss_first=float(hist[0]),
                   loss_last=float(hist[-1]))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

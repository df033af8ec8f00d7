"""Launch each hot kernel of the library a fixed number of times, for profilers.

    python -m eqcycle_amd.profile_kernels [--config cfg2] [--reps 20] [--kernels cheb,ku,...]

Kernels: ku (K u), cheb (one fine-level Chebyshev step), cheb2 (two steps in
one launch), sweep / sweep_post (a whole tile sweep before / after the coarse
correction), restrict (residual restricted to the element vertices), cg_dir,
vcycle, cg_iter, newmark.  Each kernel runs on the
library's stream through eq_time_kernel (one warm-up launch, then --reps timed
launches on well-defined inputs), and the average
launch time measured with HIP events is printed.  A device-to-device copy of
512 MiB (src and dst each that size, the library's convert_kernel<double,
double>) runs --reps times at the end: the bytes it must move are known, which
calibrates the FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc run of
this program.
"""
from __future__ import annotations

import argparse
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

KERNELS = ("ku", "cheb", "cheb2", "sweep", "sweep_post", "restrict", "cg_dir", "vcycle", "cg_iter", "newmark")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--kernels", default="cheb,ku,newmark")
    ap.add_argument("--prec", default="fp64")
    ap.add_argument("--copy", type=int, default=1, help="1: also time the 512 MiB calibration copy")
    ap.add_argument("--smoother", default="tiles", help="fine-level smoother: tiles | chebyshev")
    ap.add_argument("--degree", type=int, default=5, help="fine-level Chebyshev degree")
    ap.add_argument("--cheb-pairs", type=int, default=1,
                    help="V-cycle and CG iteration with two Chebyshev steps per launch (0: one per launch)")
    a = ap.parse_args()
    from eqcycle_amd.api import EqContext
    from eqcycle_amd.configs import CONFIGS
    from eqcycle_amd.params import set_parameters

    cfg = CONFIGS[a.config]
    P = set_parameters(cfg)
    ctx = EqContext(P, amg={"hierarchy": "pmg", "degree": a.degree, "smoother": a.smoother})
    ctx.set_solver(tol=cfg.pcg_tol, pc="amg", prec="fp64", cheb_pairs=bool(a.cheb_pairs))
    ctx.state_init()
    out = {"config": a.config, "nglob": P.nglob, "prec": a.prec, "reps": a.reps, "smoother": a.smoother,
           "degree": a.degree, "avg_us": {}}
    for k in a.kernels.split(","):
        if k not in KERNELS:
            raise SystemExit(f"unknown kernel {k!r}; choose from {KERNELS}")
        ms = ctx.time_kernel(k, a.prec if k != "newmark" else "fp64", reps=a.reps)
        out["avg_us"][k] = ms * 1e3
        print(f"[profile_kernels] {k}: {ms * 1e3:.2f} us/launch", flush=True)
    if a.copy:
        # 512 MiB read + 512 MiB written per launch
        ms = ctx.time_kernel("copy", "fp64", reps=a.reps)
        nb = 512 << 20
        out["copy"] = {"bytes_read": nb, "bytes_written": nb, "avg_us": ms * 1e3, "GBps": 2 * nb / (ms * 1e-3) / 1e9}
        print(f"[profile_kernels] copy 512 MiB: {ms * 1e3:.1f} us/launch, {out['copy']['GBps']:.0f} GB/s", flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

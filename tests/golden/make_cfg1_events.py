"""Write tests/golden/cfg1_oracle_events.json: the event catalogue of the
oracle's main(P) run on config 1 (40 x 40 elements, 25,921 nodes, 200
simulated years), quasi-static steps solved with a sparse direct solver.

Recorded: every tStart / tEnd / hypo (src/main.jl:361-381), the step count,
the final time, the fault row of the final displacement, and per completed
event the catalogue of post/earthquake_cycles.jl (moment magnitude, mean
stress drop, mean coseismic slip; eqcycle_amd.catalogue, itself pinned by
closed forms in tests/test_catalogue_cpu.py) from the run's taubefore,
tauafter and delfafter.  Takes about 5 minutes on one core.

    python tests/golden/make_cfg1_events.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from eqcycle_amd.catalogue import catalogue  # noqa: E402
from eqcycle_amd.configs import CONFIGS  # noqa: E402
from eqcycle_amd.params import set_parameters  # noqa: E402
from oracle import oracle as O  # noqa: E402

YR = 365 * 24 * 60 * 60


def main():
    O.build()
    P = set_parameters(CONFIGS["cfg1"])
    sim = O.Simulation(P, solver="direct")
    r = sim.run(total_time=200 * YR)
    iflt = P.iFlt - 1
    cfg = CONFIGS["cfg1"]
    ne = len(r.tEnd)
    res = {"tStart": r.tStart[:ne], "tEnd": r.tEnd, "hypo": r.hypo[:ne], "taubefore": r.taubefore[:ne],
           "tauafter": r.tauafter, "delfafter": r.delfafter}
    cat = catalogue(res, P.FltX, cfg.rho * cfg.vs ** 2)
    out = {
        "config": "cfg1", "total_years": 200, "solver": "direct (scipy splu) on K_II",
        "n_steps": int(sim.it), "t_end": float(sim.t),
        "tStart": [float(x) for x in r.tStart], "tEnd": [float(x) for x in r.tEnd],
        "hypo": [float(x) for x in r.hypo],
        "catalogue": [{k: e[k] for k in ("hypo_m", "Mw", "stress_drop", "mean_slip_m")} for e in cat],
        "d_fault": [float(x) for x in sim.d[iflt]],
    }
    with open(os.path.join(HERE, "cfg1_oracle_events.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()

"""Write tests/golden/cfg1_oracle_events.json: the event catalogue of the
oracle's main(P) run on config 1 (40 x 40 elements, 25,921 nodes, 200
simulated years), quasi-static steps solved with a sparse direct solver.

Recorded: every tStart / tEnd (src/main.jl:361-381), the step count, the final
time and the fault row of the final displacement.  Takes about 5 minutes on
one core.

    python tests/golden/make_cfg1_events.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

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
    out = {
        "config": "cfg1", "total_years": 200, "solver": "direct (scipy splu) on K_II",
        "n_steps": int(sim.it), "t_end": float(sim.t),
        "tStart": [float(x) for x in r.tStart], "tEnd": [float(x) for x in r.tEnd],
        "d_fault": [float(x) for x in sim.d[iflt]],
    }
    with open(os.path.join(HERE, "cfg1_oracle_events.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()

"""Run output on disk: the reference's `results` struct, the simulation time
and the parameters, in one compressed NumPy archive.

The reference serialises `O, simulation_time, P` with Julia's Serialization
(run.jl:41-46) and reads them back in that order (output.jl, analyze_results.jl:17-22).
``save`` writes the same three items to an ``.npz`` (no pickles: plain arrays
and scalars only) and ``load`` returns them in the same order, with the
`results` fields (src/main.jl:479-498) under the reference's names and
orientation: fault matrices [FltNglob x n] (one column per snapshot or
event), receiver traces [n x receivers].
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np

# struct results (src/main.jl:479-498)
RESULT_FIELDS = ("seismic_stress", "seismic_slipvel", "seismic_slip", "index_eq", "is_stress", "is_slipvel",
                 "is_slip", "dSeis", "vSeis", "aSeis", "tStart", "tEnd", "taubefore", "tauafter", "delfafter",
                 "hypo", "time_", "Vfmax")
FAULT_MATRICES = ("seismic_stress", "seismic_slipvel", "seismic_slip", "is_stress", "is_slipvel", "is_slip",
                  "taubefore", "tauafter", "delfafter")
PARAM_FIELDS = ("FltNglob", "nglob", "NFBC", "Vpl", "Vthres", "Vevne", "dt0", "yr2sec", "Total_time", "FltX",
                "iFlt", "out_seis", "xout", "yout", "fo", "Vo", "xLf", "cca", "ccb", "Seff", "tauo", "XiLf")


def save(path, results, P, simulation_time):
    """Write results (dict from EqContext.run, fault matrices [n x FltNglob]),
    simulation_time (s) and the parameters of P to ``path`` (.npz)."""
    F = int(P.FltNglob)
    arrays = {}
    for k in RESULT_FIELDS:
        if k == "index_eq":
            arrays["O_index_eq"] = np.zeros(0)   # not filled by the reference either (src/main.jl:390-416)
            continue
        if k not in results:
            continue
        a = np.asarray(results[k])
        if k in FAULT_MATRICES:
            a = a.reshape(-1, F).T                # -> [FltNglob x n]
        arrays["O_" + k] = a
    arrays["simulation_time"] = np.float64(simulation_time)
    for k in PARAM_FIELDS:
        if hasattr(P, k):
            arrays["P_" + k] = np.asarray(getattr(P, k))
    np.savez_compressed(path, **arrays)


def load(path):
    """(O, simulation_time, P) as analyze_results.jl:17-22 reads them: O has the
    `results` fields, P the saved parameters."""
    with np.load(path, allow_pickle=False) as z:
        O = SimpleNamespace(**{k[2:]: z[k] for k in z.files if k.startswith("O_")})
        P = SimpleNamespace(**{k[2:]: (z[k].item() if z[k].ndim == 0 else z[k]) for k in z.files
                               if k.startswith("P_")})
        sim_time = float(z["simulation_time"])
    return O, sim_time, P

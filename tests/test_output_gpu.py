"""A GPU run written to the output archive and read back (run.jl:41-46,
output.jl), and its event catalogue (post/earthquake_cycles.jl) from either
copy."""
import numpy as np
import pytest

from eqcycle_amd import catalogue as C
from eqcycle_amd import output
from eqcycle_amd.configs import small
from eqcycle_amd.params import set_parameters

pytestmark = pytest.mark.gpu


def test_run_archive_round_trip_and_catalogue(tmp_path):
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(32, 8, material="trapezoid", total_years=3))
    ctx = EqContext(P, amg=True)
    ctx.set_solver(tol=1e-8, pc="amg")
    ctx.state_init()
    R = ctx.run(snapshots=True)
    print(f"[gpu] run: {R['n_steps']} steps, {len(R['tStart'])} starts, {len(R['tEnd'])} ends", flush=True)
    assert len(R["tEnd"]) >= 1
    path = tmp_path / "run.npz"
    output.save(path, R, P, simulation_time=12.5)
    O, sim_time, P2 = output.load(path)
    assert sim_time == 12.5
    F = P.FltNglob
    for k in output.FAULT_MATRICES:
        a = getattr(O, k)
        assert a.shape == (F, len(R[k])), k                       # [FltNglob x n], as results stores them
        assert np.array_equal(a, np.asarray(R[k]).reshape(-1, F).T), k
    for k in ("tStart", "tEnd", "hypo", "time_", "Vfmax", "dSeis", "vSeis", "aSeis"):
        assert np.array_equal(getattr(O, k), R[k]), k
    assert np.array_equal(P2.FltX, P.FltX) and P2.FltNglob == F
    mu = P.cfg.rho * P.cfg.vs ** 2
    cat_run = C.catalogue(R, P.FltX, mu)
    cat_file = C.catalogue(vars(O), P2.FltX, mu)
    assert cat_run == cat_file
    assert len(cat_run) == len(R["tEnd"])
    for e in cat_run:
        assert np.isfinite(e["Mw"]) and e["t_end_yr"] > e["t_start_yr"] and e["mean_slip_m"] > 0
    # each event's initial stress has a nearest seismic snapshot (get_index),
    # close to it: both are taken at the event's first steps
    idx = C.get_index(O.seismic_stress, O.taubefore)
    assert len(idx) == O.taubefore.shape[1] and idx.min() >= 0 and idx.max() < O.seismic_stress.shape[1]
    for i, k in enumerate(idx):
        d = np.linalg.norm(O.seismic_stress[:, k] - O.taubefore[:, i])
        assert d <= 1e-2 * np.linalg.norm(O.taubefore[:, i])

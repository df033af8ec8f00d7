"""Event catalogue and output archive (post/earthquake_cycles.jl, run.jl:41-46, output.jl)."""
import numpy as np
import pytest

from eqcycle_amd import catalogue as C
from eqcycle_amd import output
from eqcycle_amd.configs import small
from eqcycle_amd.params import set_parameters


def synthetic_series():
    """Two events on a 10-node fault: slip rates above 1.01e-3 in snapshots
    2-4 (peak at node 6) and 7-8 (peak at node 3)."""
    F, n = 10, 11
    FltX = -np.linspace(9000.0, 0.0, F)            # deep to shallow
    vel = np.full((F, n), 1e-9)
    vel[6, 2:5] = [2e-3, 5e-2, 2e-3]
    vel[3, 7:9] = [1e-1, 3e-3]
    slip = np.cumsum(vel * 10.0, axis=1)
    stress = np.tile(np.linspace(30.0, 31.0, n), (F, 1))
    time_ = np.arange(n) * 10.0
    return FltX, slip, vel, stress, time_


def test_coslip_events_hypocentres_and_slip():
    FltX, slip, vel, stress, time_ = synthetic_series()
    D, drops, tS, tE, vh, hy = C.coslip(FltX, slip, vel, stress, time_)
    assert np.allclose(tS, [20.0, 70.0]) and np.allclose(tE, [50.0, 90.0])
    assert np.allclose(hy, [FltX[6], FltX[3]]) and np.allclose(vh, [2e-3, 1e-1])
    assert np.allclose(D[:, 0], slip[:, 5] - slip[:, 2]) and np.allclose(D[:, 1], slip[:, 9] - slip[:, 7])
    assert np.allclose(drops[:, 0], stress[:, 2] - stress[:, 5])


def test_moment_magnitude_uniform_patch():
    """Uniform slip s on k nodes spaced h: M0 = mu (s k h)(k h); Mw from the
    Hanks-Kanamori relation in N m, Mw = 2/3 log10(M0) - 6.0333."""
    F, h, s, mu, k = 30, 100.0, 2.0, 3.2e10, 12
    FltX = -h * np.arange(F - 1, -1, -1, dtype=float)
    D = np.zeros((F, 1))
    D[10:10 + k, 0] = s
    D[5, 0] = 0.001 * s                           # below the 1% threshold
    drops = np.full((F, 1), 3.0)
    Mw, dsig, fslip = C.moment_magnitude(mu, F, FltX, D, drops)
    M0 = mu * (s * k * h) * (k * h)
    assert abs(Mw[0] - (2.0 / 3.0 * np.log10(M0) - 6.0333333333)) < 1e-9
    assert abs(dsig[0] - 3.0) < 1e-12 and abs(fslip[0] - s * k / (k * h)) < 1e-12
    D2 = D.copy()
    D2[0, 0] = s                                  # the deepest node has no interval below it
    with pytest.raises(IndexError):
        C.moment_magnitude(mu, F, FltX, D2, drops)


def test_get_index_nearest_snapshot():
    S = np.array([[0.0, 1.0, 2.0, 3.0], [0.0, 1.0, 2.0, 3.0]])
    T = np.array([[1.9, 0.2], [2.1, -0.1]])
    assert list(C.get_index(S, T)) == [2, 0]


def test_output_round_trip(tmp_path):
    P = set_parameters(small(4, 4))
    F = P.FltNglob
    rng = np.random.default_rng(3)
    R = {"time_": np.arange(5.0), "Vfmax": rng.random(5), "tStart": np.array([1.0]), "tEnd": np.array([2.0]),
         "hypo": np.array([-3.0]), "taubefore": rng.random((1, F)), "tauafter": rng.random((1, F)),
         "delfafter": rng.random((1, F)), "seismic_slip": rng.random((3, F)), "seismic_slipvel": rng.random((3, F)),
         "seismic_stress": rng.random((3, F)), "is_slip": rng.random((2, F)), "is_slipvel": rng.random((2, F)),
         "is_stress": rng.random((2, F)), "dSeis": rng.random((4, 6)), "vSeis": rng.random((4, 6)),
         "aSeis": rng.random((4, 6))}
    R["delfafter"][:, 0] = 0.0                    # the deepest node creeps (no coseismic slip)
    path = tmp_path / "run.npz"
    output.save(path, R, P, 12.5)
    O, t, Q = output.load(path)
    assert t == 12.5 and Q.FltNglob == F and np.array_equal(Q.FltX, P.FltX)
    assert O.seismic_slip.shape == (F, 3) and np.array_equal(O.seismic_slip, R["seismic_slip"].T)
    assert np.array_equal(O.delfafter[:, 0], R["delfafter"][0]) and O.index_eq.size == 0
    assert np.array_equal(O.dSeis, R["dSeis"]) and np.array_equal(O.time_, R["time_"])
    # the catalogue reads either orientation
    ev = C.catalogue({"delfafter": O.delfafter, "taubefore": O.taubefore, "tauafter": O.tauafter,
                      "tStart": O.tStart, "tEnd": O.tEnd, "hypo": O.hypo}, P.FltX, 3e10)
    assert len(ev) == 1 and ev[0]["duration_s"] == 1.0

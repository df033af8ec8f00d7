"""Earthquake catalogue of a run: events, hypocentres, coseismic slip, stress
drops and moment magnitudes.

Same quantities and conventions as the reference's post-processing
(post/earthquake_cycles.jl): fault matrices are [FltNglob x nsnap] (one
column per snapshot, the orientation of the reference's `results` fields;
`output.load` returns them so), fault index 0 is the deepest node.

* ``get_index``          -- post/earthquake_cycles.jl:14-30
* ``coslip``             -- post/earthquake_cycles.jl:35-76
* ``moment_magnitude``   -- post/earthquake_cycles.jl:84-127 (moment_magnitude_new)
* ``catalogue``          -- one record per event from a run's results
"""
from __future__ import annotations

import numpy as np


def get_index(seismic_stress, taubefore):
    """Index of the seismic snapshot closest (2-norm) to each event's initial
    stress (post/earthquake_cycles.jl:14-30); 0-based."""
    S = np.asarray(seismic_stress, dtype=np.float64)
    T = np.asarray(taubefore, dtype=np.float64)
    out = np.zeros(T.shape[1], dtype=np.int64)
    for i in range(T.shape[1]):
        d = np.linalg.norm(S - T[:, i:i + 1], axis=0)
        out[i] = int(np.argmin(d))
    return out


def coslip(FltX, slip, slipvel, stress, time_, Vthres=0.001):
    """Events from a snapshot series (post/earthquake_cycles.jl:35-76).

    slip, slipvel, stress: [FltNglob x n]; time_: [n].  An event starts at the
    first snapshot with max slip rate > 1.01 Vthres and ends at the next one
    below 0.99 Vthres.  Returns (delfafter [F x ne_end], stressdrops
    [F x ne_end], tStart, tEnd, vhypo, hypo)."""
    slip = np.asarray(slip, dtype=np.float64)
    slipvel = np.asarray(slipvel, dtype=np.float64)
    stress = np.asarray(stress, dtype=np.float64)
    time_ = np.asarray(time_, dtype=np.float64)
    FltX = np.asarray(FltX, dtype=np.float64)
    F, n = slip.shape
    Vfmax = slipvel.max(axis=0)
    delfafter, taubefore, tauafter = [], [], []
    tStart, tEnd, vhypo, hypo = [], [], [], []
    slipstart = 0
    delfref = np.zeros(F)
    for i in range(n):
        if Vfmax[i] > 1.01 * Vthres and slipstart == 0:
            delfref = slip[:, i].copy()
            slipstart = 1
            tStart.append(time_[i])
            taubefore.append(stress[:, i].copy())
            indx = int(np.argmax(slipvel[:, i]))
            vhypo.append(slipvel[indx, i])
            hypo.append(FltX[indx])
        if Vfmax[i] < 0.99 * Vthres and slipstart == 1:
            delfafter.append(slip[:, i] - delfref)
            tauafter.append(stress[:, i].copy())
            tEnd.append(time_[i])
            slipstart = 0
    ne = len(tEnd)
    D = np.array(delfafter).T if ne else np.zeros((F, 0))
    drops = (np.array(taubefore[:ne]).T - np.array(tauafter).T) if ne else np.zeros((F, 0))
    return D, drops, np.array(tStart), np.array(tEnd), np.array(vhypo), np.array(hypo)


def moment_magnitude(mu, FltNglob, FltX, delfafter, stressdrops):
    """Mw, mean stress drop and mean slip of each event
    (post/earthquake_cycles.jl:84-127): nodes with slip >= 1% of the event's
    maximum form the rupture; area = sum(slip dx), along-dip extent
    zdim = sum(dx), moment M0 = mu area zdim (square rupture),
    Mw = 2/3 log10(M0 1e7) - 10.7.  dx[j-1] = FltX[j] - FltX[j-1] weights node j
    (1-based j in the reference); the deepest node has no such interval and
    must not belong to a rupture, as there."""
    D = np.asarray(delfafter, dtype=np.float64)
    S = np.asarray(stressdrops, dtype=np.float64)
    dx = np.diff(np.asarray(FltX, dtype=np.float64))
    ne = D.shape[1]
    M0 = np.zeros(ne)
    del_sigma = np.zeros(ne)
    fault_slip = np.zeros(ne)
    for i in range(ne):
        thres = 0.01 * D[:, i].max()
        area = zdim = sig = sl = 0.0
        for j in range(FltNglob):
            if D[j, i] >= thres:
                if j == 0:
                    raise IndexError("the deepest fault node lies in the rupture (no interval below it)")
                area += D[j, i] * dx[j - 1]
                zdim += dx[j - 1]
                sl += D[j, i]
                sig += S[j, i] * dx[j - 1]
        M0[i] = mu * area * zdim
        del_sigma[i] = sig / zdim
        fault_slip[i] = sl / zdim
    with np.errstate(divide="ignore"):
        Mw = (2.0 / 3.0) * np.log10(M0 * 1e7) - 10.7
    return Mw, del_sigma, fault_slip


def catalogue(results, FltX, mu, yr2sec=365 * 24 * 60 * 60):
    """One record per completed event of a run (results as returned by
    EqContext.run / output.load, fault matrices [events x F] or [F x events]):
    start and end (years), duration (s), hypocentre depth, Mw, mean stress
    drop (MPa, the units of tau*) and mean slip."""
    F = len(FltX)

    def cols(a):
        a = np.asarray(a, dtype=np.float64)
        return a if a.ndim == 2 and a.shape[0] == F else a.T
    D = cols(results["delfafter"])
    drops = cols(results["taubefore"])[:, :D.shape[1]] - cols(results["tauafter"])
    Mw, dsig, fslip = moment_magnitude(mu, F, FltX, D, drops)
    tS, tE, hy = results["tStart"], results["tEnd"], results["hypo"]
    return [{"t_start_yr": float(tS[i]) / yr2sec, "t_end_yr": float(tE[i]) / yr2sec,
             "duration_s": float(tE[i] - tS[i]), "hypo_m": float(hy[i]), "Mw": float(Mw[i]),
             "stress_drop": float(dsig[i]), "mean_slip_m": float(fslip[i])} for i in range(D.shape[1])]

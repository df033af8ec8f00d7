"""Write tests/golden/nr_ids_cases.json: Newton-Raphson and state-update vectors.

NRsearch! iterates in BigFloat (256-bit MPFR) in the reference
(src/NRsearch_serial.jl:9,51-54); the expected outputs here come from the same
Newton sequence evaluated with mpmath at 256 bits (oracle.nrsearch_mp).  The
IDS!/IDS2! expected values come from mpmath evaluation of the closed-form
aging-law update at 100 digits.  Inputs are drawn from splitmix64 (seed 7).

``nrsearch_fast`` covers fast slip (free slip velocities of 1-130 m/s in both
directions, FltZ up to 2.4e7 as on the 640 x 320 mesh): there the first step
from tau = 0 lands hundreds of a*sigma above the root, which an FP64 exp
overflows and the BigFloat sequence descends step by step.  Odd cases start
the second search of FBC! (src/NRsearch_serial.jl:37-38) from the first
search's tau, with the state moved as IDS2! would.  Named cases: the surface
node of the 640 x 320 mesh at 38 m/s, a search that ends on k == 1000 with
an infinite slip rate, and one that ends on |delta| > 1e10.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from eqcycle_amd.params import uniform01  # noqa: E402
from oracle.oracle import nrsearch_mp  # noqa: E402


def main():
    import mpmath as mp
    u = uniform01(7, 4000).reshape(-1, 8)
    cases = []
    for k in range(40):
        r = u[k]
        fo, Vo, b = 0.6, 1e-6, 0.019
        a = 0.011 + 0.008 * r[0]           # a-b from -0.008 to 0
        Seff = 50e6
        Vss = 10 ** (-12 + 11 * r[1])      # 1e-12 .. 1e-1 m/s
        psi = np.log(Vo / Vss) + (r[2] - 0.5)
        tauo = Seff * (fo + (a - b) * np.log(1e-9 / Vo)) * (1 + 0.01 * (2 * r[3] - 1))
        FltZ = 4e6 + 6e6 * r[4]
        Vfree = 10 ** (-10 + 10.5 * r[5])
        tau0 = 0.0 if k % 2 == 0 else 0.8 * tauo * r[6]
        Vf, tau, it = nrsearch_mp(fo, Vo, a, b, Seff, tau0, tauo, psi, FltZ, Vfree)
        cases.append(dict(fo=fo, Vo=Vo, cca=a, ccb=b, Seff=Seff, tau0=tau0, tauo=tauo, psi=psi, FltZ=FltZ,
                          FltVfree=Vfree, Vf=Vf, tau=tau, k=it))
    fast = []

    def add_fast(fo, Vo, a, b, Seff, tau0, tauo, psi, FltZ, Vfree, name=""):
        Vf, tau, it = nrsearch_mp(fo, Vo, a, b, Seff, tau0, tauo, psi, FltZ, Vfree)
        fast.append(dict(name=name, fo=fo, Vo=Vo, cca=a, ccb=b, Seff=Seff, tau0=tau0, tauo=tauo, psi=psi,
                         FltZ=FltZ, FltVfree=Vfree, Vf=Vf, tau=tau, k=it))
        return Vf, tau, it

    for k in range(48):
        r = u[200 + k]
        fo, Vo, b, Seff = 0.6, 1e-6, 0.019, 50e6
        a = 0.011 + 0.008 * r[0]
        tauo = Seff * (fo + (a - b) * np.log(1e-9 / Vo)) * (1 + 0.01 * (2 * r[3] - 1))
        FltZ = 4e6 + 2e7 * r[4]
        Vfree = 10 ** (2.114 * r[5]) * (1.0 if r[6] < 0.8 else -1.0)   # 1 .. 130 m/s
        psi = -2.0 + 14.0 * r[2]
        _, tau1, _ = add_fast(fo, Vo, a, b, Seff, 0.0, tauo, psi, FltZ, Vfree)
        # second search of FBC!: from tau1, state moved by the step's slip
        add_fast(fo, Vo, a, b, Seff, tau1, tauo, psi - 0.5 * r[7], FltZ, Vfree * (1 + 0.05 * (r[1] - 0.5)))
    # surface node of the 640 x 320 mesh (a = 0.015, steady-state tauo)
    tauo5 = 50e6 * (0.6 + (0.015 - 0.019) * np.log(1e-9 / 1e-6))
    add_fast(0.6, 1e-6, 0.015, 0.019, 50e6, 0.0, tauo5, 5.0, 1.44106406484e7, 38.0, "surface node, 38 m/s")
    add_fast(0.6, 1e-6, 0.015, 0.019, 50e6, 0.0, tauo5, 5.0, 5e6, 130.0, "130 m/s")
    Vf_, tau_, _ = add_fast(0.6, 1e-6, 0.015, 0.019, 50e6, 0.0, tauo5, 3.0, 2e7, 100.0, "k == 1000 exit")
    add_fast(0.6, 1e-6, 0.015, 0.019, 50e6, tau_, tauo5, 3.0, 2e7, 100.0, "from a k == 1000 exit")
    add_fast(0.6, 1e-6, 0.015, 0.019, 50e6, 0.0, tauo5, 5.0, 1.5e7, 1000.0, "|delta| > 1e10 exit")
    add_fast(0.6, 1e-6, 0.015, 0.019, 50e6, 1.0e10, tauo5, 4.0, 9e6, 1.1e3, "from tau = 1e10")
    ids = []
    for k in range(40):
        r = u[100 + k]
        xLf, Vo = 8e-3, 1e-6
        psi = 2 + 20 * r[0]
        V = 10 ** (-12 + 12.5 * r[1])
        dt = 10 ** (-3 + 10 * r[2])
        mp.mp.dps = 100
        VdtL = mp.mpf(abs(V)) * mp.mpf(dt) / mp.mpf(xLf)
        exact = mp.log(mp.exp(mp.mpf(psi) - VdtL) + (mp.mpf(Vo) / mp.mpf(abs(V))) * (1 - mp.exp(-VdtL)))
        ids.append(dict(xLf=xLf, Vo=Vo, psi=psi, Vf=V, dt=dt, psi1=float(exact)))
    out = dict(source="tests/golden/make_golden.py (mpmath 256-bit Newton sequence, 100-digit aging law)",
               nrsearch=cases, nrsearch_fast=fast, ids=ids)
    with open(os.path.join(HERE, "nr_ids_cases.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()

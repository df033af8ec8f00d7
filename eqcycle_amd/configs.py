"""Synthetic fault models: the five benchmark set-ups plus small test meshes.

Every model is the 2D antiplane spectral-element problem of ``setParameters``
(par.jl:15-230): a rectangular half-domain, a vertical strike-slip fault on the
y = 0 edge (``BoundaryMatrix!`` side 'B', src/BoundaryMatrix.jl:39-45), free
surface at x = 0, absorbing edges at x = -LX ('L') and y = LY ('T'), NGLL = 5.

Parameters that the reference hard-codes in ``setParameters`` and
``src/initialConditions/defaultInitialConditions.jl`` are fields here; the
profiles differ from the reference defaults where the model definitions say so
(friction, constant normal stress, steady-state initial stress).
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

YR2SEC = 365 * 24 * 60 * 60  # par.jl:42


@dataclass(frozen=True)
class Config:
    name: str
    nelx: int                      # elements along dip (x), fault length
    nely: int                      # elements fault-normal (y)
    LX: float = 48e3               # along-dip extent (m)
    LY: float = 48e3               # fault-normal extent (m)
    total_years: float = 200.0
    # medium
    rho: float = 2670.0
    vs: float = 3464.0
    material: str = "homogeneous"  # homogeneous | trapezoid | gaussian_lognormal
    dz_halfwidth: float = 500.0    # damage zone half-width (m)
    dz_depth: float = 8e3          # full reduction down to this depth (m)
    dz_taper: float = 2e3          # reduction tapers linearly to 0 over this depth range
    dz_vs_reduction: float = 0.4   # vs *= (1 - reduction) inside the zone
    gauss_sigma: float = 400.0     # Gaussian damage zone width (m)
    gauss_vs_reduction: float = 0.3
    lognormal_sigma: float = 0.0   # vs *= exp(sigma * Z), Z ~ N(0,1) per node
    # friction (rate-state, aging law)
    f0: float = 0.6
    V0: float = 1e-6
    Dc: float = 8e-3
    sigma_n: float = 50e6
    b: float = 0.019
    a_vw: float = 0.015            # a in the velocity-weakening zone
    vw_depth: float = 15e3         # a-b = a_vw - b down to this depth
    ab_deep: float = 0.004         # a-b reached at vs_depth and held below
    vs_depth: float = 18e3
    creep_depth: float = 24e3      # fault below this depth slides at Vpl (FltIglobBC)
    Vpl: float = 35e-3 / YR2SEC    # par.jl:81
    Vthres: float = 1e-3           # par.jl:87
    # initial state: steady state at Vinit, tau0 perturbed by +-tau_perturb (uniform)
    Vinit: float = 1e-9
    tau_perturb: float = 0.01
    seed_tau: int = 1
    seed_vs: int = 2
    # numerics
    CFL: float = 0.6               # par.jl:46
    IDstate: int = 2               # par.jl:48
    pcg_tol: float = 1e-8
    precision: str = "fp64"        # fp64 | mixed (fp32 inner PCG with fp64 refinement) | f32pc (fp32 V-cycle)
    learned: bool = False          # learned initial guess for the pass-1 quasi-static solve
    n_snapshots: int = 0           # training pairs recorded from the plain solver

    @property
    def total_time(self) -> float:
        return self.total_years * YR2SEC

    def with_(self, **kw) -> "Config":
        return replace(self, **kw)


CONFIGS = {
    # 1. plain method, CPU-runnable
    "cfg1": Config("cfg1", 40, 40, total_years=200.0, pcg_tol=1e-8),
    # 2. plain method on GPU, trapezoid damage zone
    "cfg2": Config("cfg2", 160, 160, total_years=500.0, material="trapezoid", pcg_tol=1e-8),
    # 3. mixed precision
    "cfg3": Config("cfg3", 320, 320, total_years=300.0, material="trapezoid", pcg_tol=1e-8,
                   precision="mixed"),
    # 4. learned initial guess on the config-2 mesh, FP32 V-cycle, compared with
    #    plain FP64 at the same 1e-6 relative residual
    "cfg4": Config("cfg4", 160, 160, total_years=500.0, material="trapezoid", pcg_tol=1e-6,
                   precision="f32pc", learned=True, n_snapshots=2000),
    # 5. combined: Gaussian damage zone + lognormal heterogeneity, 20 cycles
    "cfg5": Config("cfg5", 640, 320, total_years=2000.0, material="gaussian_lognormal",
                   lognormal_sigma=0.1, pcg_tol=1e-6, precision="f32pc", learned=True, n_snapshots=3000),
}


def small(nelx: int = 4, nely: int = 4, **kw) -> Config:
    """Tiny meshes for brute-force tests (same physics, 48x48 km)."""
    base = dict(total_years=1.0)
    base.update(kw)
    return Config(f"small{nelx}x{nely}", nelx, nely, **base)

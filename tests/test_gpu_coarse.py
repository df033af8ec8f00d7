"""GPU: the coarse levels of the V-cycle collapsed at set-up into one sparse
product down (every coarse right-hand side from the first: [M_1; M_2 M_1; ...])
and one up (the first level's correction from every right-hand side and the
dense level's correction: [U_1, V_1 U_2, ..., V_1 ... V_k]) against the same
condensed levels applied one by one (eq_amg_coarse_mode).  The two are the
same linear operator; only the summation order differs."""
import numpy as np
import pytest

from eqcycle_amd.configs import small
from eqcycle_amd.params import normal, set_parameters

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def deep():
    """96 x 64 elements with a small coarsest level: the bilinear level and
    three or more Ruge-Stueben levels below it."""
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(96, 64, material="trapezoid", lognormal_sigma=0.3))
    ctx = EqContext(P, amg={"hierarchy": "pmg", "max_coarse": 200})
    levels = ctx.amg_info()
    print(f"[gpu] levels {levels}", flush=True)
    assert len(levels) >= 5
    return P, ctx


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_collapsed_coarse_levels_match_level_by_level(deep, prec, tol):
    P, ctx = deep
    n = P.nglob - P.FltNglob
    for seed in (3, 5):
        b = normal(seed, n)
        ctx.amg_coarse_mode(True)
        xc = ctx.amg_ldiv(b, prec)
        ctx.amg_coarse_mode(False)
        xl = ctx.amg_ldiv(b, prec)
        print(f"[gpu] {prec} seed {seed}: collapsed vs level by level {rel(xc, xl):.2e}", flush=True)
        assert rel(xc, xl) <= tol
    ctx.amg_coarse_mode(True)


def test_collapsed_coarse_levels_keep_cg_iterations(deep):
    """CG to 1e-10 with either form: the same iteration count to within one,
    and the same solution to the tolerance."""
    P, ctx = deep
    n = P.nglob - P.FltNglob
    b = normal(7, n)
    out = {}
    for mode in (True, False):
        ctx.amg_coarse_mode(mode)
        x, it, rr, ok = ctx.cg(np.zeros(n), b, 1e-10)
        assert ok
        out[mode] = (x, it)
    ctx.amg_coarse_mode(True)
    print(f"[gpu] CG iterations collapsed {out[True][1]}, level by level {out[False][1]}", flush=True)
    assert abs(out[True][1] - out[False][1]) <= 1
    assert rel(out[True][0], out[False][0]) <= 1e-8


def test_coarse_mode_rejects_a_hierarchy_without_collapsed_form():
    """Three levels (one condensed level above the dense one) have nothing to
    collapse: the level-by-level form is the only one."""
    from eqcycle_amd.api import EqContext, EqError
    P = set_parameters(small(40, 24, material="trapezoid"))
    ctx = EqContext(P, amg={"hierarchy": "pmg"})
    assert len(ctx.amg_info()) <= 3
    with pytest.raises(EqError):
        ctx.amg_coarse_mode(True)
    ctx.amg_coarse_mode(False)

"""eq_newmark_step_noalloc (the allocation-free Newmark step) against
eq_newmark_step through the C ABI: bit for bit on two calls in a row, the
second on a smaller mesh (the shared scratch array then holds the first call's
larger contents), and the size limit's error."""
import numpy as np
import pytest

from eqcycle_amd.configs import small
from eqcycle_amd.params import set_parameters


def _state(ctx, P, seed):
    """A state with the fault slipping: a few quasi-static steps, then v on
    the fault raised to metres per second so that the Newton searches run."""
    ctx.set_solver(tol=1e-10, pc="amg")
    ctx.state_init()
    for _ in range(20):
        ctx.step()
    s = ctx.state_get()
    rng = np.random.default_rng(seed)
    v = s["v"].copy()
    v[P.iFlt - 1] += rng.uniform(0.1, 2.0, P.FltNglob)
    return s["d"], v, s["a"], s["psi"]


@pytest.mark.gpu
def test_newmark_step_noalloc_equals_newmark_step():
    from eqcycle_amd.api import EqContext
    for k, (nelx, nely) in enumerate([(32, 8), (16, 4)]):
        P = set_parameters(small(nelx, nely, material="trapezoid", total_years=3))
        ctx = EqContext(P, amg=True)
        d, v, a, psi = _state(ctx, P, k)
        rc1, o1 = ctx.newmark_step(d, v, a, psi)
        rc2, o2 = ctx.newmark_step_noalloc(d, v, a, psi)
        assert rc1 == rc2
        for key in o1:
            assert np.array_equal(o1[key], o2[key]), (nelx, nely, key)
        assert np.abs(o1["Vf"]).max() > 0.1 and not np.array_equal(o1["d"], d)


@pytest.mark.gpu
def test_newmark_step_noalloc_rejects_meshes_above_its_limit():
    from eqcycle_amd.api import EqContext, EqError
    P = set_parameters(small(512, 128, material="homogeneous", total_years=1))
    assert P.nglob > 1 << 20
    ctx = EqContext(P, amg=False)
    ctx.state_init()
    s = ctx.state_get()
    with pytest.raises(EqError, match="exceeds the allocation-free step's scratch: at most 1048576 nodes"):
        ctx.newmark_step_noalloc(s["d"], s["v"], s["a"], s["psi"])

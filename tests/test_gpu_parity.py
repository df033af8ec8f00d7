"""GPU parity: every C-ABI entry point against the CPU FP64 oracle on the same inputs.

Tolerances: FP64 kernels are compared to the oracle at 1e-12 relative (the
summation order differs; exp/log come from different libraries), FP32 K u at
1e-5, linear solves against a sparse direct solve at the CG tolerance times
a condition allowance, time steps at 1e-10 relative per step.
"""
import os

import numpy as np
import pytest

from eqcycle_amd.configs import small
from eqcycle_amd.params import normal, set_parameters

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


@pytest.fixture(scope="module")
def mesh(oracle):
    from eqcycle_amd.api import EqContext
    # strongly heterogeneous nodal rigidity exercises the mu handling
    P = set_parameters(small(12, 8, material="trapezoid", lognormal_sigma=0.3))
    print(f"[gpu] mesh {P.NelX}x{P.NelY} nglob={P.nglob}", flush=True)
    ctx = EqContext(P, amg={"max_coarse": 10})
    K = oracle.assemble_K(P.NelX, P.NelY, P.nglob, P.iglob, P.W, P.H, P.dxe, P.dye)
    kni = K.submatrix(P.FltNI - 1, -1.0)
    return P, ctx, K, kni


def test_ku_fp64(mesh):
    P, ctx, K, _ = mesh
    u = normal(3, P.nglob)
    y = K.matvec(u)
    assert rel(ctx.ku(u), y) < 1e-13
    assert rel(ctx.ku(u, mask="neg"), -y) < 1e-13
    ui = u.copy()
    ui[P.iFlt - 1] = 0
    yi = K.matvec(ui)
    yi[P.iFlt - 1] = 0
    assert rel(ctx.ku(u, mask="interior"), yi) < 1e-13


def test_ku_fault_rows(mesh):
    """mask EQ_MASK_FAULT_ROWS: K u on the element row along the fault (all
    that (K F)[FltNI] and a[iFlt] read, src/main.jl:246,267-272), 0 elsewhere."""
    P, ctx, K, _ = mesh
    u = normal(9, P.nglob)
    y = K.matvec(u)
    rows = P.y <= P.dye * (1 + 1e-9)
    assert rows.sum() == 5 * P.FltNglob
    for prec, tol in (("fp64", 1e-13), ("fp32", 1e-6)):
        yf = ctx.ku(u, mask="fault_rows", prec=prec)
        assert np.all(yf[~rows] == 0.0)
        assert np.linalg.norm(yf[rows] - y[rows]) / np.linalg.norm(y[rows]) < tol
    F = np.zeros(P.nglob)
    F[P.iFlt - 1] = normal(11, P.FltNglob)
    rhs = ctx.ku(F, mask="fault_rows")[P.FltNI - 1]
    assert rel(rhs, K.matvec(F)[P.FltNI - 1]) < 1e-13


def test_ku_constant_nullspace(mesh):
    P, ctx, _, _ = mesh
    y = ctx.ku(np.ones(P.nglob))
    scale = np.max(np.abs(ctx.ku(normal(5, P.nglob))))
    assert np.max(np.abs(y)) < 1e-12 * scale


def test_ku_fp32(mesh):
    """FP32 K u (FP32 operands and sums) against the oracle's FP64 product: 1e-6
    relative in the 2-norm, as on the benchmark meshes (test_gpu_configs)."""
    P, ctx, K, _ = mesh
    u = normal(3, P.nglob)
    y = K.matvec(u)
    y32 = ctx.ku(u, prec="fp32")
    err = np.linalg.norm(y32 - y) / np.linalg.norm(y)
    print(f"[gpu] FP32 K u rel err {err:.2e}", flush=True)
    assert err < 1e-6


def test_kni(mesh):
    P, ctx, _, kni = mesh
    x = normal(7, P.nglob - P.FltNglob)
    assert rel(ctx.kni(x), kni.matvec(x)) < 1e-13


def test_element_computation(mesh, oracle):
    P, ctx, _, _ = mesh
    u = normal(11, P.nglob)
    ae = ctx.element_computation(u)
    c1 = (0.5 * P.dxe * 0.5 * P.dye) / (0.5 * P.dxe) ** 2
    c2 = (0.5 * P.dxe * 0.5 * P.dye) / (0.5 * P.dye) ** 2
    for e in [0, 5, P.Nel // 2, P.Nel - 1]:
        ig = P.iglob[e] - 1          # [j, i]
        ue = u[ig].T                  # [i, j]
        We = P.W[e].T
        ref = oracle.element_ku(ue, We, P.H, c1, c2)
        assert rel(ae[e].T, ref) < 1e-13


@pytest.mark.parametrize("pc", ["amg", "jacobi", "none"])
def test_cg_matches_direct_solve(mesh, pc):
    import scipy.sparse.linalg as spla
    P, ctx, _, kni = mesh
    A = kni.to_scipy().tocsc()
    b = normal(13, A.shape[0])
    xs = spla.spsolve(A, b)
    x, it, rr, ok = ctx.cg(np.zeros_like(b), b, tol=1e-11, pc=pc)
    assert ok and rr <= 1e-11
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-7
    # the returned residual is the true one
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 2e-11


@pytest.mark.parametrize("hierarchy", ["rs", "pmg"])
def test_cg_hierarchies(mesh, hierarchy):
    """ruge_stuben-structured hierarchy and the p-multigrid one both solve to tolerance."""
    import scipy.sparse.linalg as spla
    from eqcycle_amd.api import EqContext
    P, _, _, kni = mesh
    ctx = EqContext(P, amg={"hierarchy": hierarchy})
    A = kni.to_scipy().tocsc()
    b = normal(31, A.shape[0])
    xs = spla.spsolve(A, b)
    for prec in ("fp64", "mixed", "f32pc"):
        x, it, rr, ok = ctx.cg(np.zeros_like(b), b, tol=1e-10, pc="amg", prec=prec)
        print(f"[gpu] {hierarchy} {prec} iterations {it}", flush=True)
        assert ok
        assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-10
        assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-6


@pytest.mark.parametrize("nelx, nely, bound", [(60, 30, 12), (30, 60, 12), (160, 80, 12), (96, 24, 20)])
def test_pmg_elongated_elements(nelx, nely, bound):
    """p-MG on elongated elements (aspect 2 either way, and 4): the Q1 level
    is semi-coarsened geometrically before Ruge-Stueben, and CG from zero
    needs about as many iterations as on square elements (6 there; RS alone
    on the anisotropic Q1 operator needed 27 on 80 x 40 and 151 on 320 x 160).
    Aspect 4 (Q1 couplings of opposite sign, 1.2 against -2.6) stays harder."""
    import scipy.sparse.linalg as spla
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(nelx, nely))
    ctx = EqContext(P, amg={"hierarchy": "pmg"})
    n = P.nglob - P.FltNglob
    b = normal(5, n)
    x, it, rr, ok = ctx.cg(np.zeros(n), b, tol=1e-6, pc="amg")
    print(f"[gpu] pmg {nelx}x{nely}: {it} iterations, levels {[l[0] for l in ctx.amg_info()]}", flush=True)
    assert ok and it <= bound


def test_cg_mixed_precision(mesh):
    import scipy.sparse.linalg as spla
    P, ctx, _, kni = mesh
    A = kni.to_scipy().tocsc()
    b = normal(17, A.shape[0])
    xs = spla.spsolve(A, b)
    x, it, rr, ok = ctx.cg(np.zeros_like(b), b, tol=1e-10, pc="amg", prec="mixed")
    assert ok
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-10
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-6


@pytest.mark.parametrize("degree", [2, 3, 5, 6])
def test_cheb_pairs_equal_single_steps(degree):
    """Two level-0 Chebyshev steps per launch give the V-cycle and the CG
    solves of one launch per step bit for bit (FP64 and FP32 V-cycles), on a
    mesh with several tiles in both directions and partial edge tiles."""
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(40, 13, material="trapezoid", lognormal_sigma=0.3))
    print(f"[gpu] cheb pairs degree {degree} nglob={P.nglob}", flush=True)
    ctx = EqContext(P, amg={"hierarchy": "pmg", "degree": degree, "smoother": "chebyshev"})
    b = normal(41, P.nglob - P.FltNglob)
    out = {}
    for pairs in (True, False):
        ctx.set_solver(tol=1e-10, pc="amg", cheb_pairs=pairs)
        out[pairs] = [ctx.amg_ldiv(b, prec) for prec in ("fp64", "fp32")]
        for prec in ("fp64", "f32pc"):
            x, it, rr, ok = ctx.cg(np.zeros_like(b), b, tol=1e-10, pc="amg", prec=prec)
            assert ok
            out[pairs].append(x)
    for k, (a, c) in enumerate(zip(out[True], out[False])):
        assert np.array_equal(a, c), k


def test_cheb_pairs_run_equals_single_steps():
    """main(P) through the first event with paired and with single Chebyshev
    launches (CG batches replayed as graphs): identical histories and state."""
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(32, 8, material="trapezoid", total_years=3))
    ctx = EqContext(P, amg={"hierarchy": "pmg", "degree": 5, "smoother": "chebyshev"})
    res = {}
    for pairs in (True, False):
        ctx.set_solver(tol=1e-8, pc="amg", prec="f32pc", cheb_pairs=pairs)
        ctx.state_init()
        R = ctx.run(snapshots=False, events=True)
        res[pairs] = (R, ctx.state_get())
    (Ra, sa), (Rb, sb) = res[True], res[False]
    assert len(Ra["tStart"]) >= 1
    for k in ("time_", "Vfmax", "pcg_iters", "tStart", "tEnd", "delfafter"):
        assert np.array_equal(Ra[k], Rb[k]), k
    for k in ("d", "v", "psi", "tau"):
        assert np.array_equal(sa[k], sb[k]), k


@pytest.mark.parametrize("prec", ["fp64", "f32pc"])
def test_early_tail_run_equals_synchronous(prec):
    """main(P) through the first event with the second quasi-static solve's
    fault update and status read-back sent ahead of its convergence test (and
    redone when the first batch fell short) and with the host waiting for
    convergence first: identical histories, iteration counts and state."""
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(32, 8, material="trapezoid", total_years=3))
    ctx = EqContext(P, amg={"hierarchy": "pmg", "degree": 5})
    res = {}
    for early in (True, False):
        ctx.set_solver(tol=1e-8, pc="amg", prec=prec, early_tail=early)
        ctx.state_init()
        R = ctx.run(snapshots=False, events=True)
        res[early] = (R, ctx.state_get())
    (Ra, sa), (Rb, sb) = res[True], res[False]
    assert len(Ra["tStart"]) >= 1
    # the second solve's count rose from one QS step to the next somewhere, so
    # a first batch (sized by the previous count) fell short and was redone
    it2 = Ra["pcg_iters"][:, 1]
    assert np.any(np.diff(it2[it2 > 0]) > 0)
    for k in ("time_", "Vfmax", "pcg_iters", "tStart", "tEnd", "delfafter", "n_qs", "n_dyn"):
        assert np.array_equal(Ra[k], Rb[k]), k
    for k in ("d", "v", "a", "psi", "tau", "Vf"):
        assert np.array_equal(sa[k], sb[k]), k


def test_amg_upload_matches_numpy_vcycle(mesh, oracle):
    """eq_amg_upload with the oracle's Ruge-Stueben hierarchy of kni (what the
    reference's ruge_stuben(kni) hands over, src/main.jl:200): the device
    V-cycle equals the same V-cycle in numpy (one Jacobi sweep of weight
    omega_l before and after each coarse correction, dense solve on the last
    level) to 1e-12, and CG with it, or with Chebyshev on level 0, solves to
    tolerance."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from eqcycle_amd.api import EqContext
    P, _, _, kni = mesh
    ctx = EqContext(P)
    ml = oracle.Amg(kni)
    L = ml.nlev
    assert L >= 3
    A = [ml.level_csr(l, 0) for l in range(L)]
    Pm = [ml.level_csr(l, 1) for l in range(L - 1)]
    om = []
    for Al in A:
        s = sp.diags(1.0 / np.sqrt(np.abs(Al.diagonal())))
        S = (s @ (-Al) @ s).toarray()                 # similar to D^-1 A (A negative definite)
        om.append(4.0 / (3.0 * np.linalg.eigvalsh(S)[-1]))
    ctx.amg_upload(A, Pm, omega=om)
    b = normal(43, A[0].shape[0])

    def vcycle(l, rhs):
        if l == L - 1:
            return np.linalg.solve(A[l].toarray(), rhs)
        dinv = 1.0 / A[l].diagonal()
        x = om[l] * dinv * rhs
        e = vcycle(l + 1, Pm[l].T @ (rhs - A[l] @ x))
        x = x + Pm[l] @ e
        return x + om[l] * dinv * (rhs - A[l] @ x)

    assert rel(ctx.amg_ldiv(b), vcycle(0, b)) < 1e-12
    Ad = kni.to_scipy().tocsc()
    for smoother, degree in (("jacobi", 1), ("chebyshev", 3)):
        ctx.amg_upload(A, Pm, omega=om if smoother == "jacobi" else None, smoother=smoother, degree=degree)
        x, it, rr, ok = ctx.cg(np.zeros_like(b), b, tol=1e-10, pc="amg")
        print(f"[gpu] uploaded hierarchy ({L} levels), {smoother}: {it} iterations", flush=True)
        assert ok and np.linalg.norm(b - Ad @ x) <= 1e-10 * np.linalg.norm(b)
    # a level-0 operator that is not this mesh's kni is refused
    with pytest.raises(Exception, match="not kni"):
        ctx.amg_upload([1.01 * A[0]] + A[1:], Pm)


@pytest.mark.parametrize("smoother", ["chebyshev", "tiles"])
def test_amg_ldiv_is_spd(smoother):
    """kni is negative definite, so is its preconditioner.  With whole-mesh
    Chebyshev sweeps the V-cycle is symmetric to rounding; with tile sweeps
    (restricted additive Schwarz, the same sweep before and after the coarse
    correction) only nearly so: it is pinned to 1e-2 here."""
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(40, 40, material="trapezoid", lognormal_sigma=0.3))
    ctx = EqContext(P, amg={"smoother": smoother})
    n = P.nglob - P.FltNglob
    x, y = normal(19, n), normal(23, n)
    Mx, My = ctx.amg_ldiv(x), ctx.amg_ldiv(y)
    asym = abs(x @ My - y @ Mx) / abs(x @ My)
    print(f"[gpu] {smoother} V-cycle asymmetry {asym:.2e}", flush=True)
    assert asym <= (1e-10 if smoother == "chebyshev" else 1e-2)
    for seed in (19, 23, 29):
        z = normal(seed, n)
        assert z @ ctx.amg_ldiv(z) < 0


def test_pcg_jacobi(mesh, oracle):
    P, ctx, K, _ = mesh
    F = np.zeros(P.nglob)
    F[P.iFlt - 1] = normal(29, P.FltNglob)
    Kii = K.submatrix(P.FltNI - 1, 1.0)
    Fnew = -K.matvec(F)[P.FltNI - 1]
    diag = Kii.to_scipy().diagonal().copy()
    x0 = np.zeros(P.nglob - P.FltNglob)
    xo, ito, oko = oracle.pcg_jacobi(Kii, diag, x0.copy(), Fnew, tol=1e-9, maxit=8000)
    xg, itg, okg = ctx.pcg_jacobi(F, x0, tol=1e-9, maxit=8000)
    assert oko and okg
    assert abs(itg - ito) <= 2
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-7


def _fault_inputs(P, seed):
    rng = np.random.default_rng(seed)
    F = P.FltNglob
    psi = np.log(P.Vo / 10 ** rng.uniform(-12, -1, F))
    Vf = 10 ** rng.uniform(-12, 0, F)
    return psi, Vf, rng


def test_ids_ids2(mesh, oracle):
    P, ctx, _, _ = mesh
    psi, Vf, rng = _fault_inputs(P, 1)
    for dt in [1e-3, 0.05, 3600.0, 4e6]:
        for cnd in [1e-6, 1e-5]:
            g = ctx.ids(P.xLf, P.Vo, psi, dt, Vf, cnd, 2)
            o = np.array([oracle.ids(P.xLf[j], P.Vo[j], psi[j], dt, Vf[j], cnd, 2) for j in range(len(psi))])
            assert rel(g, o) < 1e-13
        psi1 = psi + rng.normal(0, 0.1, len(psi))
        Vf1 = Vf * 10 ** rng.normal(0, 0.5, len(psi))
        g2 = ctx.ids2(P.xLf, P.Vo, psi, psi1, dt, Vf, Vf1, 2)
        o2 = np.array([oracle.ids2(P.xLf[j], P.Vo[j], psi[j], psi1[j], dt, Vf[j], Vf1[j], 2)
                       for j in range(len(psi))])
        assert rel(g2, o2) < 1e-13


def test_nrsearch(mesh, oracle):
    P, ctx, _, _ = mesh
    psi, Vf, rng = _fault_inputs(P, 2)
    n = P.FltNglob
    Vfree = 10 ** rng.uniform(-10, 0.5, n)
    tau0 = np.zeros(n)
    Vg, tg, kg = ctx.nrsearch(P.fo, P.Vo, P.cca, P.ccb, P.Seff, tau0, P.tauo, psi, P.FltZ, Vfree)
    for j in range(n):
        vo, to, ko = oracle.nrsearch(P.fo[j], P.Vo[j], P.cca[j], P.ccb[j], P.Seff[j], 0.0, P.tauo[j], psi[j],
                                     P.FltZ[j], Vfree[j])
        assert kg[j] == ko
        assert abs(tg[j] - to) <= 1e-12 * max(abs(to), 1.0)
        assert abs(Vg[j] - vo) <= 1e-11 * max(abs(vo), 1e-300)


def test_nrsearch_fast_slip_matches_bigfloat_sequence(mesh):
    """Device NRsearch! on the 256-bit fast-slip fixtures
    (tests/golden/nr_ids_cases.json, nrsearch_fast: 1-130 m/s, FltZ up to
    2.4e7, the k == 1000 and |delta| > 1e10 exits): the same iteration count
    and exit as the BigFloat sequence of src/NRsearch_serial.jl:47-98, tau to
    1e-11, the slip rate to 1e-9 (infinite where the reference's Float64
    conversion is)."""
    import json
    P, ctx, _, _ = mesh
    G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "nr_ids_cases.json")))
    cases = G["nrsearch_fast"]
    col = {k: np.array([c[k] for c in cases], dtype=np.float64)
           for k in ("fo", "Vo", "cca", "ccb", "Seff", "tau0", "tauo", "psi", "FltZ", "FltVfree")}
    Vg, tg, kg = ctx.nrsearch(col["fo"], col["Vo"], col["cca"], col["ccb"], col["Seff"], col["tau0"], col["tauo"],
                              col["psi"], col["FltZ"], col["FltVfree"])
    for i, c in enumerate(cases):
        assert kg[i] == c["k"], (i, c)
        assert abs(tg[i] - c["tau"]) <= 1e-11 * max(abs(c["tau"]), 1.0), (i, c)
        if np.isinf(c["Vf"]) or np.isinf(Vg[i]):
            assert Vg[i] == c["Vf"], (i, c)
        else:
            assert abs(Vg[i] - c["Vf"]) <= 1e-9 * abs(c["Vf"]), (i, c)


def test_fbc_and_slr(mesh, oracle):
    P, ctx, _, _ = mesh
    psi, Vf, rng = _fault_inputs(P, 3)
    n = P.FltNglob
    Vfree = 10 ** rng.uniform(-10, 0, n)
    dt = 0.01
    outs_o = [np.zeros(n) for _ in range(6)]
    st_o, _ = oracle.fbc(2, P, P.NFBC, n, *outs_o, psi, Vf, Vfree, dt)
    rc, fail, outs_g = ctx.fbc(2, P.NFBC, n, *[np.zeros(n) for _ in range(6)], psi, Vf, Vfree, dt)
    assert st_o == 0 and fail == 0
    for a, b in zip(outs_g, outs_o):
        assert rel(a, b) < 1e-11
    tau1 = rng.normal(0, 1e6, n)
    p1o, v1o = np.zeros(n), np.zeros(n)
    oracle.slr(P, P.NFBC, n, psi, p1o, Vf, v1o, 2, tau1, 3600.0)
    p1g, v1g = ctx.slr(P.NFBC, n, psi, np.zeros(n), Vf, np.zeros(n), 2, tau1, 3600.0)
    assert rel(p1g, p1o) < 1e-13
    assert rel(v1g, v1o) < 1e-11


def test_zero_length_fault_routines(mesh):
    """Empty inputs to the element-wise fault routines return empty outputs
    (no launch, no error), as the reference's loops over empty ranges do."""
    P, ctx, _, _ = mesh
    e = np.zeros(0)
    assert ctx.ids(e, e, e, 1.0, e, 1e-5).shape == (0,)
    assert ctx.ids2(e, e, e, e, 1.0, e, e).shape == (0,)
    out = ctx.nrsearch(e, e, e, e, e, e, e, e, e, e)
    assert all(np.asarray(o).shape == (0,) for o in out)


def test_fbc_guard_returns_nothing(mesh):
    """The FBC! NaN guard (src/NRsearch_serial.jl:19-32) maps to None."""
    from eqcycle_amd.api import FBC
    P, ctx, _, _ = mesh
    psi, Vf, rng = _fault_inputs(P, 4)
    n = P.FltNglob
    Vf = Vf.copy()
    Vf[P.NFBC + 2] = np.nan
    from eqcycle_amd.api import use_context
    use_context(ctx)
    out = FBC(2, P, P.NFBC, n, *[np.zeros(n) for _ in range(6)], psi, Vf, np.full(n, 1e-9), 0.01)
    assert out is None


def test_fault_routines_read_the_callers_params(mesh, oracle):
    """FBC!(IDstate, P, ...) and slrFunc!(P, ...) with the reference's
    arguments read the caller's P (src/NRsearch_serial.jl:6,
    src/otherFunctions_serial.jl:62): a P whose tauo, cca and FltZ differ
    from the context's gives the oracle's result for that P, not the one
    for the parameters given at set-up."""
    import dataclasses

    from eqcycle_amd.api import FBC, slrFunc, use_context
    P, ctx, _, _ = mesh
    use_context(ctx)
    psi, Vf, rng = _fault_inputs(P, 21)
    n = P.FltNglob
    Vfree = 10 ** rng.uniform(-10, 0, n)
    P2 = dataclasses.replace(P, tauo=P.tauo * 1.05, cca=P.cca * 1.1, FltZ=P.FltZ * 0.8)
    dt = 0.01
    want = [np.zeros(n) for _ in range(6)]
    assert oracle.fbc(2, P2, P2.NFBC, n, *want, psi, Vf, Vfree, dt)[0] == 0
    got = FBC(2, P2, P.NFBC, n, *[np.zeros(n) for _ in range(6)], psi, Vf, Vfree, dt)
    for a, b in zip(got, want):
        assert rel(a, b) < 1e-11
    old = [np.zeros(n) for _ in range(6)]
    oracle.fbc(2, P, P.NFBC, n, *old, psi, Vf, Vfree, dt)
    assert rel(got[2], old[2]) > 1e-3
    tau1 = rng.normal(0, 1e6, n)
    p1o, v1o = np.zeros(n), np.zeros(n)
    oracle.slr(P2, P2.NFBC, n, psi, p1o, Vf, v1o, 2, tau1, 3600.0)
    p1g, v1g = slrFunc(P2, P.NFBC, n, psi, np.zeros(n), Vf, np.zeros(n), 2, tau1, 3600.0)
    assert rel(p1g, p1o) < 1e-13 and rel(v1g, v1o) < 1e-11
    # the parameters given at set-up are untouched for calls without P
    p1c, v1c = ctx.slr(P.NFBC, n, psi, np.zeros(n), Vf, np.zeros(n), 2, tau1, 3600.0)
    p1r, v1r = np.zeros(n), np.zeros(n)
    oracle.slr(P, P.NFBC, n, psi, p1r, Vf, v1r, 2, tau1, 3600.0)
    assert rel(v1c, v1r) < 1e-11


def test_dtevol(mesh, oracle):
    P, ctx, _, _ = mesh
    rng = np.random.default_rng(5)
    for isolver in (1, 2):
        for _ in range(5):
            Vf = 10 ** rng.uniform(-12, 0, P.FltNglob)
            dt = 10 ** rng.uniform(-2, 6)
            g = ctx.dtevol(dt, P.dt0, P.XiLf, P.FltNglob, P.NFBC, Vf, isolver)
            o = oracle.dtevol(dt, P.dt0, P.XiLf, P.FltNglob, P.NFBC, Vf, isolver)
            assert g == o


def _compare_state(ctx, sim, tol):
    s = ctx.state_get()
    assert rel(s["d"], sim.d) < tol
    assert rel(s["v"], sim.v) < tol
    assert rel(s["psi"], sim.psi) < tol
    assert rel(s["Vf"], sim.Vf) < tol
    assert abs(s["dt"] - sim.dt) <= tol * sim.dt
    assert s["isolver"] == sim.isolver


@pytest.mark.parametrize("hierarchy", ["rs", "pmg"])
def test_quasi_static_steps_match_oracle(oracle, hierarchy):
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(32, 8, material="trapezoid"))
    ctx = EqContext(P, amg={"hierarchy": hierarchy})
    ctx.set_solver(tol=1e-13, pc="amg")
    ctx.state_init()
    sim = oracle.Simulation(P, solver="direct")
    for k in range(12):
        info = ctx.step()
        sim.step()
        assert info.isolver_used == 1
        _compare_state(ctx, sim, 1e-9)


@pytest.mark.parametrize("nel", [(1, 1), (2, 3), (1, 6), (6, 1), (17, 5), (65, 3)])
def test_small_and_ragged_meshes(oracle, nel):
    """Meshes of one element, one element column or row, and sizes that leave
    every tile or tile row partial (a 64 x 16 node tile covers 16 x 4
    elements): K u against the assembled K, CG against a direct solve, and
    quasi-static steps of the loop against the oracle's."""
    import scipy.sparse.linalg as spla
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(*nel, material="trapezoid", lognormal_sigma=0.3))
    print(f"[gpu] mesh {nel} nglob={P.nglob}", flush=True)
    ctx = EqContext(P, amg={"hierarchy": "pmg"})
    K = oracle.assemble_K(P.NelX, P.NelY, P.nglob, P.iglob, P.W, P.H, P.dxe, P.dye)
    u = normal(21, P.nglob)
    assert rel(ctx.ku(u), K.matvec(u)) < 1e-13
    A = K.submatrix(P.FltNI - 1, -1.0).to_scipy().tocsc()
    b = normal(22, A.shape[0])
    xs = spla.spsolve(A, b)
    for prec in ("fp64", "f32pc"):
        x, it, rr, ok = ctx.cg(np.zeros_like(b), b, tol=1e-11, pc="amg", prec=prec)
        assert ok and np.linalg.norm(x - xs) <= 1e-7 * np.linalg.norm(xs)
    ctx.set_solver(tol=1e-13, pc="amg")
    ctx.state_init()
    sim = oracle.Simulation(P, solver="direct")
    for k in range(8):
        info = ctx.step()
        sim.step()
        assert info.isolver_used == sim.res.isolver[-1]
        _compare_state(ctx, sim, 1e-9)


def test_dynamic_steps_match_oracle(oracle):
    """Step both sides into the first earthquake (about step 256 on this mesh)
    and compare every Newmark step."""
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(32, 8, material="trapezoid", total_years=3))
    ctx = EqContext(P, amg=True)
    ctx.set_solver(tol=1e-13, pc="amg")
    ctx.state_init()
    sim = oracle.Simulation(P, solver="direct")
    ndyn = 0
    for k in range(600):
        info = ctx.step()
        sim.step()
        assert info.isolver_used == sim.res.isolver[-1]
        if info.isolver_used == 2:
            ndyn += 1
            _compare_state(ctx, sim, 1e-8)
            assert rel(ctx.state_get()["a"], sim.a) < 1e-7
        elif k % 50 == 0:
            _compare_state(ctx, sim, 1e-9)
        if ndyn >= 60:
            break
    assert ndyn >= 60


def test_run_matches_oracle(oracle):
    """main(P) through the first event: same steps, event times, slip-rate history."""
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(32, 8, material="trapezoid", total_years=3))
    ctx = EqContext(P, amg=True)
    ctx.set_solver(tol=1e-13, pc="amg")
    ctx.state_init()
    R = ctx.run()
    sim = oracle.Simulation(P, tol=1e-12, solver="direct")
    sim.run()
    r = sim.res
    # Oracle direct vs oracle AMG-PCG (tol 1e-12) on this run differ by 5e-9 in
    # time_, 1.5e-8 in Vfmax, 1.5e-9 in d and coseismic slip, 1e-14 in event
    # times; two LU orderings of the oracle's direct solve differ by 5.0e-9 in
    # d and 4.9e-9 in coseismic slip.  The bounds below are the parity
    # targets (1e-8 displacement, 1e-6 slip rate, event times well inside 0.1%).
    assert R["n_steps"] == len(r.time_)
    assert R["n_dyn"] > 1000
    m = dict(time_=rel(R["time_"], r.time_), Vfmax=rel(R["Vfmax"], r.Vfmax),
             delfafter=rel(R["delfafter"], np.array(r.delfafter)), d=rel(ctx.state_get()["d"], sim.d))
    print("[gpu] run parity", m, flush=True)
    assert m["time_"] < 1e-7
    assert m["Vfmax"] < 1e-6
    assert len(R["tStart"]) == len(r.tStart) >= 1
    assert np.allclose(R["tStart"], r.tStart, rtol=1e-9, atol=0)
    assert np.allclose(R["tEnd"], r.tEnd, rtol=1e-9, atol=0)
    assert m["delfafter"] < 1e-8
    assert m["d"] < 1e-8
    # receiver traces every 10 steps and the seismic fault snapshots; slip
    # 2 d + Vpl t carries the time_ difference (measured 4.6e-8, time_ 2.3e-8)
    dS = np.array(r.dSeis)
    assert R["dSeis"].shape == dS.shape
    assert rel(R["dSeis"], dS) < 1e-8
    assert len(R["seismic_slip"]) == len(r.seismic_slip) >= 1
    assert rel(R["seismic_slip"], np.array(r.seismic_slip)) < 1e-7


def test_fast_path_run_matches_oracle(oracle):
    """The bench's fast path through the first event against the oracle's
    direct-solve loop: Galerkin start on a POD basis learned from this mesh's
    own quasi-static steps, FP32 V-cycle, change-of-b starts, early step tail,
    Chebyshev pairs, graph-replayed CG batches and device-batched dynamic
    steps, at CG tolerance 1e-12."""
    from eqcycle_amd import learned
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(32, 8, material="trapezoid", total_years=3))
    ctx = EqContext(P, amg={"hierarchy": "pmg"})
    ctx.set_solver(tol=1e-12, pc="amg")
    ctx.state_init()
    learned.record(ctx, max_steps=150, capacity=160, nbasis=24)
    learned.fit_basis(ctx, tol=1e-12)
    ctx.set_solver(tol=1e-12, pc="amg", prec="f32pc", guess="galerkin")
    ctx.state_init()
    R = ctx.run()
    sim = oracle.Simulation(P, tol=1e-12, solver="direct")
    sim.run()
    r = sim.res
    m = dict(time_=rel(R["time_"], r.time_), Vfmax=rel(R["Vfmax"], r.Vfmax),
             delfafter=rel(R["delfafter"], np.array(r.delfafter)), d=rel(ctx.state_get()["d"], sim.d))
    print(f"[gpu] fast-path run parity {m}, {R['n_qs']} QS / {R['n_dyn']} dynamic steps, "
          f"{R['total_pcg_iters'] / max(1, 2 * R['n_qs']):.2f} CG iterations per solve", flush=True)
    assert R["n_steps"] == len(r.time_)
    assert len(R["tStart"]) == len(r.tStart) >= 1
    assert np.allclose(R["tStart"], r.tStart, rtol=1e-8, atol=0)
    assert np.allclose(R["tEnd"], r.tEnd, rtol=1e-8, atol=0)
    assert m["time_"] < 1e-6
    assert m["Vfmax"] < 1e-5
    assert m["delfafter"] < 1e-6
    assert m["d"] < 1e-6


def test_run_batched_dynamic_steps_equal_single_steps():
    """eq_run advances dynamic steps in device-side batches that stop at the
    regime switch; the trajectory must equal, bit for bit, one eq_step call per
    step (same kernels, same order), including a max_steps cut inside a batch."""
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(32, 8, material="trapezoid", total_years=3))
    a = EqContext(P, amg=True)
    a.set_solver(tol=1e-12, pc="amg")
    a.state_init()
    R1 = a.run(max_steps=1000)            # stops inside the first event's batches
    R2 = a.run()                          # rest of the 3 years, through the event end
    b = EqContext(P, amg=True)
    b.set_solver(tol=1e-12, pc="amg")
    b.state_init()
    n = R1["n_steps"] + R2["n_steps"]
    infos = [b.step() for _ in range(n)]
    assert R1["n_steps"] == 1000
    iso = np.array([i.isolver_used for i in infos])
    assert np.array_equal(np.concatenate([R1["isolver"], R2["isolver"]]), iso)
    assert (iso == 2).sum() > 1000 and iso[-1] == 1      # the event ended inside the window
    assert np.array_equal(np.concatenate([R1["Vfmax"], R2["Vfmax"]]), [i.Vfmax for i in infos])
    assert np.array_equal(np.concatenate([R1["time_"], R2["time_"]]), [i.t for i in infos])
    sa, sb = a.state_get(), b.state_get()
    for k in ("d", "v", "a", "psi", "tau", "Vf"):
        assert np.array_equal(sa[k], sb[k]), k
    assert (sa["t"], sa["dt"], sa["it"], sa["isolver"]) == (sb["t"], sb["dt"], sb["it"], sb["isolver"])


def test_state_init_after_dynamic_steps(oracle):
    """Regression: re-initialising after an odd number of Newmark steps (swapped
    state buffers) restores the initial state and reproduces a fresh run."""
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(32, 8, material="trapezoid", total_years=3))
    ctx = EqContext(P, amg=True)
    ctx.set_solver(tol=1e-12, pc="amg")
    ctx.state_init()
    ndyn = 0
    for _ in range(3000):
        ndyn += ctx.step().isolver_used == 2
        if ndyn % 2 == 1:
            break
    assert ndyn % 2 == 1
    ctx.state_init()
    s = ctx.state_get()
    assert np.all(s["d"] == 0) and np.all(s["a"] == 0) and s["t"] == 0
    a = [ctx.step().Vfmax for _ in range(30)]
    fresh = EqContext(P, amg=True)
    fresh.set_solver(tol=1e-12, pc="amg")
    fresh.state_init()
    b = [fresh.step().Vfmax for _ in range(30)]
    assert a == b


def test_config1_event_catalogue_matches_oracle():
    """main(P) on config 1 (40 x 40 elements, 25,921 nodes) through its first
    two earthquake clusters (160 of 200 simulated years) against the oracle's
    catalogue (direct solves, tests/golden/cfg1_oracle_events.json): every
    sub-event of the first cluster to 1e-9 relative in tStart and tEnd, the
    second cluster's onset to 1e-6 (solver tolerance 1e-12), and for every
    completed event of the first cluster the catalogue of
    post/earthquake_cycles.jl: the same hypocentre, Mw to 1e-6, mean stress
    drop and mean coseismic slip to 1e-6 relative.  The FP32 V-cycle at
    tolerance 1e-8 keeps both onsets within 0.1%."""
    import json
    import os

    from eqcycle_amd.api import EqContext
    from eqcycle_amd.catalogue import catalogue
    from eqcycle_amd.configs import CONFIGS
    G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "cfg1_oracle_events.json")))
    ts_o, te_o = np.array(G["tStart"]), np.array(G["tEnd"])
    yr = 365 * 24 * 3600.0
    split = 100 * yr                       # the first cluster is near 20 years, the second near 149
    n1 = int((ts_o < split).sum())
    assert n1 >= 1 and (ts_o > split).any()
    P = set_parameters(CONFIGS["cfg1"])
    ctx = EqContext(P, amg={"hierarchy": "pmg"})
    for tol, prec, r1, r2 in ((1e-12, "fp64", 1e-9, 1e-6), (1e-8, "f32pc", 1e-6, 1e-3)):
        ctx.set_solver(tol=tol, pc="amg", prec=prec)
        ctx.state_init()
        R = ctx.run(total_time=160 * yr, snapshots=False, events=True)
        ts, te = R["tStart"], R["tEnd"]
        print(f"[gpu] cfg1 {prec} tol {tol:g}: {R['n_steps']} steps, {len(ts)} event starts", flush=True)
        assert int((ts < split).sum()) == n1
        if prec == "fp64":
            assert np.allclose(ts[:n1], ts_o[:n1], rtol=r1, atol=0)
            assert np.allclose(te[:n1], te_o[:n1], rtol=r1, atol=0)
            cfg = CONFIGS["cfg1"]
            ne = len(te)
            cat = catalogue({k: R[k][:ne] if k in ("tStart", "hypo", "taubefore") else R[k]
                             for k in ("tStart", "tEnd", "hypo", "taubefore", "tauafter", "delfafter")},
                            P.FltX, cfg.rho * cfg.vs ** 2)
            for e, eo in zip(cat[:n1], G["catalogue"][:n1]):
                assert e["hypo_m"] == eo["hypo_m"]
                assert abs(e["Mw"] - eo["Mw"]) <= 1e-6
                assert abs(e["stress_drop"] - eo["stress_drop"]) <= 1e-6 * abs(eo["stress_drop"])
                assert abs(e["mean_slip_m"] - eo["mean_slip_m"]) <= 1e-6 * abs(eo["mean_slip_m"])
            print(f"[gpu] cfg1 first cluster: Mw {[round(e['Mw'], 3) for e in cat[:n1]]}", flush=True)
        else:
            assert abs(ts[0] - ts_o[0]) <= r1 * ts_o[0]
        s2, s2o = ts[ts > split][0], ts_o[ts_o > split][0]
        assert abs(s2 - s2o) <= r2 * s2o


@pytest.mark.parametrize("prec", ["fp64", "f32pc"])
def test_cg_graph_replay_equals_eager_launch(prec):
    """CG iteration batches replayed from captured HIP graphs give the same
    trajectory, bit for bit, as launching every kernel: 400 steps replay each
    graph many times, with host synchronisations and other kernels between
    the replays."""
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(32, 8, material="trapezoid", total_years=3))
    out = []
    for graphs in (False, True):
        ctx = EqContext(P, amg={"hierarchy": "pmg"})
        ctx.set_solver(tol=1e-10, pc="amg", prec=prec, graphs=graphs)
        ctx.state_init()
        R = ctx.run(max_steps=400)
        out.append((R, ctx.state_get()))
    (Ra, sa), (Rb, sb) = out
    assert Ra["total_pcg_iters"] > 400
    assert np.array_equal(Ra["pcg_iters"], Rb["pcg_iters"])
    assert np.array_equal(Ra["Vfmax"], Rb["Vfmax"])
    for k in ("d", "v", "a", "psi", "tau"):
        assert np.array_equal(sa[k], sb[k]), k


def test_consecutive_runs_continue_one_run():
    """main(P) split over several eq_run calls gives the event list and step
    history of one call: the bookkeeping carries over, the split point lies
    inside the first event."""
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(32, 8, material="trapezoid", total_years=3))
    a = EqContext(P, amg=True)
    a.set_solver(tol=1e-12, pc="amg")
    a.state_init()
    whole = a.run(snapshots=False, events=True)
    a.state_init()
    parts = [a.run(max_steps=n, snapshots=False, events=True) for n in (700, 300)]
    parts.append(a.run(snapshots=False, events=True))
    cat = {k: np.concatenate([p[k] for p in parts]) for k in ("time_", "Vfmax", "tStart", "tEnd", "hypo")}
    assert sum(p["n_steps"] for p in parts) == whole["n_steps"]
    assert len(whole["tStart"]) >= 1 and whole["tStart"][0] < parts[0]["time_"][-1] < whole["tEnd"][0]
    for k in cat:
        assert np.array_equal(cat[k], whole[k]), k
    assert np.array_equal(np.concatenate([p["delfafter"] for p in parts]), whole["delfafter"])


def test_nr_failure_stops_run_like_oracle(oracle):
    """A surface slip rate the rate-state search cannot follow, as the BigFloat
    search of src/NRsearch_serial.jl:47-98 behaves on it: free slip velocity
    ~390 m/s at the surface node.  The first search from tau = 0 takes the
    k == 1000 exit with an infinite rate and a finite tau, so the guard of
    :19 (which tests the step's incoming Vf and tau1) lets it through;
    IDS2! then gives psi2 = -inf and the second search NaN.  The step ends
    with a NaN traction at that node and Vfmax = NaN, which keeps the dynamic
    regime (src/main.jl:439-443); the next step's guard fires at the first
    node, in loop order, that the NaN reached through K -- the lowest node
    of the surface element.  GPU and oracle take the same two steps;
    keep_on_nr_failure returns the results up to the failing step."""
    import re
    from eqcycle_amd.api import EqContext, EqError
    P = set_parameters(small(32, 8, material="trapezoid", total_years=3))
    F = P.FltNglob
    ctx = EqContext(P, amg=True)
    ctx.set_solver(tol=1e-12, pc="amg")
    ctx.state_init()
    s = ctx.state_get()
    v = s["v"].copy()
    j = F - 1
    v[P.iFlt[j] - 1] = 0.5 * 1e10 / P.FltZ[j]
    sim = oracle.Simulation(P, solver="direct")
    sim.d, sim.v, sim.a = s["d"].copy(), v.copy(), s["a"].copy()
    sim.psi, sim.isolver, sim.t, sim.dt = s["psi"].copy(), 2, 1e7, P.dt0
    sim.step()
    assert np.isnan(sim.tau[j]) and sim.isolver == 2
    assert np.isfinite(sim.tau[:j]).all()
    with pytest.raises(FloatingPointError) as ei:
        sim.step()
    loc = int(re.search(r"fault node (\d+)$", str(ei.value)).group(1))
    assert loc == F - 4     # 1-based: the surface element spans nodes F-4 .. F
    kw = dict(d=s["d"], v=v, a=s["a"], psi=s["psi"], tau=s["tau"], t=1e7, dt=P.dt0, isolver=2)
    ctx.state_set(**kw)
    R = ctx.run(max_steps=5, snapshots=False, keep_on_nr_failure=True)
    assert R["error"] is not None and R["error"].endswith(f"(fault location {loc})")
    assert R["n_steps"] == 1 and np.isnan(R["Vfmax"][0])
    ctx.state_set(**kw)
    info = ctx.step()
    assert np.isnan(info.Vfmax) and info.isolver_next == 2
    g = ctx.state_get()
    assert np.isnan(g["tau"][j]) and np.isfinite(g["tau"][:j]).all()
    with pytest.raises(EqError, match=f"NR SEARCH FAILED! \\(fault location {loc}\\)"):
        ctx.step()
    # an ordinary run reports no error
    ctx.state_init()
    assert ctx.run(max_steps=3, snapshots=False, keep_on_nr_failure=True)["error"] is None


def test_newmark_step_entry_point_matches_loop_and_oracle(oracle):
    """eq_newmark_step (the dynamic branch, src/main.jl:297-344, on caller
    arrays) from a mid-event state: bit for bit the loop's own next step, and
    the oracle's Newmark step from the same state to 1e-10."""
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(32, 8, material="trapezoid", total_years=3))
    ctx = EqContext(P, amg=True)
    ctx.set_solver(tol=1e-12, pc="amg")
    ctx.state_init()
    ndyn = 0
    for _ in range(800):
        info = ctx.step()
        ndyn += info.isolver_used == 2
        if ndyn >= 20 and info.isolver_next == 2:
            break
    s = ctx.state_get()
    assert s["isolver"] == 2 and ndyn >= 20
    rc, out = ctx.newmark_step(s["d"], s["v"], s["a"], s["psi"], dt=s["dt"])
    assert rc == 0
    info = ctx.step()
    assert info.isolver_used == 2
    s2 = ctx.state_get()
    j0 = P.NFBC - 1
    for k in ("d", "v", "a"):
        assert np.array_equal(out[k], s2[k]), k
    assert np.array_equal(out["psi"][j0:], s2["psi"][j0:])
    assert np.array_equal(out["tau"], s2["tau"])
    sim = oracle.Simulation(P, solver="direct")
    sim.d, sim.v, sim.a = s["d"].copy(), s["v"].copy(), s["a"].copy()
    sim.psi, sim.dt = s["psi"].copy(), s["dt"]
    sim.dyn_step()
    for k in ("d", "v"):
        assert rel(out[k], getattr(sim, k)) < 1e-10, k
    assert rel(out["a"], sim.a) < 1e-9
    assert rel(out["psi"][j0:], sim.psi[j0:]) < 1e-10
    assert rel(out["tau"], sim.tau) < 1e-10
    assert rel(out["FltVfree"], sim.FltVfree) < 1e-10

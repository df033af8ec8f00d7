"""GPU parity at the benchmark configurations' full sizes (configs 1-5), each
against the CPU oracle on the same mesh.

* config 1 mesh (40 x 40): K u in FP64 and FP32;
* config 2 (160 x 160, 410,881 nodes, trapezoid damage zone): K u in FP64
  and FP32 against the oracle's assembled K, the p-multigrid CG against a
  direct solve, and the time loop across the onset of the first event
  against the oracle's direct-solve loop from the same state;
* config 3 (320 x 320, 1,640,961 nodes): the mixed-precision solves (FP32
  inner CG with FP64 refinement, FP32 V-cycle) to 1e-8, checked with the
  oracle's FP64 SpMV residual;
* config 4 (the config-2 mesh, learned start): 2,000 snapshots, the
  64-vector POD basis and the network's gradient against the FP64 CPU
  network, the Galerkin start's iteration gain;
* config 5 (640 x 320, 3,280,641 nodes, Gaussian + lognormal medium): K u,
  the FP32-V-cycle CG to 1e-6 checked with the oracle's SpMV residual, and
  FBC! in fast slip (38 m/s at the surface node) against the oracle.
"""
import numpy as np
import pytest

from eqcycle_amd.configs import CONFIGS
from eqcycle_amd.params import normal, set_parameters

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


def _fault_rhs(P, K, seed):
    """rhs = (K F)[FltNI] for a smooth random slip F on the fault (src/main.jl:237-246)."""
    F = np.zeros(P.nglob)
    x = np.linspace(0, 1, P.FltNglob)
    r = normal(seed, 4)
    F[P.iFlt - 1] = r[0] + r[1] * np.sin(3 * np.pi * x) + r[2] * np.cos(7 * np.pi * x) + 0.1 * r[3] * x
    return K.matvec(F)[P.FltNI - 1]


def test_config1_mesh_ku_fp64_and_fp32(oracle):
    """K u on the config-1 mesh (40 x 40, 25,921 nodes): FP64 to 1e-13 and FP32
    to 1e-6 of the oracle's assembled product, for the full operator and for
    the element row along the fault (src/main.jl:246,267)."""
    from eqcycle_amd.api import EqContext
    P = set_parameters(CONFIGS["cfg1"])
    ctx = EqContext(P)
    K = oracle.assemble_K(P.NelX, P.NelY, P.nglob, P.iglob, P.W, P.H, P.dxe, P.dye)
    u = normal(3, P.nglob)
    y = K.matvec(u)
    e64, e32 = rel(ctx.ku(u), y), rel(ctx.ku(u, prec="fp32"), y)
    rows = P.y <= P.dye * (1 + 1e-9)
    e32f = rel(ctx.ku(u, mask="fault_rows", prec="fp32")[rows], y[rows])
    print(f"[gpu] cfg1 K u fp64 {e64:.2e} fp32 {e32:.2e} fault rows fp32 {e32f:.2e}", flush=True)
    assert e64 < 1e-13 and e32 < 1e-6 and e32f < 1e-6


def test_config2_full_size_products_solve_and_event_onset(oracle):
    from eqcycle_amd.api import EqContext
    P = set_parameters(CONFIGS["cfg2"])
    print(f"[gpu] cfg2 nglob={P.nglob}", flush=True)
    ctx = EqContext(P, amg=True)
    sim = oracle.Simulation(P, tol=1e-12, solver="direct")
    K = sim.K
    u = normal(3, P.nglob)
    y = K.matvec(u)
    assert rel(ctx.ku(u), y) < 1e-13
    assert rel(ctx.ku(u, prec="fp32"), y) < 1e-6
    rhs = _fault_rhs(P, K, 4)
    x, it, rr, ok = ctx.cg(np.zeros(P.nglob - P.FltNglob), rhs, 1e-12)
    xs = sim.lu.solve(rhs)
    print(f"[gpu] cfg2 CG {it} iterations, relres {rr:.2e}, error {rel(x, xs):.2e}", flush=True)
    assert ok and rel(x, xs) < 1e-8
    # the loop across the first event's onset: GPU from the seeded state to 60
    # steps before its first dynamic step, then GPU and oracle side by side
    ctx.set_solver(tol=1e-12, pc="amg")
    ctx.state_init()
    R = ctx.run(max_steps=4000, snapshots=False)
    i_dyn = int(np.argmax(R["isolver"] == 2))
    assert R["isolver"][i_dyn] == 2 and i_dyn > 60
    ctx.state_init()
    ctx.run(max_steps=i_dyn - 60, snapshots=False)
    s = ctx.state_get()
    sim.d, sim.v, sim.a = s["d"].copy(), s["v"].copy(), s["a"].copy()
    sim.psi, sim.tau = s["psi"].copy(), s["tau"].copy()
    sim.t, sim.dt, sim.isolver = s["t"], s["dt"], s["isolver"]
    kinds_g, kinds_o = [], []
    for _ in range(360):
        kinds_g.append(ctx.step().isolver_used)
        kinds_o.append(sim.isolver)
        sim.step()
    g = ctx.state_get()
    print(f"[gpu] cfg2 onset window: {kinds_g.count(2)} dynamic steps, d {rel(g['d'], sim.d):.2e}, "
          f"v {rel(g['v'], sim.v):.2e}", flush=True)
    assert kinds_g == kinds_o and kinds_g.count(2) > 200
    assert abs(g["t"] - sim.t) <= 1e-12 * sim.t
    assert rel(g["d"], sim.d) < 1e-8
    assert rel(g["v"], sim.v) < 1e-6
    assert rel(g["psi"][P.NFBC - 1:], sim.psi[P.NFBC - 1:]) < 1e-8


def test_config3_full_size_mixed_precision_solves(oracle):
    from eqcycle_amd.api import EqContext
    P = set_parameters(CONFIGS["cfg3"])
    print(f"[gpu] cfg3 nglob={P.nglob}", flush=True)
    ctx = EqContext(P, amg=True)
    K = oracle.assemble_K(P.NelX, P.NelY, P.nglob, P.iglob, P.W, P.H, P.dxe, P.dye)
    u = normal(3, P.nglob)
    y = K.matvec(u)
    assert rel(ctx.ku(u), y) < 1e-13
    assert rel(ctx.ku(u, prec="fp32"), y) < 1e-6
    kni = K.submatrix(P.FltNI - 1, -1.0)
    rhs = _fault_rhs(P, K, 5)
    nb = np.linalg.norm(rhs)
    xs = {}
    for prec in ("fp64", "mixed", "f32pc"):
        x, it, rr, ok = ctx.cg(np.zeros(kni.n), rhs, 1e-8, prec=prec)
        true = np.linalg.norm(rhs - kni.matvec(x)) / nb
        print(f"[gpu] cfg3 {prec}: {it} iterations, reported {rr:.2e}, oracle residual {true:.2e}", flush=True)
        assert ok and true <= 1.05e-8
        xs[prec] = x
    assert rel(xs["mixed"], xs["fp64"]) < 1e-5 and rel(xs["f32pc"], xs["fp64"]) < 1e-5


def test_config4_learned_path_full_size():
    from eqcycle_amd import learned
    from eqcycle_amd.api import EqContext
    from oracle.learned_ref import NetRef
    cfg = CONFIGS["cfg4"]
    P = set_parameters(cfg)
    ctx = EqContext(P, amg=True)
    ctx.set_solver(tol=cfg.pcg_tol, pc="amg")
    ctx.state_init()
    ctx.learn_init(capacity=cfg.n_snapshots, nbasis=64)
    ctx.learn_record(True)
    while ctx.learn_info()["nsnap"] < cfg.n_snapshots:
        ctx.run(total_time=1e30, max_steps=400, snapshots=False)
    ctx.learn_record(False)
    fit = learned.fit_basis(ctx, tol=cfg.pcg_tol)
    print(f"[gpu] cfg4 {cfg.n_snapshots} snapshots, {fit['used']} used, energy {fit['energy_captured']:.8f}",
          flush=True)
    data = ctx.learn_get()
    Phi = data["Phi"].astype(np.float64)
    assert Phi.shape == (64, P.nglob - P.FltNglob)       # rows: basis vectors
    assert np.max(np.abs(Phi @ Phi.T - np.eye(64))) < 1e-3
    net = NetRef(P.FltNglob, C=32, NL=4, H=256, NB=64)
    assert net.nparams == len(data["params"])
    idx = np.arange(0, 128, 8, dtype=np.int32)
    loss_g, grad_g = ctx.learn_loss_grad(data["params"].copy(), idx, use_bf16=False)
    loss_r, grad_r = net.loss_grad(data["params"].astype(np.float64), data["X"][idx].astype(np.float64),
                                   data["Y"][idx].astype(np.float64))
    err = np.linalg.norm(grad_g - grad_r) / np.linalg.norm(grad_r)
    print(f"[gpu] cfg4 gradient rel err {err:.2e}", flush=True)
    assert abs(loss_g - loss_r) <= 1e-5 * abs(loss_r) and err < 1e-4
    # the Galerkin start on the learned basis against the previous-step guess,
    # 60 steps from the same state at the bench tolerance
    s0 = ctx.state_get()
    its = {}
    for guess in ("previous", "galerkin"):
        ctx.state_set(d=s0["d"], v=s0["v"], a=s0["a"], psi=s0["psi"], tau=s0["tau"], t=s0["t"], dt=s0["dt"],
                      isolver=s0["isolver"])
        ctx.set_solver(tol=cfg.pcg_tol, pc="amg", guess=guess)
        its[guess] = sum(sum(ctx.step().pcg_iters) for _ in range(60))
    print(f"[gpu] cfg4 iterations over 60 steps: {its}", flush=True)
    assert its["galerkin"] < its["previous"]


def test_config5_full_size_products_solve_and_fast_slip(oracle):
    from eqcycle_amd.api import EqContext
    P = set_parameters(CONFIGS["cfg5"])
    print(f"[gpu] cfg5 nglob={P.nglob}", flush=True)
    ctx = EqContext(P, amg=True)
    K = oracle.assemble_K(P.NelX, P.NelY, P.nglob, P.iglob, P.W, P.H, P.dxe, P.dye)
    u = normal(3, P.nglob)
    assert rel(ctx.ku(u), K.matvec(u)) < 1e-13
    kni = K.submatrix(P.FltNI - 1, -1.0)
    rhs = _fault_rhs(P, K, 6)
    x, it, rr, ok = ctx.cg(np.zeros(kni.n), rhs, 1e-6, prec="f32pc")
    true = np.linalg.norm(rhs - kni.matvec(x)) / np.linalg.norm(rhs)
    print(f"[gpu] cfg5 CG {it} iterations, oracle residual {true:.2e}", flush=True)
    assert ok and true <= 1.05e-6
    # FBC! in fast slip: free slip velocities up to 38 m/s at the surface node,
    # where the first Newton step from tau = 0 lands ~700 a*sigma above the root
    n = P.FltNglob
    depth = np.abs(P.FltX)
    Vfree = 38.0 * np.exp(-(depth / 6e3) ** 2)
    psi = np.full(n, 5.0)
    Vf = 0.5 * Vfree
    dt = P.dt0
    outs_o = [np.zeros(n) for _ in range(6)]
    st_o, _ = oracle.fbc(2, P, P.NFBC, n, *outs_o, psi, Vf, Vfree, dt)
    rc, fail, outs_g = ctx.fbc(2, P.NFBC, n, *[np.zeros(n) for _ in range(6)], psi, Vf, Vfree, dt)
    assert st_o == 0 and fail == 0
    for a, b in zip(outs_g, outs_o):
        assert rel(a, b) < 1e-11
    j = n - 1
    _, _, k = oracle.nrsearch(P.fo[j], P.Vo[j], P.cca[j], P.ccb[j], P.Seff[j], 0.0, P.tauo[j], outs_o[0][j],
                              P.FltZ[j], Vfree[j])
    assert k > 500 and np.isfinite(outs_g[2][j]) and outs_g[1][j] > 30.0

"""GPU checks of the learned initial guess: POD basis, gradients against the
FP64 reference network, BF16 training, and what the network gains over the
previous-step guess."""
import numpy as np
import pytest

from eqcycle_amd.configs import small
from eqcycle_amd.params import set_parameters

pytestmark = pytest.mark.gpu

NB = 24


@pytest.fixture(scope="module")
def trained():
    from eqcycle_amd import learned
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(32, 16, material="trapezoid", total_years=50))
    ctx = EqContext(P, amg=True)
    ctx.set_solver(tol=1e-10, pc="amg")
    ctx.state_init()
    learned.record(ctx, max_steps=160, capacity=200, nbasis=NB)
    info = ctx.learn_info()
    print(f"[gpu] recorded {info['nsnap']} snapshots, {info['nparams']} parameters", flush=True)
    fit = learned.fit_basis(ctx, tol=1e-10)
    print(f"[gpu] POD energy captured {fit['energy_captured']:.6f}", flush=True)
    # a quasi-static state for the tests that compare starts (the tests before
    # them may leave the context inside an event, where no solve runs)
    fit["qs_state"] = ctx.state_get()
    assert fit["qs_state"]["isolver"] == 1
    return P, ctx, fit


def test_pod_basis_is_orthonormal(trained):
    P, ctx, fit = trained
    Phi = ctx.learn_get()["Phi"].astype(np.float64)
    G = Phi @ Phi.T
    assert np.max(np.abs(G - np.eye(NB))) < 1e-3
    assert fit["energy_captured"] > 0.99


def test_gram_on_matrix_cores_matches_fp64_product(trained):
    """eq_learn_gram (FP64 MFMA over the widened FP32 snapshots) against the
    FP64 product S^T S: every product is exact, so only the summation order
    differs; the snapshot count spans several 64-column tiles, so off-diagonal
    tiles and the operand layout are exercised."""
    P, ctx, fit = trained
    S = ctx.learn_snapshots().astype(np.float64)
    G = ctx.learn_gram()
    ref = S @ S.T
    print(f"[gpu] Gram {G.shape}, max rel diff {np.max(np.abs(G - ref)) / np.max(np.abs(ref)):.2e}", flush=True)
    assert S.shape[0] > 64
    assert np.max(np.abs(G - ref)) <= 1e-13 * np.max(np.abs(ref))
    assert np.array_equal(G, G.T)


def test_gpu_gradient_matches_fp64_reference(trained):
    from oracle.learned_ref import NetRef
    P, ctx, _ = trained
    data = ctx.learn_get()
    net = NetRef(P.FltNglob, C=32, NL=4, H=256, NB=NB)
    assert net.nparams == len(data["params"])
    idx = np.arange(0, 64, 2, dtype=np.int32)
    params = data["params"].copy()
    loss_g, grad_g = ctx.learn_loss_grad(params, idx, use_bf16=False)
    loss_r, grad_r = net.loss_grad(params.astype(np.float64), data["X"][idx].astype(np.float64),
                                   data["Y"][idx].astype(np.float64))
    assert abs(loss_g - loss_r) <= 1e-5 * abs(loss_r)
    err = np.linalg.norm(grad_g - grad_r) / np.linalg.norm(grad_r)
    print(f"[gpu] gradient rel err fp32 {err:.2e}", flush=True)
    assert err < 1e-4
    # BF16 weights: same loss to bf16 accuracy
    loss_b, grad_b = ctx.learn_loss_grad(None, idx, use_bf16=True)
    assert abs(loss_b - loss_r) <= 3e-2 * abs(loss_r)
    assert np.linalg.norm(grad_b - grad_r) / np.linalg.norm(grad_r) < 5e-2
    # and exactly the FP64 network at the BF16-rounded weights (biases stay FP32)
    pw = net.unpack(params)
    for name, _ in net.layout:
        if not name.startswith("b") and not name.endswith("b"):
            pw[name] = _bf16(pw[name].astype(np.float32)).astype(np.float64)
    loss_rb, grad_rb = net.loss_grad(net.pack(pw), data["X"][idx].astype(np.float64),
                                     data["Y"][idx].astype(np.float64))
    errb = np.linalg.norm(grad_b - grad_rb) / np.linalg.norm(grad_rb)
    print(f"[gpu] gradient rel err bf16 weights {errb:.2e}", flush=True)
    assert abs(loss_b - loss_rb) <= 1e-5 * abs(loss_rb) and errb < 1e-4


def test_training_and_learned_guess_gain(trained):
    from eqcycle_amd import learned
    P, ctx, _ = trained
    hist = learned.train(ctx, epochs=300, batch=32, lr=1e-3)
    print(f"[gpu] loss {hist[0]:.3e} -> {hist[-1]:.3e}", flush=True)
    assert hist[-1] < 0.05 * hist[0]
    s0 = ctx.state_get()
    # plain previous-step guess
    ctx.set_solver(tol=1e-10, pc="amg", guess="previous")
    plain = [ctx.step().pcg_iters[0] for _ in range(40)]
    d_plain = ctx.state_get()["d"]
    # same 40 steps from the same state with the learned guess
    ctx.state_set(d=s0["d"], v=s0["v"], a=s0["a"], psi=s0["psi"], tau=s0["tau"], t=s0["t"], dt=s0["dt"],
                  isolver=s0["isolver"])
    learned.enable(ctx, tol=1e-10)
    lrn = [ctx.step().pcg_iters[0] for _ in range(40)]
    d_lrn = ctx.state_get()["d"]
    print(f"[gpu] pass-1 iterations plain {sum(plain)} learned {sum(lrn)}", flush=True)
    assert sum(lrn) < sum(plain)
    assert np.max(np.abs(d_lrn - d_plain)) <= 1e-7 * np.max(np.abs(d_plain))


def test_galerkin_start_is_energy_optimal(trained, oracle):
    """x_out = x + Phi (Phi^T K_II Phi)^-1 Phi^T (b - K_II x): applying it again
    changes nothing (Phi^T r_out = 0), and the K_II-norm error never grows."""
    P, ctx, _ = trained
    K = oracle.assemble_K(P.NelX, P.NelY, P.nglob, P.iglob, P.W, P.H, P.dxe, P.dye)
    kni = K.submatrix(P.FltNI - 1, -1.0)          # kni = -K_II (src/main.jl:195)
    rng = np.random.default_rng(5)
    n = ctx.nI
    xs = rng.standard_normal(n)
    b = -kni.matvec(xs)                           # K_II xs = b
    x = xs + 0.1 * rng.standard_normal(n)
    x1 = ctx.learn_galerkin(x, b)
    x2 = ctx.learn_galerkin(x1, b)
    # rounding of the residual is amplified by cond(Phi^T K_II Phi); a wrong
    # inverse would leave a correction of the size of the first one
    assert np.max(np.abs(x2 - x1)) <= 1e-6 * np.max(np.abs(x1 - x))

    def enorm(e):
        return float(np.sqrt(-e @ kni.matvec(e)))
    assert enorm(x1 - xs) <= enorm(x - xs) * (1 + 1e-12)


def test_galerkin_guess_gain(trained):
    """Same 40 quasi-static steps with the Galerkin start on the POD basis
    (both passes) against the previous-step guess: fewer CG iterations, same
    displacements to the solver tolerance."""
    P, ctx, _ = trained
    s0 = ctx.state_get()
    ctx.set_solver(tol=1e-10, pc="amg", guess="previous")
    plain = [sum(ctx.step().pcg_iters) for _ in range(40)]
    d_plain = ctx.state_get()["d"]
    ctx.state_set(d=s0["d"], v=s0["v"], a=s0["a"], psi=s0["psi"], tau=s0["tau"], t=s0["t"], dt=s0["dt"],
                  isolver=s0["isolver"])
    ctx.set_solver(tol=1e-10, pc="amg", guess="galerkin")
    gal = [sum(ctx.step().pcg_iters) for _ in range(40)]
    d_gal = ctx.state_get()["d"]
    print(f"[gpu] CG iterations (both passes) plain {sum(plain)} galerkin {sum(gal)}", flush=True)
    assert sum(gal) < sum(plain)
    assert np.max(np.abs(d_gal - d_plain)) <= 1e-7 * np.max(np.abs(d_plain))


@pytest.mark.parametrize("prec", ["fp64", "f32pc"])
def test_galerkin_change_of_b_start_matches_full_projection(trained, prec):
    """When the previous solve left d solving K_II d = b_old to tolerance and
    only rows 1..4 of b have changed, the Galerkin start projects just that
    change of b (4 NX rows of Phi) instead of the full residual.  The same 40
    steps with the full residual and projection in every solve: iteration
    counts within 1 % and displacements that agree far below the solver
    tolerance."""
    P, ctx, _ = trained
    s0 = ctx.state_get()

    def run(shortcut):
        ctx.state_set(d=s0["d"], v=s0["v"], a=s0["a"], psi=s0["psi"], tau=s0["tau"], t=s0["t"], dt=s0["dt"],
                      isolver=s0["isolver"])
        ctx.set_solver(tol=1e-10, pc="amg", prec=prec, guess="galerkin", galerkin_shortcut=shortcut)
        its = [sum(ctx.step().pcg_iters) for _ in range(40)]
        return its, ctx.state_get()["d"]

    it_a, d_a = run(True)
    it_b, d_b = run(False)
    print(f"[gpu] {prec}: iterations change-of-b start {np.sum(it_a)}, full projection {np.sum(it_b)}", flush=True)
    assert abs(np.sum(it_a) - np.sum(it_b)) <= 0.01 * np.sum(it_b)
    assert np.max(np.abs(d_a - d_b)) <= 1e-9 * np.max(np.abs(d_b))


@pytest.mark.parametrize("tol", [1e-6, 1e-10])
def test_fp16_basis_expansion(trained, tol):
    """The Galerkin expansion x += Phi c from the FP16 copy of the basis
    (eq_solver_opts.basis_fp16): 40 steps from the same state with and
    without it need the same CG iterations within 5 %, and, since each solve
    still runs to its tolerance, the displacements agree to that tolerance
    (the expansion's rounding, 5e-4 of the increment, is what CG removes)."""
    P, ctx, _ = trained
    s0 = ctx.state_get()

    def run(fp16):
        ctx.state_set(d=s0["d"], v=s0["v"], a=s0["a"], psi=s0["psi"], tau=s0["tau"], t=s0["t"], dt=s0["dt"],
                      isolver=s0["isolver"])
        ctx.set_solver(tol=tol, pc="amg", prec="f32pc", guess="galerkin", basis_fp16=fp16)
        its = [sum(ctx.step().pcg_iters) for _ in range(40)]
        return its, ctx.state_get()["d"]

    it_a, d_a = run(False)
    it_b, d_b = run(True)
    print(f"[gpu] tol {tol:g}: iterations FP32 basis {np.sum(it_a)}, FP16 basis {np.sum(it_b)}, "
          f"d rel diff {np.max(np.abs(d_a - d_b)) / np.max(np.abs(d_a)):.2e}", flush=True)
    assert np.sum(it_b) <= 1.05 * np.sum(it_a) + 1
    assert np.max(np.abs(d_a - d_b)) <= 20 * tol * np.max(np.abs(d_a))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("f32pc", 1e-6)])
def test_extrapolated_galerkin_start(trained, prec, tol):
    """The first pass's Galerkin start from the last increment extrapolated,
    d + (dt/dt_prev)(d - d_prev) (eq_solver_opts.extrapolate): the same 40
    steps from the same state without it need more first-pass iterations,
    and the displacements agree to the solver tolerance.  With it, projecting
    the change of b less (dt/dt_prev) times the last step's (the start then
    solves the extrapolated b off rows 1..4) gives the start that projecting
    the full residual of the extrapolated displacements gives: iteration
    counts within 1 % (and one iteration) and displacements to 10 tol.  A wrong
    sign or factor in either the displacement or the b extrapolation breaks
    the first or the second comparison."""
    P, ctx, fit = trained
    s0 = fit["qs_state"]

    def run(extrapolate, shortcut=True):
        ctx.state_set(d=s0["d"], v=s0["v"], a=s0["a"], psi=s0["psi"], tau=s0["tau"], t=s0["t"], dt=s0["dt"],
                      isolver=s0["isolver"])
        ctx.set_solver(tol=tol, pc="amg", prec=prec, guess="galerkin", galerkin_shortcut=shortcut,
                       extrapolate=extrapolate)
        its = np.array([ctx.step().pcg_iters for _ in range(40)])
        return its, ctx.state_get()["d"]

    it_0, d_0 = run(False)
    it_x, d_x = run(True)
    it_f, d_f = run(True, shortcut=False)
    print(f"[gpu] {prec}: first-pass iterations from d {it_0[:, 0].sum()}, extrapolated {it_x[:, 0].sum()}, "
          f"extrapolated with the full residual {it_f[:, 0].sum()}", flush=True)
    # the first two steps cannot extrapolate (no consecutive solves yet)
    assert np.array_equal(it_x[:2], it_0[:2])
    assert it_x[:, 0].sum() < it_0[:, 0].sum()
    assert it_0[:, 0].min() > 0   # every step solved
    assert np.max(np.abs(d_x - d_0)) <= 20 * tol * np.max(np.abs(d_0))
    assert abs(it_x.sum() - it_f.sum()) <= 0.01 * it_f.sum() + 1
    assert np.max(np.abs(d_x - d_f)) <= 10 * tol * np.max(np.abs(d_f))


@pytest.mark.parametrize("prec", ["fp64", "f32pc"])
def test_fused_galerkin_start_is_bit_identical(trained, prec):
    """The change-of-b Galerkin start from the FP16 basis with the start
    residual in the same launch (eq_solver_opts.galerkin_split = 0: the window
    of the residual's tiles loads d + Phi c, summed in the order of the
    separate expansion, and x lands in the second displacement buffer) against
    the expansion and the residual in two launches: 40 steps from the same
    state, with the extrapolated start, give the same iteration counts and the
    same state bit for bit.  The fused form also sends the second pass's
    begin and start with the first solve's batch, gated on its convergence
    test, and enqueues them again when the batch falls short (a first solve
    that needs more iterations than the step before it): both cases occur in
    these 40 steps."""
    P, ctx, fit = trained
    s0 = fit["qs_state"]

    def run(fused):
        ctx.state_set(d=s0["d"], v=s0["v"], a=s0["a"], psi=s0["psi"], tau=s0["tau"], t=s0["t"], dt=s0["dt"],
                      isolver=s0["isolver"])
        ctx.set_solver(tol=1e-8, pc="amg", prec=prec, guess="galerkin", basis_fp16=True, extrapolate=True,
                       galerkin_fused=fused)
        its = np.array([ctx.step().pcg_iters for _ in range(40)])
        return its, ctx.state_get()

    it_s, st_s = run(False)
    it_f, st_f = run(True)
    # a first solve that takes more iterations than the step before's first
    # solve outruns its batch (the first step starts unfused: not counted)
    short = int(np.sum(it_f[1:, 0] > it_f[:-1, 0]))
    print(f"[gpu] {prec}: iterations split {it_s.sum()}, fused {it_f.sum()}, short first batches {short}",
          flush=True)
    assert it_s[:, 0].min() > 0 and it_s[:, 1].min() > 0   # every step solved both passes
    assert short > 0   # the re-enqueue after a short first batch ran
    assert np.array_equal(it_s, it_f)
    for k in ("d", "v", "psi", "tau"):
        assert np.array_equal(st_s[k], st_f[k]), k
    assert st_s["t"] == st_f["t"] and st_s["dt"] == st_f["dt"]


@pytest.mark.parametrize("prec", ["fp64", "f32pc"])
def test_run_sends_next_begin_ahead_bit_identical(trained, prec):
    """Inside run(), the end of a quasi-static step sends the next step's
    begin and Galerkin start before the host has read the step's status, and
    they do nothing unless the status lets the next step go (quasi-static, the
    searches sound, the second solve converged in its first batch).  40 steps
    of run() against 40 calls of step(), which send nothing ahead; then 400
    steps of run() through an event (the last quasi-static step before it
    sends a begin that its status stops) against run() with early_tail=False
    (every read-back waited for): the same iteration counts and the same
    state bit for bit."""
    P, ctx, fit = trained
    s0 = fit["qs_state"]

    def reset(early=True):
        ctx.state_set(d=s0["d"], v=s0["v"], a=s0["a"], psi=s0["psi"], tau=s0["tau"], t=s0["t"], dt=s0["dt"],
                      isolver=s0["isolver"])
        ctx.set_solver(tol=1e-8, pc="amg", prec=prec, guess="galerkin", basis_fp16=True, extrapolate=True,
                       early_tail=early)

    def same(a, b):
        for k in ("d", "v", "a", "psi", "tau"):
            assert np.array_equal(a[k], b[k]), k
        assert a["t"] == b["t"] and a["dt"] == b["dt"] and a["isolver"] == b["isolver"]

    reset()
    its = np.array([ctx.step().pcg_iters for _ in range(40)])
    st_step = ctx.state_get()
    reset()
    out = ctx.run(max_steps=40)
    print(f"[gpu] {prec}: 40 steps, iterations step() {its.sum()}, run() {out['pcg_iters'].sum()}", flush=True)
    assert out["n_qs"] == 40 and its[:, 1].min() > 0
    assert np.array_equal(out["pcg_iters"], its)
    same(ctx.state_get(), st_step)

    reset(early=True)
    a = ctx.run(max_steps=400)
    sa = ctx.state_get()
    reset(early=False)
    b = ctx.run(max_steps=400)
    print(f"[gpu] {prec}: 400 steps, {a['n_qs']} quasi-static, {a['n_dyn']} dynamic, iterations "
          f"{a['total_pcg_iters']} / {b['total_pcg_iters']}", flush=True)
    assert a["n_dyn"] > 0 and a["n_qs"] > 40
    assert np.array_equal(a["pcg_iters"], b["pcg_iters"])
    assert np.array_equal(a["time_"], b["time_"])
    same(sa, ctx.state_get())


def _bf16(x):
    """Round FP32 values to BF16 (nearest even) and back, as the kernels store weights."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def test_one_launch_network_inference_matches_fp64_reference(trained):
    """The network's whole forward in one launch (eq_learn_predict): |dF| times
    the FP64 reference network on dF/|dF|, with the weights rounded to BF16 as
    the kernels hold them (biases FP32), to FP32 accumulation accuracy."""
    from oracle.learned_ref import NetRef
    P, ctx, _ = trained
    data = ctx.learn_get()
    net = NetRef(P.FltNglob, C=32, NL=4, H=256, NB=NB)
    p = data["params"].astype(np.float64)
    u = net.unpack(p)
    for k in u:
        if "w" in k:
            u[k] = _bf16(u[k]).astype(np.float64)
    pb = net.pack(u)
    rng = np.random.default_rng(8)
    for _ in range(3):
        dF = rng.standard_normal(P.FltNglob) * 1e-3
        got = ctx.learn_predict(dF)
        n = np.linalg.norm(dF)
        want = n * net.forward(pb, (dF / n)[None, :])[0]
        assert np.linalg.norm(got - want) <= 1e-5 * np.linalg.norm(want)


def test_linear_predictor_against_network_and_previous_guess(trained):
    """The linear least-squares predictor of the POD coefficients from the
    network's input (the bar the network must beat): its one-launch device
    output equals |dF| (A dF/|dF| + b) computed here; and over the same 40
    steps from the same state, the pass-1 iterations of the previous-step
    guess, the network start and the linear start (both energy-safeguarded,
    so neither needs more iterations than the guess it corrects)."""
    from eqcycle_amd import learned
    P, ctx, fit = trained
    lin = learned.fit_linear(ctx, fit["use"])
    print(f"[gpu] linear fit: {lin}", flush=True)
    data = ctx.learn_get()
    X, Y = data["X"][fit["use"]].astype(np.float64), data["Y"][fit["use"]].astype(np.float64)
    Xa = np.hstack([X, np.ones((X.shape[0], 1))])
    beta = np.linalg.lstsq(Xa, Y, rcond=None)[0]
    A, b = beta[:-1].T.astype(np.float32).astype(np.float64), beta[-1].astype(np.float32).astype(np.float64)
    dF = np.random.default_rng(9).standard_normal(P.FltNglob) * 1e-3
    n = np.linalg.norm(dF)
    want = n * (A @ (dF / n).astype(np.float32).astype(np.float64) + b)
    got = ctx.learn_predict_linear(dF)
    assert np.linalg.norm(got - want) <= 1e-5 * np.linalg.norm(want)
    s0 = fit["qs_state"]
    its, dfin = {}, {}
    for guess in ("previous", "learned", "linear"):
        ctx.state_set(d=s0["d"], v=s0["v"], a=s0["a"], psi=s0["psi"], tau=s0["tau"], t=s0["t"], dt=s0["dt"],
                      isolver=s0["isolver"])
        ctx.set_solver(tol=1e-10, pc="amg", guess=guess)
        its[guess] = sum(ctx.step().pcg_iters[0] for _ in range(40))
        dfin[guess] = ctx.state_get()["d"]
    print(f"[gpu] pass-1 iterations over 40 steps: {its}", flush=True)
    assert its["previous"] > 0
    assert its["linear"] <= its["previous"] and its["learned"] <= its["previous"]
    dref = np.max(np.abs(dfin["previous"]))
    assert np.max(np.abs(dfin["linear"] - dfin["previous"])) <= 1e-7 * dref



@pytest.mark.parametrize("nb", [16, 32, 64])
def test_full_group_galerkin_start_is_bit_identical(nb):
    """The fused start's variants without a basis-index test (16, 32 and 64
    vectors: two, four and eight 8-vector groups) against the split expansion and
    residual: 20 steps from the same state give the same iteration counts and
    the same state bit for bit."""
    from eqcycle_amd import learned
    from eqcycle_amd.api import EqContext
    P = set_parameters(small(32, 16, material="trapezoid", total_years=50))
    ctx = EqContext(P, amg=True)
    ctx.set_solver(tol=1e-10, pc="amg")
    ctx.state_init()
    learned.record(ctx, max_steps=120, capacity=160, nbasis=nb)
    learned.fit_basis(ctx, tol=1e-10)
    s0 = ctx.state_get()
    assert s0["isolver"] == 1

    def run(fused):
        ctx.state_set(d=s0["d"], v=s0["v"], a=s0["a"], psi=s0["psi"], tau=s0["tau"], t=s0["t"], dt=s0["dt"],
                      isolver=s0["isolver"])
        ctx.set_solver(tol=1e-8, pc="amg", prec="f32pc", guess="galerkin", basis_fp16=True, extrapolate=True,
                       galerkin_fused=fused)
        its = np.array([ctx.step().pcg_iters for _ in range(20)])
        return its, ctx.state_get()

    it_s, st_s = run(False)
    it_f, st_f = run(True)
    print(f"[gpu] nb {nb}: iterations split {it_s.sum()}, fused {it_f.sum()}", flush=True)
    assert it_s[:, 0].min() > 0 and it_s[:, 1].min() > 0
    assert np.array_equal(it_s, it_f)
    for k in ("d", "v", "psi", "tau"):
        assert np.array_equal(st_s[k], st_f[k]), k

"""Host-side mirror of the reference routines on the hot path, backed by the HIP library.

Each function keeps the reference's name (``!`` dropped), argument order and
meaning, and its error behaviour: where the reference logs ``@error`` and
returns ``nothing`` (src/NRsearch_serial.jl:19-32, src/PCG.jl:55-63) the
mirror returns ``None``; anything else that fails raises :class:`EqError`.

    ctx = EqContext(P)                       # P from eqcycle_amd.params.set_parameters
    psi1, Vf1, tau1, psi2, Vf2, tau2 = FBC(IDstate, P, NFBC, ...)   # on the current context
    dnew = cg(ctx, dnew, kni_rhs, tol=1e-6)  # cg!(dnew, kni, rhs; Pl=p, tol)

FBC, NRsearch, IDS, IDS2, slrFunc and dtevol take exactly the reference's
arguments and run on the current context (the last EqContext created, or the
one given to use_context); FBC and slrFunc read the fault parameters from
their P argument on every call.
"""
from __future__ import annotations

import ctypes as C
import logging
from dataclasses import dataclass

import numpy as np

from . import lib as L
from .lib import f64, i64, ptr

log = logging.getLogger("eqcycle_amd")

PREC = {"fp64": L.EQ_PREC_FP64, "fp32": L.EQ_PREC_FP32, "mixed": L.EQ_PREC_MIXED, "f32pc": L.EQ_PREC_F32PC}
SMOOTHERS = {"jacobi": L.EQ_SMOOTH_JACOBI, "chebyshev": L.EQ_SMOOTH_CHEBYSHEV,
             "tiles": L.EQ_SMOOTH_CHEBYSHEV_TILES}
PC = {"none": L.EQ_PC_NONE, "jacobi": L.EQ_PC_JACOBI, "amg": L.EQ_PC_AMG}
MASK = {"all": L.EQ_MASK_ALL, "interior": L.EQ_MASK_INTERIOR, "neg": L.EQ_MASK_NEG,
        "fault_rows": L.EQ_MASK_FAULT_ROWS}


class EqError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{L.STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class EqContext:
    """One mesh on one GPU: owns the device buffers (eq_create / eq_destroy)."""

    def __init__(self, P, amg: bool | dict = False):
        self.lib = L.load()
        self.P = P
        self._keep = []
        iglob = i64(P.iglob)
        xg, wg = f64(P.xgll), f64(P.wgll)
        Hc = f64(np.asarray(P.H).T)        # column-major H[i,j]
        self._keep += [iglob, xg, wg, Hc]
        desc = L.MeshDesc(5, P.NelX, P.NelY, P.dxe, P.dye, ptr(iglob, C.c_int64), ptr(xg), ptr(wg), ptr(Hc))
        h = C.c_void_p()
        rc = self.lib.eq_create(C.byref(desc), C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.lib.eq_last_error(h).decode() if h.value else "eq_create failed"
            if h.value:
                self.lib.eq_destroy(h)
                self.h = C.c_void_p()
            raise EqError(rc, msg)
        self._check(self.lib.eq_set_material(self.h, ptr(f64(P.W))))
        self._set_params(P)
        n = [C.c_int64(), C.c_int64(), C.c_int64()]
        self.lib.eq_sizes(self.h, *[C.byref(x) for x in n])
        self.nglob, self.nI, self.nfault = (x.value for x in n)
        if amg:
            self.amg_setup(**(amg if isinstance(amg, dict) else {}))
        use_context(self)

    def _set_params(self, P):
        arrs = {k: f64(getattr(P, k)) for k in ["fo", "Vo", "xLf", "M", "BcLC", "BcTC", "FltB", "FltZ", "FltX",
                                                "cca", "ccb", "Seff", "tauo", "XiLf"]}
        iarrs = {k: i64(getattr(P, k)) for k in ["iFlt", "iBcL", "iBcT", "FltIglobBC", "out_seis"]}
        self._keep += list(arrs.values()) + list(iarrs.values())
        d = L.ParamsDesc()
        d.FltNglob, d.nglob, d.IDstate = P.FltNglob, P.nglob, P.IDstate
        d.Vpl, d.Vthres, d.Vevne, d.dt0 = P.Vpl, P.Vthres, P.Vevne, P.dt0
        for k, v in arrs.items():
            setattr(d, k, ptr(v))
        d.iFlt = ptr(iarrs["iFlt"], C.c_int64)
        d.iBcL, d.n_iBcL = ptr(iarrs["iBcL"], C.c_int64), len(iarrs["iBcL"])
        d.iBcT, d.n_iBcT = ptr(iarrs["iBcT"], C.c_int64), len(iarrs["iBcT"])
        d.FltIglobBC, d.n_FltIglobBC = ptr(iarrs["FltIglobBC"], C.c_int64), len(iarrs["FltIglobBC"])
        d.out_seis, d.n_out_seis = ptr(iarrs["out_seis"], C.c_int64), len(iarrs["out_seis"])
        self._check(self.lib.eq_set_params(self.h, C.byref(d)))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.eq_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self):
        return self.lib.eq_last_error(self.h).decode()

    def _check(self, rc, allow=()):
        if rc != 0 and rc not in allow:
            raise EqError(rc, self._err())
        return rc

    # ---- operators -------------------------------------------------------
    def ku(self, u, mask="all", prec="fp64"):
        """y = K u (mul!(y, Ksparse', u)), reference node numbering."""
        u = f64(u)
        y = np.empty_like(u)
        self._check(self.lib.eq_ku_apply(self.h, ptr(u), ptr(y), MASK[mask], PREC[prec]))
        return y

    def kni(self, x, prec="fp64"):
        """y = kni x with kni = -K[FltNI, FltNI]."""
        x = f64(x)
        y = np.empty_like(x)
        self._check(self.lib.eq_kni_apply(self.h, ptr(x), ptr(y), PREC[prec]))
        return y

    def element_computation(self, u):
        """a_elem[Nel, 5(j), 5(i)] of element_computation! (src/PCG.jl:71-86)."""
        u = f64(u)
        out = np.empty((self.P.Nel, 5, 5))
        self._check(self.lib.eq_element_computation(self.h, ptr(u), ptr(out)))
        return out

    # ---- AMG / CG ----------------------------------------------------------
    def amg_setup(self, hierarchy="pmg", theta=0.25, max_levels=10, max_coarse=None, strength_abs=1,
                  smoother=None, degree=5):
        """hierarchy "rs": ruge_stuben on K_II with Jacobi smoothing (the reference's
        preconditioner structure); "pmg": degree-4 -> bilinear first level with
        Chebyshev smoothing, then Ruge-Stueben (the default, far fewer iterations).
        smoother "tiles" (pmg default): each Chebyshev sweep in one launch per tile
        and its 2-node ring; "chebyshev": the same sweep on the whole mesh, one
        launch per one or two steps; "jacobi"."""
        pc = {"rs": 0, "pmg": 1}[hierarchy]
        if smoother is None:
            smoother = "tiles" if pc else "jacobi"
        if max_coarse is None:
            max_coarse = 2000 if pc else 10
        sm = SMOOTHERS[smoother]
        self._check(self.lib.eq_amg_setup(self.h, theta, max_levels, max_coarse, strength_abs, sm, degree, pc))
        self.hierarchy = hierarchy
        return self.amg_info()

    def amg_coarse_mode(self, collapsed=True):
        """Apply the coarse levels collapsed (one product down, one up) or level
        by level (eq_amg_coarse_mode)."""
        self._check(self.lib.eq_amg_coarse_mode(self.h, 1 if collapsed else 0))

    def amg_l0_info(self):
        """(smoother name, Chebyshev degree, lo, hi): the fine-level smoother and its
        target interval of D^-1 K_II."""
        sm, deg, lo, hi = C.c_int(), C.c_int(), C.c_double(), C.c_double()
        self._check(self.lib.eq_amg_l0_info(self.h, C.byref(sm), C.byref(deg), C.byref(lo), C.byref(hi)))
        name = {v: k for k, v in SMOOTHERS.items()}[sm.value]
        return name, deg.value, lo.value, hi.value

    def amg_info(self):
        nl = C.c_int()
        n = np.zeros(32, np.int64)
        nnz = np.zeros(32, np.int64)
        self._check(self.lib.eq_amg_info(self.h, C.byref(nl), ptr(n, C.c_int64), ptr(nnz, C.c_int64), 32))
        return list(zip(n[:nl.value].tolist(), nnz[:nl.value].tolist()))

    def amg_upload(self, A, P, omega=None, smoother="jacobi", degree=1):
        """Use a host-built hierarchy (ruge_stuben(kni), src/main.jl:200-201):
        A = [A_0 .. A_{L-1}] and P = [P_0 .. P_{L-2}] as scipy sparse matrices,
        level 0 in FltNI order; omega: Jacobi weights per level (None:
        estimated).  See eq_amg_upload."""
        import scipy.sparse as sp
        keep, descs = [], []
        for M in list(A) + list(P):
            M = sp.csr_matrix(M)
            M.sort_indices()
            rp = np.ascontiguousarray(M.indptr, dtype=np.int64)
            cl = np.ascontiguousarray(M.indices, dtype=np.int64)
            vl = np.ascontiguousarray(M.data, dtype=np.float64)
            keep += [rp, cl, vl]
            descs.append(L.CsrDesc(M.shape[0], M.shape[1], ptr(rp, C.c_int64), ptr(cl, C.c_int64), ptr(vl)))
        nA = len(A)
        Aarr = (L.CsrDesc * nA)(*descs[:nA])
        Parr = (L.CsrDesc * max(1, len(P)))(*descs[nA:])
        om = None if omega is None else f64(omega)
        sm = SMOOTHERS[smoother]
        self._check(self.lib.eq_amg_upload(self.h, nA, Aarr, Parr, ptr(om), sm, degree))
        self.amg_levels = nA

    def amg_ldiv(self, b, prec="fp64"):
        b = f64(b)
        x = np.empty_like(b)
        self._check(self.lib.eq_amg_ldiv(self.h, ptr(b), ptr(x), PREC[prec]))
        return x

    def cg(self, x, b, tol, maxiter=0, pc="amg", prec="fp64"):
        """cg!(x, kni, b; Pl, tol, maxiter) on the GPU.  Returns (x, iters, relres, converged)."""
        x = f64(x).copy()
        b = f64(b)
        it = C.c_int64()
        rr = C.c_double()
        rc = self._check(self.lib.eq_cg(self.h, ptr(x), ptr(b), tol, maxiter, PC[pc], PREC[prec],
                                        C.byref(it), C.byref(rr)), allow=(L.EQ_PCG_NOCONV,))
        return x, it.value, rr.value, rc == 0

    def pcg_jacobi(self, F, dnew, tol=1e-5, maxit=8000):
        F = f64(F)
        d = f64(dnew).copy()
        it = C.c_int64()
        rc = self._check(self.lib.eq_pcg_jacobi(self.h, ptr(F), ptr(d), tol, maxit, C.byref(it)),
                         allow=(L.EQ_PCG_NOCONV,))
        return d, it.value, rc == 0

    # ---- fault routines -------------------------------------------------------
    def ids(self, xLf, Vo, psi, dt, Vf, cnd, IDstate=2):
        a = [f64(np.atleast_1d(z)) for z in (xLf, Vo, psi, Vf)]
        out = np.empty_like(a[0])
        self._check(self.lib.eq_ids(self.h, len(out), ptr(a[0]), ptr(a[1]), ptr(a[2]), dt, ptr(a[3]), cnd,
                                    IDstate, ptr(out)))
        return out

    def ids2(self, xLf, Vo, psi, psi1, dt, Vf, Vf1, IDstate=2):
        a = [f64(np.atleast_1d(z)) for z in (xLf, Vo, psi, psi1, Vf, Vf1)]
        out = np.empty_like(a[0])
        self._check(self.lib.eq_ids2(self.h, len(out), ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(a[3]), dt, ptr(a[4]),
                                     ptr(a[5]), IDstate, ptr(out)))
        return out

    def nrsearch(self, fo, Vo, cca, ccb, Seff, tau, tauo, psi, FltZ, FltVfree):
        a = [f64(np.atleast_1d(z)) for z in (fo, Vo, cca, ccb, Seff, tau, tauo, psi, FltZ, FltVfree)]
        n = len(a[0])
        Vf = np.empty(n)
        t = np.empty(n)
        k = np.empty(n, np.int32)
        self._check(self.lib.eq_nrsearch(self.h, n, *[ptr(z) for z in a], ptr(Vf), ptr(t), ptr(k, C.c_int32)),
                    allow=(L.EQ_NR_NOCONV,))
        return Vf, t, k

    @staticmethod
    def _fault_params(P):
        """eq_fault_params of a params_farray-like object P (None: the context's)."""
        if P is None:
            return None, []
        keep = [f64(getattr(P, k)) for k in ("fo", "Vo", "xLf", "cca", "ccb", "Seff", "tauo", "FltZ")]
        return C.byref(L.FaultParams(*[ptr(a) for a in keep])), keep

    def fbc(self, IDstate, NFBC, FltNglob, psi1, Vf1, tau1, psi2, Vf2, tau2, psi, Vf, FltVfree, dt, P=None):
        """FBC! with the fault parameters of P (None: those given at set-up)."""
        outs = [f64(z).copy() for z in (psi1, Vf1, tau1, psi2, Vf2, tau2)]
        ins = [f64(z) for z in (psi, Vf, FltVfree)]
        fail = C.c_int64()
        pp, keep = self._fault_params(P)
        rc = self._check(self.lib.eq_fbc(self.h, IDstate, pp, NFBC, FltNglob, *[ptr(z) for z in outs],
                                         *[ptr(z) for z in ins], dt, C.byref(fail)),
                         allow=(L.EQ_NR_NAN, L.EQ_NR_NOCONV))
        del keep
        return rc, fail.value, outs

    def slr(self, NFBC, FltNglob, psi, psi1, Vf, Vf1, IDstate, tau1, dt, P=None):
        """slrFunc! with the fault parameters of P (None: those given at set-up)."""
        p1 = f64(psi1).copy()
        v1 = f64(Vf1).copy()
        pp, keep = self._fault_params(P)
        self._check(self.lib.eq_slr(self.h, pp, NFBC, FltNglob, ptr(f64(psi)), ptr(p1), ptr(f64(Vf)), ptr(v1),
                                    IDstate, ptr(f64(tau1)), dt))
        del keep
        return p1, v1

    def newmark_step(self, d, v, a, psi, dt=None):
        """The dynamic branch of the loop body (src/main.jl:297-344) on copies
        of d, v, a (reference numbering) and psi; returns (rc, dict with d, v,
        a, psi, tau, Vf, FltVfree after the step).  rc is EQ_NR_NAN when the
        FBC! guard fired (the message names the fault location)."""
        P = self.P
        d, v, a, psi = (f64(z).copy() for z in (d, v, a, psi))
        tau, Vf, Vfree = (np.empty(P.FltNglob) for _ in range(3))
        rc = self._check(self.lib.eq_newmark_step(self.h, ptr(d), ptr(v), ptr(a), ptr(psi), ptr(tau), ptr(Vf),
                                                  ptr(Vfree), P.dt0 if dt is None else dt),
                         allow=(L.EQ_NR_NAN, L.EQ_NR_NOCONV))
        return rc, dict(d=d, v=v, a=a, psi=psi, tau=tau, Vf=Vf, FltVfree=Vfree)

    def newmark_step_noalloc(self, d, v, a, psi, dt=None):
        """newmark_step through eq_newmark_step_noalloc (no device allocation:
        its scratch is the library's file-scope device array)."""
        P = self.P
        d, v, a, psi = (f64(z).copy() for z in (d, v, a, psi))
        tau, Vf, Vfree = (np.empty(P.FltNglob) for _ in range(3))
        rc = self._check(self.lib.eq_newmark_step_noalloc(self.h, ptr(d), ptr(v), ptr(a), ptr(psi), ptr(tau),
                                                          ptr(Vf), ptr(Vfree), P.dt0 if dt is None else dt),
                         allow=(L.EQ_NR_NAN, L.EQ_NR_NOCONV))
        return rc, dict(d=d, v=v, a=a, psi=psi, tau=tau, Vf=Vf, FltVfree=Vfree)

    def dtevol(self, dt, dtmin, XiLf, FaultNglob, NFBC, Vf, isolver):
        out = C.c_double()
        self._check(self.lib.eq_dtevol(self.h, dt, dtmin, ptr(f64(XiLf)), FaultNglob, NFBC, ptr(f64(Vf)),
                                       isolver, C.byref(out)))
        return out.value

    # ---- time loop ------------------------------------------------------------
    def set_solver(self, tol=1e-6, maxiter=0, pc="amg", prec="fp64", guess="previous", graphs=False,
                   cheb_pairs=True, early_tail=True, galerkin_shortcut=True, basis_fp16=False, extrapolate=False,
                   galerkin_fused=True):
        """early_tail: the second quasi-static solve's fault update and the
        step's status read-back go out behind its first batch of iterations
        and are redone when that batch fell short (same results, one host
        wait fewer per step); with the fused Galerkin start also the first
        solve's fault update and the second pass's start, and in run() the
        next step's begin and start, go out before the host has read the test
        each waits on, and do nothing unless it passed (same results).  galerkin_shortcut: a Galerkin start after a
        solve of the same x projects only the change of b (rows 1..4).
        basis_fp16: the Galerkin expansion reads the FP16 copy of the basis.
        extrapolate: the Galerkin start of the first pass begins from the last
        step's increment extrapolated (after two quasi-static steps in a row).
        galerkin_fused: the FP16 expansion and the start residual in one launch
        (False: two, the same results bit for bit)."""
        o = L.SolverOpts(tol, maxiter, PC[pc], PREC[prec],
                         {"previous": L.EQ_GUESS_PREVIOUS, "learned": L.EQ_GUESS_LEARNED,
                          "galerkin": L.EQ_GUESS_GALERKIN, "linear": L.EQ_GUESS_LINEAR}[guess], 1,
                         1 if graphs else 0,
                         0 if cheb_pairs else 1, 0 if early_tail else 1, 0 if galerkin_shortcut else 1,
                         1 if basis_fp16 else 0, 1 if extrapolate else 0, 0 if galerkin_fused else 1)
        self._check(self.lib.eq_set_solver(self.h, C.byref(o)))

    def state_init(self, v_init=None, psi_init=None):
        P = self.P
        v0 = P.v_init if v_init is None else v_init
        ps = f64(P.psi_init) if psi_init is None else (None if psi_init is False else f64(psi_init))
        self._check(self.lib.eq_state_init(self.h, v0, ptr(ps)))

    def state_get(self):
        P = self.P
        d, v, a = (np.empty(P.nglob) for _ in range(3))
        psi, tau, Vf = (np.empty(P.FltNglob) for _ in range(3))
        t, dt = C.c_double(), C.c_double()
        it, iso = C.c_int64(), C.c_int64()
        self._check(self.lib.eq_state_get(self.h, ptr(d), ptr(v), ptr(a), ptr(psi), ptr(tau), ptr(Vf),
                                          C.byref(t), C.byref(dt), C.byref(it), C.byref(iso)))
        return dict(d=d, v=v, a=a, psi=psi, tau=tau, Vf=Vf, t=t.value, dt=dt.value, it=it.value,
                    isolver=iso.value)

    def state_set(self, d=None, v=None, a=None, psi=None, tau=None, t=0.0, dt=None, isolver=1):
        arrs = [None if z is None else f64(z) for z in (d, v, a, psi, tau)]
        self._check(self.lib.eq_state_set(self.h, *[ptr(z) for z in arrs], t,
                                          self.P.dt0 if dt is None else dt, isolver))

    def step(self):
        info = L.StepInfo()
        self._check(self.lib.eq_step(self.h, C.byref(info)))
        return StepInfo(info.t, info.dt_used, info.dt_next, info.Vfmax, info.it, info.isolver_used,
                        info.isolver_next, (info.pcg_iters[0], info.pcg_iters[1]), info.nr_noconv)

    def run(self, total_time=None, max_steps=0, cap_steps=200000, cap_events=400, cap_is=4000, cap_seis=30000,
            cap_rec=190000, snapshots=True, events=False, keep_on_nr_failure=False):
        """main(P) on the GPU; returns a dict with the fields of struct results.
        Consecutive calls continue one run (the event bookkeeping carries
        over until state_init / state_set).  events=True keeps the per-event
        fault matrices (taubefore, tauafter, delfafter) without the snapshot
        series.  A failed rate-state search (NaN slip rate or stress,
        FBC! in src/NRsearch.jl:15-54) raises EqError, unless
        keep_on_nr_failure: then the results up to the failing step come back
        with the message under "error" (None when the run completed)."""
        P = self.P
        F = P.FltNglob
        ns = len(P.out_seis)
        cap_steps = min(cap_steps, max_steps) if max_steps > 0 else cap_steps
        bufs = dict(time_=np.zeros(cap_steps), Vfmax=np.zeros(cap_steps), isolver=np.zeros(cap_steps, np.int64),
                    pcg_iters=np.zeros(2 * cap_steps, np.int64), tStart=np.zeros(cap_events),
                    tEnd=np.zeros(cap_events), hypo=np.zeros(cap_events))
        if snapshots or events:
            for k in ("taubefore", "tauafter", "delfafter"):
                bufs[k] = np.zeros((cap_events, F))
        if snapshots:
            for k in ("is_slip", "is_slipvel", "is_stress"):
                bufs[k] = np.zeros((cap_is, F))
            for k in ("seismic_slip", "seismic_slipvel", "seismic_stress"):
                bufs[k] = np.zeros((cap_seis, F))
            for k in ("dSeis", "vSeis", "aSeis"):
                bufs[k] = np.zeros((ns, cap_rec))
        R = L.Results()
        R.cap_steps, R.cap_events, R.cap_is, R.cap_seis, R.cap_rec = cap_steps, cap_events, cap_is, cap_seis, cap_rec
        for k, v in bufs.items():
            setattr(R, k, ptr(v, C.c_int64 if v.dtype == np.int64 else C.c_double))
        T = P.Total_time if total_time is None else total_time
        rc = self.lib.eq_run(self.h, T, max_steps, C.byref(R))
        self._check(rc, allow=(L.EQ_NR_NAN,) if keep_on_nr_failure else ())
        n = R.n_steps
        out = dict(n_steps=n, n_qs=R.n_qs, n_dyn=R.n_dyn, total_pcg_iters=R.total_pcg_iters,
                   nr_noconv=R.nr_noconv, wall_qs=R.wall_qs, wall_dyn=R.wall_dyn,
                   time_=bufs["time_"][:min(n, cap_steps)], Vfmax=bufs["Vfmax"][:min(n, cap_steps)],
                   isolver=bufs["isolver"][:min(n, cap_steps)],
                   pcg_iters=bufs["pcg_iters"][:2 * min(n, cap_steps)].reshape(-1, 2),
                   tStart=bufs["tStart"][:min(R.n_start, cap_events)], tEnd=bufs["tEnd"][:min(R.n_end, cap_events)],
                   hypo=bufs["hypo"][:min(R.n_start, cap_events)],
                   error=str(EqError(rc, self._err())) if rc else None)
        if snapshots or events:
            out.update(taubefore=bufs["taubefore"][:min(R.n_start, cap_events)],
                       tauafter=bufs["tauafter"][:min(R.n_end, cap_events)],
                       delfafter=bufs["delfafter"][:min(R.n_end, cap_events)])
        if snapshots:
            out.update(is_slip=bufs["is_slip"][:min(R.n_is, cap_is)],
                       is_slipvel=bufs["is_slipvel"][:min(R.n_is, cap_is)],
                       is_stress=bufs["is_stress"][:min(R.n_is, cap_is)],
                       seismic_slip=bufs["seismic_slip"][:min(R.n_seis, cap_seis)],
                       seismic_slipvel=bufs["seismic_slipvel"][:min(R.n_seis, cap_seis)],
                       seismic_stress=bufs["seismic_stress"][:min(R.n_seis, cap_seis)],
                       dSeis=bufs["dSeis"][:, :min(R.n_rec, cap_rec)].T, vSeis=bufs["vSeis"][:, :min(R.n_rec, cap_rec)].T,
                       aSeis=bufs["aSeis"][:, :min(R.n_rec, cap_rec)].T)
        return out


    # ---- learned initial guess ---------------------------------------------------
    def learn_init(self, capacity=2000, channels=32, hidden=256, nbasis=64, seed=4):
        self._check(self.lib.eq_learn_init(self.h, capacity, channels, hidden, nbasis, seed))
        self.nbasis = nbasis

    def learn_record(self, on=True):
        self._check(self.lib.eq_learn_record(self.h, 1 if on else 0))

    def learn_info(self):
        ns, npar, nl = C.c_int64(), C.c_int64(), C.c_int64()
        tr = C.c_int()
        self._check(self.lib.eq_learn_info(self.h, C.byref(ns), C.byref(npar), C.byref(tr), C.byref(nl)))
        return dict(nsnap=ns.value, nparams=npar.value, trained=bool(tr.value), learned_solves=nl.value)

    def learn_gram(self):
        m = self.learn_info()["nsnap"]
        G = np.empty((m, m))
        self._check(self.lib.eq_learn_gram(self.h, ptr(G)))
        return G

    def learn_set_basis(self, V, sigma, use=None):
        V = f64(V)
        sigma = f64(sigma)
        u = None if use is None else np.ascontiguousarray(use, dtype=np.int32)
        self._check(self.lib.eq_learn_set_basis(self.h, ptr(V), ptr(sigma), ptr(u, C.c_int32)))

    def learn_norms(self):
        m = self.learn_info()["nsnap"]
        out = np.empty(2 * m)
        self._check(self.lib.eq_learn_norms(self.h, ptr(out)))
        return out.reshape(m, 2)

    def learn_train(self, epochs=200, batch=64, lr=1e-3, seed=11):
        hist = np.zeros(max(epochs, 1))
        self._check(self.lib.eq_learn_train(self.h, epochs, batch, lr, seed, ptr(hist)))
        return hist[:epochs]

    def learn_loss_grad(self, params, idx, use_bf16=False):
        info = self.learn_info()
        prm = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        g = np.empty(info["nparams"], np.float32)
        loss = C.c_double()
        self._check(self.lib.eq_learn_loss_grad(self.h, ptr(prm, C.c_float), len(idx), ptr(idx, C.c_int32),
                                                1 if use_bf16 else 0, C.byref(loss), ptr(g, C.c_float)))
        return loss.value, g

    def learn_get(self):
        info = self.learn_info()
        m, F = info["nsnap"], self.P.FltNglob
        nb = self.nbasis
        params = np.empty(info["nparams"], np.float32)
        Phi = np.empty((nb, self.nI), np.float32)
        X = np.empty((m, F), np.float32)
        Y = np.empty((m, nb), np.float32)
        self._check(self.lib.eq_learn_get(self.h, ptr(params, C.c_float), ptr(Phi, C.c_float), ptr(X, C.c_float),
                                          ptr(Y, C.c_float)))
        return dict(params=params, Phi=Phi, X=X, Y=Y)

    def learn_snapshots(self):
        """The recorded interior increments, [nsnap x nI] (one snapshot a row)."""
        m = self.learn_info()["nsnap"]
        S = np.empty((m, self.nI), np.float32)
        self._check(self.lib.eq_learn_snapshots(self.h, ptr(S, C.c_float)))
        return S

    def learn_predict(self, dF):
        dF = f64(dF)
        out = np.empty(self.nbasis)
        self._check(self.lib.eq_learn_predict(self.h, ptr(dF), ptr(out)))
        return out

    def learn_set_linear(self, A, b):
        """The linear predictor coef = |dF| (A dF/|dF| + b), A [nbasis x FltNglob]."""
        A = np.ascontiguousarray(A, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        assert A.shape == (self.nbasis, self.P.FltNglob) and b.shape == (self.nbasis,)
        self._check(self.lib.eq_learn_set_linear(self.h, ptr(A, C.c_float), ptr(b, C.c_float)))

    def learn_predict_linear(self, dF):
        dF = f64(dF)
        out = np.empty(self.nbasis)
        self._check(self.lib.eq_learn_predict_linear(self.h, ptr(dF), ptr(out)))
        return out

    def learn_galerkin(self, x, b):
        """x + Phi (Phi^T K_II Phi)^-1 Phi^T (b - K_II x); vectors in FltNI order."""
        x, b = f64(x), f64(b)
        out = np.empty(self.nI)
        self._check(self.lib.eq_learn_galerkin(self.h, ptr(x), ptr(b), ptr(out)))
        return out

    def launch_timing(self, on: bool):
        """Bracket every launch of the fine-level kernels with HIP events on the
        library's stream (CG batches launch eagerly meanwhile); off stops."""
        self._check(self.lib.eq_launch_timing(self.h, 1 if on else 0))

    def launch_timing_get(self):
        """{kernel name: (launches, total device ms)} since launch_timing(True)."""
        n = np.zeros(L.EQ_TK_COUNT, np.int64)
        ms = np.zeros(L.EQ_TK_COUNT)
        self._check(self.lib.eq_launch_timing_get(self.h, ptr(n, C.c_int64), ptr(ms)))
        return {name: (int(n[k]), float(ms[k])) for k, name in enumerate(L.TK_NAMES)}

    def time_kernel(self, which="ku", prec="fp64", reps=50):
        """Average device ms per launch, HIP events on the library's stream."""
        w = {"ku": L.EQ_TIME_KU, "cg_dir": L.EQ_TIME_CG_DIR, "vcycle": L.EQ_TIME_VCYCLE,
             "cg_iter": L.EQ_TIME_CG_ITER, "newmark": L.EQ_TIME_NEWMARK, "cheb": L.EQ_TIME_CHEB,
             "copy": L.EQ_TIME_COPY, "cheb2": L.EQ_TIME_CHEB2, "sweep": L.EQ_TIME_SWEEP,
             "sweep_post": L.EQ_TIME_SWEEP_POST, "restrict": L.EQ_TIME_RESTRICT}[which]
        out = C.c_double()
        self._check(self.lib.eq_time_kernel(self.h, w, PREC[prec], reps, C.byref(out)))
        return out.value


@dataclass
class StepInfo:
    t: float
    dt_used: float
    dt_next: float
    Vfmax: float
    it: int
    isolver_used: int
    isolver_next: int
    pcg_iters: tuple
    nr_noconv: int


# ---------------------------------------------------------------------------
# host-side Ruge-Stueben set-up (no GPU)
# ---------------------------------------------------------------------------
def rs_hierarchy(A, theta=0.25, max_levels=10, max_coarse=10, negate=True):
    """The Ruge-Stueben hierarchy the library's "rs" set-up builds from the SPD
    matrix A (coarsening -A as ruge_stuben(kni) does, src/main.jl:195-200, or A
    itself with negate=False; see eq_rs_build): {"A": [A_l], "P": [P_l],
    "split": [C mask of level l]}, scipy CSR matrices with A's sign."""
    import scipy.sparse as sp
    lib = L.load()
    M = sp.csr_matrix(A)
    M.sort_indices()
    rp = np.ascontiguousarray(M.indptr, dtype=np.int64)
    cl = np.ascontiguousarray(M.indices, dtype=np.int64)
    vl = np.ascontiguousarray(M.data, dtype=np.float64)
    desc = L.CsrDesc(M.shape[0], M.shape[1], ptr(rp, C.c_int64), ptr(cl, C.c_int64), ptr(vl))
    h = C.c_void_p()
    rc = lib.eq_rs_build(C.byref(desc), theta, max_levels, max_coarse, 1 if negate else 0, C.byref(h))
    if rc:
        raise EqError(rc, f"eq_rs_build: {L.STATUS_NAMES.get(rc, rc)}")
    try:
        nlev = C.c_int()
        cap = 64
        n, nna, nnp = (np.zeros(cap, np.int64) for _ in range(3))
        lib.eq_rs_info(h, C.byref(nlev), ptr(n, C.c_int64), ptr(nna, C.c_int64), ptr(nnp, C.c_int64), cap)
        out = {"A": [], "P": [], "split": []}
        for l in range(nlev.value):
            for which, nnz in ((0, nna[l]), (1, nnp[l])):
                if which == 1 and l + 1 >= nlev.value:
                    continue
                r = np.zeros(n[l] + 1, np.int64)
                c = np.zeros(max(1, nnz), np.int64)
                v = np.zeros(max(1, nnz))
                lib.eq_rs_level(h, l, which, ptr(r, C.c_int64), ptr(c, C.c_int64), ptr(v))
                m = n[l] if which == 0 else n[l + 1]
                out["A" if which == 0 else "P"].append(sp.csr_matrix((v[:nnz], c[:nnz], r), shape=(n[l], m)))
            if l + 1 < nlev.value:
                s_ = np.zeros(n[l], np.int32)
                lib.eq_rs_split(h, l, ptr(s_, C.c_int32))
                out["split"].append(s_.astype(bool))
        return out
    finally:
        lib.eq_rs_free(h)


# ---------------------------------------------------------------------------
# reference-named functions: the positional signatures of the Julia routines,
# argument for argument, on the current context (use_context; the last
# EqContext created otherwise), so that a call site needs no edit
# ---------------------------------------------------------------------------
_current: list = [None]


def use_context(ctx: EqContext) -> EqContext:
    """Make ctx the context the reference-named functions run on."""
    _current[0] = ctx
    return ctx


def current_context() -> EqContext:
    if _current[0] is None:
        raise EqError(L.EQ_ENOTREADY, "no EqContext: create one (or call use_context) first")
    return _current[0]


def FBC(IDstate, P, NFBC, FltNglob, psi1, Vf1, tau1, psi2, Vf2, tau2, psi, Vf, FltVfree, dt):
    """FBC!(IDstate, P::params_farray, NFBC, FltNglob, psi1, Vf1, tau1, psi2, Vf2, tau2, psi, Vf, FltVfree,
    dt) (src/NRsearch_serial.jl:6-45), with P's fault parameters.  Returns the 6-tuple, or None when the
    NaN / Vf guard fires (the reference's @error and bare return)."""
    rc, fail, outs = current_context().fbc(IDstate, NFBC, FltNglob, psi1, Vf1, tau1, psi2, Vf2, tau2, psi, Vf,
                                           FltVfree, dt, P=P)
    if rc == L.EQ_NR_NAN:
        log.error("NR SEARCH FAILED! (fault location %d)", fail)
        return None
    for dst, src in zip((psi1, Vf1, tau1, psi2, Vf2, tau2), outs):
        if isinstance(dst, np.ndarray) and dst.dtype == np.float64:
            dst[:] = src
    return tuple(outs)


def NRsearch(fo, Vo, cca, ccb, Seff, tau, tauo, psi, FltZ, FltVfree):
    """NRsearch!(fo, Vo, cca, ccb, Seff, tau, tauo, psi, FltZ, FltVfree) -> (Vf, tau) (src/NRsearch_serial.jl:47-98)."""
    Vf, t, k = current_context().nrsearch(fo, Vo, cca, ccb, Seff, tau, tauo, psi, FltZ, FltVfree)
    for kk in np.atleast_1d(k)[np.atleast_1d(k) < 0]:
        log.info("k = %d", -kk)
    if np.ndim(fo) == 0:
        return float(Vf[0]), float(t[0])
    return Vf, t


def IDS(xLf, Vo, psi, dt, Vf, cnd, IDstate=2):
    """IDS!(xLf, Vo, psi, dt, Vf, cnd, IDstate) (src/otherFunctions_serial.jl:5-32)."""
    r = current_context().ids(xLf, Vo, psi, dt, Vf, cnd, IDstate)
    return float(r[0]) if np.ndim(psi) == 0 else r


def IDS2(xLf, Vo, psi, psi1, dt, Vf, Vf1, IDstate=2):
    """IDS2!(xLf, Vo, psi, psi1, dt, Vf, Vf1, IDstate) (src/otherFunctions_serial.jl:35-59)."""
    r = current_context().ids2(xLf, Vo, psi, psi1, dt, Vf, Vf1, IDstate)
    return float(r[0]) if np.ndim(psi) == 0 else r


def slrFunc(P, NFBC, FltNglob, psi, psi1, Vf, Vf1, IDstate, tau1, dt):
    """slrFunc!(P::params_farray, NFBC, FltNglob, psi, psi1, Vf, Vf1, IDstate, tau1, dt) -> (psi1, Vf1)
    (src/otherFunctions_serial.jl:62-83), with P's fault parameters."""
    p1, v1 = current_context().slr(NFBC, FltNglob, psi, psi1, Vf, Vf1, IDstate, tau1, dt, P=P)
    if isinstance(psi1, np.ndarray):
        psi1[:] = p1
    if isinstance(Vf1, np.ndarray):
        Vf1[:] = v1
    return p1, v1


def dtevol(dt, dtmin, XiLf, FaultNglob, NFBC, Vf, isolver):
    """dtevol!(dt, dtmin, XiLf, FaultNglob, NFBC, Vf, isolver) -> dt (src/dtevol.jl:5-44)."""
    return current_context().dtevol(dt, dtmin, XiLf, FaultNglob, NFBC, Vf, isolver)


def cg(ctx: EqContext, x, b, tol, maxiter=0, pc="amg", prec="fp64"):
    """cg!(x, kni, b; Pl, tol) (src/main.jl:256): solves kni x = b in place, returns x."""
    xn, it, rr, ok = ctx.cg(x, b, tol, maxiter, pc, prec)
    if isinstance(x, np.ndarray) and x.dtype == np.float64:
        x[:] = xn
    return xn


def PCG(ctx: EqContext, F, dnew, tol=1e-5, maxit=8000):
    """PCG! (src/PCG.jl:8-67): Jacobi CG; returns dnew, or None (with @error) when it does not converge."""
    d, it, ok = ctx.pcg_jacobi(F, dnew, tol, maxit)
    if not ok:
        log.error("PCG did not converge")
        return None
    return d


def main(P, *, tol=1e-6, pc="amg", prec="fp64", total_time=None, max_steps=0, amg=None):
    """main(P) (src/main.jl:67-476) on the GPU; returns the results dict."""
    ctx = EqContext(P, amg=(amg or {}) if pc == "amg" else False)
    ctx.set_solver(tol=tol, pc=pc, prec=prec)
    ctx.state_init()
    return ctx.run(total_time=total_time, max_steps=max_steps)

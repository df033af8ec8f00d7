"""CPU FP64 oracle for the eqcycle time-stepping hot path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``, and only to check results.  The
product library never imports, links or calls anything here.

What it restates (reference file:line):

* ``meshbox``            -- MeshBox! numbering, literal loop (src/MeshBox.jl:27-88)
* ``assemble_K``         -- K_element + FEMSparse assembly (src/Kassemble.jl:3-63), in C
* ``Amg`` / ``cg``       -- ruge_stuben + aspreconditioner + IterativeSolvers.cg!
                            (src/main.jl:195-201,256), in C
* ``pcg_jacobi``         -- dormant PCG! (src/PCG.jl:8-67), in C
* ``ids``/``ids2``/``slr``/``nrsearch``/``fbc``/``dtevol``
                         -- src/otherFunctions_serial.jl, src/NRsearch_serial.jl,
                            src/dtevol.jl, in C (FP64; the live NRsearch! iterates in
                            BigFloat, see ``nrsearch_mp``)
* ``nrsearch_mp``        -- the same Newton sequence in 256-bit mpmath (BigFloat stand-in)
* ``Simulation``         -- the main(P) time loop (src/main.jl:67-476)
* ``element_ku``         -- element_computation! (src/PCG.jl:71-86), numpy

Pinning: see tests/test_oracle.py (GLL table, closed forms, brute force,
direct sparse solves, mpmath replays of the BigFloat Newton sequence up to
130 m/s, the Newmark step against the central-difference recurrence on an
eigenmode and against the equation of motion, slrFunc! against the inverse
friction law and the exact aging law, the quasi-static step against
equilibrium, traction and the average-rate rule; the Ruge-Stueben set-up against
the first-pass split invariants, constant-preserving direct interpolation,
Galerkin coarse operators and the V-cycle's spectrum in tests/test_amg_cpu.py).
Parity unpinned: which of the admissible C/F splits AlgebraicMultigrid.jl's
tie-breaking picks, and CG iteration counts (AlgebraicMultigrid.jl /
IterativeSolvers.jl are absent and unversioned), and reference full-run
trajectories (none are stored).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libeqoracle.so")

_lib = None

i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


class CSR(C.Structure):
    _fields_ = [("n", C.c_int64), ("m", C.c_int64), ("rowptr", C.POINTER(C.c_int64)),
                ("col", C.POINTER(C.c_int64)), ("val", C.POINTER(C.c_double))]


def build(verbose: bool = False) -> str:
    """Compile eqoracle.c with gcc into oracle/_build/libeqoracle.so."""
    os.makedirs(os.path.dirname(LIB_PATH), exist_ok=True)
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-fopenmp", "-o", LIB_PATH,
           os.path.join(HERE, "eqoracle.c"), "-lm"]
    subprocess.run(cmd, check=True, capture_output=not verbose)
    return LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    L.orc_assemble_K.argtypes = [C.c_int64, C.c_int64, i64p, f64p, f64p, C.c_double, C.c_double, C.POINTER(CSR)]
    L.orc_spmv.argtypes = [C.POINTER(CSR), f64p, f64p]
    L.orc_csr_submatrix.argtypes = [C.POINTER(CSR), i64p, C.c_int64, C.c_double, C.POINTER(CSR)]
    L.orc_csr_free.argtypes = [C.POINTER(CSR)]
    L.orc_amg_build.argtypes = [C.POINTER(CSR), C.c_double, C.c_int, C.c_int64]
    L.orc_amg_build.restype = C.c_void_p
    L.orc_amg_free.argtypes = [C.c_void_p]
    L.orc_amg_apply.argtypes = [C.c_void_p, f64p, f64p]
    L.orc_amg_nlev.argtypes = [C.c_void_p]
    L.orc_amg_level_n.argtypes = [C.c_void_p, C.c_int]
    L.orc_amg_level_n.restype = C.c_int64
    L.orc_amg_level_nnz.argtypes = [C.c_void_p, C.c_int]
    L.orc_amg_level_nnz.restype = C.c_int64
    L.orc_amg_level_pnnz.argtypes = [C.c_void_p, C.c_int]
    L.orc_amg_level_pnnz.restype = C.c_int64
    L.orc_amg_get_csr.argtypes = [C.c_void_p, C.c_int, C.c_int, i64p, i64p, f64p]
    L.orc_amg_get_split.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    L.orc_amg_get_split.restype = None
    L.orc_cg.argtypes = [C.POINTER(CSR), f64p, f64p, C.c_void_p, C.c_double, C.c_int64,
                         C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.orc_pcg_jacobi.argtypes = [C.POINTER(CSR), f64p, f64p, f64p, C.c_double, C.c_int64, C.POINTER(C.c_int64)]
    L.orc_ids.argtypes = [C.c_double] * 6 + [C.c_int]
    L.orc_ids.restype = C.c_double
    L.orc_ids2.argtypes = [C.c_double] * 7 + [C.c_int]
    L.orc_ids2.restype = C.c_double
    L.orc_nrsearch.argtypes = [C.c_double] * 10 + [C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_fbc.argtypes = ([C.c_int] + [f64p] * 8 + [C.c_int64, C.c_int64] + [f64p] * 9 +
                          [C.c_double, C.POINTER(C.c_int64)])
    L.orc_slr.argtypes = [C.c_int] + [f64p] * 7 + [C.c_int64, C.c_int64] + [f64p] * 5 + [C.c_double]
    L.orc_dtevol.argtypes = [C.c_double, C.c_double, f64p, C.c_int64, C.c_int64, f64p, C.c_int]
    L.orc_dtevol.restype = C.c_double
    L.orc_splitmix64.argtypes = [C.c_uint64, C.c_int64, u64p]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_set_threads(1)
    _lib = L
    return L


def set_threads(n: int) -> None:
    """Threads of the C restatement's OpenMP loops (1 = serial order, the default)."""
    lib().orc_set_threads(int(n))


# ---------------------------------------------------------------------------
# sparse matrices
# ---------------------------------------------------------------------------
class Csr:
    """Owning wrapper of a C CSR matrix."""

    def __init__(self):
        self.s = CSR()

    def __del__(self):
        if _lib is not None and self.s.rowptr:
            _lib.orc_csr_free(C.byref(self.s))

    @property
    def n(self):
        return self.s.n

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.s.n)
        lib().orc_spmv(C.byref(self.s), x, y)
        return y

    def to_scipy(self):
        import scipy.sparse as sp
        n, m = self.s.n, self.s.m
        nnz = self.s.rowptr[n]
        rp = np.ctypeslib.as_array(self.s.rowptr, shape=(n + 1,)).copy()
        cl = np.ctypeslib.as_array(self.s.col, shape=(nnz,)).copy() if nnz else np.zeros(0, np.int64)
        vl = np.ctypeslib.as_array(self.s.val, shape=(nnz,)).copy() if nnz else np.zeros(0)
        return sp.csr_matrix((vl, cl, rp), shape=(n, m))

    def submatrix(self, idx0, scale=1.0):
        out = Csr()
        idx0 = np.ascontiguousarray(idx0, dtype=np.int64)
        rc = lib().orc_csr_submatrix(C.byref(self.s), idx0, len(idx0), scale, C.byref(out.s))
        assert rc == 0
        return out


def assemble_K(nelx, nely, nglob, iglob, W, H, dxe, dye) -> Csr:
    """Assembled stiffness K from K_element (src/Kassemble.jl:3-63).

    iglob, W: [Nel, 5(j), 5(i)] (column-major iglob[i,j,e] memory), iglob 1-based.
    H: H[i, j] = h'_i(x_j) as a (5, 5) array.
    """
    K = Csr()
    ig0 = np.ascontiguousarray(iglob, dtype=np.int64).reshape(-1) - 1
    Wc = np.ascontiguousarray(W, dtype=np.float64).reshape(-1)
    Hc = np.ascontiguousarray(np.asarray(H, dtype=np.float64).T).reshape(-1)   # column-major
    rc = lib().orc_assemble_K(nelx * nely, nglob, ig0, Wc, Hc, float(dxe), float(dye), C.byref(K.s))
    assert rc == 0
    return K


def element_ku(u_e, W_e, H, c1, c2):
    """element_computation! for one element (src/PCG.jl:82): a = c1 H(W.(H'u)) + c2 (W.(uH))H'.

    u_e, W_e indexed [i, j] (x index first)."""
    Ht = H.T
    return c1 * H @ (W_e * (Ht @ u_e)) + c2 * (W_e * (u_e @ H)) @ Ht


def meshbox(NGLL, Nel, NelX, NelY, FltNglob, dxe, dye, XGLL):
    """Literal restatement of MeshBox! (src/MeshBox.jl:27-88); returns iglob[i,j,e] (1-based), x, y."""
    iglob = np.zeros((NGLL, NGLL, Nel), dtype=np.int64)
    nglob = FltNglob * (NelY * (NGLL - 1) + 1)
    x = np.zeros(nglob)
    y = np.zeros(nglob)
    et = 0
    last_iglob = 0
    ig = np.arange(1, NGLL * NGLL + 1).reshape(NGLL, NGLL, order="F")
    igL = np.arange(1, NGLL * (NGLL - 1) + 1).reshape(NGLL - 1, NGLL, order="F")
    igB = np.arange(1, NGLL * (NGLL - 1) + 1).reshape(NGLL, NGLL - 1, order="F")
    igLB = np.arange(1, (NGLL - 1) * (NGLL - 1) + 1).reshape(NGLL - 1, NGLL - 1, order="F")
    xgll = np.tile((0.5 * (1 + np.asarray(XGLL)))[:, None], (1, NGLL))
    ygll = dye * xgll.T
    xgll = dxe * xgll
    for ey in range(1, NelY + 1):
        for ex in range(1, NelX + 1):
            et += 1
            if et == 1:
                ig = np.arange(1, NGLL * NGLL + 1).reshape(NGLL, NGLL, order="F")
            else:
                ig = ig.copy()
                if ey == 1:
                    ig[0, :] = iglob[NGLL - 1, :, et - 2]
                    ig[1:, :] = last_iglob + igL
                elif ex == 1:
                    ig[:, 0] = iglob[:, NGLL - 1, et - NelX - 1]
                    ig[:, 1:] = last_iglob + igB
                else:
                    ig[0, :] = iglob[NGLL - 1, :, et - 2]
                    ig[:, 0] = iglob[:, NGLL - 1, et - NelX - 1]
                    ig[1:, 1:] = last_iglob + igLB
            iglob[:, :, et - 1] = ig
            last_iglob = ig[NGLL - 1, NGLL - 1]
            x[ig - 1] = dxe * (ex - 1) + xgll
            y[ig - 1] = dye * (ey - 1) + ygll
    return iglob, x, y


# ---------------------------------------------------------------------------
# AMG + CG
# ---------------------------------------------------------------------------
class Amg:
    """ruge_stuben(A) + aspreconditioner(ml) (src/main.jl:200-201)."""

    def __init__(self, A: Csr, theta=0.25, max_levels=10, max_coarse=10):
        self.A = A
        self.h = lib().orc_amg_build(C.byref(A.s), theta, max_levels, max_coarse)
        if not self.h:
            raise MemoryError("AMG build failed")

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.orc_amg_free(self.h)
            self.h = None

    @property
    def nlev(self):
        return lib().orc_amg_nlev(self.h)

    def level_sizes(self):
        L = lib()
        return [(L.orc_amg_level_n(self.h, l), L.orc_amg_level_nnz(self.h, l)) for l in range(self.nlev)]

    def level_csr(self, l, which):
        """which = 0: A_l, 1: P_l (scipy csr)."""
        import scipy.sparse as sp
        L = lib()
        n = L.orc_amg_level_n(self.h, l)
        if which == 0:
            nnz = L.orc_amg_level_nnz(self.h, l)
            m = n
        else:
            nnz = L.orc_amg_level_pnnz(self.h, l)
            m = L.orc_amg_level_n(self.h, l + 1)
        rp = np.zeros(n + 1, np.int64)
        cl = np.zeros(max(nnz, 1), np.int64)
        vl = np.zeros(max(nnz, 1))
        L.orc_amg_get_csr(self.h, l, which, rp, cl, vl)
        return sp.csr_matrix((vl[:nnz], cl[:nnz], rp), shape=(n, m))

    def split(self, l):
        """C mask of level l's C/F split (l < nlev - 1)."""
        isc = np.zeros(lib().orc_amg_level_n(self.h, l), np.intc)
        lib().orc_amg_get_split(self.h, l, isc.ctypes.data_as(C.POINTER(C.c_int)))
        return isc.astype(bool)

    def apply(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        lib().orc_amg_apply(self.h, b, x)
        return x


def cg(A: Csr, x, b, Pl: Amg | None, tol, maxiter=None):
    """cg!(x, A, b; Pl, tol): in place on x; returns (x, iters, relres, converged)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    it = C.c_int64()
    rr = C.c_double()
    rc = lib().orc_cg(C.byref(A.s), x, b, Pl.h if Pl is not None else None, tol,
                      A.n if maxiter is None else maxiter, C.byref(it), C.byref(rr))
    return x, it.value, rr.value, rc == 0


def pcg_jacobi(Kii: Csr, diag, x, Fnew, tol=1e-5, maxit=8000):
    x = np.ascontiguousarray(x, dtype=np.float64)
    it = C.c_int64()
    rc = lib().orc_pcg_jacobi(C.byref(Kii.s), np.ascontiguousarray(diag, np.float64), x,
                              np.ascontiguousarray(Fnew, np.float64), tol, maxit, C.byref(it))
    return x, it.value, rc == 0


# ---------------------------------------------------------------------------
# fault functions (vectorised wrappers over the C restatements)
# ---------------------------------------------------------------------------
def ids(xLf, Vo, psi, dt, Vf, cnd, IDstate=2):
    return lib().orc_ids(xLf, Vo, psi, dt, Vf, cnd, IDstate)


def ids2(xLf, Vo, psi, psi1, dt, Vf, Vf1, IDstate=2):
    return lib().orc_ids2(xLf, Vo, psi, psi1, dt, Vf, Vf1, IDstate)


def nrsearch(fo, Vo, cca, ccb, Seff, tau, tauo, psi, FltZ, FltVfree):
    """FP64 NRsearch!: returns (Vf, tau, k); k < 0 flags non-convergence."""
    v = C.c_double()
    t = C.c_double()
    k = lib().orc_nrsearch(fo, Vo, cca, ccb, Seff, tau, tauo, psi, FltZ, FltVfree, C.byref(v), C.byref(t))
    return v.value, t.value, k


def nrsearch_mp(fo, Vo, cca, ccb, Seff, tau, tauo, psi, FltZ, FltVfree, prec=256):
    """NRsearch! with the BigFloat variables at 256-bit precision (src/NRsearch_serial.jl:47-98).

    fact, eps and help stay Float64 as in the reference; fa, help1, help2,
    delta and tau are multi-precision (mpmath stands in for MPFR).  A Float64
    starting tau (the second call, tau0 = tau1, :37-38) stays Float64 until the
    first update, so the first fa is formed in Float64 and then widened, as the
    typed assignment ``fa::BigFloat = fact*tau/(Seff*cca)`` does; the first
    call's tau0 is BigFloat 0 (:9,13)."""
    import mpmath as mp
    with mp.workprec(prec):
        Vw = 1e10
        fact = 1.0 + (Vo / Vw) * float(np.exp(-psi))
        eps = 0.001 * cca * Seff
        k = 0
        tau = float(tau)
        delta = mp.inf
        Vf = mp.mpf(0)
        help_ = -(fo + ccb * psi) / cca
        while abs(delta) > eps:
            fa = mp.mpf(fact * tau / (Seff * cca))
            help1 = mp.exp(help_ + fa)
            help2 = mp.exp(help_ - fa)
            Vf = Vo * (help1 - help2)
            Vfprime = fact * (Vo / (cca * Seff)) * (help1 + help2)
            delta = (FltZ * FltVfree - FltZ * Vf + tauo - tau) / (1 + FltZ * Vfprime)
            tau = tau + delta
            k += 1
            if abs(delta) > 1e10 or k == 1000:
                return float(Vf), float(tau), -k
        fa = fact * tau / (Seff * cca)
        Vf = Vo * (mp.exp(help_ + fa) - mp.exp(help_ - fa))
        return float(Vf), float(tau), k


def fbc(IDstate, P, NFBC, FltNglob, psi1, Vf1, tau1, psi2, Vf2, tau2, psi, Vf, FltVfree, dt):
    """FBC! in place; returns (status, nr_noconv) with status 0 or 1 + failing index (0-based)."""
    nc = C.c_int64()
    st = lib().orc_fbc(IDstate, P.fo, P.Vo, P.xLf, P.cca, P.ccb, P.Seff, P.tauo, P.FltZ, NFBC - 1, FltNglob,
                       psi1, Vf1, tau1, psi2, Vf2, tau2, psi, Vf, FltVfree, dt, C.byref(nc))
    return st, nc.value


def slr(P, NFBC, FltNglob, psi, psi1, Vf, Vf1, IDstate, tau1, dt):
    lib().orc_slr(IDstate, P.fo, P.Vo, P.xLf, P.cca, P.ccb, P.Seff, P.tauo, NFBC - 1, FltNglob,
                  psi, psi1, Vf, Vf1, tau1, dt)
    return psi1, Vf1


def dtevol(dt, dtmin, XiLf, FaultNglob, NFBC, Vf, isolver):
    return lib().orc_dtevol(dt, dtmin, np.ascontiguousarray(XiLf), FaultNglob, NFBC - 1,
                            np.ascontiguousarray(Vf, dtype=np.float64), isolver)


def splitmix64(seed, n):
    out = np.zeros(n, np.uint64)
    lib().orc_splitmix64(seed, n, out)
    return out


# ---------------------------------------------------------------------------
# time loop: main(P) (src/main.jl:67-476)
# ---------------------------------------------------------------------------
@dataclass
class Results:
    time_: list = field(default_factory=list)
    Vfmax: list = field(default_factory=list)
    isolver: list = field(default_factory=list)
    pcg_iters: list = field(default_factory=list)
    tStart: list = field(default_factory=list)
    tEnd: list = field(default_factory=list)
    hypo: list = field(default_factory=list)
    taubefore: list = field(default_factory=list)
    tauafter: list = field(default_factory=list)
    delfafter: list = field(default_factory=list)
    is_slip: list = field(default_factory=list)
    is_slipvel: list = field(default_factory=list)
    is_stress: list = field(default_factory=list)
    seismic_slip: list = field(default_factory=list)
    seismic_slipvel: list = field(default_factory=list)
    seismic_stress: list = field(default_factory=list)
    dSeis: list = field(default_factory=list)
    vSeis: list = field(default_factory=list)
    aSeis: list = field(default_factory=list)


class Simulation:
    """main(P) restated step by step on the CPU.

    ``P`` is any object with the params_* fields (see eqcycle_amd.params.Params);
    ``solver``: "amg" (ruge_stuben + cg!, the live path) or "direct" (scipy
    sparse LU of K_II, for pinning)."""

    def __init__(self, P, tol=1e-6, solver="amg", amg_kw=None, record=True):
        self.P = P
        self.tol = tol
        self.solver = solver
        self.record = record
        nglob = P.nglob
        self.K = assemble_K(P.NelX, P.NelY, nglob, P.iglob, P.W, P.H, P.dxe, P.dye)
        self.iFlt = P.iFlt - 1
        self.FltNI = P.FltNI - 1
        self.FltIglobBC = P.FltIglobBC - 1
        self.iBcL = P.iBcL - 1
        self.iBcT = P.iBcT - 1
        self.out_seis = P.out_seis - 1
        self.kni = self.K.submatrix(self.FltNI, -1.0)
        if solver == "amg":
            self.ml = Amg(self.kni, **(amg_kw or {}))
        elif solver == "direct":
            # SuperLU's default column ordering up to 100k unknowns; above,
            # K_II being symmetric positive definite, a symmetric minimum-degree
            # ordering without pivoting keeps the fill near Cholesky's (71 M
            # entries instead of 400 M on the 160 x 160 mesh).  Either is a
            # direct solve; through an event the two orderings' trajectories
            # part by ~5e-9 in d (the 32 x 8 test mesh), the loop's own
            # round-off sensitivity.
            import scipy.sparse.linalg as spla
            A = self.kni.to_scipy().tocsc()
            if A.shape[0] <= 100_000:
                self.lu = spla.splu(A)
            else:
                self.lu = spla.splu(A, permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0,
                                    options=dict(SymmetricMode=True))
        else:
            raise ValueError(solver)
        # state (src/main.jl:88-190)
        self.dt = P.dt0
        self.dtmin = P.dt0
        self.half_dt = 0.5 * self.dtmin
        self.half_dt_sq = 0.5 * self.dtmin ** 2
        self.d = np.zeros(nglob)
        self.v = np.full(nglob, P.v_init)
        self.a = np.zeros(nglob)
        F = P.FltNglob
        self.tau = np.zeros(F)
        self.Vf1 = np.zeros(F)
        self.Vf2 = np.zeros(F)
        self.Vf0 = np.zeros(F)
        self.FltVfree = np.zeros(F)
        self.psi1 = np.zeros(F)
        self.psi2 = np.zeros(F)
        self.tau1 = np.zeros(F)
        self.tau2 = np.zeros(F)
        self.psi = np.array(P.psi_init, dtype=np.float64, copy=True)
        self.isolver = 1
        self.dPre = np.zeros(nglob)
        self.vPre = np.zeros(nglob)
        self.v = self.v - 0.5 * P.Vpl
        self.Vf = 2 * self.v[self.iFlt]
        self.iFBC = np.arange(P.NFBC - 1)
        self.NFBC = P.NFBC
        self.Vf[self.iFBC] = 0.0
        self.v[self.FltIglobBC] = 0.0
        self.it = 0
        self.t = 0.0
        self.res = Results()
        self.tvsx = 5 * P.yr2sec
        self.tvsxinc = self.tvsx
        self.tevneinc = 0.5
        self.delfref = np.zeros(F)
        self.idelevne = 0
        self.tevneb = 0.0
        self.tevne = 0.0
        self.slipstart = 0
        self.nr_noconv = 0
        self.step_kind = []

    # -- linear solve of kni * x = rhs --------------------------------------
    def _solve(self, x0, rhs):
        if self.solver == "direct":
            return self.lu.solve(rhs), 0
        x, it, rr, ok = cg(self.kni, x0.copy(), rhs, self.ml, self.tol)
        return x, it

    def qs_step(self):
        """Quasi-static branch (src/main.jl:226-295)."""
        P = self.P
        iFlt, FltNI = self.iFlt, self.FltNI
        dt = self.dt
        self.vPre[:] = self.v
        self.dPre[:] = self.d
        self.Vf0[:] = 2 * self.v[iFlt] + P.Vpl
        self.Vf[:] = self.Vf0
        iters = []
        for _ in range(2):
            F = np.zeros(P.nglob)
            F[iFlt] = self.dPre[iFlt] + self.v[iFlt] * dt
            dnew = self.d[FltNI].copy()
            rhs = self.K.matvec(F)[FltNI]
            dnew, it = self._solve(dnew, rhs)
            iters.append(it)
            self.d[FltNI] = dnew
            self.d[iFlt] = F[iFlt]
            self.a = self.K.matvec(self.d)
            self.a[self.FltIglobBC] = 0.0
            self.tau1[:] = -self.a[iFlt] / P.FltB
            slr(P, self.NFBC, P.FltNglob, self.psi, self.psi1, self.Vf, self.Vf1, P.IDstate, self.tau1, dt)
            self.Vf1[self.iFBC] = P.Vpl
            self.Vf[:] = (self.Vf0 + self.Vf1) / 2
            self.v[iFlt] = 0.5 * (self.Vf - P.Vpl)
        self.psi[:] = self.psi1
        self.tau[:] = self.tau1
        self.tau[self.iFBC] = 0.0
        self.Vf1[self.iFBC] = P.Vpl
        self.v[iFlt] = 0.5 * (self.Vf1 - P.Vpl)
        self.v[FltNI] = (self.d[FltNI] - self.dPre[FltNI]) / dt
        self.a[:] = 0.0
        self.d[self.FltIglobBC] = 0.0
        self.v[self.FltIglobBC] = 0.0
        return iters

    def dyn_step(self):
        """Newmark branch (src/main.jl:297-344)."""
        P = self.P
        iFlt = self.iFlt
        dt = self.dt
        self.dPre[:] = self.d
        self.vPre[:] = self.v
        self.d = self.d + dt * self.v + self.half_dt_sq * self.a
        self.v = self.v + self.half_dt * self.a
        self.a = -self.K.matvec(self.d)
        self.a[self.FltIglobBC] = 0.0
        self.a[self.iBcL] = self.a[self.iBcL] - P.BcLC * self.v[self.iBcL]
        self.a[self.iBcT] = self.a[self.iBcT] - P.BcTC * self.v[self.iBcT]
        self.FltVfree[:] = 2 * self.v[iFlt] + 2 * self.half_dt * self.a[iFlt] / P.M[iFlt]
        self.Vf[:] = 2 * self.vPre[iFlt] + P.Vpl
        st, nc = fbc(P.IDstate, P, self.NFBC, P.FltNglob, self.psi1, self.Vf1, self.tau1, self.psi2,
                     self.Vf2, self.tau2, self.psi, self.Vf, self.FltVfree, dt)
        self.nr_noconv += nc
        if st != 0:
            raise FloatingPointError(f"NR SEARCH FAILED at fault node {st}")
        self.tau[:] = self.tau2 - P.tauo
        self.tau[self.iFBC] = 0.0
        self.psi[:] = self.psi2
        self.a[iFlt] = self.a[iFlt] - P.FltB * self.tau
        self.a = self.a / P.M
        self.v = self.v + self.half_dt * self.a
        self.v[self.FltIglobBC] = 0.0
        self.a[self.FltIglobBC] = 0.0
        return []

    def step(self):
        """One pass of the while-loop body (src/main.jl:221-451)."""
        P = self.P
        self.it += 1
        self.t += self.dt
        kind = self.isolver
        iters = self.qs_step() if self.isolver == 1 else self.dyn_step()
        self.step_kind.append(kind)
        Vfmax = 2 * np.max(self.v[self.iFlt]) + P.Vpl
        if self.record:
            self._bookkeeping(Vfmax)
        r = self.res
        r.time_.append(self.t)
        r.Vfmax.append(Vfmax)
        r.isolver.append(kind)
        r.pcg_iters.append(iters)
        if (self.isolver == 1 and Vfmax < 5e-3) or (self.isolver == 2 and Vfmax < 2e-3):
            self.isolver = 1
        else:
            self.isolver = 2
        current = 2 * self.v[self.iFlt] + P.Vpl
        self.dt = dtevol(self.dt, self.dtmin, P.XiLf, P.FltNglob, self.NFBC, current, self.isolver)
        return Vfmax

    def _bookkeeping(self, Vfmax):
        """Event and snapshot outputs (src/main.jl:361-436)."""
        P, r, t = self.P, self.res, self.t
        iFlt = self.iFlt
        if Vfmax > 1.01 * P.Vthres and self.slipstart == 0:
            self.delfref = 2 * self.d[iFlt] + P.Vpl * t
            self.slipstart = 1
            r.tStart.append(t)
            r.taubefore.append((self.tau + P.tauo) / 1e6)
            indx = int(np.argmax(2 * self.v[iFlt] + P.Vpl))
            r.hypo.append(P.FltX[indx])
        if Vfmax < 0.99 * P.Vthres and self.slipstart == 1:
            r.delfafter.append(2 * self.d[iFlt] + P.Vpl * t - self.delfref)
            r.tauafter.append((self.tau + P.tauo) / 1e6)
            r.tEnd.append(t)
            self.slipstart = 0
        if t > self.tvsx:
            r.is_slip.append(2 * self.d[iFlt] + P.Vpl * t)
            r.is_slipvel.append(2 * self.v[iFlt] + P.Vpl)
            r.is_stress.append((self.tau + P.tauo) / 1e6)
            self.tvsx += self.tvsxinc
        if Vfmax > P.Vevne:
            if self.idelevne == 0:
                self.idelevne = 1
                self.tevneb = t
                self.tevne = self.tevneinc
                r.seismic_slip.append(2 * self.d[iFlt] + P.Vpl * t)
                r.seismic_slipvel.append(2 * self.v[iFlt] + P.Vpl)
                r.seismic_stress.append((self.tau + P.tauo) / 1e6)
            if self.idelevne == 1 and (t - self.tevneb) > self.tevne:
                r.seismic_slip.append(2 * self.d[iFlt] + P.Vpl * t)
                r.seismic_slipvel.append(2 * self.v[iFlt] + P.Vpl)
                r.seismic_stress.append((self.tau + P.tauo) / 1e6)
                self.tevne += self.tevneinc
        else:
            self.idelevne = 0
        if self.it % 10 == 0:
            r.dSeis.append(self.d[self.out_seis].copy())
            r.vSeis.append(self.v[self.out_seis].copy())
            r.aSeis.append(self.a[self.out_seis].copy())

    def run(self, total_time=None, max_steps=None):
        T = self.P.Total_time if total_time is None else total_time
        while self.t < T and (max_steps is None or self.it < max_steps):
            self.step()
        return self.res

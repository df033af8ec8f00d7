/*
 * eqcycle_hip.h -- C ABI of libeqcycle_hip.so, the gfx950 implementation of
 * eqcycle's time-stepping hot path.
 *
 * Every entry point replaces one reference routine or operator call; the
 * reference file:line it stands in for is given beside it (paths relative to
 * the eqcycle repository).  Conventions follow the reference:
 *   - arrays are Float64 / Int64, index arrays 1-based, 3-D arrays column-major
 *     (iglob[i,j,e] is element i + 5 j + 25 e);
 *   - nodal vectors use the reference's MeshBox global numbering;
 *   - the caller owns every host array; calls are synchronous on return;
 *   - a context is not thread-safe: one host thread drives it.
 * Every call returns an eq status; eq_last_error() gives the message.  The
 * Julia binding maps a non-zero status to `@error` + `return nothing`, as the
 * reference does; EQ_NR_NOCONV is informational (NRsearch! returns its last
 * iterate, src/NRsearch_serial.jl:77-85).
 */
#ifndef EQCYCLE_HIP_H
#define EQCYCLE_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  EQ_OK = 0,
  EQ_EINVAL = 1,      /* bad argument, unsupported mesh or material */
  EQ_HIP_ERR = 2,     /* HIP runtime error */
  EQ_NR_NOCONV = 3,   /* a Newton search ended at |delta| > 1e10 or k == 1000 */
  EQ_NR_NAN = 4,      /* FBC! guard: Vf > 1e10 or NaN (src/NRsearch_serial.jl:19-32) */
  EQ_PCG_NOCONV = 5,  /* CG stopped at maxiter above tolerance */
  EQ_ENOTREADY = 6,   /* a required set-up call is missing */
  EQ_NOMEM = 7
};

enum { EQ_MASK_ALL = 0, EQ_MASK_INTERIOR = 1, EQ_MASK_NEG = 2, EQ_MASK_FAULT_ROWS = 3 };
enum { EQ_PREC_FP64 = 0, EQ_PREC_FP32 = 1, EQ_PREC_MIXED = 2, EQ_PREC_F32PC = 3 };
enum { EQ_PC_NONE = 0, EQ_PC_JACOBI = 1, EQ_PC_AMG = 2 };
enum { EQ_SMOOTH_JACOBI = 0, EQ_SMOOTH_CHEBYSHEV = 1, EQ_SMOOTH_CHEBYSHEV_TILES = 2 };
enum { EQ_GUESS_PREVIOUS = 0, EQ_GUESS_LEARNED = 1, EQ_GUESS_GALERKIN = 2, EQ_GUESS_LINEAR = 3 };

typedef struct eq_ctx eq_ctx;

/* Mesh of setParameters (par.jl:20-36, 99-106). */
typedef struct {
  int64_t ngll;            /* NGLL; must be 5 */
  int64_t nelx, nely;      /* NelX, NelY */
  double dxe, dye;         /* element sizes */
  const int64_t *iglob;    /* [ngll*ngll*Nel] MeshBox! numbering (src/MeshBox.jl:27-88) */
  const double *xgll;      /* GetGLL(5) nodes   (src/GetGLL.jl:14-24) */
  const double *wgll;      /* GetGLL(5) weights */
  const double *H;         /* GetGLL(5) H, column-major, H[i,j] = h'_i(xgll[j]) */
} eq_mesh_desc;

/* params_int / params_float / params_farray / params_iarray (par.jl:236-301). */
typedef struct {
  int64_t FltNglob, nglob, IDstate;
  double Vpl, Vthres, Vevne, dt0;
  const double *fo, *Vo, *xLf;     /* [FltNglob] */
  const double *M;                 /* [nglob] lumped mass incl. absorbing terms */
  const double *BcLC;              /* [n_iBcL] */
  const double *BcTC;              /* [n_iBcT] */
  const double *FltB, *FltZ, *FltX, *cca, *ccb, *Seff, *tauo, *XiLf;  /* [FltNglob] */
  const int64_t *iFlt;             /* [FltNglob] */
  const int64_t *iBcL; int64_t n_iBcL;
  const int64_t *iBcT; int64_t n_iBcT;
  const int64_t *FltIglobBC; int64_t n_FltIglobBC;
  const int64_t *out_seis; int64_t n_out_seis;
} eq_params_desc;

/* ---- context ------------------------------------------------------------ */
/* Builds the device mesh from setParameters' output (par.jl:99-106). */
int eq_create(const eq_mesh_desc *mesh, eq_ctx **out);
void eq_destroy(eq_ctx *ctx);
const char *eq_last_error(const eq_ctx *ctx);
/* Number of global nodes, interior (non-fault) nodes. */
int eq_sizes(const eq_ctx *ctx, int64_t *nglob, int64_t *nI, int64_t *nfault);

/* W[5,5,Nel] = wgll2 .* mu (src/MaterialProperties.jl:3-21,77-128).  mu must
 * be nodal (equal on nodes shared by elements), as in mat_trap / mat_gauss and
 * any homogeneous medium; otherwise EQ_EINVAL. */
int eq_set_material(eq_ctx *ctx, const double *W);
/* The params_* arrays used by the fault and time-stepping routines. */
int eq_set_params(eq_ctx *ctx, const eq_params_desc *p);

/* ---- stiffness operator ------------------------------------------------- */
/* y = K u with K = stiffness_assembly (src/Kassemble.jl:3-63), matrix free;
 * replaces Ksparse'*u (src/main.jl:246,267).  mask: EQ_MASK_ALL (K u),
 * EQ_MASK_NEG (-K u = nKsparse'*u, src/main.jl:311), EQ_MASK_INTERIOR (fault
 * entries of u read as 0 and of y set to 0), EQ_MASK_FAULT_ROWS (K u on the
 * nodes of the element row along the fault only, 0 elsewhere: all that
 * (K F)[FltNI] with F on the fault, src/main.jl:246, and a[iFlt] of K d,
 * :267-272, read).  prec: EQ_PREC_FP64 / _FP32. */
int eq_ku_apply(eq_ctx *ctx, const double *u, double *y, int mask, int prec);
/* y = kni * x, kni = -K[FltNI,FltNI] (src/main.jl:195,211); x, y in FltNI order. */
int eq_kni_apply(eq_ctx *ctx, const double *x, double *y, int prec);
/* a_elem[5,5,Nel] = element_computation!(u) (src/PCG.jl:71-86). */
int eq_element_computation(eq_ctx *ctx, const double *u, double *a_elem);

/* ---- AMG preconditioner -------------------------------------------------- */
/* ruge_stuben(kni) + aspreconditioner(ml) (src/main.jl:200-201): multigrid
 * hierarchy built on the host, uploaded, applied as one V-cycle on the GPU.
 * pcoarsen = 0: Ruge-Stueben on K_II itself, as ruge_stuben(kni) does.
 * pcoarsen = 1: first level is the degree-4 -> bilinear (element vertex)
 * prolongation, then Ruge-Stueben on the Galerkin bilinear operator.
 * strength_abs = 1 is the |a_ij| strength test of AlgebraicMultigrid.jl's
 * Classical(theta); 0 uses -a_ij.  smoother: EQ_SMOOTH_JACOBI on every level,
 * or EQ_SMOOTH_CHEBYSHEV of smooth_degree on the matrix-free fine level
 * (Jacobi on the assembled coarse levels), one launch per one or two steps;
 * EQ_SMOOTH_CHEBYSHEV_TILES applies the same polynomial per 64 x 16 tile on
 * the tile and a ring of 2 nodes (restricted additive Schwarz: the sweep's
 * products are those of K_II restricted to that window, only the tile's own
 * nodes are kept), one launch per sweep. */
int eq_amg_setup(eq_ctx *ctx, double theta, int max_levels, int64_t max_coarse, int strength_abs,
                 int smoother, int smooth_degree, int pcoarsen);
/* A sparse matrix in 0-based CSR: rowptr[n + 1], col and val[rowptr[n]]. */
typedef struct {
  int64_t n, m;
  const int64_t *rowptr;
  const int64_t *col;
  const double *val;
} eq_csr;
/* Use a hierarchy built on the host by the caller, as the reference's
 * ruge_stuben(kni) builds it (src/main.jl:200-201), instead of eq_amg_setup's
 * own: A[0 .. nlev-1] the level operators, P[0 .. nlev-2] the prolongations
 * (A[l+1] = P[l]^T A[l] P[l]), rows and columns of level 0 in FltNI order.
 * A[0] must be kni (= -K[FltNI, FltNI]) or K[FltNI, FltNI] of this mesh (checked
 * by a product): the device applies level 0 matrix-free and the coarser
 * levels as given (negated with A[0]), with one Jacobi sweep (weight omega[l],
 * or 4 / (3 lambda_max(D^-1 A)) estimated when omega is NULL) before and after
 * each coarse correction, or Chebyshev of smooth_degree on level 0, and a
 * dense solve on the last level.  From a Julia ml = ruge_stuben(kni): A[l] is
 * ml.levels[l].A (or ml.final_A), and the CSC arrays of ml.levels[l].R are the
 * CSR arrays of P[l]. */
int eq_amg_upload(eq_ctx *ctx, int nlev, const eq_csr *A, const eq_csr *P, const double *omega, int smoother,
                  int smooth_degree);
/* Host-side Ruge-Stueben set-up, for inspection; needs no GPU and no context.
 * Builds the hierarchy eq_amg_setup(pcoarsen = 0) builds from K_II, from the
 * caller's SPD matrix A (n x n CSR): with negate = 1 on -A, as
 * ruge_stuben(kni) coarsens kni = -K[FltNI, FltNI] (src/main.jl:195-200), with
 * negate = 0 on A itself; the level operators come back with A's sign.  No
 * counterpart in the reference (its hierarchy lives in AlgebraicMultigrid.jl's
 * ml object); exposes what the set-up chose. */
typedef struct eq_rs eq_rs;
int eq_rs_build(const eq_csr *A, double theta, int max_levels, int64_t max_coarse, int negate, eq_rs **out);
/* nlev, and per level l < cap: n[l], nnz of A_l, nnz of P_l (0 on the last) */
int eq_rs_info(const eq_rs *h, int *nlev, int64_t *n, int64_t *nnz_a, int64_t *nnz_p, int cap);
/* which = 0: A_l (n_l x n_l), 1: P_l (n_l x n_{l+1}), as CSR into the caller's
 * arrays (rowptr[n_l + 1], col and val of the nnz eq_rs_info reports) */
int eq_rs_level(const eq_rs *h, int l, int which, int64_t *rowptr, int64_t *col, double *val);
/* is_c[i] = 1 for the C points of level l (l < nlev - 1), 0 for F points */
int eq_rs_split(const eq_rs *h, int l, int *is_c);
void eq_rs_free(eq_rs *h);
/* Per-level sizes: n[l], nnz[l] for l < *nlev (arrays of capacity cap). */
int eq_amg_info(const eq_ctx *ctx, int *nlev, int64_t *n, int64_t *nnz, int cap);
/* Fine-level smoother of the current hierarchy (no reference counterpart;
 * for tests that restate the V-cycle): EQ_SMOOTH_*, the Chebyshev degree and
 * the target interval [lo, hi] of D^-1 K_II the polynomial is built on. */
int eq_amg_l0_info(const eq_ctx *ctx, int *smoother, int *degree, double *lo, double *hi);
/* How the V-cycle applies the coarse levels (no reference counterpart; for
 * tests that compare the two forms).  With two or more condensed coarse levels
 * (one Jacobi step before and after each coarse correction, as linear maps),
 * the set-up also composes them into one sparse product down (every coarse
 * right-hand side from the first) and one up (the first level's correction
 * from every right-hand side and the dense level's correction), three launches
 * instead of 2 L - 3, unless their fill would stream more than 8 MB more per
 * launch saved.  collapsed = 1 (the default when built): those; 0: level by
 * level.  EQ_EINVAL when the hierarchy has no collapsed form. */
int eq_amg_coarse_mode(eq_ctx *ctx, int collapsed);
/* x = Pl \ b for kni (ldiv!(x, p, b)); x, b in FltNI order. */
int eq_amg_ldiv(eq_ctx *ctx, const double *b, double *x, int prec);

/* ---- conjugate gradients ------------------------------------------------- */
/* cg!(x, kni, b; Pl, tol, maxiter) (src/main.jl:256): stops when
 * ||b - kni x||_2 <= tol ||b||_2 or after maxiter iterations.  x, b in FltNI
 * order.  pc: EQ_PC_NONE / _JACOBI / _AMG; prec: _FP64, _FP32, _MIXED (FP32
 * inner CG inside FP64 iterative refinement) or _F32PC (FP64 CG whose AMG
 * V-cycle, i.e. most of the K u applications, runs in FP32). */
int eq_cg(eq_ctx *ctx, double *x, const double *b, double tol, int64_t maxiter, int pc, int prec,
          int64_t *iters, double *relres);
/* PCG!(...) (src/PCG.jl:8-67): Jacobi CG on K_II with rhs -(K F)[FltNI];
 * F is [nglob] (fault entries used), dnew [nI] in FltNI order, in/out. */
int eq_pcg_jacobi(eq_ctx *ctx, const double *F, double *dnew, double tol, int64_t maxit, int64_t *iters);

/* ---- fault physics (arrays in fault order, length FltNglob unless n) ---- */
/* IDS! (src/otherFunctions_serial.jl:5-32), element-wise over n nodes. */
int eq_ids(eq_ctx *ctx, int64_t n, const double *xLf, const double *Vo, const double *psi, double dt,
           const double *Vf, double cnd, int64_t IDstate, double *psi1);
/* IDS2! (src/otherFunctions_serial.jl:35-59). */
int eq_ids2(eq_ctx *ctx, int64_t n, const double *xLf, const double *Vo, const double *psi,
            const double *psi1, double dt, const double *Vf, const double *Vf1, int64_t IDstate,
            double *psi2);
/* NRsearch! (src/NRsearch_serial.jl:47-98), element-wise; k_out[i] = Newton
 * iterations (negative when the non-convergence exit was taken). */
int eq_nrsearch(eq_ctx *ctx, int64_t n, const double *fo, const double *Vo, const double *cca,
                const double *ccb, const double *Seff, const double *tau, const double *tauo,
                const double *psi, const double *FltZ, const double *FltVfree, double *Vf_out,
                double *tau_out, int32_t *k_out);
/* The params_farray fields (par.jl:262-289) that FBC! and slrFunc! read,
 * as the caller holds them (FltNglob values each); the argument P::params_farray
 * of FBC! and slrFunc! (src/NRsearch_serial.jl:6, src/otherFunctions_serial.jl:62).
 * slrFunc! does not read FltZ (may be NULL there). */
typedef struct {
  const double *fo, *Vo, *xLf, *cca, *ccb, *Seff, *tauo, *FltZ;
} eq_fault_params;
/* FBC!(IDstate, P, NFBC, FltNglob, psi1, Vf1, tau1, psi2, Vf2, tau2, psi, Vf,
 * FltVfree, dt) (src/NRsearch_serial.jl:6-45) over j = NFBC..FltNglob with the
 * caller's P (copied to the device on each call; NULL: the context's copy
 * from eq_set_params); on the guard returns EQ_NR_NAN and *fail_node = j. */
int eq_fbc(eq_ctx *ctx, int64_t IDstate, const eq_fault_params *P, int64_t NFBC, int64_t FltNglob, double *psi1,
           double *Vf1, double *tau1, double *psi2, double *Vf2, double *tau2, const double *psi, const double *Vf,
           const double *FltVfree, double dt, int64_t *fail_node);
/* slrFunc!(P, NFBC, FltNglob, psi, psi1, Vf, Vf1, IDstate, tau1, dt)
 * (src/otherFunctions_serial.jl:62-83), P as for eq_fbc. */
int eq_slr(eq_ctx *ctx, const eq_fault_params *P, int64_t NFBC, int64_t FltNglob, const double *psi, double *psi1,
           const double *Vf, double *Vf1, int64_t IDstate, const double *tau1, double dt);
/* dtevol! (src/dtevol.jl:5-44). */
int eq_dtevol(eq_ctx *ctx, double dt, double dtmin, const double *XiLf, int64_t FaultNglob, int64_t NFBC,
              const double *Vf, int64_t isolver, double *dt_out);

/* ---- device-resident time loop (main(P), src/main.jl:67-476) ----------- */
typedef struct {
  double tol;              /* cg! tol (src/main.jl:256 uses 1e-6) */
  int64_t maxiter;         /* 0: nI, as cg! */
  int pc;                  /* EQ_PC_* */
  int prec;                /* EQ_PREC_FP64 or EQ_PREC_MIXED */
  int guess;               /* EQ_GUESS_PREVIOUS (src/main.jl:241), _LEARNED, _GALERKIN or _LINEAR */
  int record_snapshots;    /* copy fault fields at events (src/main.jl:361-422) */
  int graphs;              /* 1: CG iteration batches replay captured HIP graphs (same kernels) */
  int cheb_single;         /* 1: one launch per level-0 Chebyshev step; 0 (default): two steps per
                              launch where they follow each other, in FP32 and on FP64 meshes of up
                              to 1024 tiles of 64 x 16 nodes (same results) */
  int sync_pass2;          /* 1: the host waits for the second quasi-static solve to converge before
                              the fault update; 0 (default, FP64 and F32PC): the fault update, the
                              step finish and the status read-back follow the solve's first batch of
                              iterations and are redone when that batch fell short (same results).
                              With 0 and the fused Galerkin start also: the first solve's fault
                              update and the second pass's start follow its first batch, and inside
                              eq_run the next step's begin and start follow the step's end, each
                              doing nothing unless the test it waits on passed (same results) */
  int galerkin_full;       /* 1: every Galerkin start projects the full residual b - K_II x; 0
                              (default): when the previous solve left x solving the old b and only
                              b's rows 1..4 changed, just that change is projected (the old
                              residual, at tolerance level, is left to CG) */
  int basis_fp16;          /* 1: the Galerkin start's expansion x += Phi c reads an FP16 copy of the
                              basis (half the bytes; the rounding, 5e-4 of the increment Phi c, is
                              left to CG, which runs to the same tolerance); 0: the FP32 basis */
  int extrapolate;         /* 1: after two Galerkin-started quasi-static steps in a row, the first
                              pass's start begins from d + (dt / dt_prev) (d - d_prev), the last
                              step's increment extrapolated, and projects the change of b less the
                              same multiple of the last step's (with galerkin_full, the full
                              residual of that start); 0 (default): from d.  Off the Galerkin
                              start it has no effect */
  int galerkin_split;      /* 1: the Galerkin start's expansion x += Phi c and the start residual
                              b - K_II x in two launches; 0 (default): in one (the same x and r bit
                              for bit) wherever the change-of-b start applies with the FP16 basis
                              and an FP64 or F32PC solve */
} eq_solver_opts;
int eq_set_solver(eq_ctx *ctx, const eq_solver_opts *opts);

/* The dynamic branch of the loop body (src/main.jl:297-344) on caller
 * arrays: predictor, a = -K d with the absorbing edges, FBC! on the fault row
 * (src/NRsearch_serial.jl:6-45), corrector.  d, v, a [nglob] (reference
 * numbering) are advanced by one step of length dt (the loop's dtmin); psi
 * [FltNglob] becomes psi2 (nodes deeper than NFBC keep theirs); tau receives
 * tau2 - tauo, Vf the slip rate 2 v[iFlt] + Vpl before the step, FltVfree
 * the free slip velocity (each of the three may be NULL).  Status as
 * eq_fbc. */
int eq_newmark_step(eq_ctx *ctx, double *d, double *v, double *a, double *psi, double *tau, double *Vf,
                    double *FltVfree, double dt);

/* eq_newmark_step for callers that cannot allocate device memory (code that
 * runs under HIP graph capture, serving loops, harnesses that forbid
 * allocation).  No counterpart in the reference.  Same arguments, results and
 * status as eq_newmark_step; every device buffer it needs beyond the context's
 * own set-up data (the double-buffered nodal state, the fault staging arrays,
 * the FBC! guard counters) lies in one file-scope device array of 48.7 MiB,
 * which each GPU that loads the library holds until the process ends.  Limit:
 * meshes of at most 1,048,576 nodes and 8,192 fault nodes (larger ones return
 * EQ_EINVAL with the limit in eq_last_error).  The array is shared by every
 * call on one GPU, so calls that overlap on a device race. */
int eq_newmark_step_noalloc(eq_ctx *ctx, double *d, double *v, double *a, double *psi, double *tau, double *Vf,
                            double *FltVfree, double dt);

/* Initial state (src/main.jl:99-190): v = v_init everywhere then -Vpl/2,
 * d = a = 0, psi = psi_init (NULL: the formula of src/main.jl:125). */
int eq_state_init(eq_ctx *ctx, double v_init, const double *psi_init);
/* Copy state out (reference numbering); any pointer may be NULL. */
int eq_state_get(eq_ctx *ctx, double *d, double *v, double *a, double *psi, double *tau, double *Vf,
                 double *t, double *dt, int64_t *it, int64_t *isolver);
/* Replace the nodal state (reference numbering); any pointer may be NULL. */
int eq_state_set(eq_ctx *ctx, const double *d, const double *v, const double *a, const double *psi,
                 const double *tau, double t, double dt, int64_t isolver);

typedef struct {
  double t;                /* time after the step */
  double dt_used;          /* step length */
  double dt_next;          /* dtevol! result */
  double Vfmax;            /* 2 max(v[iFlt]) + Vpl (src/main.jl:348) */
  int64_t it;              /* step counter */
  int64_t isolver_used;    /* 1 quasi-static, 2 dynamic */
  int64_t isolver_next;
  int64_t pcg_iters[2];    /* CG iterations of the two QS passes */
  int64_t nr_noconv;       /* Newton searches that hit the failure exit */
} eq_step_info;
/* One pass of the while-loop body (src/main.jl:221-451) without output
 * bookkeeping; host reads back only the step scalars. */
int eq_step(eq_ctx *ctx, eq_step_info *info);

/* Output arrays of main(P) (struct results, src/main.jl:479-498): the caller
 * allocates each with the given capacity; counts are returned.  Fault
 * snapshot matrices are [FltNglob x cap] column-major and may be NULL. */
typedef struct {
  int64_t cap_steps, cap_events, cap_is, cap_seis, cap_rec;
  double *time_, *Vfmax; int64_t *isolver; int64_t *pcg_iters;  /* [cap_steps] ([2*cap_steps] iters) */
  double *tStart, *tEnd, *hypo;                       /* [cap_events] */
  double *taubefore, *tauafter, *delfafter;           /* [FltNglob x cap_events] */
  double *is_slip, *is_slipvel, *is_stress;           /* [FltNglob x cap_is] */
  double *seismic_slip, *seismic_slipvel, *seismic_stress;  /* [FltNglob x cap_seis] */
  double *dSeis, *vSeis, *aSeis;                      /* [cap_rec x n_out_seis] column-major */
  int64_t n_steps, n_start, n_end, n_is, n_seis, n_rec;
  int64_t n_qs, n_dyn, total_pcg_iters, nr_noconv;
  double wall_qs, wall_dyn;                           /* seconds spent per step type */
} eq_results;
/* while t < total_time (and it < max_steps if > 0): eq_step + bookkeeping. */
int eq_run(eq_ctx *ctx, double total_time, int64_t max_steps, eq_results *res);

/* ---- learned initial guess (no reference counterpart; the plain guess is
 * src/main.jl:241, dnew = d[FltNI]) ----------------------------------------
 * Snapshots of the pass-1 quasi-static solves -- fault increment dF and
 * interior displacement increment -- are recorded on the device; a POD basis
 * Phi comes from the snapshot Gram matrix (eigen-decomposed by the caller);
 * a conv1d + MLP network (BF16 weights, FP32 master copy, hand-written
 * forward / backward / Adam kernels) maps dF/|dF| to POD coefficients.  With
 * eq_solver_opts.guess = EQ_GUESS_LEARNED the pass-1 start becomes
 * d + gamma |dF| Phi c, gamma the energy-optimal step, so it never has a
 * larger K_II-norm error than the previous-step guess. */
int eq_learn_init(eq_ctx *ctx, int64_t capacity, int channels, int hidden, int nbasis, uint64_t seed);
int eq_learn_record(eq_ctx *ctx, int on);
int eq_learn_info(const eq_ctx *ctx, int64_t *nsnap, int64_t *nparams, int *trained, int64_t *nlearned);
/* G[nsnap x nsnap] = S^T S of the recorded interior increments (row-major). */
int eq_learn_gram(eq_ctx *ctx, double *G);
/* |dF| and |F| (fault data of the solve) per snapshot: out[2 nsnap]. */
int eq_learn_norms(eq_ctx *ctx, double *out);
/* V[nsnap x nbasis] (row-major) and sigma[nbasis] from eig(G): Phi = S V / sigma;
 * use[nsnap] (NULL: all) selects the snapshots that enter the training set. */
int eq_learn_set_basis(eq_ctx *ctx, const double *V, const double *sigma, const int32_t *use);
/* Adam training (batch, lr); hist[epochs] receives the mean loss per epoch. */
int eq_learn_train(eq_ctx *ctx, int epochs, int batch, double lr, uint64_t seed, double *hist);
/* Loss and gradient on samples idx with the given FP32 parameters (NULL: the
 * current ones); use_bf16 = 0 evaluates with the FP32 weights.  Overwrites the
 * network parameters when params is given. */
int eq_learn_loss_grad(eq_ctx *ctx, const float *params, int nsamp, const int32_t *idx, int use_bf16,
                       double *loss, float *grad);
/* Copies of parameters [nparams], basis [nI x nbasis] column-major, dataset
 * inputs [nsnap x FltNglob] and targets [nsnap x nbasis]; any may be NULL. */
int eq_learn_get(eq_ctx *ctx, float *params, float *Phi, float *X, float *Y);
/* Copy of the recorded interior increments S [nI x nsnap], column-major (one
 * snapshot after the other), whose Gram eq_learn_gram returns. */
int eq_learn_snapshots(eq_ctx *ctx, float *S);
/* Network output (POD coefficients times |dF|) for one fault increment. */
int eq_learn_predict(eq_ctx *ctx, const double *dF, double *coef);
/* The linear least-squares predictor of the same coefficients from the same
 * input, the bar the network has to beat: coef = |dF| (A dF/|dF| + b), A
 * [nbasis x FltNglob] row-major and b [nbasis] fitted by the caller on the
 * dataset (eq_learn_get).  With eq_solver_opts.guess = EQ_GUESS_LINEAR the
 * pass-1 start is that of EQ_GUESS_LEARNED with these coefficients. */
int eq_learn_set_linear(eq_ctx *ctx, const float *A, const float *b);
/* The linear predictor's output for one fault increment. */
int eq_learn_predict_linear(eq_ctx *ctx, const double *dF, double *coef);
/* Galerkin start on the POD basis (guess EQ_GUESS_GALERKIN, both quasi-static
 * passes): x_out = x + Phi c with c = (Phi^T K_II Phi)^-1 Phi^T (b - K_II x),
 * the K_II-norm best correction of x within span(Phi).  Phi^T K_II Phi is
 * formed and inverted by eq_learn_set_basis.  x, b, x_out in FltNI order. */
int eq_learn_galerkin(eq_ctx *ctx, const double *x, const double *b, double *x_out);

/* ---- measurement hook (no reference counterpart) ------------------------ */
enum { EQ_TIME_KU = 0, EQ_TIME_CG_DIR = 1, EQ_TIME_VCYCLE = 2, EQ_TIME_CG_ITER = 3, EQ_TIME_NEWMARK = 4,
       EQ_TIME_CHEB = 5, EQ_TIME_COPY = 6, EQ_TIME_CHEB2 = 7, EQ_TIME_SWEEP = 8, EQ_TIME_SWEEP_POST = 9,
       EQ_TIME_RESTRICT = 10 };
/* Average device time (ms) of one launch of the given kernel group over reps
 * back-to-back launches, bracketed by HIP events on the context's stream.
 * Runs on scratch vectors; EQ_TIME_NEWMARK advances a copy of the state and
 * must not be called in the middle of a run.  EQ_TIME_COPY copies a 512 MiB
 * FP64 buffer into another (allocated for the call, then freed): a kernel of
 * known traffic that calibrates hardware byte counters.  EQ_TIME_CHEB is one
 * level-0 Chebyshev step, EQ_TIME_CHEB2 two steps in one launch, EQ_TIME_SWEEP
 * and EQ_TIME_SWEEP_POST one whole tile sweep before / after the coarse
 * correction at the hierarchy's degree, EQ_TIME_RESTRICT the residual
 * restricted to the element vertices (the last three need a p-multigrid
 * hierarchy with Chebyshev smoothing). */
int eq_time_kernel(eq_ctx *ctx, int which, int prec, int reps, double *avg_ms);

/* Kinds of launch that eq_launch_timing brackets: the fine-level kernels of
 * the time loop (the Newmark step and every kernel that streams the whole
 * mesh inside a CG iteration). */
enum { EQ_TK_NEWMARK = 0,     /* newmark_kernel: one dynamic step */
       EQ_TK_CHEB_PAIR32 = 1, /* cheb_pair_kernel<float>: two fine-level Chebyshev steps */
       EQ_TK_CHEB_PAIR64 = 2, /* cheb_pair_kernel<double> */
       EQ_TK_CHEB_STEP32 = 3, /* one fine-level Chebyshev step, FP32 */
       EQ_TK_CHEB_STEP64 = 4, /* one fine-level Chebyshev step, FP64 */
       EQ_TK_CG_DIR = 5,      /* p = c + beta p, q = K_II p (any precision) */
       EQ_TK_CG_UPDATE = 6,   /* x += alpha p, r -= alpha q (any precision) */
       EQ_TK_SWEEP32 = 7,     /* cheb_sweep_kernel<float>: a whole fine-level smoothing sweep */
       EQ_TK_SWEEP64 = 8,     /* cheb_sweep_kernel<double> */
       EQ_TK_COUNT = 9 };
/* on = 1: clear the counters and give every launch of the kinds above a pair
 * of HIP events on the context's stream until on = 0.  The launch goes
 * through hipExtLaunchKernelGGL, so the kernel's own dispatch stamps its start
 * and end (no marker packets between the kernels).  While on, CG iteration
 * batches are launched one kernel at a time instead of replaying captured
 * graphs (the same kernels), and the host waits on the events whenever their
 * pool of 4096 pairs fills. */
int eq_launch_timing(eq_ctx *ctx, int on);
/* Launches bracketed and their summed device time (ms) per kind since the
 * last eq_launch_timing(ctx, 1): counts[EQ_TK_COUNT], total_ms[EQ_TK_COUNT].
 * Waits for the bracketed launches to finish. */
int eq_launch_timing_get(eq_ctx *ctx, int64_t *counts, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif

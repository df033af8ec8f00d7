// Kernel launchers of libeqcycle_hip (device code lives in eq_kernels.hip).
#pragma once
#include "eq_internal.h"

namespace eq {

// Optional events of one launch: when set, the launch goes through
// hipExtLaunchKernelGGL and the kernel's own dispatch stamps its start and
// end (no marker packets around it).
struct LaunchEvents {
  hipEvent_t start = nullptr, stop = nullptr;
};

constexpr int BLOCK = 256;

// 8 FP16 basis coefficients of one node (one 16-byte load)
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
#ifndef EQ_TILE_I
#define EQ_TILE_I 64
#endif
#ifndef EQ_TILE_NPT
#define EQ_TILE_NPT 4
#endif
constexpr int TILE_I = EQ_TILE_I;   // owned nodes per tile along x (64 or 32)
constexpr int TILE_J = 16;          // owned nodes per tile along y
constexpr int TILE_NPT = EQ_TILE_NPT;                // owned nodes per thread in tile_kernel (1, 2 or 4)
constexpr int TBLOCK = TILE_I * TILE_J / TILE_NPT;   // threads of a tile_kernel block
constexpr int PBLOCK = TILE_I * TILE_J / 4;          // threads of a cheb_pair_kernel block (4 nodes each)

// Device scalars of one CG solve (all FP64).
struct CgScal {
  double rho_prev;   // rho of the previous iteration (1 before the first)
  double resid;      // ||r||_2 at the last convergence check
  double bnorm;      // ||b||_2
  double reltol;     // tol * ||b||_2
  int iter;          // completed iterations
  int done;          // 1 once ||r|| <= reltol or iter >= maxiter
  int maxiter;
  int pad;
};

// Device fault parameters (fault order j = I), FP64.
struct FaultDev {
  const double *fo, *Vo, *xLf, *cca, *ccb, *Seff, *tauo, *XiLf, *FltB, *FltZ, *Mf;
  const double *BcLC, *BcTC;  // [NY], [NX]
  int j0;       // NFBC - 1: first rate-state node (0-based)
  int nflt;     // FltNglob = NX
  int IDstate;
  double Vpl;
};

// Persistent fault state arrays (reference main(P) variables).
struct FaultState {
  double *psi, *psi1, *psi2, *Vf, *Vf0, *Vf1, *Vf2, *tau, *tau1, *tau2, *FltVfree;
};

struct StepStatus {   // written by the step tail kernel
  double Vfmax;
  double dt_next;
  int isolver_next;
  int hypo;          // argmax of the slip rate (0-based fault index)
  int nr_flag;       // 1 + first failing node of the FBC! guard, or 0
  int nr_noconv;     // Newton searches that took the failure exit
  int go;            // 1: the next step is quasi-static and this one ended well (NR searches,
                     // and the second solve's convergence test where the step end ran it)
};

// Per-step record written by the step tail kernel besides its StepStatus:
// fault rows of d and v and the fault traction (3 x nflt), and d, v, a at the
// receiver nodes (3 x nseis).
struct StepCapture {
  double* fault = nullptr;            // [3 * nflt] or null
  const int64_t* seis_idx = nullptr;  // structured node index of each receiver
  int nseis = 0;
  double* seis = nullptr;             // [3 * nseis] or null
  double* seis_mirror = nullptr;      // pinned host copy of seis, or null
  StepStatus* st_mirror = nullptr;    // pinned host copy of the status, or null
};

inline dim3 tile_grid(const Geom& g) {
  return dim3((g.NX + TILE_I - 1) / TILE_I, (g.NY + TILE_J - 1) / TILE_J);
}
inline int tile_count(const Geom& g) {
  dim3 d = tile_grid(g);
  return (int)(d.x * d.y);
}
inline int vec_blocks(int64_t n) { return (int)((n + BLOCK * 4 - 1) / (BLOCK * 4)); }

// ---- matrix-free operator -----------------------------------------------
// y = sign * K u on the structured grid; interior != 0 masks the fault row.
template <typename T>
void launch_ku(const Geom& g, const T* u, const T* mu, T* y, int interior, int negate, hipStream_t s);
void launch_diag(const Geom& g, const double* mu, double* diag, hipStream_t s);
void launch_element_ku(const Geom& g, const double* u, const double* mu, double* a_elem, hipStream_t s);

// ---- CG pieces (K_II form, vectors of length N with zero fault row) ------
template <typename T>
void launch_cg_residual(const Geom& g, const T* x, const T* mu, const T* b, T* r, double* part_rr,
                        double* part_bb, int npart, hipStream_t s, float* r32 = nullptr);
void launch_cg_init_scal(CgScal* sc, const double* part_rr, const double* part_bb, int nparts, double tol,
                         int maxiter, hipStream_t s);
// Galerkin start x = d + Phi c (interior; Phi the blocked FP16 basis
// [ceil(nb / 8)][nI][8]) into x_out, and r = b - K_II x with the partials of
// launch_cg_residual, in one launch.  x_out must not alias d.  gate (or
// null): the launch does nothing unless *gate.
void launch_gal_residual(const Geom& g, const double* d, const half8* phi16, int64_t nI, int nb, const double* c,
                         const double* mu, const double* b, double* r, float* r32, double* x_out, double* part_rr,
                         double* part_bb, int npart, hipStream_t s, const int* gate = nullptr);
// mirror: pinned host copy of the scalars, written after the check (or null)
void launch_cg_check(CgScal* sc, const double* part_rr, int nparts, hipStream_t s, CgScal* mirror = nullptr);
template <typename T>
void launch_cg_dir(const Geom& g, const T* c, const T* u_old, const T* mu, T* u_new, T* q, const double* part_rho,
                   int nrho, const CgScal* sc, double* part_uq, hipStream_t s, LaunchEvents le = {});
template <typename T>
void launch_cg_update(int64_t n, T* x, T* r, const T* u, const T* q, const double* part_rho, int nrho,
                      const double* part_uq, int nuq, CgScal* sc, double* part_rr, int npart, float* r32,
                      hipStream_t s, LaunchEvents le = {});
void launch_guess_dot(const Geom& g, const double* p, const double* mu, const double* r0, double* q,
                      double* part_pr, double* part_pq, hipStream_t s);
void launch_guess_update(int64_t n, double* x, double* r, const double* p, const double* q, const double* part_pr,
                         const double* part_pq, int nparts, double* part_rr, int npart_out, hipStream_t s);
void launch_widen_dot(int64_t n, const float* c32, const double* r, double* c, const CgScal* sc, double* part,
                      hipStream_t s);
template <typename T>
void launch_jacobi_pc(int64_t n, const T* dinv, const T* r, T* c, const double* part_rr, int nrr, CgScal* sc,
                      double* part_rho, hipStream_t s);

// ---- AMG V-cycle ----------------------------------------------------------
template <typename T>
void launch_l0_down(const Geom& g, const T* r, const T* mu, const T* dinv, T omega, T* x0, T* res,
                    const double* part_rr, int nrr, CgScal* sc, hipStream_t s);
template <typename T>
void launch_l0_post(const Geom& g, const T* x0, const T* r, const T* mu, const T* dinv, T omega, T* c,
                    double* part_rho, const CgScal* sc, hipStream_t s);

// One Chebyshev smoothing step on the matrix-free level (see OpCheb).
template <typename T>
struct ChebStep {
  const T* src = nullptr;
  const T* b = nullptr;
  T s0 = 0;
  int src_mode = 0;
  const T* rin = nullptr;
  T* rout = nullptr;
  T c1 = 0, c2 = 0;
  const T* xin = nullptr;
  int xsrc_center = 0;
  T* dout = nullptr;
  T* xout = nullptr;
  int mode = 0;
  int write_x_center = 0;
  const T* rdot = nullptr;
  int check = 0;
  const T* coarse = nullptr;   // src_mode 2: coarse (vertex) correction added on the fly
  // mode 2 with an FP32 step: c widened to FP64 goes to xout64 (instead of
  // xout) and the partial is <c, rdot64> with the FP64 residual
  double* xout64 = nullptr;
  const double* rdot64 = nullptr;
};
template <typename T>
void launch_cheb(const Geom& g, const ChebStep<T>& c, const T* mu, const T* dinv, const double* part_rr, int nrr,
                 CgScal* sc, double* part_rho, hipStream_t s, LaunchEvents le = {});
// y = K u on the first nrows grid rows (the element row along the fault for
// nrows = 5), from global memory; other entries of y untouched.
template <typename T>
void launch_ku_rows(const Geom& g, const T* u, const T* mu, T* y, int nrows, hipStream_t s);
// Steps a and b (a in mode 0, b reading a's outputs) in one launch; b's
// outputs are written, a's stay on chip except a.xout when b is in mode 1.
// b.rout and b.dout must not alias a.rin or a.src.
template <typename T>
void launch_cheb_pair(const Geom& g, const ChebStep<T>& a, const ChebStep<T>& b, const T* mu, const T* dinv,
                      const double* part_rr, int nrr, CgScal* sc, double* part_rho, hipStream_t s,
                      LaunchEvents le = {});

// A whole level-0 Chebyshev sweep of degree `degree` per tile in one launch,
// on the tile and a ring of 2 nodes (restricted additive Schwarz; see
// cheb_sweep_kernel).  Step m = 1 .. degree-1 uses c1[m], c2[m]; the first
// direction is s0 D^-1 r.
constexpr int kSweepMaxDeg = 8;
template <typename T>
struct ChebSweep {
  int post = 0;                  // 0: from x = 0 (pre-smoothing); 1: from src (+ P0 coarse)
  int degree = 1;
  T s0 = 0;
  T c1[kSweepMaxDeg] = {}, c2[kSweepMaxDeg] = {};
  const T* b = nullptr;          // right-hand side of the smoothed system
  const T* src = nullptr;        // post: x before the sweep
  const T* coarse = nullptr;     // post: vertex correction added to src on the fly (or null)
  int nelx = 0, nely = 0;
  T gx[5] = {};                  // GLL nodes (bilinear prolongation weights)
  // D[p][i] D[p][k] for i <= k (15 per p, row-major in (i, k)): a segment's operator
  // D^T diag(wm) D is then sum_p wm_p dd[p] (FP32 sweeps)
  float dd[5][15] = {};
  const T* mu = nullptr;
  const T* dinv = nullptr;
  T* xout = nullptr;             // pre: x; post: c (unless xout64)
  const T* rdot = nullptr;       // post: partial <c, rdot> per tile (or <c, rdot64>)
  double* xout64 = nullptr;      // post of an FP32 sweep: c widened to FP64
  const double* rdot64 = nullptr;
  int check = 0;                 // carries the cg! convergence test (first kernel of a V-cycle)
  // check with init: the first kernel of a solve also sets the solve's scalars up from the
  // partials of ||r||^2 and ||b||^2 (what cg_init_scal_kernel does), tol and maxiter
  int init = 0;
  const double* part_bb = nullptr;
  double tol = 0;
  int maxiter = 0;
  const double* part_rr = nullptr;
  int nrr = 0;
  CgScal* sc = nullptr;
  double* part0 = nullptr;       // post: one partial per tile
};
template <typename T>
void launch_cheb_sweep(const Geom& g, const ChebSweep<T>& op, hipStream_t s, LaunchEvents le = {});
// b1 = P0^T (b - K_II x): the residual of x restricted to the element vertices by
// the transpose of the bilinear prolongation (vertex rows 1 .. nely), matrix-free.
template <typename T>
void launch_restrict(const Geom& g, const T* x, const T* b, const T* mu, T* b1, const CgScal* sc, hipStream_t s);

struct DevCsr {
  int64_t n = 0, m = 0, nnz = 0;
  int32_t* rowptr = nullptr;
  int32_t* col = nullptr;
  double* val = nullptr;
  float* val32 = nullptr;
  int lpr = 1;   // lanes per row
};
enum CsrMode { CSR_SPMV = 0, CSR_DOWN = 1, CSR_PROLONG_ADD = 2, CSR_POST = 3 };
// SPMV: y = M x.  DOWN: x = w dinv b, y(res) = b - M x.  PROLONG_ADD: y += M x.
// POST: y = x + w dinv (b - M x).
template <typename T>
void launch_csr(CsrMode mode, const DevCsr& M, const T* x, const T* b, const T* dinv, T omega, T* y, T* xout,
                const CgScal* sc, hipStream_t s);
// y = U b + V e (condensed up-sweep of a coarse level; U and V have n rows)
template <typename T>
void launch_csr_up(const DevCsr& U, const DevCsr& V, const T* b, const T* e, T* y, const CgScal* sc, hipStream_t s);
template <typename T>
void launch_dense(int64_t n, const T* inv, const T* b, T* x, const CgScal* sc, hipStream_t s);



// ---- time stepping --------------------------------------------------------
// Step record of a Newmark step inside a batch (see OpNewmark); all null
// outside a batch.
struct NewmarkRecord {
  const double* vprev = nullptr;
  const int* nrprev = nullptr;
  double* vpart_out = nullptr;
  double* cap_fault = nullptr;
  const int64_t* seis_idx = nullptr;
  const int* seis_tile = nullptr;     // tile index of each receiver (tile_grid order)
  int nseis = 0;
  double* cap_seis = nullptr;
};
// One Newmark step (src/main.jl:297-344) from the predicted displacement P
// (= d + dt v + dt^2/2 a, the step's new d) and the previous v, a; writes
// the new v, a into v2, a2 and the next step's predictor into P2.
void launch_newmark(const Geom& g, const FaultDev& fp, FaultState fs, const double* mu, const double* M,
                    const double* P, const double* v, const double* a, double* P2, double* v2, double* a2,
                    double dt, double half_dt, double half_dt_sq, int* flags, const NewmarkRecord& rec,
                    hipStream_t s, LaunchEvents le = {});
// Reference (MeshBox) order <-> structured order on the device:
// to_structured: dst[perm[k]] = src[k]; else dst[k] = src[perm[k]]
void launch_permute(int64_t n, const int32_t* perm, const double* src, double* dst, int to_structured,
                    hipStream_t s);
// P = d + dt v + dt^2/2 a on n nodes
void launch_newmark_predict(int64_t n, const double* d, const double* v, const double* a, double dt,
                            double half_dt_sq, double* P, hipStream_t s);
// db: the change of b on rows 1..4 since the last solve; bold: those rows of b
// before this pass (saved); theta != 0 (first pass): the start off the fault
// becomes d + theta (d - dPre) and db the change less theta times bold's.
// st_next (or null; first pass only): the begin of the step after the one
// whose step tail wrote *st_next, enqueued before the host has read it: the
// launch does nothing unless st_next->go, dt is st_next->dt_next and a
// nonzero theta becomes dt / dt_prev.
void launch_qs_begin(const Geom& g, const FaultDev& fp, FaultState fs, const double* mu, double* d, double* dPre,
                     const double* v, double* b, double dt, int first_pass, hipStream_t s,
                     double* db = nullptr, double theta = 0.0, double* bold = nullptr,
                     const StepStatus* st_next = nullptr, double dt_prev = 0.0);
// launch_qs_fault (no save) then the fault-row part of launch_qs_begin for the
// second pass (first_pass = 0, theta = 0, no bold), in one single-block launch;
// sc (or null): the first solve's convergence test first (cg_check's, with
// sc_mirror), and nothing more while that solve has not converged
void launch_qs_mid(const Geom& g, const FaultDev& fp, FaultState fs, const double* mu, double* d, const double* dPre,
                   double* v, double* b, double dt, double* db, hipStream_t s, CgScal* sc = nullptr,
                   const double* part_rr = nullptr, int nparts = 0, CgScal* sc_mirror = nullptr);
// save: psi, Vf and the fault row of d before the update (3 nflt), or null
void launch_qs_fault(const Geom& g, const FaultDev& fp, FaultState fs, const double* mu, const double* d,
                     double* v, double dt, hipStream_t s, double* save = nullptr);
// The end of a quasi-static step in one launch: the interior of the step
// finish (v = (d - dPre) / dt off the fault, a = 0) beside one block with, when
// sc is set, the second solve's convergence test (scalars to sc_mirror), the
// second pass's fault update (save: psi, Vf and d of the fault row first), the
// fault row of the finish and the step tail (status, flags, cap): what
// launch_cg_check, launch_qs_fault and launch_step_tail do, and the finish.
void launch_qs_end(const Geom& g, const FaultDev& fp, FaultState fs, const double* mu, double* d, const double* dPre,
                   double* v, double* a, double dt, double dtmin, double* save, CgScal* sc, const double* part_rr,
                   int nparts, CgScal* sc_mirror, StepStatus* st, int* flags, const StepCapture& cap, hipStream_t s);
// Vfmax, regime switch, dtevol!, Newton flags (then reset for the next step),
// and the step record of cap.
void launch_step_tail(const FaultDev& fp, const double* d, const double* v, const double* a, const double* tau,
                      double dt, double dtmin, int isolver, StepStatus* st, int* flags, const StepCapture& cap,
                      hipStream_t s);

// ---- fault routines on their own (API parity entry points) ---------------
void launch_ids(int64_t n, const double* xLf, const double* Vo, const double* psi, double dt, const double* Vf,
                double cnd, int IDstate, double* psi1, hipStream_t s);
void launch_ids2(int64_t n, const double* xLf, const double* Vo, const double* psi, const double* psi1, double dt,
                 const double* Vf, const double* Vf1, int IDstate, double* psi2, hipStream_t s);
void launch_nrsearch(int64_t n, const double* fo, const double* Vo, const double* cca, const double* ccb,
                     const double* Seff, const double* tau, const double* tauo, const double* psi,
                     const double* FltZ, const double* FltVfree, double* Vf, double* tau_out, int* k,
                     hipStream_t s);
void launch_fbc(const FaultDev& fp, FaultState fs, const double* FltVfree, double dt, int* flags, hipStream_t s);
void launch_slr(const FaultDev& fp, FaultState fs, double dt, hipStream_t s);
void launch_dtevol(const double* XiLf, const double* Vf, int j0, int n, double dt, double dtmin, int isolver,
                   double* out, hipStream_t s);

template <typename T, typename U>
void launch_convert(int64_t n, const T* x, U* y, hipStream_t s);
template <typename T>
void launch_fill(int64_t n, T* y, T val, hipStream_t s);
// y[k * stride] = val for k < n
template <typename T>
void launch_fill_strided(int64_t n, T* y, int stride, T val, hipStream_t s);
// gather: y = [x0[0..n) | x1[0..n) | x2[0..n)]; otherwise the reverse (one launch)
void launch_copy3(int64_t n, double* x0, double* x1, double* x2, double* y, bool gather, hipStream_t s);
void launch_axpy(int64_t n, double alpha, const double* x, double* y, hipStream_t s);
// y = sign * x on N-vectors, zeroing the fault row (first NX entries)
void launch_scale_mask(int64_t n, int nx, const double* x, double sign, double* y, hipStream_t s);

}  // namespace eq

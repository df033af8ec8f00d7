// C ABI of libeqcycle_hip: context, boundary conversion (reference MeshBox
// numbering <-> structured device layout), solver orchestration and the
// device-resident time loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <memory>
#include <tuple>
#include <string>
#include <vector>

#include "eq_internal.h"
#include "eq_kernels.h"
#include "eq_learn.h"

using namespace eq;

#ifndef EQ_CHEB_LOWER
#define EQ_CHEB_LOWER 10.0
#endif

namespace {

constexpr int kMaxChebDegree = 8;   // largest level-0 Chebyshev degree eq_amg_setup accepts
static_assert(kMaxChebDegree <= kSweepMaxDeg, "tile sweeps hold every step's coefficients");

template <typename T>
struct CgWork {
  T *r = nullptr, *c = nullptr, *u[2] = {nullptr, nullptr}, *q = nullptr, *x0 = nullptr, *res0 = nullptr;
  T *b = nullptr, *x = nullptr;   // solve-level vectors for the mixed-precision inner solve
  T *chx = nullptr, *chx2 = nullptr;   // level-0 Chebyshev iterates
  T *chr[3] = {nullptr, nullptr, nullptr}, *chd[3] = {nullptr, nullptr, nullptr};   // residuals, directions
  int cur = 0;
};

template <typename T>
struct DevLevelVecs {
  T *b = nullptr, *x = nullptr, *xp = nullptr, *res = nullptr;
};

struct DevLevel {
  int64_t n = 0;
  DevCsr A, P, R;   // P, R link this level to the next
  DevCsr M, U, V;   // condensed sweeps (condense_level) of levels 1 .. L-2
  double* dinv = nullptr;
  float* dinv32 = nullptr;
  double omega = 0.0;
  DevLevelVecs<double> v64;
  DevLevelVecs<float> v32;
};

}  // namespace

struct eq_ctx {
  std::string err;
  hipStream_t s = nullptr;
  Geom g{};
  int NX = 0, NY = 0;
  int64_t N = 0, nI = 0, Nel = 0;
  std::vector<int64_t> perm;          // reference index (0-based) -> structured node
  std::vector<int64_t> fltni;         // FltNI position -> structured node
  std::vector<void*> allocs;
  // material
  bool have_material = false;
  std::vector<double> mu_host;        // structured
  double* mu = nullptr;
  float* mu32 = nullptr;
  double* dinv = nullptr;             // 1 / diag(K_II) (fault row 0)
  float* dinv32 = nullptr;
  double* ones = nullptr;
  float* ones32 = nullptr;
  // params
  bool have_params = false;
  int64_t FltNglob = 0, NFBC = 1, IDstate = 2;
  double Vpl = 0, Vthres = 1e-3, Vevne = 1e-3, dt0 = 0;
  FaultDev fp{};
  double* M = nullptr;
  std::vector<double> FltX, tauo;
  std::vector<int64_t> out_seis;      // structured
  // state
  bool have_state = false;
  double* state_blk = nullptr;        // [8 N]: d, v, a, d2, v2, a2, dPre, bq (d/v/a swap with d2/v2/a2)
  double *d = nullptr, *v = nullptr, *a = nullptr, *d2 = nullptr, *v2 = nullptr, *a2 = nullptr;
  double *dPre = nullptr, *bq = nullptr;
  FaultState fs{}, fs_api{};
  double* fp_api_blk = nullptr;       // [8 NX] a caller's eq_fault_params on the device
  double* api_nodal = nullptr;        // [6 N] d, v, a in and out of eq_newmark_step
  int32_t* perm32 = nullptr;          // perm on the device (eq_newmark_step_noalloc)
  void* scratch_dev = nullptr;        // this device's address of the file-scope scratch array
  // quasi-static pass 2 run ahead of its convergence read-back: its CG batch,
  // the pass's fault update, the finish and the step tail go out together;
  // psi, Vf and the fault row of d before the fault update are kept to redo
  // them if the batch fell short
  double* spec_save = nullptr;        // [3 NX]
  bool spec_pending = false;
  // d's interior solves K_II d = bq of the last quasi-static solve to its
  // tolerance, and neither has changed since (any other entry point or a
  // dynamic step clears it): the next step's Galerkin start may project
  // just the change of b
  bool qs_r_valid = false;
  double* gal_db = nullptr;           // [4 NX] change of b on rows 1..4
  double* gal_bold = nullptr;         // [4 NX] rows 1..4 of b of the step before (extrapolated start)
  bool qs_prev_r_valid = false;       // the last quasi-static step also followed one
  double qs_prev_dt = 0;              // its dt
  bool spec_f32pc = false;
  int64_t spec_batch = 0;
  // inside eq_run: the run goes on after the quasi-static step being taken,
  // so its end may send the next step's begin and Galerkin start ahead of
  // the read-back, gated on the step tail's status (StepStatus::go)
  bool run_more = false;
  bool begin_sent = false;            // sent, and the status said go
  bool begin_r_fused = false;
  bool spin_recorded = false;         // ev_spin marks the step's end (work queued after it)
  double t = 0, dt = 0;
  int64_t it = 0;
  int isolver = 1;
  // solver
  eq_solver_opts opts{1e-6, 0, EQ_PC_AMG, EQ_PREC_FP64, EQ_GUESS_PREVIOUS, 1, 0, 0, 0, 0};
  // captured CG iteration batches: (precision, f32pc, pc, iterations, parity, x) -> executable graph
  std::map<std::tuple<int, int, int, int64_t, int, const void*, double, int64_t>, hipGraphExec_t> cg_graphs;
  // the next V-cycle's first kernel sets the solve's scalars up (see sweep_pre); tol, maxiter
  bool cg_init_pending = false;
  bool tail_done = false;   // this step's tail already ran (the quasi-static end kernel)
  double cg_init_tol = 0;
  int64_t cg_init_maxiter = 0;
  CgWork<double> w64;
  CgWork<float> w32;
  double *part_rr = nullptr, *part_bb = nullptr, *part_rho = nullptr, *part_uq = nullptr;
  int npart = 0;
  CgScal* sc = nullptr;
  CgScal* sc_host = nullptr;
  // CG iterations enqueued before the first convergence read-back: the last
  // count of the same kind of solve (0: quasi-static pass 1, 1: pass 2, 2: other)
  int64_t pred_iters[3] = {8, 8, 8};
  int cg_slot = 2;
  // AMG
  bool have_amg = false;
  std::vector<DevLevel> lev;          // lev[0]: matrix-free fine level (A unused)
  double* coarse_inv = nullptr;
  float* coarse_inv32 = nullptr;      // FP32 copy for the FP32 V-cycle
  int64_t ncoarse = 0;
  // the condensed coarse levels collapsed to one product down and one up
  // (collapse_coarse), on [b_1 | ... | b_(L-1) | e_(L-1)] in cvec64 / cvec32,
  // where the levels' b (and the last level's xp) then live
  DevCsr coarse_down, coarse_up;
  bool coarse_have = false;           // the collapsed operators exist (L >= 4, fill bounded)
  bool coarse_collapsed = false;      // the V-cycle applies them (eq_amg_coarse_mode)
  double* cvec64 = nullptr;
  float* cvec32 = nullptr;
  std::vector<int64_t> amg_n, amg_nnz;
  int pmg = 0;                        // level-0 prolongation is the bilinear vertex interpolation
  int l0_cheb = 0;                    // level-0 smoother: 0 Jacobi, 1 Chebyshev
  int l0_sweep = 0;                   // Chebyshev as whole tile-local sweeps (one launch per sweep)
  int l0_degree = 1;
  double l0_theta = 1, l0_delta = 1;  // Chebyshev interval centre / half-width of D^-1 K_II
  // step records: slot 0 serves single steps, a batch of dynamic steps uses
  // slots 0..kBatch-1 (status, fault rows of d / v / tau, receiver samples)
  static constexpr int kBatch = 256;
  StepStatus* st = nullptr;           // [kBatch]
  StepStatus* st_host = nullptr;      // [kBatch], pinned
  double* rec_fault = nullptr;        // [kBatch x 3 x NX]
  double* rec_fault_host = nullptr;   // pinned, same shape
  int64_t* seis_idx = nullptr;        // [seis_cap]
  int* seis_tile = nullptr;           // [seis_cap] tile of each receiver
  double* rec_seis = nullptr;         // [kBatch x 3 x seis_cap]
  double* rec_seis_host = nullptr;    // pinned
  int64_t seis_cap = 0;
  // dynamic batches: per step (max, index) of the fault velocity per bottom
  // tile and the FBC! flags, written by the Newmark kernels themselves
  double* vpart = nullptr;            // [kBatch x 2 nbx]
  double* vpart_host = nullptr;       // pinned
  int* nrflags = nullptr;             // [kBatch x 2]
  int* nrflags_host = nullptr;        // pinned
  int* flags = nullptr;
  hipEvent_t ev_spin = nullptr;       // polled by spin_sync
  bool flags_dirty = true;            // flags not known to hold (INT_MAX, 0)
  int64_t nr_noconv_total = 0;
  // results bookkeeping of main(P) (src/main.jl:361-436), carried across eq_run
  // calls so that consecutive calls continue one run; reset by eq_state_init
  struct RunBook {
    double tvsx = 0, tevneb = 0, tevne = 0;
    int idelevne = 0, slipstart = 0;
    std::vector<double> delfref;
  } book;
  int64_t pcg_noconv = 0;
  // learned initial guess
  LearnState learn;
  bool learn_on = false;
  int learn_record_on = 0;
  double *gp = nullptr, *gq = nullptr;
  int64_t n_learned_solves = 0;
  // launch timing (eq_launch_timing): event pairs around launches of the
  // fine-level kernels, summed per kind when the pool fills or is read
  struct LaunchTiming {
    static constexpr int kSlots = 4096;
    bool on = false;
    std::vector<hipEvent_t> ev;       // [2 kSlots]
    int kind[kSlots] = {};
    int used = 0;
    int64_t count[EQ_TK_COUNT] = {};
    double ms[EQ_TK_COUNT] = {};
  } lt;

  ~eq_ctx() {
    for (hipEvent_t e : lt.ev) (void)hipEventDestroy(e);
    for (auto& kv : cg_graphs) (void)hipGraphExecDestroy(kv.second);
    learn_free(&learn);
    for (void* p : allocs) (void)hipFree(p);
    if (sc_host) (void)hipHostFree(sc_host);
    if (st_host) (void)hipHostFree(st_host);
    if (rec_seis_host) (void)hipHostFree(rec_seis_host);
    if (rec_fault_host) (void)hipHostFree(rec_fault_host);
    if (vpart_host) (void)hipHostFree(vpart_host);
    if (nrflags_host) (void)hipHostFree(nrflags_host);
    if (ev_spin) (void)hipEventDestroy(ev_spin);
    if (s) (void)hipStreamDestroy(s);
  }
  template <typename T>
  T* alloc(int64_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(T) * (size_t)std::max<int64_t>(n, 1)) != hipSuccess) return nullptr;
    allocs.push_back(p);
    return (T*)p;
  }
};

#define HIPCHK(ctx, call)                                                            \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                \
      return EQ_HIP_ERR;                                                             \
    }                                                                                \
  } while (0)

#define NEED(ptr)                                                 \
  do {                                                            \
    if (!(ptr)) {                                                 \
      ctx->err = "device allocation failed";                      \
      return EQ_NOMEM;                                            \
    }                                                             \
  } while (0)

// Launch timing: each bracketed launch carries an event pair that its
// dispatch stamps (hipExtLaunchKernelGGL).  Pool full or read: wait for the
// last event and sum per kind.
static int lt_harvest(eq_ctx* ctx) {
  auto& L = ctx->lt;
  if (L.used == 0) return EQ_OK;
  HIPCHK(ctx, hipEventSynchronize(L.ev[2 * L.used - 1]));
  for (int i = 0; i < L.used; i++) {
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, L.ev[2 * i], L.ev[2 * i + 1]));
    L.count[L.kind[i]]++;
    L.ms[L.kind[i]] += ms;
  }
  L.used = 0;
  return EQ_OK;
}
// The events of the next bracketed launch ({} when timing is off)
static LaunchEvents lt_slot(eq_ctx* ctx, int kind) {
  auto& L = ctx->lt;
  if (!L.on) return {};
  if (L.used == L.kSlots) (void)lt_harvest(ctx);
  L.kind[L.used] = kind;
  LaunchEvents e{L.ev[2 * L.used], L.ev[2 * L.used + 1]};
  L.used++;
  return e;
}

// Order of Julia's maximum/findmax (src/main.jl:348,367): NaN is the largest
// value, and among equal values (NaN included) the lowest index wins.
static bool findmax_before_host(double o, double oi, double m, double mi) {
  if (std::isnan(o)) return !std::isnan(m) || oi < mi;
  if (std::isnan(m)) return false;
  return o > m || (o == m && oi < mi);
}

static int fail(eq_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

static int sync_check(eq_ctx* ctx) {
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->s));
  return EQ_OK;
}

// Wait for the stream by polling an event: the time loop reads a few scalars
// back several times per step, and a busy poll answers sooner than a
// blocking synchronisation.  spin_wait polls ev_spin as recorded earlier on
// the stream (work queued after it keeps the device busy meanwhile);
// spin_sync records it at the stream's end first.
static int spin_wait(eq_ctx* ctx) {
  HIPCHK(ctx, hipGetLastError());
  for (;;) {
    const hipError_t e = hipEventQuery(ctx->ev_spin);
    if (e == hipSuccess) return EQ_OK;
    if (e != hipErrorNotReady) {
      ctx->err = std::string("hipEventQuery: ") + hipGetErrorString(e);
      return EQ_HIP_ERR;
    }
  }
}
static int spin_sync(eq_ctx* ctx) {
  HIPCHK(ctx, hipEventRecord(ctx->ev_spin, ctx->s));
  return spin_wait(ctx);
}

template <typename T>
static int h2d(eq_ctx* ctx, T* dst, const T* src, int64_t n) {
  HIPCHK(ctx, hipMemcpyAsync(dst, src, sizeof(T) * (size_t)n, hipMemcpyHostToDevice, ctx->s));
  return EQ_OK;
}
template <typename T>
static int d2h(eq_ctx* ctx, T* dst, const T* src, int64_t n) {
  HIPCHK(ctx, hipMemcpyAsync(dst, src, sizeof(T) * (size_t)n, hipMemcpyDeviceToHost, ctx->s));
  return EQ_OK;
}

// reference-ordered host vector -> structured device vector
static int upload_ref(eq_ctx* ctx, const double* ref, double* dev) {
  std::vector<double> tmp(ctx->N);
  for (int64_t k = 0; k < ctx->N; k++) tmp[ctx->perm[k]] = ref[k];
  if (int rc = h2d(ctx, dev, tmp.data(), ctx->N)) return rc;
  return sync_check(ctx);
}
static int download_ref(eq_ctx* ctx, const double* dev, double* ref) {
  std::vector<double> tmp(ctx->N);
  if (int rc = d2h(ctx, tmp.data(), dev, ctx->N)) return rc;
  if (int rc = sync_check(ctx)) return rc;
  for (int64_t k = 0; k < ctx->N; k++) ref[k] = tmp[ctx->perm[k]];
  return EQ_OK;
}
// FltNI-ordered host vector (length nI) <-> structured device vector (fault row 0)
static int upload_ni(eq_ctx* ctx, const double* x, double* dev) {
  std::vector<double> tmp(ctx->N, 0.0);
  for (int64_t k = 0; k < ctx->nI; k++) tmp[ctx->fltni[k]] = x[k];
  if (int rc = h2d(ctx, dev, tmp.data(), ctx->N)) return rc;
  return sync_check(ctx);
}
static int download_ni(eq_ctx* ctx, const double* dev, double* x) {
  std::vector<double> tmp(ctx->N);
  if (int rc = d2h(ctx, tmp.data(), dev, ctx->N)) return rc;
  if (int rc = sync_check(ctx)) return rc;
  for (int64_t k = 0; k < ctx->nI; k++) x[k] = tmp[ctx->fltni[k]];
  return EQ_OK;
}

// ===========================================================================
// context
// ===========================================================================
// ---------------------------------------------------------------------------
// Scratch of eq_newmark_step_noalloc: one file-scope device array, which each
// GPU that loads the library holds until the process ends.  Layout: the
// double-buffered nodal state (6 x kNoallocMaxNodes doubles), the fault
// staging block (11 x kNoallocMaxFault doubles), the FBC! guard counters.
// In host code `scratch` names a host-side stand-in; the device address comes
// from hipGetSymbolAddress, once per device, in scratch_address.
// ---------------------------------------------------------------------------
constexpr int64_t kNoallocMaxNodes = 1 << 20;   // 1,048,576 nodes
constexpr int64_t kNoallocMaxFault = 1 << 13;   // 8,192 fault nodes
constexpr size_t kNoallocFaultOff = sizeof(double) * 6 * kNoallocMaxNodes;
constexpr size_t kNoallocFlagsOff = kNoallocFaultOff + sizeof(double) * 11 * kNoallocMaxFault;
constexpr size_t kNoallocBytes = kNoallocFlagsOff + 256;
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_addr[64] = {};   // per device ordinal

// The current device's address of `scratch` (looked up on the first call per
// device, with that device current).
static int scratch_address(eq_ctx* ctx) {
  int dev = -1;
  HIPCHK(ctx, hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(ctx, EQ_HIP_ERR, "device ordinal out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(ctx, EQ_HIP_ERR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_addr[dev] = p;
  }
  ctx->scratch_dev = g_scratch_addr[dev];
  return EQ_OK;
}

extern "C" int eq_create(const eq_mesh_desc* mesh, eq_ctx** out) {
  if (!out) return EQ_EINVAL;
  *out = nullptr;
  if (!mesh || !mesh->iglob || !mesh->xgll || !mesh->wgll || !mesh->H) return EQ_EINVAL;
  std::unique_ptr<eq_ctx> ctx(new eq_ctx());
  if (mesh->ngll != 5) { *out = ctx.release(); return fail(*out, EQ_EINVAL, "only NGLL = 5 is supported"); }
  if (mesh->nelx < 1 || mesh->nely < 1 || !(mesh->dxe > 0) || !(mesh->dye > 0)) {
    *out = ctx.release();
    return fail(*out, EQ_EINVAL, "bad mesh sizes");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    *out = ctx.release();
    return fail(*out, EQ_HIP_ERR, "no HIP device visible");
  }
  ctx->NX = (int)(4 * mesh->nelx + 1);
  ctx->NY = (int)(4 * mesh->nely + 1);
  ctx->N = (int64_t)ctx->NX * ctx->NY;
  ctx->nI = ctx->N - ctx->NX;
  ctx->Nel = mesh->nelx * mesh->nely;
  Geom& g = ctx->g;
  g.NX = ctx->NX;
  g.NY = ctx->NY;
  g.nelx = (int)mesh->nelx;
  g.nely = (int)mesh->nely;
  const double dx_dxi = 0.5 * mesh->dxe, dy_deta = 0.5 * mesh->dye, jac = dx_dxi * dy_deta;
  g.c1 = jac / (dx_dxi * dx_dxi);
  g.c2 = jac / (dy_deta * dy_deta);
  for (int p = 0; p < 5; p++)
    for (int i = 0; i < 5; i++) g.D[p * 5 + i] = mesh->H[p * 5 + i];   // H[i + 5 p] = h'_i(x_p)
  for (int p = 0; p < 5; p++) g.w[p] = mesh->wgll[p];
  for (int p = 0; p < 5; p++) g.x[p] = mesh->xgll[p];
  // numbering: iglob[i,j,e] -> node (4 ex + i, 4 ey + j)
  ctx->perm.assign(ctx->N, -1);
  std::vector<char> seen(ctx->N, 0);
  for (int64_t e = 0; e < ctx->Nel; e++) {
    const int64_t ex = e % mesh->nelx, ey = e / mesh->nelx;
    for (int j = 0; j < 5; j++)
      for (int i = 0; i < 5; i++) {
        const int64_t ref = mesh->iglob[e * 25 + j * 5 + i] - 1;
        const int64_t n = (4 * ey + j) * ctx->NX + 4 * ex + i;
        if (ref < 0 || ref >= ctx->N) { *out = ctx.release(); return fail(*out, EQ_EINVAL, "iglob out of range"); }
        if (ctx->perm[ref] == -1) {
          ctx->perm[ref] = n;
          if (seen[n]) { *out = ctx.release(); return fail(*out, EQ_EINVAL, "iglob is not a box mesh"); }
          seen[n] = 1;
        } else if (ctx->perm[ref] != n) {
          *out = ctx.release();
          return fail(*out, EQ_EINVAL, "iglob is not a structured box mesh (MeshBox!)");
        }
      }
  }
  for (int64_t k = 0; k < ctx->N; k++)
    if (ctx->perm[k] < 0) { *out = ctx.release(); return fail(*out, EQ_EINVAL, "iglob does not cover all nodes"); }
  // FltNI: non-fault nodes in increasing reference index (par.jl:209)
  ctx->fltni.reserve(ctx->nI);
  for (int64_t k = 0; k < ctx->N; k++)
    if (ctx->perm[k] >= ctx->NX) ctx->fltni.push_back(ctx->perm[k]);
  if (hipStreamCreate(&ctx->s) != hipSuccess) { *out = ctx.release(); return fail(*out, EQ_HIP_ERR, "stream"); }
  if (hipEventCreateWithFlags(&ctx->ev_spin, hipEventDisableTiming) != hipSuccess) {
    *out = ctx.release();
    return fail(*out, EQ_HIP_ERR, "event");
  }
  eq_ctx* c = ctx.get();
  const int64_t N = c->N;
  c->npart = std::max(tile_count(g), vec_blocks(N)) + 1;
  c->part_rr = c->alloc<double>(c->npart);
  c->part_bb = c->alloc<double>(c->npart);
  c->part_rho = c->alloc<double>(c->npart);
  c->part_uq = c->alloc<double>(c->npart);
  c->sc = c->alloc<CgScal>(1);
  c->st = c->alloc<StepStatus>(eq_ctx::kBatch);
  c->rec_fault = c->alloc<double>((int64_t)eq_ctx::kBatch * 3 * c->NX);
  const int nbx = (int)tile_grid(g).x;
  c->vpart = c->alloc<double>((int64_t)eq_ctx::kBatch * 2 * nbx);
  c->nrflags = c->alloc<int>(2 * eq_ctx::kBatch);
  c->flags = c->alloc<int>(4);
  if (!c->part_rr || !c->part_bb || !c->part_rho || !c->part_uq || !c->sc || !c->st || !c->flags || !c->rec_fault ||
      !c->vpart || !c->nrflags) {
    *out = ctx.release();
    return fail(*out, EQ_NOMEM, "device allocation failed");
  }
  // sc_host and st_host (and rec_seis_host) are also written by kernels
  // directly (coherent: no GPU caching), read by the host after a
  // synchronisation
  if (hipHostMalloc((void**)&c->sc_host, sizeof(CgScal), hipHostMallocCoherent) != hipSuccess ||
      hipHostMalloc((void**)&c->st_host, sizeof(StepStatus) * eq_ctx::kBatch, hipHostMallocCoherent) != hipSuccess ||
      hipHostMalloc((void**)&c->vpart_host, sizeof(double) * eq_ctx::kBatch * 2 * nbx, hipHostMallocDefault) !=
          hipSuccess ||
      hipHostMalloc((void**)&c->nrflags_host, sizeof(int) * 2 * eq_ctx::kBatch, hipHostMallocDefault) !=
          hipSuccess ||
      hipHostMalloc((void**)&c->rec_fault_host, sizeof(double) * eq_ctx::kBatch * 3 * c->NX, hipHostMallocDefault) !=
          hipSuccess) {
    *out = ctx.release();
    return fail(*out, EQ_NOMEM, "pinned allocation failed");
  }
  // the reference -> structured permutation on the device, and the scratch
  // address of eq_newmark_step_noalloc (which allocates nothing itself)
  if (N < INT32_MAX) {
    std::vector<int32_t> p32(c->perm.begin(), c->perm.end());
    c->perm32 = c->alloc<int32_t>(N);
    if (!c->perm32) { *out = ctx.release(); return fail(*out, EQ_NOMEM, "device allocation failed"); }
    if (int rc = h2d(c, c->perm32, p32.data(), N)) { *out = ctx.release(); return rc; }
    if (int rc = sync_check(c)) { *out = ctx.release(); return rc; }
  }
  if (int rc = scratch_address(c)) { *out = ctx.release(); return rc; }
  *out = ctx.release();
  return EQ_OK;
}

extern "C" void eq_destroy(eq_ctx* ctx) { delete ctx; }

extern "C" const char* eq_last_error(const eq_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int eq_sizes(const eq_ctx* ctx, int64_t* nglob, int64_t* nI, int64_t* nfault) {
  if (!ctx) return EQ_EINVAL;
  if (nglob) *nglob = ctx->N;
  if (nI) *nI = ctx->nI;
  if (nfault) *nfault = ctx->NX;
  return EQ_OK;
}

template <typename T>
static int alloc_work(eq_ctx* ctx, CgWork<T>& w) {
  if (w.r) return EQ_OK;
  const int64_t N = ctx->N;
  w.r = ctx->alloc<T>(N); w.c = ctx->alloc<T>(N); w.u[0] = ctx->alloc<T>(N); w.u[1] = ctx->alloc<T>(N);
  w.q = ctx->alloc<T>(N); w.x0 = ctx->alloc<T>(N); w.res0 = ctx->alloc<T>(N); w.b = ctx->alloc<T>(N);
  w.x = ctx->alloc<T>(N);
  w.chx = ctx->alloc<T>(N); w.chx2 = ctx->alloc<T>(N);
  for (int i = 0; i < 3; i++) { w.chr[i] = ctx->alloc<T>(N); w.chd[i] = ctx->alloc<T>(N); }
  if (!w.r || !w.c || !w.u[0] || !w.u[1] || !w.q || !w.x0 || !w.res0 || !w.b || !w.x || !w.chx || !w.chx2 ||
      !w.chr[0] || !w.chr[1] || !w.chr[2] || !w.chd[0] || !w.chd[1] || !w.chd[2])
    return fail(ctx, EQ_NOMEM, "device allocation failed");
  launch_fill<T>(N, w.x0, T(0), ctx->s);
  launch_fill<T>(N, w.res0, T(0), ctx->s);
  return EQ_OK;
}

extern "C" int eq_set_material(eq_ctx* ctx, const double* W) {
  if (!ctx || !W) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  const int NX = ctx->NX;
  std::vector<double> mu(ctx->N, NAN);
  const double* w = ctx->g.w;
  for (int64_t e = 0; e < ctx->Nel; e++) {
    const int64_t ex = e % ctx->g.nelx, ey = e / ctx->g.nelx;
    for (int j = 0; j < 5; j++)
      for (int i = 0; i < 5; i++) {
        const int64_t n = (4 * ey + j) * NX + 4 * ex + i;
        const double m = W[e * 25 + j * 5 + i] / (w[i] * w[j]);
        if (std::isnan(mu[n])) mu[n] = m;
        else if (std::fabs(mu[n] - m) > 1e-12 * std::fabs(m))
          return fail(ctx, EQ_EINVAL, "W is not wgll2 .* nodal mu (rigidity jumps across an element edge)");
      }
  }
  ctx->mu_host = mu;
  if (!ctx->mu) {
    ctx->mu = ctx->alloc<double>(ctx->N); ctx->mu32 = ctx->alloc<float>(ctx->N);
    ctx->dinv = ctx->alloc<double>(ctx->N); ctx->dinv32 = ctx->alloc<float>(ctx->N);
    ctx->ones = ctx->alloc<double>(ctx->N); ctx->ones32 = ctx->alloc<float>(ctx->N);
    NEED(ctx->mu); NEED(ctx->mu32); NEED(ctx->dinv); NEED(ctx->dinv32); NEED(ctx->ones); NEED(ctx->ones32);
  }
  if (int rc = h2d(ctx, ctx->mu, mu.data(), ctx->N)) return rc;
  launch_convert<double, float>(ctx->N, ctx->mu, ctx->mu32, ctx->s);
  // Jacobi diagonal of K_II, fault row zeroed
  launch_diag(ctx->g, ctx->mu, ctx->dinv, ctx->s);
  std::vector<double> dg(ctx->N);
  if (int rc = d2h(ctx, dg.data(), ctx->dinv, ctx->N)) return rc;
  if (int rc = sync_check(ctx)) return rc;
  for (int64_t n = 0; n < ctx->N; n++) dg[n] = (n < NX) ? 0.0 : 1.0 / dg[n];
  if (int rc = h2d(ctx, ctx->dinv, dg.data(), ctx->N)) return rc;
  launch_convert<double, float>(ctx->N, ctx->dinv, ctx->dinv32, ctx->s);
  for (int64_t n = 0; n < ctx->N; n++) dg[n] = (n < NX) ? 0.0 : 1.0;
  if (int rc = h2d(ctx, ctx->ones, dg.data(), ctx->N)) return rc;
  launch_convert<double, float>(ctx->N, ctx->ones, ctx->ones32, ctx->s);
  if (int rc = alloc_work(ctx, ctx->w64)) return rc;
  if (int rc = alloc_work(ctx, ctx->w32)) return rc;
  ctx->have_material = true;
  ctx->have_amg = false;
  return sync_check(ctx);
}

static FaultState alloc_fault_state(eq_ctx* ctx, int64_t n) {
  FaultState f{};
  double* blk = ctx->alloc<double>(11 * n);
  if (!blk) return f;
  f.psi = blk; f.psi1 = blk + n; f.psi2 = blk + 2 * n; f.Vf = blk + 3 * n; f.Vf0 = blk + 4 * n;
  f.Vf1 = blk + 5 * n; f.Vf2 = blk + 6 * n; f.tau = blk + 7 * n; f.tau1 = blk + 8 * n; f.tau2 = blk + 9 * n;
  f.FltVfree = blk + 10 * n;
  launch_fill<double>(11 * n, blk, 0.0, ctx->s);
  return f;
}

extern "C" int eq_set_params(eq_ctx* ctx, const eq_params_desc* p) {
  if (!ctx || !p) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  const int NX = ctx->NX, NY = ctx->NY;
  if (p->FltNglob != NX || p->nglob != ctx->N) return fail(ctx, EQ_EINVAL, "params do not match the mesh");
  const double* need[] = {p->fo, p->Vo, p->xLf, p->M, p->FltB, p->FltZ, p->FltX, p->cca, p->ccb, p->Seff, p->tauo,
                          p->XiLf};
  for (const double* q : need)
    if (!q) return fail(ctx, EQ_EINVAL, "missing params array");
  if (!p->iFlt) return fail(ctx, EQ_EINVAL, "missing iFlt");
  for (int64_t j = 0; j < NX; j++)
    if (p->iFlt[j] < 1 || p->iFlt[j] > ctx->N || ctx->perm[p->iFlt[j] - 1] != j)
      return fail(ctx, EQ_EINVAL, "iFlt is not the y = 0 edge in depth order (BoundaryMatrix! side 'B')");
  if (p->n_iBcL != NY || (p->n_iBcL && !p->iBcL) || p->n_iBcT != NX || (p->n_iBcT && !p->iBcT) || !p->BcLC ||
      !p->BcTC)
    return fail(ctx, EQ_EINVAL, "absorbing edges must be the full 'L' and 'T' sides");
  for (int64_t k = 0; k < NY; k++)
    if (p->iBcL[k] < 1 || p->iBcL[k] > ctx->N || ctx->perm[p->iBcL[k] - 1] != k * NX)
      return fail(ctx, EQ_EINVAL, "iBcL is not side 'L'");
  for (int64_t k = 0; k < NX; k++)
    if (p->iBcT[k] < 1 || p->iBcT[k] > ctx->N || ctx->perm[p->iBcT[k] - 1] != (int64_t)(NY - 1) * NX + k)
      return fail(ctx, EQ_EINVAL, "iBcT is not side 'T'");
  // creeping nodes: FltIglobBC must be the leading (deepest) fault nodes
  std::vector<char> creep(NX, 0);
  for (int64_t k = 0; k < p->n_FltIglobBC; k++) {
    const int64_t r = p->FltIglobBC[k] - 1;
    if (r < 0 || r >= ctx->N || ctx->perm[r] >= NX) return fail(ctx, EQ_EINVAL, "FltIglobBC must be fault nodes");
    creep[ctx->perm[r]] = 1;
  }
  int64_t ncreep = 0;
  while (ncreep < NX && creep[ncreep]) ncreep++;
  for (int64_t j = ncreep; j < NX; j++)
    if (creep[j]) return fail(ctx, EQ_EINVAL, "FltIglobBC must be the deepest fault nodes");
  ctx->FltNglob = NX;
  ctx->NFBC = ncreep + 1;
  ctx->IDstate = p->IDstate;
  ctx->Vpl = p->Vpl; ctx->Vthres = p->Vthres; ctx->Vevne = p->Vevne; ctx->dt0 = p->dt0;
  ctx->FltX.assign(p->FltX, p->FltX + NX);
  ctx->tauo.assign(p->tauo, p->tauo + NX);
  ctx->out_seis.clear();
  for (int64_t k = 0; k < p->n_out_seis; k++) {
    const int64_t r = p->out_seis[k] - 1;
    if (r < 0 || r >= ctx->N) return fail(ctx, EQ_EINVAL, "out_seis out of range");
    ctx->out_seis.push_back(ctx->perm[r]);
  }
  const int64_t nseis = (int64_t)ctx->out_seis.size();
  if (nseis > 31) return fail(ctx, EQ_EINVAL, "at most 31 receivers (out_seis)");
  if (nseis > ctx->seis_cap) {
    ctx->seis_idx = ctx->alloc<int64_t>(nseis);
    ctx->seis_tile = ctx->alloc<int>(nseis);
    ctx->rec_seis = ctx->alloc<double>((int64_t)eq_ctx::kBatch * 3 * nseis);
    NEED(ctx->seis_idx);
    NEED(ctx->seis_tile);
    NEED(ctx->rec_seis);
    if (ctx->rec_seis_host) (void)hipHostFree(ctx->rec_seis_host);
    ctx->rec_seis_host = nullptr;
    if (hipHostMalloc((void**)&ctx->rec_seis_host, sizeof(double) * eq_ctx::kBatch * 3 * nseis,
                      hipHostMallocCoherent) != hipSuccess)
      return fail(ctx, EQ_NOMEM, "pinned allocation failed");
    ctx->seis_cap = nseis;
  }
  if (nseis > 0) {
    if (int rc = h2d(ctx, ctx->seis_idx, ctx->out_seis.data(), nseis)) return rc;
    std::vector<int> tiles(nseis);
    const int ntx = (int)tile_grid(ctx->g).x;
    for (int64_t q = 0; q < nseis; q++) {
      const int64_t n = ctx->out_seis[q];
      tiles[q] = (int)(n / NX / TILE_J) * ntx + (int)(n % NX / TILE_I);
    }
    if (int rc = h2d(ctx, ctx->seis_tile, tiles.data(), nseis)) return rc;
    if (int rc = sync_check(ctx)) return rc;
  }
  // device copies
  double* blk = ctx->alloc<double>(11 * (int64_t)NX + NY + NX);
  NEED(blk);
  const double* src[11] = {p->fo, p->Vo, p->xLf, p->cca, p->ccb, p->Seff, p->tauo, p->XiLf, p->FltB, p->FltZ, nullptr};
  std::vector<double> Mf(NX);
  std::vector<double> Ms(ctx->N);
  for (int64_t k = 0; k < ctx->N; k++) Ms[ctx->perm[k]] = p->M[k];
  for (int j = 0; j < NX; j++) Mf[j] = Ms[j];
  src[10] = Mf.data();
  for (int k = 0; k < 11; k++)
    if (int rc = h2d(ctx, blk + (int64_t)k * NX, src[k], NX)) return rc;
  if (int rc = h2d(ctx, blk + 11 * (int64_t)NX, p->BcLC, NY)) return rc;
  if (int rc = h2d(ctx, blk + 11 * (int64_t)NX + NY, p->BcTC, NX)) return rc;
  FaultDev& f = ctx->fp;
  f.fo = blk; f.Vo = blk + NX; f.xLf = blk + 2 * NX; f.cca = blk + 3 * NX; f.ccb = blk + 4 * NX;
  f.Seff = blk + 5 * NX; f.tauo = blk + 6 * NX; f.XiLf = blk + 7 * NX; f.FltB = blk + 8 * NX;
  f.FltZ = blk + 9 * NX; f.Mf = blk + 10 * NX; f.BcLC = blk + 11 * NX; f.BcTC = blk + 11 * NX + NY;
  f.j0 = (int)(ctx->NFBC - 1);
  f.nflt = NX;
  f.IDstate = (int)p->IDstate;
  f.Vpl = p->Vpl;
  if (!ctx->M) {
    ctx->M = ctx->alloc<double>(ctx->N);
    NEED(ctx->M);
  }
  if (int rc = h2d(ctx, ctx->M, Ms.data(), ctx->N)) return rc;
  if (!ctx->fs.psi) {
    ctx->fs = alloc_fault_state(ctx, NX);
    ctx->fs_api = alloc_fault_state(ctx, NX);
    ctx->fp_api_blk = ctx->alloc<double>(8 * NX);
    if (!ctx->fs.psi || !ctx->fs_api.psi || !ctx->fp_api_blk) return fail(ctx, EQ_NOMEM, "device allocation failed");
  }
  ctx->have_params = true;
  return sync_check(ctx);
}

// ===========================================================================
// operators
// ===========================================================================
extern "C" int eq_ku_apply(eq_ctx* ctx, const double* u, double* y, int mask, int prec) {
  if (!ctx || !u || !y) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (!ctx->have_material) return fail(ctx, EQ_ENOTREADY, "eq_set_material first");
  if (mask < EQ_MASK_ALL || mask > EQ_MASK_FAULT_ROWS) return fail(ctx, EQ_EINVAL, "bad mask");
  const int interior = mask == EQ_MASK_INTERIOR, neg = mask == EQ_MASK_NEG;
  double* du = ctx->w64.x;
  double* dy = ctx->w64.q;
  if (int rc = upload_ref(ctx, u, du)) return rc;
  if (mask == EQ_MASK_FAULT_ROWS) {
    launch_fill<double>(ctx->N, dy, 0.0, ctx->s);
    if (prec == EQ_PREC_FP32) {
      launch_convert<double, float>(ctx->N, du, ctx->w32.x, ctx->s);
      launch_fill<float>(ctx->N, ctx->w32.q, 0.0f, ctx->s);
      launch_ku_rows<float>(ctx->g, ctx->w32.x, ctx->mu32, ctx->w32.q, 5, ctx->s);
      launch_convert<float, double>(ctx->N, ctx->w32.q, dy, ctx->s);
    } else {
      launch_ku_rows<double>(ctx->g, du, ctx->mu, dy, 5, ctx->s);
    }
  } else if (prec == EQ_PREC_FP32) {
    launch_convert<double, float>(ctx->N, du, ctx->w32.x, ctx->s);
    launch_ku<float>(ctx->g, ctx->w32.x, ctx->mu32, ctx->w32.q, interior, neg, ctx->s);
    launch_convert<float, double>(ctx->N, ctx->w32.q, dy, ctx->s);
  } else {
    launch_ku<double>(ctx->g, du, ctx->mu, dy, interior, neg, ctx->s);
  }
  if (int rc = sync_check(ctx)) return rc;
  return download_ref(ctx, dy, y);
}

extern "C" int eq_kni_apply(eq_ctx* ctx, const double* x, double* y, int prec) {
  if (!ctx || !x || !y) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (!ctx->have_material) return fail(ctx, EQ_ENOTREADY, "eq_set_material first");
  double* dx = ctx->w64.x;
  double* dy = ctx->w64.q;
  if (int rc = upload_ni(ctx, x, dx)) return rc;
  if (prec == EQ_PREC_FP32) {
    launch_convert<double, float>(ctx->N, dx, ctx->w32.x, ctx->s);
    launch_ku<float>(ctx->g, ctx->w32.x, ctx->mu32, ctx->w32.q, 1, 1, ctx->s);
    launch_convert<float, double>(ctx->N, ctx->w32.q, dy, ctx->s);
  } else {
    launch_ku<double>(ctx->g, dx, ctx->mu, dy, 1, 1, ctx->s);
  }
  if (int rc = sync_check(ctx)) return rc;
  return download_ni(ctx, dy, y);
}

extern "C" int eq_element_computation(eq_ctx* ctx, const double* u, double* a_elem) {
  if (!ctx || !u || !a_elem) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (!ctx->have_material) return fail(ctx, EQ_ENOTREADY, "eq_set_material first");
  double* du = ctx->w64.x;
  if (int rc = upload_ref(ctx, u, du)) return rc;
  double* ae = ctx->alloc<double>(25 * ctx->Nel);
  NEED(ae);
  launch_element_ku(ctx->g, du, ctx->mu, ae, ctx->s);
  int rc = d2h(ctx, a_elem, ae, 25 * ctx->Nel);
  if (!rc) rc = sync_check(ctx);
  ctx->allocs.pop_back();
  (void)hipFree(ae);
  return rc;
}

// ===========================================================================
// AMG
// ===========================================================================
static int upload_csr(eq_ctx* ctx, const HostCsr& H, DevCsr* D) {
  D->n = H.n;
  D->m = H.m;
  D->nnz = H.nnz();
  if (D->nnz > INT_MAX) return fail(ctx, EQ_EINVAL, "level too large for 32-bit CSR offsets");
  std::vector<int32_t> rp(H.n + 1);
  for (int64_t i = 0; i <= H.n; i++) rp[i] = (int32_t)H.rowptr[i];
  D->rowptr = ctx->alloc<int32_t>(H.n + 1);
  D->col = ctx->alloc<int32_t>(D->nnz);
  D->val = ctx->alloc<double>(D->nnz);
  D->val32 = ctx->alloc<float>(D->nnz);
  NEED(D->rowptr); NEED(D->col); NEED(D->val); NEED(D->val32);
  if (int rc = h2d(ctx, D->rowptr, rp.data(), H.n + 1)) return rc;
  if (D->nnz) {
    if (int rc = h2d(ctx, D->col, H.col.data(), D->nnz)) return rc;
    if (int rc = h2d(ctx, D->val, H.val.data(), D->nnz)) return rc;
    launch_convert<double, float>(D->nnz, D->val, D->val32, ctx->s);
  }
  const double avg = H.n ? (double)D->nnz / (double)H.n : 1.0;
  D->lpr = avg <= 2.5 ? 1 : avg <= 5 ? 2 : avg <= 10 ? 4 : avg <= 24 ? 8 : 16;
  return EQ_OK;
}

template <typename T>
static int alloc_level_vecs(eq_ctx* ctx, DevLevelVecs<T>& v, int64_t n) {
  v.b = ctx->alloc<T>(n); v.x = ctx->alloc<T>(n); v.xp = ctx->alloc<T>(n); v.res = ctx->alloc<T>(n);
  NEED(v.b); NEED(v.x); NEED(v.xp); NEED(v.res);
  return EQ_OK;
}

// Device copy of a host hierarchy: level 0 stays matrix-free (K_II from mu)
// with Jacobi or Chebyshev smoothing; levels 1 .. L-2 as condensed sweeps;
// the last level's dense inverse.
static int upload_hierarchy(eq_ctx* ctx, const HostHierarchy& H, int smoother, int smooth_degree, int pcoarsen) {
  if (H.lev.size() < 2) return fail(ctx, EQ_EINVAL, "AMG hierarchy has a single level; use the Jacobi preconditioner");
  ctx->have_amg = false;
  for (auto& kv : ctx->cg_graphs) (void)hipGraphExecDestroy(kv.second);   // they hold the old levels
  ctx->cg_graphs.clear();
  ctx->l0_cheb = smoother == EQ_SMOOTH_CHEBYSHEV || smoother == EQ_SMOOTH_CHEBYSHEV_TILES;
  ctx->l0_sweep = smoother == EQ_SMOOTH_CHEBYSHEV_TILES;
  ctx->l0_degree = ctx->l0_cheb ? smooth_degree : 1;
  ctx->coarse_have = ctx->coarse_collapsed = false;
  std::vector<HostCsr> cM, cU, cV;   // the condensed sweeps of levels 1 .. L-2
  {
    // Chebyshev target interval [lambda_max / EQ_CHEB_LOWER, 1.1 lambda_max] of D^-1 K_II
    const double lmax = 1.1 * H.lev[0].lambda_max, lmin = H.lev[0].lambda_max / EQ_CHEB_LOWER;
    ctx->l0_theta = 0.5 * (lmax + lmin);
    ctx->l0_delta = 0.5 * (lmax - lmin);
  }
  ctx->lev.clear();
  ctx->lev.resize(H.lev.size());
  ctx->amg_n.clear();
  ctx->amg_nnz.clear();
  for (size_t l = 0; l < H.lev.size(); l++) {
    DevLevel& D = ctx->lev[l];
    const HostLevel& L = H.lev[l];
    D.n = L.A.n;
    ctx->amg_n.push_back(L.A.n);
    ctx->amg_nnz.push_back(L.A.nnz());
    D.omega = 4.0 / (3.0 * L.lambda_max);
    if (l > 0 && l + 1 < H.lev.size()) {
      if (int rc = upload_csr(ctx, L.A, &D.A)) return rc;  // the coarsest level uses its dense inverse
      D.dinv = ctx->alloc<double>(D.n);
      D.dinv32 = ctx->alloc<float>(D.n);
      NEED(D.dinv); NEED(D.dinv32);
      if (int rc = h2d(ctx, D.dinv, L.dinv.data(), D.n)) return rc;
      launch_convert<double, float>(D.n, D.dinv, D.dinv32, ctx->s);
    }
    if (l + 1 < H.lev.size()) {
      if (int rc = upload_csr(ctx, L.P, &D.P)) return rc;
      if (int rc = upload_csr(ctx, L.R, &D.R)) return rc;
    }
    if (l > 0 && l + 1 < H.lev.size()) {
      HostCsr M, U, V;
      condense_level(L, D.omega, &M, &U, &V);
      if (int rc = upload_csr(ctx, M, &D.M)) return rc;
      if (int rc = upload_csr(ctx, U, &D.U)) return rc;
      if (int rc = upload_csr(ctx, V, &D.V)) return rc;
      cM.push_back(std::move(M));
      cU.push_back(std::move(U));
      cV.push_back(std::move(V));
    }
    if (l > 0) {
      if (int rc = alloc_level_vecs(ctx, D.v64, D.n)) return rc;
      if (int rc = alloc_level_vecs(ctx, D.v32, D.n)) return rc;
    }
  }
  // dense coarsest inverse (row-major; one wavefront per row applies it)
  ctx->ncoarse = H.ncoarse;
  ctx->coarse_inv = ctx->alloc<double>(H.ncoarse * H.ncoarse);
  NEED(ctx->coarse_inv);
  if (int rc = h2d(ctx, ctx->coarse_inv, H.coarse_inv.data(), H.ncoarse * H.ncoarse)) return rc;
  ctx->coarse_inv32 = ctx->alloc<float>(H.ncoarse * H.ncoarse);
  NEED(ctx->coarse_inv32);
  launch_convert<double, float>(H.ncoarse * H.ncoarse, ctx->coarse_inv, ctx->coarse_inv32, ctx->s);
  // two or more condensed levels: collapse them (one launch down, one up for 2 k
  // launches), unless the fill costs more HBM time than the launches saved: at
  // most 8 MB more to stream (8 bytes per entry: FP32 value and column) per
  // launch saved, about 2 us at the rate these short products reach
  const int nl = (int)H.lev.size();
  if (nl >= 4) {
    std::vector<int64_t> off(nl + 1, 0);   // off[l]: b_l (l = 1 .. nl-1); off[nl]: e_(nl-1)
    for (int l = 1; l < nl; l++) off[l + 1] = off[l] + ctx->lev[l].n;
    const int64_t total = off[nl] + ctx->lev[nl - 1].n;
    std::vector<int64_t> col_off;
    for (int l = 1; l <= nl - 2; l++) col_off.push_back(off[l]);
    col_off.push_back(off[nl]);
    HostCsr S, Cu;
    collapse_coarse(cM, cU, cV, col_off, total, &S, &Cu);
    int64_t per_level = 0;
    for (size_t j = 0; j < cM.size(); j++) per_level += cM[j].nnz() + cU[j].nnz() + cV[j].nnz();
    const int64_t extra_bytes = 8 * (S.nnz() + Cu.nnz() - per_level);
    const int64_t saved = 2 * (int64_t)cM.size() - 2;
    if (extra_bytes <= saved * (8ll << 20)) {
      if (int rc = upload_csr(ctx, S, &ctx->coarse_down)) return rc;
      if (int rc = upload_csr(ctx, Cu, &ctx->coarse_up)) return rc;
      ctx->cvec64 = ctx->alloc<double>(total);
      ctx->cvec32 = ctx->alloc<float>(total);
      NEED(ctx->cvec64); NEED(ctx->cvec32);
      for (int l = 1; l < nl; l++) {
        ctx->lev[l].v64.b = ctx->cvec64 + off[l];
        ctx->lev[l].v32.b = ctx->cvec32 + off[l];
      }
      ctx->lev[nl - 1].v64.xp = ctx->cvec64 + off[nl];
      ctx->lev[nl - 1].v32.xp = ctx->cvec32 + off[nl];
      ctx->coarse_have = ctx->coarse_collapsed = true;
    }
  }
  ctx->pmg = pcoarsen ? 1 : 0;
  ctx->have_amg = true;
  return sync_check(ctx);
}

// Which form of the coarse levels the V-cycle applies (no reference
// counterpart; for tests that compare them): 1 collapsed, 0 level by level.
extern "C" int eq_amg_coarse_mode(eq_ctx* ctx, int collapsed) {
  if (!ctx) return EQ_EINVAL;
  if (!ctx->have_amg) return fail(ctx, EQ_ENOTREADY, "eq_amg_setup first");
  if (collapsed && !ctx->coarse_have) return fail(ctx, EQ_EINVAL, "no collapsed coarse levels in this hierarchy");
  if ((collapsed != 0) != ctx->coarse_collapsed) {
    for (auto& kv : ctx->cg_graphs) (void)hipGraphExecDestroy(kv.second);   // they hold the other form
    ctx->cg_graphs.clear();
    ctx->coarse_collapsed = collapsed != 0;
  }
  return EQ_OK;
}

extern "C" int eq_amg_setup(eq_ctx* ctx, double theta, int max_levels, int64_t max_coarse, int strength_abs,
                            int smoother, int smooth_degree, int pcoarsen) {
  if (!ctx) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (!ctx->have_material) return fail(ctx, EQ_ENOTREADY, "eq_set_material first");
  if (smoother != EQ_SMOOTH_JACOBI && smoother != EQ_SMOOTH_CHEBYSHEV && smoother != EQ_SMOOTH_CHEBYSHEV_TILES)
    return fail(ctx, EQ_EINVAL, "bad smoother");
  if (smoother != EQ_SMOOTH_JACOBI && (smooth_degree < 1 || smooth_degree > kMaxChebDegree))
    return fail(ctx, EQ_EINVAL, "Chebyshev degree must be 1..8");
  if (max_levels < 2) return fail(ctx, EQ_EINVAL, "max_levels must be >= 2");
  HostCsr Kii;
  assemble_kii(ctx->g, ctx->mu_host.data(), &Kii);
  AmgOptions opt;
  opt.theta = theta;
  opt.max_levels = max_levels;
  opt.max_coarse = max_coarse;
  opt.strength_abs = strength_abs;
  HostHierarchy H;
  std::string err;
  if (pcoarsen) {
    if (int rc = pmg_build(ctx->g, Kii, opt, &H, &err)) return fail(ctx, rc, err);
  } else {
    if (int rc = amg_build(Kii, opt, &H, &err)) return fail(ctx, rc, err);
  }
  return upload_hierarchy(ctx, H, smoother, smooth_degree, pcoarsen);
}

// A hierarchy the caller built (the reference's ruge_stuben(kni)).
extern "C" int eq_amg_upload(eq_ctx* ctx, int nlev, const eq_csr* A, const eq_csr* P, const double* omega,
                             int smoother, int smooth_degree) {
  if (!ctx || !A || !P) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (!ctx->have_material) return fail(ctx, EQ_ENOTREADY, "eq_set_material first");
  if (nlev < 2 || nlev > 32) return fail(ctx, EQ_EINVAL, "nlev must be 2..32");
  if (smoother != EQ_SMOOTH_JACOBI && smoother != EQ_SMOOTH_CHEBYSHEV && smoother != EQ_SMOOTH_CHEBYSHEV_TILES)
    return fail(ctx, EQ_EINVAL, "bad smoother");
  if (smoother != EQ_SMOOTH_JACOBI && (smooth_degree < 1 || smooth_degree > kMaxChebDegree))
    return fail(ctx, EQ_EINVAL, "Chebyshev degree must be 1..8");
  if (omega)
    for (int l = 0; l < nlev; l++)
      if (!(omega[l] > 0) || !std::isfinite(omega[l])) return fail(ctx, EQ_EINVAL, "omega must be positive");
  const int64_t nI = ctx->nI, NX = ctx->NX;
  auto take = [&](const eq_csr& c, int64_t n, int64_t m, const std::string& what, HostCsr* out) -> int {
    if (c.n != n || c.m != m || !c.rowptr) return fail(ctx, EQ_EINVAL, what + ": wrong shape");
    if (c.rowptr[0] != 0) return fail(ctx, EQ_EINVAL, what + ": rowptr[0] must be 0");
    for (int64_t i = 0; i < n; i++)
      if (c.rowptr[i + 1] < c.rowptr[i]) return fail(ctx, EQ_EINVAL, what + ": rowptr not monotone");
    const int64_t nnz = c.rowptr[n];
    if (nnz > 0 && (!c.col || !c.val)) return fail(ctx, EQ_EINVAL, what + ": missing col / val");
    if (m > INT32_MAX) return fail(ctx, EQ_EINVAL, what + ": too many columns");
    out->n = n;
    out->m = m;
    out->rowptr.assign(c.rowptr, c.rowptr + n + 1);
    out->col.resize(nnz);
    out->val.assign(c.val, c.val + nnz);
    for (int64_t q = 0; q < nnz; q++) {
      if (c.col[q] < 0 || c.col[q] >= m) return fail(ctx, EQ_EINVAL, what + ": column index out of range");
      out->col[q] = (int32_t)c.col[q];
    }
    return EQ_OK;
  };
  std::vector<int64_t> sz(nlev);
  for (int l = 0; l < nlev; l++) sz[l] = A[l].n;
  if (sz[0] != nI) return fail(ctx, EQ_EINVAL, "A[0] must be kni, of size nI");
  HostHierarchy H;
  H.lev.resize(nlev);
  // level 0: the caller's A_0 (FltNI order) must be kni = -K_II or K_II of this mesh
  HostCsr A0, Kii;
  if (int rc = take(A[0], nI, nI, "A[0]", &A0)) return rc;
  assemble_kii(ctx->g, ctx->mu_host.data(), &Kii);
  std::vector<int64_t> rmap(nI);   // FltNI index -> structured interior index
  for (int64_t k = 0; k < nI; k++) rmap[k] = ctx->fltni[k] - NX;
  double sgn = 1.0;
  {
    std::vector<double> x(nI), xs(nI), y0(nI, 0.0), y1(nI, 0.0);
    uint64_t st = 0x9e3779b97f4a7c15ull;
    for (int64_t k = 0; k < nI; k++) {
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      x[k] = (double)(st >> 11) / 9007199254740992.0 - 0.5;
      xs[rmap[k]] = x[k];
    }
    for (int64_t i = 0; i < nI; i++) {
      for (int64_t q = A0.rowptr[i]; q < A0.rowptr[i + 1]; q++) y0[i] += A0.val[q] * x[A0.col[q]];
      for (int64_t q = Kii.rowptr[i]; q < Kii.rowptr[i + 1]; q++) y1[i] += Kii.val[q] * xs[Kii.col[q]];
    }
    double dot = 0.0, ymax = 0.0;
    for (int64_t k = 0; k < nI; k++) {
      dot += y0[k] * y1[rmap[k]];
      ymax = std::max(ymax, std::fabs(y1[k]));
    }
    sgn = dot < 0 ? -1.0 : 1.0;
    double e = 0.0;
    for (int64_t k = 0; k < nI; k++) e = std::max(e, std::fabs(sgn * y0[k] - y1[rmap[k]]));
    if (!(e <= 1e-10 * ymax))
      return fail(ctx, EQ_EINVAL, "A[0] is not kni (= -K[FltNI, FltNI]) or K[FltNI, FltNI] of this mesh");
  }
  H.lev[0].A = std::move(Kii);
  {
    HostCsr P0;
    if (int rc = take(P[0], nI, sz[1], "P[0]", &P0)) return rc;
    // rows to the structured interior order
    HostCsr& Q = H.lev[0].P;
    Q.n = nI;
    Q.m = P0.m;
    std::vector<int64_t> inv(nI);
    for (int64_t k = 0; k < nI; k++) inv[rmap[k]] = k;
    Q.rowptr.assign(nI + 1, 0);
    for (int64_t r = 0; r < nI; r++) Q.rowptr[r + 1] = Q.rowptr[r] + (P0.rowptr[inv[r] + 1] - P0.rowptr[inv[r]]);
    Q.col.resize(P0.col.size());
    Q.val.resize(P0.val.size());
    for (int64_t r = 0; r < nI; r++) {
      const int64_t k = inv[r];
      std::copy(P0.col.begin() + P0.rowptr[k], P0.col.begin() + P0.rowptr[k + 1], Q.col.begin() + Q.rowptr[r]);
      std::copy(P0.val.begin() + P0.rowptr[k], P0.val.begin() + P0.rowptr[k + 1], Q.val.begin() + Q.rowptr[r]);
    }
    csr_transpose(Q, &H.lev[0].R);
  }
  for (int l = 1; l < nlev; l++) {
    if (int rc = take(A[l], sz[l], sz[l], "A[" + std::to_string(l) + "]", &H.lev[l].A)) return rc;
    for (auto& v : H.lev[l].A.val) v *= sgn;
    if (l + 1 < nlev) {
      if (int rc = take(P[l], sz[l], sz[l + 1], "P[" + std::to_string(l) + "]", &H.lev[l].P)) return rc;
      csr_transpose(H.lev[l].P, &H.lev[l].R);
    }
  }
  std::string err;
  if (int rc = finish_hierarchy(&H, omega, &err)) return fail(ctx, rc, err);
  return upload_hierarchy(ctx, H, smoother, smooth_degree, 0);
}

extern "C" int eq_amg_l0_info(const eq_ctx* ctx, int* smoother, int* degree, double* lo, double* hi) {
  if (!ctx || !smoother || !degree || !lo || !hi) return EQ_EINVAL;
  if (!ctx->have_amg) return EQ_ENOTREADY;
  *smoother = ctx->l0_sweep ? EQ_SMOOTH_CHEBYSHEV_TILES : ctx->l0_cheb ? EQ_SMOOTH_CHEBYSHEV : EQ_SMOOTH_JACOBI;
  *degree = ctx->l0_degree;
  *lo = ctx->l0_theta - ctx->l0_delta;
  *hi = ctx->l0_theta + ctx->l0_delta;
  return EQ_OK;
}

extern "C" int eq_amg_info(const eq_ctx* ctx, int* nlev, int64_t* n, int64_t* nnz, int cap) {
  if (!ctx || !nlev) return EQ_EINVAL;
  *nlev = (int)ctx->amg_n.size();
  for (int l = 0; l < *nlev && l < cap; l++) {
    if (n) n[l] = ctx->amg_n[l];
    if (nnz) nnz[l] = ctx->amg_nnz[l];
  }
  return EQ_OK;
}

template <typename T>
static DevLevelVecs<T>& lvecs(DevLevel& L);
template <>
DevLevelVecs<double>& lvecs<double>(DevLevel& L) { return L.v64; }
template <>
DevLevelVecs<float>& lvecs<float>(DevLevel& L) { return L.v32; }
template <typename T>
static CgWork<T>& work(eq_ctx* c);
template <>
CgWork<double>& work<double>(eq_ctx* c) { return c->w64; }
template <>
CgWork<float>& work<float>(eq_ctx* c) { return c->w32; }
template <typename T>
static const T* mu_of(eq_ctx* c);
template <>
const double* mu_of<double>(eq_ctx* c) { return c->mu; }
template <>
const float* mu_of<float>(eq_ctx* c) { return c->mu32; }
template <typename T>
static const T* dinv_of(eq_ctx* c);
template <>
const double* dinv_of<double>(eq_ctx* c) { return c->dinv; }
template <>
const float* dinv_of<float>(eq_ctx* c) { return c->dinv32; }
template <typename T>
static const T* coarse_inv_of(eq_ctx* c);
template <>
const double* coarse_inv_of<double>(eq_ctx* c) { return c->coarse_inv; }
template <>
const float* coarse_inv_of<float>(eq_ctx* c) { return c->coarse_inv32; }
template <typename T>
static T* cvec_of(eq_ctx* c);
template <>
double* cvec_of<double>(eq_ctx* c) { return c->cvec64; }
template <>
float* cvec_of<float>(eq_ctx* c) { return c->cvec32; }
template <typename T>
static const T* ones_of(eq_ctx* c);
template <>
const double* ones_of<double>(eq_ctx* c) { return c->ones; }
template <>
const float* ones_of<float>(eq_ctx* c) { return c->ones32; }

// One V-cycle: c = M^-1 r with partial c.r (and the cg! convergence test in
// its first kernel).
// Launch a level-0 Chebyshev sweep given as its single steps: consecutive
// steps go to one cheb_pair launch (unless opts.cheb_single or a large FP64
// mesh), an odd last step to one OpCheb launch.  The steps rotate r and d through three buffers, so a
// pair never writes a buffer it reads beyond its own tile.
template <typename T>
static void cheb_launch(eq_ctx* ctx, const ChebStep<T>* st, int k) {
  // A pair moves one step's HBM bytes but holds 2x the LDS of a step (FP64:
  // 2 blocks per CU).  It wins while one round of blocks covers the mesh and
  // in FP32; on large FP64 meshes (config 5: 92 us against 2 x 44 us) single
  // steps stream better.
  const bool pairs = !ctx->opts.cheb_single && (sizeof(T) == 4 || tile_count(ctx->g) <= 1024);
  int m = 0;
  while (m < k) {
    if (pairs && m + 1 < k && st[m].mode == 0) {
      launch_cheb_pair<T>(ctx->g, st[m], st[m + 1], mu_of<T>(ctx), dinv_of<T>(ctx), ctx->part_rr, ctx->npart,
                          ctx->sc, ctx->part_rho, ctx->s,
                          lt_slot(ctx, sizeof(T) == 4 ? EQ_TK_CHEB_PAIR32 : EQ_TK_CHEB_PAIR64));
      m += 2;
    } else {
      launch_cheb<T>(ctx->g, st[m], mu_of<T>(ctx), dinv_of<T>(ctx), ctx->part_rr, ctx->npart, ctx->sc,
                     ctx->part_rho, ctx->s, lt_slot(ctx, sizeof(T) == 4 ? EQ_TK_CHEB_STEP32 : EQ_TK_CHEB_STEP64));
      m += 1;
    }
  }
}

// Level-0 Chebyshev pre-smoothing from x = 0: k steps, leaving x_k in chx;
// returns the buffer that holds the residual b - K_II x_k.
template <typename T>
static const T* cheb_pre(eq_ctx* ctx, const T* b) {
  CgWork<T>& w = work<T>(ctx);
  const int k = ctx->l0_degree;
  const double th = ctx->l0_theta, de = ctx->l0_delta, sg = th / de;
  double rho_prev = 1.0 / sg;
  ChebStep<T> st[kMaxChebDegree];
  for (int m = 1; m <= k; m++) {
    ChebStep<T>& c = st[m - 1];
    c.check = (m == 1);
    c.rin = (m == 1) ? b : w.chr[(m - 2) % 3];
    c.rout = w.chr[(m - 1) % 3];
    if (m == 1) { c.src_mode = 1; c.b = b; c.s0 = (T)(1.0 / th); }
    else { c.src = w.chd[(m - 2) % 3]; }
    if (m == k) {
      c.mode = 1;
      c.write_x_center = (k == 1);
      c.xout = w.chx;
    } else {
      const double rho = 1.0 / (2.0 * sg - rho_prev);
      c.mode = 0;
      c.c1 = (T)(rho * rho_prev);
      c.c2 = (T)(2.0 * rho / de);
      rho_prev = rho;
      c.xsrc_center = (m == 1);
      c.xin = w.chx;
      c.xout = w.chx;
      c.dout = w.chd[(m - 1) % 3];
    }
  }
  cheb_launch<T>(ctx, st, k);
  return w.chr[(k - 1) % 3];
}

// Level-0 Chebyshev post-smoothing from x = chx: k steps; output c and
// partial <c, r> (with c64: c widened to FP64 into c64 and partial <c, r64>).
template <typename T>
static void cheb_post(eq_ctx* ctx, const T* b, T* c, const T* coarse, double* c64 = nullptr,
                      const double* r64 = nullptr) {
  CgWork<T>& w = work<T>(ctx);
  const int k = ctx->l0_degree;
  const double th = ctx->l0_theta, de = ctx->l0_delta, sg = th / de;
  double rho_prev = 1.0 / sg;
  ChebStep<T> st[kMaxChebDegree];
  for (int m = 0; m < k; m++) {
    ChebStep<T>& q = st[m];
    q.rin = (m == 0) ? b : w.chr[(m - 1) % 3];
    q.src = (m == 0) ? w.chx : w.chd[(m - 1) % 3];
    if (m == 0 && coarse) {
      q.src_mode = 2;     // x0 = chx + P0 coarse, evaluated in the window load
      q.coarse = coarse;
    }
    if (m == 0) {
      q.c1 = T(0);
      q.c2 = (T)(1.0 / th);
      q.xsrc_center = 1;
    } else {
      const double rho = 1.0 / (2.0 * sg - rho_prev);
      q.c1 = (T)(rho * rho_prev);
      q.c2 = (T)(2.0 * rho / de);
      rho_prev = rho;
      q.xin = w.chx2;
    }
    if (m == k - 1) {
      q.mode = 2;
      q.xout = c;
      q.rdot = b;
      q.xout64 = c64;
      q.rdot64 = r64;
    } else {
      q.mode = 0;
      q.rout = w.chr[m % 3];
      q.dout = w.chd[m % 3];
      q.xout = w.chx2;
    }
  }
  cheb_launch<T>(ctx, st, k);
}

// Parameters of a whole-sweep launch: the step coefficients of cheb_pre and
// cheb_post (the same recurrence), the first direction's scale 1 / theta.
template <typename T>
static ChebSweep<T> sweep_params(eq_ctx* ctx, int post) {
  ChebSweep<T> op{};
  const int k = ctx->l0_degree;
  const double th = ctx->l0_theta, de = ctx->l0_delta, sg = th / de;
  double rho_prev = 1.0 / sg;
  for (int m = 1; m < k; m++) {
    const double rho = 1.0 / (2.0 * sg - rho_prev);
    op.c1[m] = (T)(rho * rho_prev);
    op.c2[m] = (T)(2.0 * rho / de);
    rho_prev = rho;
  }
  op.post = post;
  op.degree = k;
  op.s0 = (T)(1.0 / th);
  op.nelx = ctx->g.nelx;
  op.nely = ctx->g.nely;
  for (int p = 0; p < 5; p++) op.gx[p] = (T)ctx->g.x[p];
  for (int p = 0; p < 5; p++) {
    int e = 0;
    for (int i = 0; i < 5; i++)
      for (int kk = i; kk < 5; kk++) op.dd[p][e++] = (float)(ctx->g.D[p * 5 + i] * ctx->g.D[p * 5 + kk]);
  }
  op.mu = mu_of<T>(ctx);
  op.dinv = dinv_of<T>(ctx);
  op.part_rr = ctx->part_rr;
  op.nrr = ctx->npart;
  op.sc = ctx->sc;
  op.part0 = ctx->part_rho;
  return op;
}

// Tile-sweep pre-smoothing of b from x = 0 into chx (one launch, carrying the
// cg! convergence test); with residual, then the residual b - K_II chx, whose
// buffer is returned.
template <typename T>
static const T* sweep_pre(eq_ctx* ctx, const T* b, bool residual) {
  CgWork<T>& w = work<T>(ctx);
  ChebSweep<T> op = sweep_params<T>(ctx, 0);
  op.b = b;
  op.xout = w.chx;
  op.check = 1;
  if (ctx->cg_init_pending) {
    op.init = 1;
    op.part_bb = ctx->part_bb;
    op.tol = ctx->cg_init_tol;
    op.maxiter = (int)ctx->cg_init_maxiter;
    ctx->cg_init_pending = false;
  }
  launch_cheb_sweep<T>(ctx->g, op, ctx->s, lt_slot(ctx, sizeof(T) == 4 ? EQ_TK_SWEEP32 : EQ_TK_SWEEP64));
  if (!residual) return nullptr;
  ChebStep<T> st;
  st.src = w.chx;
  st.rin = b;
  st.rout = w.chr[0];
  st.mode = 1;
  launch_cheb<T>(ctx->g, st, mu_of<T>(ctx), dinv_of<T>(ctx), ctx->part_rr, ctx->npart, ctx->sc, ctx->part_rho,
                 ctx->s, lt_slot(ctx, sizeof(T) == 4 ? EQ_TK_CHEB_STEP32 : EQ_TK_CHEB_STEP64));
  return w.chr[0];
}

// Tile-sweep post-smoothing from chx (+ P0 coarse): output c and partial
// <c, b> (with c64: c widened to FP64 into c64 and partial <c, r64>).
template <typename T>
static void sweep_post(eq_ctx* ctx, const T* b, T* c, const T* coarse, double* c64 = nullptr,
                       const double* r64 = nullptr) {
  CgWork<T>& w = work<T>(ctx);
  ChebSweep<T> op = sweep_params<T>(ctx, 1);
  op.b = b;
  op.src = w.chx;
  op.coarse = coarse;
  op.xout = c;
  op.rdot = b;
  op.xout64 = c64;
  op.rdot64 = r64;
  launch_cheb_sweep<T>(ctx->g, op, ctx->s, lt_slot(ctx, sizeof(T) == 4 ? EQ_TK_SWEEP32 : EQ_TK_SWEEP64));
}

// c64/r64 (matrix-free Chebyshev level 0 only): the last smoothing step
// writes c widened to FP64 into c64 and the partial <c, r64>.
template <typename T>
static void vcycle(eq_ctx* ctx, const T* r, T* c, double* c64 = nullptr, const double* r64 = nullptr) {
  CgWork<T>& w = work<T>(ctx);
  const Geom& g = ctx->g;
  hipStream_t s = ctx->s;
  const int L = (int)ctx->lev.size();
  DevLevel& L0 = ctx->lev[0];
  const T om0 = (T)L0.omega;
  T* x0 = ctx->l0_cheb ? w.chx : w.x0;
  if (ctx->l0_sweep && ctx->pmg) {
    // bilinear level: the residual of the pre-smoothed x restricted in one matrix-free launch
    sweep_pre<T>(ctx, r, false);
    launch_restrict<T>(g, w.chx, r, mu_of<T>(ctx), lvecs<T>(ctx->lev[1]).b, ctx->sc, s);
  } else if (ctx->l0_cheb) {
    const T* res = ctx->l0_sweep ? sweep_pre<T>(ctx, r, true) : cheb_pre<T>(ctx, r);
    launch_csr<T>(CSR_SPMV, L0.R, res + g.NX, nullptr, nullptr, T(0), lvecs<T>(ctx->lev[1]).b, nullptr, ctx->sc,
                  s);
  } else {
    launch_l0_down<T>(g, r, mu_of<T>(ctx), dinv_of<T>(ctx), om0, w.x0, w.res0, ctx->part_rr, ctx->npart, ctx->sc, s);
    launch_csr<T>(CSR_SPMV, L0.R, w.res0 + g.NX, nullptr, nullptr, T(0), lvecs<T>(ctx->lev[1]).b, nullptr, ctx->sc,
                  s);
  }
  DevLevelVecs<T>& vc = lvecs<T>(ctx->lev[L - 1]);
  if (ctx->coarse_collapsed) {
    // collapsed: every coarse b from b_1 in one product, the dense level, e_1 in one product
    T* cv = cvec_of<T>(ctx);
    launch_csr<T>(CSR_SPMV, ctx->coarse_down, cv, nullptr, nullptr, T(0), cv + ctx->lev[1].n, nullptr, ctx->sc, s);
    launch_dense<T>(ctx->ncoarse, coarse_inv_of<T>(ctx), vc.b, vc.xp, ctx->sc, s);
    launch_csr<T>(CSR_SPMV, ctx->coarse_up, cv, nullptr, nullptr, T(0), lvecs<T>(ctx->lev[1]).xp, nullptr, ctx->sc,
                  s);
  } else {
    // coarse levels: one kernel down (b_next = M b) and one up (xp = U b + V xp_next) per level
    for (int l = 1; l <= L - 2; l++) {
      DevLevel& D = ctx->lev[l];
      launch_csr<T>(CSR_SPMV, D.M, lvecs<T>(D).b, nullptr, nullptr, T(0), lvecs<T>(ctx->lev[l + 1]).b, nullptr,
                    ctx->sc, s);
    }
    launch_dense<T>(ctx->ncoarse, coarse_inv_of<T>(ctx), vc.b, vc.xp, ctx->sc, s);
    for (int l = L - 2; l >= 1; l--) {
      DevLevel& D = ctx->lev[l];
      DevLevelVecs<T>& v = lvecs<T>(D);
      launch_csr_up<T>(D.U, D.V, v.b, lvecs<T>(ctx->lev[l + 1]).xp, v.xp, ctx->sc, s);
    }
  }
  const bool fused_prolong = ctx->l0_cheb && ctx->pmg;
  if (!fused_prolong)
    launch_csr<T>(CSR_PROLONG_ADD, L0.P, lvecs<T>(ctx->lev[1]).xp, nullptr, nullptr, T(0), x0 + g.NX, nullptr,
                  ctx->sc, s);
  if (ctx->l0_sweep)
    sweep_post<T>(ctx, r, c, fused_prolong ? lvecs<T>(ctx->lev[1]).xp : nullptr, c64, r64);
  else if (ctx->l0_cheb)
    cheb_post<T>(ctx, r, c, fused_prolong ? lvecs<T>(ctx->lev[1]).xp : nullptr, c64, r64);
  else
    launch_l0_post<T>(g, w.x0, r, mu_of<T>(ctx), dinv_of<T>(ctx), om0, c, ctx->part_rho, ctx->sc, s);
}

// ===========================================================================
// CG in K_II form on device vectors (fault row ignored); x in/out, b input.
// ===========================================================================
// f32pc: apply the V-cycle in FP32 to the FP32 copy of r kept by the update
// kernel, widen c back to FP64 and take <c, r> in FP64 (T must be double).
template <typename T>
static void cg_iteration(eq_ctx* ctx, T* x, int pc, bool f32pc = false) {
  CgWork<T>& w = work<T>(ctx);
  const Geom& g = ctx->g;
  hipStream_t s = ctx->s;
  const int nt = tile_count(g);
  int nrho;
  if (pc == EQ_PC_AMG && f32pc && ctx->l0_cheb) {
    // the last smoothing step widens c and takes <c, r> in FP64 per tile
    vcycle<float>(ctx, ctx->w32.r, ctx->w32.c, (double*)w.c, (const double*)w.r);
    nrho = nt;
  } else if (pc == EQ_PC_AMG && f32pc) {
    vcycle<float>(ctx, ctx->w32.r, ctx->w32.c);
    launch_widen_dot(ctx->N, ctx->w32.c, (const double*)w.r, (double*)w.c, ctx->sc, ctx->part_rho, s);
    nrho = vec_blocks(ctx->N);
  } else if (pc == EQ_PC_AMG) {
    vcycle<T>(ctx, w.r, w.c);
    nrho = nt;
  } else {
    const T* dv = (pc == EQ_PC_JACOBI) ? dinv_of<T>(ctx) : ones_of<T>(ctx);
    launch_jacobi_pc<T>(ctx->N, dv, w.r, w.c, ctx->part_rr, ctx->npart, ctx->sc, ctx->part_rho, s);
    nrho = vec_blocks(ctx->N);
  }
  launch_cg_dir<T>(g, w.c, w.u[w.cur], mu_of<T>(ctx), w.u[1 - w.cur], w.q, ctx->part_rho, nrho, ctx->sc,
                   ctx->part_uq, s, lt_slot(ctx, EQ_TK_CG_DIR));
  launch_cg_update<T>(ctx->N, x, w.r, w.u[1 - w.cur], w.q, ctx->part_rho, nrho, ctx->part_uq, nt, ctx->sc,
                      ctx->part_rr, ctx->npart, f32pc ? ctx->w32.r : nullptr, s, lt_slot(ctx, EQ_TK_CG_UPDATE));
  w.cur = 1 - w.cur;
}

// n CG iterations.  With opts.graphs the batch is captured once per (precision,
// preconditioner, n, direction-buffer parity, x) as a HIP graph of the same
// kernels and replayed, which removes the per-kernel launch cost on the host;
// the host-side parity of the direction buffers advances as in eager launch.
template <typename T>
static int enqueue_iterations(eq_ctx* ctx, T* x, int pc, bool f32pc, int64_t n) {
  if (n <= 0) return EQ_OK;
  CgWork<T>& w = work<T>(ctx);
  if (!ctx->opts.graphs || ctx->lt.on) {
    for (int64_t k = 0; k < n; k++) cg_iteration<T>(ctx, x, pc, f32pc);
    return EQ_OK;
  }
  const bool init = ctx->cg_init_pending;
  const auto key = std::make_tuple((int)sizeof(T), f32pc ? 1 : 0, pc, n, w.cur, (const void*)x,
                                   init ? ctx->cg_init_tol : -1.0, init ? ctx->cg_init_maxiter : (int64_t)-1);
  auto it = ctx->cg_graphs.find(key);
  if (it == ctx->cg_graphs.end()) {
    const int cur0 = w.cur;
    HIPCHK(ctx, hipStreamBeginCapture(ctx->s, hipStreamCaptureModeThreadLocal));
    for (int64_t k = 0; k < n; k++) cg_iteration<T>(ctx, x, pc, f32pc);
    hipGraph_t graph = nullptr;
    HIPCHK(ctx, hipStreamEndCapture(ctx->s, &graph));
    w.cur = cur0;   // capture recorded the launches without running them
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(ctx, EQ_HIP_ERR, std::string("graph instantiate: ") + hipGetErrorString(e));
    it = ctx->cg_graphs.emplace(key, exec).first;
  }
  HIPCHK(ctx, hipGraphLaunch(it->second, ctx->s));
  ctx->cg_init_pending = false;   // the replayed first kernel carries it
  if (n & 1) w.cur = 1 - w.cur;
  return EQ_OK;
}

// CG set-up: r = b - K_II x (unless r_ready), the device scalars, and the
// FP32 copy of r for an FP32 V-cycle.  When the first batch (first_batch
// iterations) opens with a tile sweep, that kernel sets the scalars up
// instead of a launch of its own.
template <typename T>
static int cg_begin(eq_ctx* ctx, T* x, const T* b, double tol, int64_t maxiter, int pc, bool f32pc, bool r_ready,
                    int64_t first_batch, bool r32_ready = false) {
  CgWork<T>& w = work<T>(ctx);
  hipStream_t s = ctx->s;
  if (pc == EQ_PC_AMG && !ctx->have_amg) return fail(ctx, EQ_ENOTREADY, "eq_amg_setup first");
  if (maxiter <= 0) maxiter = ctx->nI;
  if (maxiter > INT_MAX) maxiter = INT_MAX;
  // (the residual kernel also leaves the FP32 copy of r for an FP32 V-cycle)
  if (!r_ready)
    launch_cg_residual<T>(ctx->g, x, mu_of<T>(ctx), b, w.r, ctx->part_rr, ctx->part_bb, ctx->npart, s,
                          f32pc ? ctx->w32.r : nullptr);
  if (first_batch > 0 && pc == EQ_PC_AMG && ctx->l0_sweep) {
    ctx->cg_init_pending = true;
    ctx->cg_init_tol = tol;
    ctx->cg_init_maxiter = maxiter;
  } else {
    launch_cg_init_scal(ctx->sc, ctx->part_rr, ctx->part_bb, ctx->npart, tol, (int)maxiter, s);
  }
  if (f32pc && r_ready && !r32_ready) launch_convert<T, float>(ctx->N, w.r, ctx->w32.r, s);
  return EQ_OK;
}

// A batch of iterations and the convergence test, which also leaves the
// scalars in the pinned host copy; nothing waits.
template <typename T>
static int cg_batch(eq_ctx* ctx, T* x, int pc, bool f32pc, int64_t batch) {
  const int rc = enqueue_iterations<T>(ctx, x, pc, f32pc, batch);
  ctx->cg_init_pending = false;   // consumed by the batch's first kernel (or nothing to set up after an error)
  if (rc) return rc;
  launch_cg_check(ctx->sc, ctx->part_rr, ctx->npart, ctx->s, ctx->sc_host);
  return EQ_OK;
}

// After the host has the scalars of a first batch of `batch` iterations:
// further, smaller batches until the test passes.  The iterates do not depend
// on where the batches end (a converged solve turns later kernels into exits).
template <typename T>
static int cg_rest(eq_ctx* ctx, T* x, int pc, bool f32pc, int64_t batch) {
  while (!ctx->sc_host->done) {
    batch = std::min<int64_t>(std::max<int64_t>(2, batch / 4), 16);
    if (int rc = cg_batch<T>(ctx, x, pc, f32pc, batch)) return rc;
    if (int rc = spin_sync(ctx)) return rc;
  }
  return EQ_OK;
}

static int cg_end(eq_ctx* ctx, int64_t* iters, double* relres) {
  const CgScal& h = *ctx->sc_host;
  if (iters) *iters = h.iter;
  if (relres) *relres = h.bnorm > 0 ? h.resid / h.bnorm : h.resid;
  ctx->pred_iters[ctx->cg_slot] = h.iter;
  return (h.resid <= h.reltol) ? EQ_OK : EQ_PCG_NOCONV;
}

// r_ready: w.r already holds b - A x with its ||r||^2 partials in part_rr and
// the ||b||^2 partials in part_bb (the learned and Galerkin starts leave them
// so); r32_ready: and w32.r its FP32 copy.
template <typename T>
static int cg_device(eq_ctx* ctx, T* x, const T* b, double tol, int64_t maxiter, int pc, int64_t* iters,
                     double* relres, bool f32pc = false, bool r_ready = false, bool r32_ready = false) {
  int64_t batch = ctx->pred_iters[ctx->cg_slot];
  if (int rc = cg_begin<T>(ctx, x, b, tol, maxiter, pc, f32pc, r_ready, batch, r32_ready)) return rc;
  if (int rc = cg_batch<T>(ctx, x, pc, f32pc, batch)) return rc;
  if (int rc = spin_sync(ctx)) return rc;
  if (int rc = cg_rest<T>(ctx, x, pc, f32pc, batch)) return rc;
  return cg_end(ctx, iters, relres);
}

// Mixed precision: FP64 residual / update around FP32 inner PCG.
static int cg_mixed(eq_ctx* ctx, double* x, const double* b, double tol, int64_t maxiter, int pc, int64_t* iters,
                    double* relres) {
  CgWork<double>& w = ctx->w64;
  CgWork<float>& f = ctx->w32;
  const Geom& g = ctx->g;
  hipStream_t s = ctx->s;
  if (maxiter <= 0) maxiter = ctx->nI;
  int64_t total = 0;
  double bnorm = 0, resid = 0;
  const double inner_tol = 1e-3;
  for (int outer = 0; outer < 50; outer++) {
    launch_cg_residual<double>(g, x, ctx->mu, b, w.r, ctx->part_rr, ctx->part_bb, ctx->npart, s);
    launch_cg_init_scal(ctx->sc, ctx->part_rr, ctx->part_bb, ctx->npart, tol, 1, s);
    HIPCHK(ctx, hipMemcpyAsync(ctx->sc_host, ctx->sc, sizeof(CgScal), hipMemcpyDeviceToHost, s));
    if (int rc = sync_check(ctx)) return rc;
    bnorm = ctx->sc_host->bnorm;
    resid = ctx->sc_host->resid;
    if (resid <= tol * bnorm || total >= maxiter) break;
    launch_convert<double, float>(ctx->N, w.r, f.b, s);
    launch_fill<float>(ctx->N, f.x, 0.0f, s);
    int64_t it = 0;
    double rr = 0;
    int rc = cg_device<float>(ctx, f.x, f.b, inner_tol, maxiter - total, pc, &it, &rr);
    if (rc != EQ_OK && rc != EQ_PCG_NOCONV) return rc;
    total += it;
    // x += e (FP32 correction, fault row of e is zero)
    launch_convert<float, double>(ctx->N, f.x, w.q, s);
    launch_axpy(ctx->N, 1.0, w.q, x, s);
  }
  if (iters) *iters = total;
  if (relres) *relres = bnorm > 0 ? resid / bnorm : resid;
  return resid <= tol * bnorm ? EQ_OK : EQ_PCG_NOCONV;
}

// r_ready: see cg_device (honoured by the FP64 and F32PC modes).
static int cg_any(eq_ctx* ctx, double* x, const double* b, double tol, int64_t maxiter, int pc, int prec,
                  int64_t* iters, double* relres, bool r_ready = false, bool r32_ready = false) {
  if (prec == EQ_PREC_MIXED) return cg_mixed(ctx, x, b, tol, maxiter, pc, iters, relres);
  if (prec == EQ_PREC_F32PC)
    return cg_device<double>(ctx, x, b, tol, maxiter, pc, iters, relres, pc == EQ_PC_AMG, r_ready, r32_ready);
  if (prec == EQ_PREC_FP32) {
    launch_convert<double, float>(ctx->N, x, ctx->w32.x, ctx->s);
    launch_convert<double, float>(ctx->N, b, ctx->w32.b, ctx->s);
    int rc = cg_device<float>(ctx, ctx->w32.x, ctx->w32.b, tol, maxiter, pc, iters, relres);
    launch_convert<float, double>(ctx->N, ctx->w32.x, x, ctx->s);
    int rc2 = sync_check(ctx);
    return rc ? rc : rc2;
  }
  return cg_device<double>(ctx, x, b, tol, maxiter, pc, iters, relres, false, r_ready, r32_ready);
}

extern "C" int eq_cg(eq_ctx* ctx, double* x, const double* b, double tol, int64_t maxiter, int pc, int prec,
                     int64_t* iters, double* relres) {
  if (!ctx || !x || !b) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (!ctx->have_material) return fail(ctx, EQ_ENOTREADY, "eq_set_material first");
  if (pc < EQ_PC_NONE || pc > EQ_PC_AMG || prec < EQ_PREC_FP64 || prec > EQ_PREC_F32PC)
    return fail(ctx, EQ_EINVAL, "bad pc / prec");
  // cg!(x, kni, b): kni = -K_II, so solve K_II x = -b (identical CG iterates)
  double* dx = ctx->w64.x;
  double* db = ctx->w64.b;
  if (int rc = upload_ni(ctx, x, dx)) return rc;
  if (int rc = upload_ni(ctx, b, db)) return rc;
  launch_scale_mask(ctx->N, ctx->NX, db, -1.0, db, ctx->s);
  int rc = cg_any(ctx, dx, db, tol, maxiter, pc, prec, iters, relres);
  if (rc != EQ_OK && rc != EQ_PCG_NOCONV) return rc;
  if (int rc2 = download_ni(ctx, dx, x)) return rc2;
  if (rc == EQ_PCG_NOCONV) ctx->err = "cg did not reach the tolerance within maxiter";
  return rc;
}

extern "C" int eq_amg_ldiv(eq_ctx* ctx, const double* b, double* x, int prec) {
  if (!ctx || !b || !x) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (!ctx->have_amg) return fail(ctx, EQ_ENOTREADY, "eq_amg_setup first");
  // x = (V-cycle of kni) b = -(V-cycle of K_II) b
  double* db = ctx->w64.b;
  if (int rc = upload_ni(ctx, b, db)) return rc;
  launch_scale_mask(ctx->N, ctx->NX, db, -1.0, db, ctx->s);
  // make the convergence test pass: reltol = -1, iter 0 < maxiter
  CgScal h{};
  h.rho_prev = 1.0; h.reltol = -1.0; h.maxiter = INT_MAX; h.done = 0; h.iter = 0;
  HIPCHK(ctx, hipMemcpyAsync(ctx->sc, &h, sizeof(CgScal), hipMemcpyHostToDevice, ctx->s));
  launch_fill<double>(ctx->npart, ctx->part_rr, 1.0, ctx->s);
  if (prec == EQ_PREC_FP32) {
    launch_convert<double, float>(ctx->N, db, ctx->w32.r, ctx->s);
    vcycle<float>(ctx, ctx->w32.r, ctx->w32.c);
    launch_convert<float, double>(ctx->N, ctx->w32.c, ctx->w64.c, ctx->s);
  } else {
    vcycle<double>(ctx, db, ctx->w64.c);
  }
  if (int rc = sync_check(ctx)) return rc;
  return download_ni(ctx, ctx->w64.c, x);
}

extern "C" int eq_pcg_jacobi(eq_ctx* ctx, const double* F, double* dnew, double tol, int64_t maxit, int64_t* iters) {
  if (!ctx || !F || !dnew) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (!ctx->have_material) return fail(ctx, EQ_ENOTREADY, "eq_set_material first");
  // Fnew = -(K F)[FltNI]
  double* dF = ctx->w64.x;
  if (int rc = upload_ref(ctx, F, dF)) return rc;
  launch_ku<double>(ctx->g, dF, ctx->mu, ctx->w64.b, 0, 1, ctx->s);
  launch_scale_mask(ctx->N, ctx->NX, ctx->w64.b, 1.0, ctx->w64.b, ctx->s);
  if (int rc = upload_ni(ctx, dnew, ctx->w64.x)) return rc;
  double rr = 0;
  int rc = cg_device<double>(ctx, ctx->w64.x, ctx->w64.b, tol, maxit, EQ_PC_JACOBI, iters, &rr);
  if (rc != EQ_OK && rc != EQ_PCG_NOCONV) return rc;
  if (int rc2 = download_ni(ctx, ctx->w64.x, dnew)) return rc2;
  if (rc == EQ_PCG_NOCONV) ctx->err = "PCG did not converge";
  return rc;
}

// ===========================================================================
// fault physics entry points
// ===========================================================================
namespace {
struct DevTmp {
  std::vector<void*> p;
  ~DevTmp() {
    for (void* q : p) (void)hipFree(q);
  }
  template <typename T>
  T* get(int64_t n) {
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(T) * (size_t)std::max<int64_t>(n, 1)) != hipSuccess) return nullptr;
    p.push_back(q);
    return (T*)q;
  }
};
}  // namespace

#define UP(dst, src, n)                                         \
  do {                                                          \
    if (!(dst)) return fail(ctx, EQ_NOMEM, "device alloc");     \
    if (int rc_ = h2d(ctx, dst, src, n)) return rc_;            \
  } while (0)

extern "C" int eq_ids(eq_ctx* ctx, int64_t n, const double* xLf, const double* Vo, const double* psi, double dt,
                      const double* Vf, double cnd, int64_t IDstate, double* psi1) {
  if (!ctx || n < 0 || !xLf || !Vo || !psi || !Vf || !psi1) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (IDstate != 1 && IDstate != 2) return fail(ctx, EQ_EINVAL, "IDstate must be 1 or 2");
  if (n == 0) return EQ_OK;
  DevTmp t;
  double *a = t.get<double>(n), *b = t.get<double>(n), *c = t.get<double>(n), *d = t.get<double>(n),
         *o = t.get<double>(n);
  UP(a, xLf, n); UP(b, Vo, n); UP(c, psi, n); UP(d, Vf, n);
  if (!o) return fail(ctx, EQ_NOMEM, "device alloc");
  launch_ids(n, a, b, c, dt, d, cnd, (int)IDstate, o, ctx->s);
  if (int rc = d2h(ctx, psi1, o, n)) return rc;
  return sync_check(ctx);
}

extern "C" int eq_ids2(eq_ctx* ctx, int64_t n, const double* xLf, const double* Vo, const double* psi,
                       const double* psi1, double dt, const double* Vf, const double* Vf1, int64_t IDstate,
                       double* psi2) {
  if (!ctx || n < 0 || !xLf || !Vo || !psi || !psi1 || !Vf || !Vf1 || !psi2) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (IDstate != 1 && IDstate != 2) return fail(ctx, EQ_EINVAL, "IDstate must be 1 or 2");
  if (n == 0) return EQ_OK;
  DevTmp t;
  double *a = t.get<double>(n), *b = t.get<double>(n), *c = t.get<double>(n), *d = t.get<double>(n),
         *e = t.get<double>(n), *f = t.get<double>(n), *o = t.get<double>(n);
  UP(a, xLf, n); UP(b, Vo, n); UP(c, psi, n); UP(d, psi1, n); UP(e, Vf, n); UP(f, Vf1, n);
  if (!o) return fail(ctx, EQ_NOMEM, "device alloc");
  launch_ids2(n, a, b, c, d, dt, e, f, (int)IDstate, o, ctx->s);
  if (int rc = d2h(ctx, psi2, o, n)) return rc;
  return sync_check(ctx);
}

extern "C" int eq_nrsearch(eq_ctx* ctx, int64_t n, const double* fo, const double* Vo, const double* cca,
                           const double* ccb, const double* Seff, const double* tau, const double* tauo,
                           const double* psi, const double* FltZ, const double* FltVfree, double* Vf_out,
                           double* tau_out, int32_t* k_out) {
  if (!ctx || n < 0 || !fo || !Vo || !cca || !ccb || !Seff || !tau || !tauo || !psi || !FltZ || !FltVfree ||
      !Vf_out || !tau_out)
    return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (n == 0) return EQ_OK;
  DevTmp t;
  const double* in[10] = {fo, Vo, cca, ccb, Seff, tau, tauo, psi, FltZ, FltVfree};
  double* din[10];
  for (int k = 0; k < 10; k++) {
    din[k] = t.get<double>(n);
    UP(din[k], in[k], n);
  }
  double *ov = t.get<double>(n), *ot = t.get<double>(n);
  int* ok = t.get<int>(n);
  if (!ov || !ot || !ok) return fail(ctx, EQ_NOMEM, "device alloc");
  launch_nrsearch(n, din[0], din[1], din[2], din[3], din[4], din[5], din[6], din[7], din[8], din[9], ov, ot, ok,
                  ctx->s);
  if (int rc = d2h(ctx, Vf_out, ov, n)) return rc;
  if (int rc = d2h(ctx, tau_out, ot, n)) return rc;
  std::vector<int> kk(n);
  if (int rc = d2h(ctx, kk.data(), ok, n)) return rc;
  if (int rc = sync_check(ctx)) return rc;
  bool noconv = false;
  for (int64_t i = 0; i < n; i++) {
    if (k_out) k_out[i] = kk[i];
    if (kk[i] < 0) noconv = true;
  }
  if (noconv) {
    ctx->err = "NRsearch! hit |delta| > 1e10 or k == 1000 (last iterate returned)";
    return EQ_NR_NOCONV;
  }
  return EQ_OK;
}

static int check_fault_call(eq_ctx* ctx, int64_t NFBC, int64_t FltNglob) {
  if (!ctx->have_params) return fail(ctx, EQ_ENOTREADY, "eq_set_params first");
  if (FltNglob != ctx->FltNglob || NFBC < 1 || NFBC > FltNglob + 1)
    return fail(ctx, EQ_EINVAL, "NFBC / FltNglob do not match the parameters");
  return EQ_OK;
}

// The FaultDev of a fault-routine call: the context's, with the caller's
// params_farray arrays (copied to the device) in place of its copies when P
// is given.  need_z: the routine reads FltZ.
static int api_fault_params(eq_ctx* ctx, const eq_fault_params* P, bool need_z, FaultDev* fp) {
  *fp = ctx->fp;
  if (!P) return EQ_OK;
  const double* src[8] = {P->fo, P->Vo, P->xLf, P->cca, P->ccb, P->Seff, P->tauo, P->FltZ};
  for (int k = 0; k < 8; k++)
    if (!src[k] && (k < 7 || need_z)) return fail(ctx, EQ_EINVAL, "eq_fault_params: missing array");
  const int64_t n = ctx->FltNglob;
  double* b = ctx->fp_api_blk;
  for (int k = 0; k < 8; k++)
    if (src[k])
      if (int rc = h2d(ctx, b + k * n, src[k], n)) return rc;
  fp->fo = b;
  fp->Vo = b + n;
  fp->xLf = b + 2 * n;
  fp->cca = b + 3 * n;
  fp->ccb = b + 4 * n;
  fp->Seff = b + 5 * n;
  fp->tauo = b + 6 * n;
  if (src[7]) fp->FltZ = b + 7 * n;
  return EQ_OK;
}

extern "C" int eq_fbc(eq_ctx* ctx, int64_t IDstate, const eq_fault_params* P, int64_t NFBC, int64_t FltNglob,
                      double* psi1, double* Vf1, double* tau1, double* psi2, double* Vf2, double* tau2,
                      const double* psi, const double* Vf, const double* FltVfree, double dt, int64_t* fail_node) {
  if (!ctx || !psi1 || !Vf1 || !tau1 || !psi2 || !Vf2 || !tau2 || !psi || !Vf || !FltVfree) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (int rc = check_fault_call(ctx, NFBC, FltNglob)) return rc;
  const int64_t n = FltNglob;
  FaultState& f = ctx->fs_api;
  if (int rc = h2d(ctx, f.psi1, psi1, n)) return rc;
  if (int rc = h2d(ctx, f.Vf1, Vf1, n)) return rc;
  if (int rc = h2d(ctx, f.tau1, tau1, n)) return rc;
  if (int rc = h2d(ctx, f.psi2, psi2, n)) return rc;
  if (int rc = h2d(ctx, f.Vf2, Vf2, n)) return rc;
  if (int rc = h2d(ctx, f.tau2, tau2, n)) return rc;
  if (int rc = h2d(ctx, f.psi, psi, n)) return rc;
  if (int rc = h2d(ctx, f.Vf, Vf, n)) return rc;
  if (int rc = h2d(ctx, f.FltVfree, FltVfree, n)) return rc;
  FaultDev fp{};
  if (int rc = api_fault_params(ctx, P, true, &fp)) return rc;
  fp.j0 = (int)(NFBC - 1);
  fp.IDstate = (int)IDstate;
  int hf[2] = {INT_MAX, 0};
  ctx->flags_dirty = true;
  if (int rc = h2d(ctx, ctx->flags, hf, 2)) return rc;
  launch_fbc(fp, f, f.FltVfree, dt, ctx->flags, ctx->s);
  if (int rc = d2h(ctx, psi1, f.psi1, n)) return rc;
  if (int rc = d2h(ctx, Vf1, f.Vf1, n)) return rc;
  if (int rc = d2h(ctx, tau1, f.tau1, n)) return rc;
  if (int rc = d2h(ctx, psi2, f.psi2, n)) return rc;
  if (int rc = d2h(ctx, Vf2, f.Vf2, n)) return rc;
  if (int rc = d2h(ctx, tau2, f.tau2, n)) return rc;
  if (int rc = d2h(ctx, hf, ctx->flags, 2)) return rc;
  if (int rc = sync_check(ctx)) return rc;
  if (fail_node) *fail_node = (hf[0] == INT_MAX) ? 0 : hf[0];
  if (hf[0] != INT_MAX) {
    char buf[128];
    snprintf(buf, sizeof buf, "NR SEARCH FAILED! (fault location %d)", hf[0]);
    ctx->err = buf;
    return EQ_NR_NAN;
  }
  if (hf[1] > 0) {
    ctx->err = "NRsearch! hit |delta| > 1e10 or k == 1000 (last iterate returned)";
    return EQ_NR_NOCONV;
  }
  return EQ_OK;
}

// The dynamic branch of the loop body on caller arrays: one launch of the
// Newmark kernel the device loop uses, on a scratch copy of the state.
extern "C" int eq_newmark_step(eq_ctx* ctx, double* d, double* v, double* a, double* psi, double* tau, double* Vf,
                               double* FltVfree, double dt) {
  if (!ctx || !d || !v || !a || !psi) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (!ctx->have_material || !ctx->have_params) return fail(ctx, EQ_ENOTREADY, "eq_set_material and eq_set_params first");
  if (!(dt > 0)) return fail(ctx, EQ_EINVAL, "dt must be positive");
  const int64_t N = ctx->N, NX = ctx->NX;
  if (!ctx->api_nodal) {
    ctx->api_nodal = ctx->alloc<double>(6 * N);
    NEED(ctx->api_nodal);
  }
  double* in = ctx->api_nodal;
  double* out = ctx->api_nodal + 3 * N;
  if (int rc = upload_ref(ctx, d, in)) return rc;
  if (int rc = upload_ref(ctx, v, in + N)) return rc;
  if (int rc = upload_ref(ctx, a, in + 2 * N)) return rc;
  FaultState& f = ctx->fs_api;
  // nodes deeper than NFBC keep their psi (FBC! does not visit them)
  if (int rc = h2d(ctx, f.psi, psi, NX)) return rc;
  if (int rc = h2d(ctx, f.psi2, psi, NX)) return rc;
  int hf[2] = {INT_MAX, 0};
  ctx->flags_dirty = true;
  if (int rc = h2d(ctx, ctx->flags, hf, 2)) return rc;
  // the new d is the predictor (into out), the next predictor is scratch (into in)
  launch_newmark_predict(N, in, in + N, in + 2 * N, dt, 0.5 * ctx->dt0 * ctx->dt0, out, ctx->s);
  launch_newmark(ctx->g, ctx->fp, f, ctx->mu, ctx->M, out, in + N, in + 2 * N, in, out + N, out + 2 * N, dt,
                 0.5 * ctx->dt0, 0.5 * ctx->dt0 * ctx->dt0, ctx->flags, NewmarkRecord{}, ctx->s);
  if (int rc = d2h(ctx, hf, ctx->flags, 2)) return rc;
  if (int rc = download_ref(ctx, out, d)) return rc;
  if (int rc = download_ref(ctx, out + N, v)) return rc;
  if (int rc = download_ref(ctx, out + 2 * N, a)) return rc;
  if (int rc = d2h(ctx, psi, f.psi, NX)) return rc;
  if (tau)
    if (int rc = d2h(ctx, tau, f.tau, NX)) return rc;
  if (Vf)
    if (int rc = d2h(ctx, Vf, f.Vf, NX)) return rc;
  if (FltVfree)
    if (int rc = d2h(ctx, FltVfree, f.FltVfree, NX)) return rc;
  if (int rc = sync_check(ctx)) return rc;
  if (hf[0] != INT_MAX) {
    char buf[128];
    snprintf(buf, sizeof buf, "NR SEARCH FAILED! (fault location %d)", hf[0]);
    ctx->err = buf;
    return EQ_NR_NAN;
  }
  if (hf[1] > 0) {
    ctx->err = "NRsearch! hit |delta| > 1e10 or k == 1000 (last iterate returned)";
    return EQ_NR_NOCONV;
  }
  return EQ_OK;
}

// eq_newmark_step with every device buffer in the file-scope scratch array:
// no allocation of any kind.  The caller's arrays (reference order) land in
// the output half of the scratch state, are permuted on the device into the
// input half, and the result goes back the same way; the fault staging block
// and the guard counters are cleared first, since the array keeps the last
// call's contents.  Calls that overlap on one GPU share the array and race.
extern "C" int eq_newmark_step_noalloc(eq_ctx* ctx, double* d, double* v, double* a, double* psi, double* tau,
                                       double* Vf, double* FltVfree, double dt) {
  if (!ctx || !d || !v || !a || !psi) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (!ctx->have_material || !ctx->have_params) return fail(ctx, EQ_ENOTREADY, "eq_set_material and eq_set_params first");
  if (!(dt > 0)) return fail(ctx, EQ_EINVAL, "dt must be positive");
  const int64_t N = ctx->N, NX = ctx->NX;
  if (N > kNoallocMaxNodes || NX > kNoallocMaxFault) {
    char buf[200];
    snprintf(buf, sizeof buf,
             "mesh of %lld nodes (%lld on the fault) exceeds the allocation-free step's scratch: at most %lld nodes "
             "and %lld fault nodes",
             (long long)N, (long long)NX, (long long)kNoallocMaxNodes, (long long)kNoallocMaxFault);
    return fail(ctx, EQ_EINVAL, buf);
  }
  if (!ctx->scratch_dev || !ctx->perm32) return fail(ctx, EQ_ENOTREADY, "scratch address not set up");
  unsigned char* base = (unsigned char*)ctx->scratch_dev;
  double* in = (double*)base;
  double* out = in + 3 * N;
  double* fblk = (double*)(base + kNoallocFaultOff);
  int* flags = (int*)(base + kNoallocFlagsOff);
  FaultState f{};
  f.psi = fblk; f.psi1 = fblk + NX; f.psi2 = fblk + 2 * NX; f.Vf = fblk + 3 * NX; f.Vf0 = fblk + 4 * NX;
  f.Vf1 = fblk + 5 * NX; f.Vf2 = fblk + 6 * NX; f.tau = fblk + 7 * NX; f.tau1 = fblk + 8 * NX;
  f.tau2 = fblk + 9 * NX; f.FltVfree = fblk + 10 * NX;
  hipStream_t s = ctx->s;
  launch_fill<double>(11 * NX, fblk, 0.0, s);
  if (int rc = h2d(ctx, out, d, N)) return rc;
  if (int rc = h2d(ctx, out + N, v, N)) return rc;
  if (int rc = h2d(ctx, out + 2 * N, a, N)) return rc;
  for (int q = 0; q < 3; q++) launch_permute(N, ctx->perm32, out + q * N, in + q * N, 1, s);
  // nodes deeper than NFBC keep their psi (FBC! does not visit them)
  if (int rc = h2d(ctx, f.psi, psi, NX)) return rc;
  if (int rc = h2d(ctx, f.psi2, psi, NX)) return rc;
  int hf[2] = {INT_MAX, 0};
  if (int rc = h2d(ctx, flags, hf, 2)) return rc;
  // as eq_newmark_step: the new d is the predictor (into out), the next
  // predictor is scratch (into in)
  launch_newmark_predict(N, in, in + N, in + 2 * N, dt, 0.5 * ctx->dt0 * ctx->dt0, out, s);
  launch_newmark(ctx->g, ctx->fp, f, ctx->mu, ctx->M, out, in + N, in + 2 * N, in, out + N, out + 2 * N, dt,
                 0.5 * ctx->dt0, 0.5 * ctx->dt0 * ctx->dt0, flags, NewmarkRecord{}, s);
  for (int q = 0; q < 3; q++) launch_permute(N, ctx->perm32, out + q * N, in + q * N, 0, s);
  if (int rc = d2h(ctx, hf, flags, 2)) return rc;
  if (int rc = d2h(ctx, d, in, N)) return rc;
  if (int rc = d2h(ctx, v, in + N, N)) return rc;
  if (int rc = d2h(ctx, a, in + 2 * N, N)) return rc;
  if (int rc = d2h(ctx, psi, f.psi, NX)) return rc;
  if (tau)
    if (int rc = d2h(ctx, tau, f.tau, NX)) return rc;
  if (Vf)
    if (int rc = d2h(ctx, Vf, f.Vf, NX)) return rc;
  if (FltVfree)
    if (int rc = d2h(ctx, FltVfree, f.FltVfree, NX)) return rc;
  if (int rc = sync_check(ctx)) return rc;
  if (hf[0] != INT_MAX) {
    char buf[128];
    snprintf(buf, sizeof buf, "NR SEARCH FAILED! (fault location %d)", hf[0]);
    ctx->err = buf;
    return EQ_NR_NAN;
  }
  if (hf[1] > 0) {
    ctx->err = "NRsearch! hit |delta| > 1e10 or k == 1000 (last iterate returned)";
    return EQ_NR_NOCONV;
  }
  return EQ_OK;
}

extern "C" int eq_slr(eq_ctx* ctx, const eq_fault_params* P, int64_t NFBC, int64_t FltNglob, const double* psi,
                      double* psi1, const double* Vf, double* Vf1, int64_t IDstate, const double* tau1, double dt) {
  if (!ctx || !psi || !psi1 || !Vf || !Vf1 || !tau1) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (int rc = check_fault_call(ctx, NFBC, FltNglob)) return rc;
  const int64_t n = FltNglob;
  FaultState& f = ctx->fs_api;
  if (int rc = h2d(ctx, f.psi, psi, n)) return rc;
  if (int rc = h2d(ctx, f.psi1, psi1, n)) return rc;
  if (int rc = h2d(ctx, f.Vf, Vf, n)) return rc;
  if (int rc = h2d(ctx, f.Vf1, Vf1, n)) return rc;
  if (int rc = h2d(ctx, f.tau1, tau1, n)) return rc;
  FaultDev fp{};
  if (int rc = api_fault_params(ctx, P, false, &fp)) return rc;
  fp.j0 = (int)(NFBC - 1);
  fp.IDstate = (int)IDstate;
  launch_slr(fp, f, dt, ctx->s);
  if (int rc = d2h(ctx, psi1, f.psi1, n)) return rc;
  if (int rc = d2h(ctx, Vf1, f.Vf1, n)) return rc;
  return sync_check(ctx);
}

extern "C" int eq_dtevol(eq_ctx* ctx, double dt, double dtmin, const double* XiLf, int64_t FaultNglob, int64_t NFBC,
                         const double* Vf, int64_t isolver, double* dt_out) {
  if (!ctx || !XiLf || !Vf || !dt_out || FaultNglob < 0 || NFBC < 1) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  DevTmp t;
  double *x = t.get<double>(FaultNglob), *v = t.get<double>(FaultNglob), *o = t.get<double>(1);
  UP(x, XiLf, FaultNglob);
  UP(v, Vf, FaultNglob);
  if (!o) return fail(ctx, EQ_NOMEM, "device alloc");
  launch_dtevol(x, v, (int)(NFBC - 1), (int)FaultNglob, dt, dtmin, (int)isolver, o, ctx->s);
  if (int rc = d2h(ctx, dt_out, o, 1)) return rc;
  return sync_check(ctx);
}

// ===========================================================================
// time loop
// ===========================================================================
extern "C" int eq_set_solver(eq_ctx* ctx, const eq_solver_opts* o) {
  if (!ctx || !o) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (o->pc < EQ_PC_NONE || o->pc > EQ_PC_AMG) return fail(ctx, EQ_EINVAL, "bad pc");
  if (o->prec < EQ_PREC_FP64 || o->prec > EQ_PREC_F32PC) return fail(ctx, EQ_EINVAL, "bad prec");
  if (o->guess != EQ_GUESS_PREVIOUS && o->guess != EQ_GUESS_LEARNED && o->guess != EQ_GUESS_GALERKIN &&
      o->guess != EQ_GUESS_LINEAR)
    return fail(ctx, EQ_EINVAL, "bad guess");
  if (o->guess == EQ_GUESS_LEARNED && !ctx->learn_on)
    return fail(ctx, EQ_ENOTREADY, "learned guess: basis and network not trained (eq_learn_train)");
  if (o->guess == EQ_GUESS_LINEAR && !ctx->learn.lin_ready)
    return fail(ctx, EQ_ENOTREADY, "linear guess: no linear predictor (eq_learn_set_linear)");
  if (o->guess == EQ_GUESS_GALERKIN && !ctx->learn.gal_ready)
    return fail(ctx, EQ_ENOTREADY, "Galerkin guess: no POD basis (eq_learn_set_basis)");
  if (o->cheb_single != ctx->opts.cheb_single) {
    // captured CG batches hold the other Chebyshev launch sequence
    for (auto& kv : ctx->cg_graphs) (void)hipGraphExecDestroy(kv.second);
    ctx->cg_graphs.clear();
  }
  ctx->opts = *o;
  return EQ_OK;
}

// (Re)sets the state pointers to the canonical layout of the state block: the
// Newmark step swaps d/v/a with d2/v2/a2, so after an odd number of dynamic
// steps ctx->d is not the block start.
static int alloc_state(eq_ctx* ctx) {
  const int64_t N = ctx->N;
  if (!ctx->state_blk) {
    ctx->state_blk = ctx->alloc<double>(8 * N);
    NEED(ctx->state_blk);
  }
  double* blk = ctx->state_blk;
  ctx->d = blk; ctx->v = blk + N; ctx->a = blk + 2 * N; ctx->d2 = blk + 3 * N; ctx->v2 = blk + 4 * N;
  ctx->a2 = blk + 5 * N; ctx->dPre = blk + 6 * N; ctx->bq = blk + 7 * N;
  if (!ctx->spec_save) {
    ctx->spec_save = ctx->alloc<double>(3 * ctx->NX);
    NEED(ctx->spec_save);
  }
  if (!ctx->gal_bold) {
    ctx->gal_bold = ctx->alloc<double>(4 * ctx->NX);
    NEED(ctx->gal_bold);
  }
  if (!ctx->gal_db) {
    ctx->gal_db = ctx->alloc<double>(4 * ctx->NX);
    NEED(ctx->gal_db);
  }
  return EQ_OK;
}

// Fresh results bookkeeping at time ctx->t: no event in progress, next
// interseismic snapshot at the next multiple of 5 years (src/main.jl:142-146).
static void reset_book(eq_ctx* ctx) {
  const double yr2sec = 365.0 * 24 * 60 * 60;
  ctx->book = eq_ctx::RunBook();
  ctx->book.tvsx = 5 * yr2sec * (std::floor(ctx->t / (5 * yr2sec)) + 1);
  ctx->book.delfref.assign(ctx->NX, 0.0);
}

extern "C" int eq_state_init(eq_ctx* ctx, double v_init, const double* psi_init) {
  if (!ctx) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (!ctx->have_material || !ctx->have_params) return fail(ctx, EQ_ENOTREADY, "material and params first");
  if (int rc = alloc_state(ctx)) return rc;
  const int64_t N = ctx->N, NX = ctx->NX;
  std::vector<double> v(N, v_init - 0.5 * ctx->Vpl);
  const int j0 = ctx->fp.j0;
  for (int j = 0; j < j0; j++) v[j] = 0.0;
  std::vector<double> hpsi(NX), hVf(NX);
  if (psi_init) {
    hpsi.assign(psi_init, psi_init + NX);
  } else {
    std::vector<double> fo(NX), Vo(NX), cca(NX), ccb(NX), Seff(NX);
    if (int rc = d2h(ctx, fo.data(), ctx->fp.fo, NX)) return rc;
    if (int rc = d2h(ctx, Vo.data(), ctx->fp.Vo, NX)) return rc;
    if (int rc = d2h(ctx, cca.data(), ctx->fp.cca, NX)) return rc;
    if (int rc = d2h(ctx, ccb.data(), ctx->fp.ccb, NX)) return rc;
    if (int rc = d2h(ctx, Seff.data(), ctx->fp.Seff, NX)) return rc;
    if (int rc = sync_check(ctx)) return rc;
    for (int j = 0; j < NX; j++)
      hpsi[j] = ctx->tauo[j] / (Seff[j] * ccb[j]) - fo[j] / ccb[j] - (cca[j] / ccb[j]) * std::log(2 * v_init / Vo[j]);
  }
  for (int j = 0; j < NX; j++) hVf[j] = (j < j0) ? 0.0 : 2 * v[j];
  launch_fill<double>(8 * N, ctx->state_blk, 0.0, ctx->s);   // d, v, a, their second buffers, dPre, rhs
  launch_fill<double>(11 * NX, ctx->fs.psi, 0.0, ctx->s);
  if (int rc = h2d(ctx, ctx->v, v.data(), N)) return rc;
  if (int rc = h2d(ctx, ctx->fs.psi, hpsi.data(), NX)) return rc;
  if (int rc = h2d(ctx, ctx->fs.Vf, hVf.data(), NX)) return rc;
  ctx->t = 0;
  ctx->dt = ctx->dt0;
  ctx->it = 0;
  ctx->isolver = 1;
  ctx->nr_noconv_total = 0;
  ctx->pcg_noconv = 0;
  reset_book(ctx);
  ctx->have_state = true;
  return sync_check(ctx);
}

extern "C" int eq_state_get(eq_ctx* ctx, double* d, double* v, double* a, double* psi, double* tau, double* Vf,
                            double* t, double* dt, int64_t* it, int64_t* isolver) {
  if (!ctx) return EQ_EINVAL;
  if (!ctx->have_state) return fail(ctx, EQ_ENOTREADY, "eq_state_init first");
  if (d) if (int rc = download_ref(ctx, ctx->d, d)) return rc;
  if (v) if (int rc = download_ref(ctx, ctx->v, v)) return rc;
  if (a) if (int rc = download_ref(ctx, ctx->a, a)) return rc;
  if (psi) if (int rc = d2h(ctx, psi, ctx->fs.psi, ctx->NX)) return rc;
  if (tau) if (int rc = d2h(ctx, tau, ctx->fs.tau, ctx->NX)) return rc;
  if (Vf) if (int rc = d2h(ctx, Vf, ctx->fs.Vf, ctx->NX)) return rc;
  if (t) *t = ctx->t;
  if (dt) *dt = ctx->dt;
  if (it) *it = ctx->it;
  if (isolver) *isolver = ctx->isolver;
  return sync_check(ctx);
}

extern "C" int eq_state_set(eq_ctx* ctx, const double* d, const double* v, const double* a, const double* psi,
                            const double* tau, double t, double dt, int64_t isolver) {
  if (!ctx) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (!ctx->have_state) return fail(ctx, EQ_ENOTREADY, "eq_state_init first");
  if (d) if (int rc = upload_ref(ctx, d, ctx->d)) return rc;
  if (v) if (int rc = upload_ref(ctx, v, ctx->v)) return rc;
  if (a) if (int rc = upload_ref(ctx, a, ctx->a)) return rc;
  if (psi) if (int rc = h2d(ctx, ctx->fs.psi, psi, ctx->NX)) return rc;
  if (tau) if (int rc = h2d(ctx, ctx->fs.tau, tau, ctx->NX)) return rc;
  ctx->t = t;
  ctx->dt = dt;
  ctx->isolver = (int)isolver;
  reset_book(ctx);
  return sync_check(ctx);
}

// Records of slots 0..k-1 (status and receiver samples) to the host.
// mirror: the step tail also writes its status and receiver samples into the
// pinned host copies (slot 0), which sync_status then leaves in place
static StepCapture capture(eq_ctx* ctx, int slot, bool mirror = false) {
  StepCapture c;
  const int64_t ns = (int64_t)ctx->out_seis.size();
  c.fault = ctx->rec_fault + (int64_t)slot * 3 * ctx->NX;
  c.seis_idx = ctx->seis_idx;
  c.nseis = (int)ns;
  c.seis = ns > 0 ? ctx->rec_seis + (int64_t)slot * 3 * ns : nullptr;
  if (mirror) {
    c.seis_mirror = ns > 0 ? ctx->rec_seis_host : nullptr;
    c.st_mirror = ctx->st_host;
  }
  return c;
}

// The status of a step whose tail wrote the pinned copies (capture with
// mirror): only the synchronisation.
static int sync_status(eq_ctx* ctx) {
  if (ctx->spin_recorded) {
    ctx->spin_recorded = false;
    return spin_wait(ctx);
  }
  return spin_sync(ctx);
}

static void clean_flags(eq_ctx* ctx) {
  if (!ctx->flags_dirty) return;
  launch_fill<int>(1, ctx->flags, INT_MAX, ctx->s);
  launch_fill<int>(1, ctx->flags + 1, 0, ctx->s);
  ctx->flags_dirty = false;
}

// Learned start for K_II x = b: x (interior = previous-step displacement,
// fault row = new fault data) += gamma p with p = |dF| Phi net(dF / |dF|)
// and gamma the energy-optimal step along p, so the guess is never worse (in
// the K_II norm) than the previous-step guess it starts from.
static void learned_correction(eq_ctx* ctx, double* x, const double* b, bool linear) {
  CgWork<double>& w = ctx->w64;
  const Geom& g = ctx->g;
  hipStream_t s = ctx->s;
  const int nt = tile_count(g);
  launch_cg_residual<double>(g, x, ctx->mu, b, w.r, ctx->part_rr, ctx->part_bb, ctx->npart, s);
  if (linear)
    learn_guess_linear(&ctx->learn, x, ctx->dPre, ctx->gp, s);
  else
    learn_guess(&ctx->learn, x, ctx->dPre, ctx->gp, s);
  launch_guess_dot(g, ctx->gp, ctx->mu, w.r, ctx->gq, ctx->part_rho, ctx->part_uq, s);
  launch_guess_update(ctx->N, x, w.r, ctx->gp, ctx->gq, ctx->part_rho, ctx->part_uq, nt, ctx->part_rr, ctx->npart,
                      s);
  ctx->n_learned_solves++;
}

// Galerkin start for K_II x = b on span(Phi): r = b - K_II x, x += Phi c with
// c = (Phi^T K_II Phi)^-1 Phi^T r, the K_II-norm best correction in span(Phi).
// db (when x solved K_II x = b_old to the solver tolerance and b - b_old
// lives on rows 1..4 only): r = (b - b_old) + (b_old - K_II x), and the
// second part, at tolerance level, is left to CG, so c comes from the 4 NX
// rows of Phi that meet b - b_old instead of a pass over the whole basis.
// The solve that follows starts from the true residual either way.
static void galerkin_correction(eq_ctx* ctx, double* x, const double* b, const double* db = nullptr) {
  if (db) {
    learn_project_rows(&ctx->learn, db, 4 * (int64_t)ctx->NX, ctx->s);
  } else {
    CgWork<double>& w = ctx->w64;
    launch_cg_residual<double>(ctx->g, x, ctx->mu, b, w.r, ctx->part_rr, ctx->part_bb, ctx->npart, ctx->s);
    learn_project(&ctx->learn, w.r, 1, ctx->s);
  }
  learn_expand_add(&ctx->learn, x, ctx->s, ctx->opts.basis_fp16 != 0);
  ctx->n_learned_solves++;
}

// The change-of-b Galerkin start from the FP16 basis with the start residual
// in the same launch (launch_gal_residual): x = d + Phi c lands in d2, which
// then becomes d (d2 is free between dynamic steps), and w64.r (with its FP32
// copy for an F32PC V-cycle), part_rr and part_bb are what cg_begin's
// residual would leave.  False when that form does not apply.  gate (or null):
// the residual launch does nothing unless *gate, and d2 is then left as it
// was (the caller swaps d and d2 back).
static bool fused_start_applies(const eq_ctx* ctx) {
  const eq_solver_opts& o = ctx->opts;
  const LearnState& ls = ctx->learn;
  return !o.galerkin_split && o.basis_fp16 && ls.Phi16 && ls.offset == ctx->NX &&
         (o.prec == EQ_PREC_FP64 || o.prec == EQ_PREC_F32PC);
}
static bool galerkin_start_fused(eq_ctx* ctx, const double* db, const int* gate = nullptr) {
  const eq_solver_opts& o = ctx->opts;
  LearnState& ls = ctx->learn;
  if (!db || !fused_start_applies(ctx)) return false;
  learn_project_rows(&ls, db, 4 * (int64_t)ctx->NX, ctx->s);
  const bool r32 = o.prec == EQ_PREC_F32PC && o.pc == EQ_PC_AMG;
  launch_gal_residual(ctx->g, ctx->d, (const half8*)ls.Phi16, ls.nI, ls.d.NB, ls.gal_c, ctx->mu, ctx->bq,
                      ctx->w64.r, r32 ? ctx->w32.r : nullptr, ctx->d2, ctx->part_rr, ctx->part_bb, ctx->npart,
                      ctx->s, gate);
  std::swap(ctx->d, ctx->d2);
  ctx->n_learned_solves++;
  return true;
}

// The end of a quasi-static step (see launch_qs_end): the second pass's fault
// update, the step finish, the step tail; with spec also the second solve's
// convergence test and the save that eq_step restores if its batch fell short.
static void qs_end(eq_ctx* ctx, bool spec) {
  launch_qs_end(ctx->g, ctx->fp, ctx->fs, ctx->mu, ctx->d, ctx->dPre, ctx->v, ctx->a, ctx->dt, ctx->dt0,
                spec ? ctx->spec_save : nullptr, spec ? ctx->sc : nullptr, ctx->part_rr, ctx->npart,
                spec ? ctx->sc_host : nullptr, ctx->st, ctx->flags, capture(ctx, 0, true), ctx->s);
  ctx->tail_done = true;
}

// r_valid: the last solve's residual may be reused (eq_ctx::qs_r_valid).
// At the end of a quasi-static step inside eq_run: the next step's begin
// (dt from the step tail's status, the extrapolation factor from it and this
// step's dt) and its change-of-b Galerkin start, sent before the host reads
// the status, so the device has them while the host turns round.  Both do
// nothing unless the status says go; eq_step then takes them back
// (qs_take_back) and the next step starts as usual.  Its conditions are the
// next step's: it follows a quasi-static solve (r_valid) and extrapolates
// when this step did follow one.
static void qs_send_next(eq_ctx* ctx) {
  const eq_solver_opts& o = ctx->opts;
  if (o.sync_pass2 || o.guess != EQ_GUESS_GALERKIN || !ctx->learn.gal_ready || o.galerkin_full ||
      !fused_start_applies(ctx) || ctx->learn_record_on)
    return;
  hipStream_t s = ctx->s;
  if (hipEventRecord(ctx->ev_spin, s) != hipSuccess) return;   // (the step's own synchronisation reports it)
  ctx->spin_recorded = true;
  const bool theta_on = o.extrapolate && ctx->qs_prev_r_valid;
  launch_qs_begin(ctx->g, ctx->fp, ctx->fs, ctx->mu, ctx->d, ctx->dPre, ctx->v, ctx->bq, 0.0, 1, s, ctx->gal_db,
                  theta_on ? 1.0 : 0.0, ctx->gal_bold, ctx->st, ctx->dt);
  ctx->begin_r_fused = galerkin_start_fused(ctx, ctx->gal_db, &ctx->st->go);
  ctx->begin_sent = true;
}

// After the step's status is on the host: a begin sent ahead that the status
// stopped did nothing, so d and d2 go back.
static void qs_take_back(eq_ctx* ctx) {
  if (!ctx->begin_sent || ctx->st_host->go) return;
  if (ctx->begin_r_fused) {
    std::swap(ctx->d, ctx->d2);
    ctx->n_learned_solves--;
  }
  ctx->begin_sent = false;
}

static int qs_step(eq_ctx* ctx, eq_step_info* info, bool r_valid) {
  const Geom& g = ctx->g;
  hipStream_t s = ctx->s;
  const double dt = ctx->dt;
  bool bold_saved = false;
  bool mid_sent = false;   // the second pass's begin and start were enqueued with the first solve
  bool r_fused_mid = false;
  const int prec = ctx->opts.prec;
  // the second solve may run ahead of its read-back (eq_step finishes it)
  const bool ahead = !ctx->opts.sync_pass2 && (prec == EQ_PREC_FP64 || prec == EQ_PREC_F32PC) && ctx->spec_save;
  const bool f32pc = prec == EQ_PREC_F32PC && ctx->opts.pc == EQ_PC_AMG;
  // the first pass's begin and start went out with the step before (qs_send_next)
  const bool sent = ctx->begin_sent;
  ctx->begin_sent = false;
  for (int pass = 0; pass < 2; pass++) {
    const bool galerkin = ctx->opts.guess == EQ_GUESS_GALERKIN && ctx->learn.gal_ready;
    // the previous solve (this step's first, or the last step's second when
    // r_valid) left d solving K_II d = b_old; only b rows 1..4 change
    const bool upd = (pass == 1 || r_valid) && galerkin && !ctx->opts.galerkin_full;
    // first pass after two Galerkin-started quasi-static steps in a row, each
    // from the solve before it: start from the last increment extrapolated
    // (d and dPre are then consecutive solutions, gal_bold the b of the older)
    const double theta = (pass == 0 && galerkin && r_valid && ctx->opts.extrapolate && ctx->qs_prev_r_valid &&
                          ctx->qs_prev_dt > 0)
                             ? dt / ctx->qs_prev_dt
                             : 0.0;
    if (pass == 0) bold_saved = galerkin;
    if (pass == 0) {
      if (!sent)
        launch_qs_begin(g, ctx->fp, ctx->fs, ctx->mu, ctx->d, ctx->dPre, ctx->v, ctx->bq, dt, 1, s,
                        upd ? ctx->gal_db : nullptr, theta, galerkin ? ctx->gal_bold : nullptr);
    } else if (!mid_sent)   // the first pass's fault update, then this pass's begin
      launch_qs_mid(g, ctx->fp, ctx->fs, ctx->mu, ctx->d, ctx->dPre, ctx->v, ctx->bq, dt,
                    upd ? ctx->gal_db : nullptr, s);
    const bool learned = pass == 0 && ctx->opts.guess == EQ_GUESS_LEARNED && ctx->learn_on;
    const bool linear = pass == 0 && ctx->opts.guess == EQ_GUESS_LINEAR && ctx->learn.lin_ready;
    if (learned || linear) learned_correction(ctx, ctx->d, ctx->bq, linear);
    // r_fused: the Galerkin start left the residual (and its FP32 copy)
    const bool r_fused = mid_sent                ? r_fused_mid
                         : (pass == 0 && sent) ? ctx->begin_r_fused
                                               : galerkin && galerkin_start_fused(ctx, upd ? ctx->gal_db : nullptr);
    if (galerkin && !r_fused) galerkin_correction(ctx, ctx->d, ctx->bq, upd ? ctx->gal_db : nullptr);
    int64_t it = 0;
    double rr = 0;
    if (pass == 0 && ahead && r_fused && !ctx->learn_record_on) {
      // The first solve's batch; then, before the host waits, the fault update
      // with the batch's convergence test (qs_mid_kernel) and the second
      // pass's start, which both change nothing unless the test passed.  The
      // host reads the test after the fault update while the start runs.
      // (The fused start applied to this pass, so it applies to the next.)
      const int64_t batch = ctx->pred_iters[0];
      ctx->cg_slot = 0;
      int rc = cg_begin<double>(ctx, ctx->d, ctx->bq, ctx->opts.tol, ctx->opts.maxiter, ctx->opts.pc, f32pc, true,
                                batch, true);
      if (!rc) rc = enqueue_iterations<double>(ctx, ctx->d, ctx->opts.pc, f32pc, batch);
      ctx->cg_init_pending = false;
      if (rc) return rc;
      launch_qs_mid(g, ctx->fp, ctx->fs, ctx->mu, ctx->d, ctx->dPre, ctx->v, ctx->bq, dt, ctx->gal_db, s, ctx->sc,
                    ctx->part_rr, ctx->npart, ctx->sc_host);
      HIPCHK(ctx, hipEventRecord(ctx->ev_spin, s));
      r_fused_mid = galerkin_start_fused(ctx, ctx->gal_db, &ctx->sc->done);
      if (int rc2 = spin_wait(ctx)) return rc2;
      if (!ctx->sc_host->done) {
        // the batch fell short: the gated launches left d, d2 and the fault
        // as they were; finish the solve, then enqueue them again
        if (r_fused_mid) {
          std::swap(ctx->d, ctx->d2);
          ctx->n_learned_solves--;
        }
        if (int rc3 = cg_rest<double>(ctx, ctx->d, ctx->opts.pc, f32pc, batch)) return rc3;
        launch_qs_mid(g, ctx->fp, ctx->fs, ctx->mu, ctx->d, ctx->dPre, ctx->v, ctx->bq, dt, ctx->gal_db, s);
        r_fused_mid = galerkin_start_fused(ctx, ctx->gal_db);
      }
      rc = cg_end(ctx, &it, &rr);
      ctx->cg_slot = 2;
      if (rc == EQ_PCG_NOCONV) ctx->pcg_noconv++;
      else if (rc) return rc;
      info->pcg_iters[0] = it;
      mid_sent = true;
      continue;
    }
    if (pass == 1 && ahead) {
      // pass 2 ahead of its read-back (eq_step finishes it)
      ctx->cg_slot = 1;
      int rc = cg_begin<double>(ctx, ctx->d, ctx->bq, ctx->opts.tol, ctx->opts.maxiter, ctx->opts.pc, f32pc, r_fused,
                                ctx->pred_iters[1], r_fused);
      // the batch's convergence test goes into the step's end kernel below
      if (!rc) rc = enqueue_iterations<double>(ctx, ctx->d, ctx->opts.pc, f32pc, ctx->pred_iters[1]);
      ctx->cg_init_pending = false;
      ctx->cg_slot = 2;
      if (rc) return rc;
      // the fault update saves psi, Vf and the fault row of d first (restored
      // if the batch falls short)
      ctx->spec_pending = true;
      ctx->spec_f32pc = f32pc;
      ctx->spec_batch = ctx->pred_iters[1];
      continue;
    }
    ctx->cg_slot = pass;
    int rc = cg_any(ctx, ctx->d, ctx->bq, ctx->opts.tol, ctx->opts.maxiter, ctx->opts.pc, ctx->opts.prec, &it, &rr,
                    learned || r_fused, r_fused);
    ctx->cg_slot = 2;
    if (rc == EQ_PCG_NOCONV) ctx->pcg_noconv++;
    else if (rc) return rc;
    info->pcg_iters[pass] = it;
    if (pass == 0 && ctx->learn_record_on) learn_record(&ctx->learn, ctx->d, ctx->dPre, s);
    // (the first pass's fault update opens the second pass: launch_qs_mid)
  }
  // second pass's fault update, the step finish and the step tail (with a pass
  // sent ahead, also its convergence test and the save of what the update reads)
  qs_end(ctx, ctx->spec_pending);
  ctx->qs_prev_r_valid = r_valid && bold_saved;
  ctx->qs_prev_dt = dt;
  if (ctx->run_more) qs_send_next(ctx);
  return EQ_OK;
}

static int dyn_step(eq_ctx* ctx) {
  const double dtmin = ctx->dt0;
  // predictor (the new d) into d2; the next predictor lands in d, which the swap retires
  launch_newmark_predict(ctx->N, ctx->d, ctx->v, ctx->a, ctx->dt, 0.5 * dtmin * dtmin, ctx->d2, ctx->s);
  launch_newmark(ctx->g, ctx->fp, ctx->fs, ctx->mu, ctx->M, ctx->d2, ctx->v, ctx->a, ctx->d, ctx->v2, ctx->a2,
                 ctx->dt, 0.5 * dtmin, 0.5 * dtmin * dtmin, ctx->flags, NewmarkRecord{}, ctx->s,
                 lt_slot(ctx, EQ_TK_NEWMARK));
  std::swap(ctx->d, ctx->d2);
  std::swap(ctx->v, ctx->v2);
  std::swap(ctx->a, ctx->a2);
  return EQ_OK;
}

extern "C" int eq_step(eq_ctx* ctx, eq_step_info* info) {
  if (!ctx || !info) return EQ_EINVAL;
  if (!ctx->have_state) return fail(ctx, EQ_ENOTREADY, "eq_state_init first");
  if (ctx->opts.pc == EQ_PC_AMG && !ctx->have_amg) return fail(ctx, EQ_ENOTREADY, "eq_amg_setup first");
  std::memset(info, 0, sizeof(*info));
  ctx->it += 1;
  ctx->t += ctx->dt;
  info->dt_used = ctx->dt;
  info->isolver_used = ctx->isolver;
  clean_flags(ctx);
  ctx->flags_dirty = true;   // until the step tail has reset them
  const bool r_valid = ctx->qs_r_valid;
  ctx->qs_r_valid = false;
  ctx->tail_done = false;
  int rc = (ctx->isolver == 1) ? qs_step(ctx, info, r_valid) : dyn_step(ctx);
  if (rc) return rc;
  if (!ctx->tail_done)
    launch_step_tail(ctx->fp, ctx->d, ctx->v, ctx->a, ctx->fs.tau, ctx->dt, ctx->dt0, ctx->isolver, ctx->st,
                     ctx->flags, capture(ctx, 0, true), ctx->s);
  const bool spec = ctx->spec_pending;
  ctx->spec_pending = false;
  if (int rc2 = sync_status(ctx)) return rc2;
  qs_take_back(ctx);
  if (spec) {
    // pass 2's scalars came back with the status
    if (!ctx->sc_host->done) {
      // the batch fell short: restore what the fault update read, finish the
      // solve, and redo the fault update, the finish and the step tail
      hipStream_t s = ctx->s;
      launch_copy3(ctx->NX, ctx->fs.psi, ctx->fs.Vf, ctx->d, ctx->spec_save, false, s);
      ctx->cg_slot = 1;
      int rc3 = cg_rest<double>(ctx, ctx->d, ctx->opts.pc, ctx->spec_f32pc, ctx->spec_batch);
      ctx->cg_slot = 2;
      if (rc3) return rc3;
      // (the first step tail left the flags cleared; nothing in this step sets them)
      qs_end(ctx, false);
      if (int rc4 = sync_status(ctx)) return rc4;
    }
    ctx->cg_slot = 1;
    int64_t it = 0;
    const int rc5 = cg_end(ctx, &it, nullptr);
    ctx->cg_slot = 2;
    info->pcg_iters[1] = it;
    if (rc5 == EQ_PCG_NOCONV) ctx->pcg_noconv++;
  }
  ctx->flags_dirty = false;
  ctx->qs_r_valid = info->isolver_used == 1;
  const StepStatus& st = *ctx->st_host;
  info->t = ctx->t;
  info->it = ctx->it;
  info->Vfmax = st.Vfmax;
  info->dt_next = st.dt_next;
  info->isolver_next = st.isolver_next;
  info->nr_noconv = st.nr_noconv;
  ctx->nr_noconv_total += st.nr_noconv;
  ctx->isolver = st.isolver_next;
  ctx->dt = st.dt_next;
  if (st.nr_flag) {
    char buf[128];
    snprintf(buf, sizeof buf, "NR SEARCH FAILED! (fault location %d)", st.nr_flag);
    ctx->err = buf;
    return EQ_NR_NAN;
  }
  return EQ_OK;
}

// Up to kBatch dynamic steps enqueued without a host synchronisation.  Each
// step is one Newmark tile kernel that also writes the step's record (fault
// velocity maxima per tile, FBC! flags, fault rows, receiver samples); the
// kernel of the next step reads it and stops, in every block alike, once a
// step has ended the dynamic regime (the step tail's switch rule), failed the
// FBC! guard or did not run, and the steps after it do nothing.  dt is dtmin
// for every dynamic step (dtevol!, src/dtevol.jl:5-10).  The host rebuilds
// the status of every step from the records; the last step taken gets the
// step tail kernel, whose status (dtevol!, regime) carries the loop on.
// Returns the number of steps taken in *m (>= 1).
static int dyn_batch(eq_ctx* ctx, double total_time, int64_t max_more, int* m, StepStatus* tail_expect) {
  ctx->qs_r_valid = false;
  int k = 0;
  for (double tt = ctx->t; k < eq_ctx::kBatch && k < max_more && tt < total_time; k++) tt += ctx->dt;
  if (k == 0) k = 1;
  hipStream_t s = ctx->s;
  const int nbx = (int)tile_grid(ctx->g).x;
  const int64_t ns = (int64_t)ctx->out_seis.size();
  const int64_t NX = ctx->NX;
  launch_fill<int>(2 * eq_ctx::kBatch, ctx->nrflags, 0, s);
  launch_fill_strided<int>(eq_ctx::kBatch, ctx->nrflags, 2, INT_MAX, s);
  const double dtmin = ctx->dt0;
  double* bufs[2][3] = {{ctx->d, ctx->v, ctx->a}, {ctx->d2, ctx->v2, ctx->a2}};
  // step q reads the predictor (its new d) from the other set and v, a from
  // its own, and writes v, a into the other set and the next predictor into
  // its own d; after q + 1 steps the state is set (q + 1) & 1, as before
  launch_newmark_predict(ctx->N, ctx->d, ctx->v, ctx->a, ctx->dt, 0.5 * dtmin * dtmin, ctx->d2, s);
  for (int q = 0; q < k; q++) {
    double** src = bufs[q & 1];
    double** dst = bufs[(q + 1) & 1];
    NewmarkRecord rec;
    rec.vprev = q > 0 ? ctx->vpart + (int64_t)(q - 1) * 2 * nbx : nullptr;
    rec.nrprev = q > 0 ? ctx->nrflags + 2 * (q - 1) : nullptr;
    rec.vpart_out = ctx->vpart + (int64_t)q * 2 * nbx;
    rec.cap_fault = ctx->rec_fault + (int64_t)q * 3 * NX;
    rec.seis_idx = ctx->seis_idx;
    rec.seis_tile = ctx->seis_tile;
    rec.nseis = (int)ns;
    rec.cap_seis = ns > 0 ? ctx->rec_seis + (int64_t)q * 3 * ns : nullptr;
    launch_newmark(ctx->g, ctx->fp, ctx->fs, ctx->mu, ctx->M, dst[0], src[1], src[2], src[0], dst[1], dst[2],
                   ctx->dt, 0.5 * dtmin, 0.5 * dtmin * dtmin, ctx->nrflags + 2 * q, rec, s,
                   lt_slot(ctx, EQ_TK_NEWMARK));
  }
  HIPCHK(ctx, hipMemcpyAsync(ctx->vpart_host, ctx->vpart, sizeof(double) * 2 * nbx * k, hipMemcpyDeviceToHost, s));
  HIPCHK(ctx, hipMemcpyAsync(ctx->nrflags_host, ctx->nrflags, sizeof(int) * 2 * k, hipMemcpyDeviceToHost, s));
  if (ns > 0)
    HIPCHK(ctx, hipMemcpyAsync(ctx->rec_seis_host, ctx->rec_seis, sizeof(double) * 3 * ns * k,
                               hipMemcpyDeviceToHost, s));
  if (int rc = spin_sync(ctx)) return rc;
  // status of each step from its record, as the step tail computes it
  int n = 0;
  for (int q = 0; q < k; q++) {
    const double* vp = ctx->vpart_host + (int64_t)q * 2 * nbx;
    double mx = -INFINITY, mi = 2147483647.0;
    bool ran = true;
    for (int b = 0; b < nbx; b++) {
      if (vp[2 * b + 1] < 0.0) { ran = false; break; }
      if (findmax_before_host(vp[2 * b], vp[2 * b + 1], mx, mi)) { mx = vp[2 * b]; mi = vp[2 * b + 1]; }
    }
    if (!ran) break;
    StepStatus& st = ctx->st_host[q];
    st.Vfmax = 2 * mx + ctx->Vpl;
    st.dt_next = dtmin;
    st.isolver_next = st.Vfmax < 2e-3 ? 1 : 2;
    st.hypo = (int)mi;
    st.nr_flag = ctx->nrflags_host[2 * q] == INT_MAX ? 0 : ctx->nrflags_host[2 * q];
    st.nr_noconv = ctx->nrflags_host[2 * q + 1];
    n = q + 1;
    if (st.isolver_next != 2 || st.nr_flag != 0) break;
  }
  if (n == 0) return fail(ctx, EQ_HIP_ERR, "dynamic batch: first step has no record");
  if (n & 1) {
    std::swap(ctx->d, ctx->d2);
    std::swap(ctx->v, ctx->v2);
    std::swap(ctx->a, ctx->a2);
  }
  // the step tail on the last step taken: dtevol! and the regime switch; its
  // status comes back with the caller's next synchronisation and must equal
  // the one rebuilt here from the record (dyn_tail_check)
  StepCapture none{};
  launch_step_tail(ctx->fp, ctx->d, ctx->v, ctx->a, ctx->fs.tau, ctx->dt, dtmin, 2, ctx->st + (n - 1),
                   ctx->nrflags + 2 * (n - 1), none, s);
  *tail_expect = ctx->st_host[n - 1];
  HIPCHK(ctx, hipMemcpyAsync(ctx->st_host + (n - 1), ctx->st + (n - 1), sizeof(StepStatus), hipMemcpyDeviceToHost,
                             s));
  *m = n;
  return EQ_OK;
}

// After the synchronisation that follows dyn_batch: the step tail's status
// against the one rebuilt from the last step's record.
static int dyn_tail_check(eq_ctx* ctx, int n, const StepStatus& mine) {
  const StepStatus& tail = ctx->st_host[n - 1];
  if (tail.isolver_next != mine.isolver_next || tail.hypo != mine.hypo || tail.nr_flag != mine.nr_flag ||
      !(tail.Vfmax == mine.Vfmax || (std::isnan(tail.Vfmax) && std::isnan(mine.Vfmax))))
    return fail(ctx, EQ_HIP_ERR, "dynamic batch: step record disagrees with the step tail");
  return EQ_OK;
}

// results bookkeeping (src/main.jl:361-436)
extern "C" int eq_run(eq_ctx* ctx, double total_time, int64_t max_steps, eq_results* R) {
  if (!ctx || !R) return EQ_EINVAL;
  if (!ctx->have_state) return fail(ctx, EQ_ENOTREADY, "eq_state_init first");
  if (ctx->opts.pc == EQ_PC_AMG && !ctx->have_amg) return fail(ctx, EQ_ENOTREADY, "eq_amg_setup first");
  const int64_t NX = ctx->NX;
  const double Vpl = ctx->Vpl;
  const double yr2sec = 365.0 * 24 * 60 * 60;
  const double tvsxinc = 5 * yr2sec;
  const double tevneinc = 0.5;
  eq_ctx::RunBook& B = ctx->book;
  double& tvsx = B.tvsx;
  double& tevneb = B.tevneb;
  double& tevne = B.tevne;
  int& idelevne = B.idelevne;
  int& slipstart = B.slipstart;
  if ((int64_t)B.delfref.size() != NX) B.delfref.assign(NX, 0.0);
  std::vector<double>& delfref = B.delfref;
  std::vector<double> tmp(NX);
  const size_t nseis = ctx->out_seis.size();
  R->n_steps = R->n_start = R->n_end = R->n_is = R->n_seis = R->n_rec = 0;
  R->n_qs = R->n_dyn = R->total_pcg_iters = R->nr_noconv = 0;
  R->wall_qs = R->wall_dyn = 0;
  // What the bookkeeping of a finished step (time t, maximum slip rate Vfmax)
  // needs from its record, with the bookkeeping state S advanced as the
  // bookkeeping advances it (src/main.jl:361-422).  Pure: a batch is first
  // run through it on a copy to learn which records to fetch.
  struct BookScalars { double tvsx, tevneb, tevne; int idelevne, slipstart; };
  struct Needs {
    bool start, end, is, seis;
    bool any() const { return start || end || is || seis; }
  };
  auto needs = [&](BookScalars& S, double t, double Vfmax) -> Needs {
    Needs n;
    n.start = Vfmax > 1.01 * ctx->Vthres && S.slipstart == 0;
    n.end = Vfmax < 0.99 * ctx->Vthres && S.slipstart == 1;
    n.is = t > S.tvsx;
    n.seis = false;
    if (Vfmax > ctx->Vevne) {
      if (S.idelevne == 0) n.seis = true;
      else if ((t - S.tevneb) > S.tevne) n.seis = true;
    }
    if (n.start) S.slipstart = 1;
    if (n.end) S.slipstart = 0;
    if (n.is) S.tvsx += tvsxinc;
    if (Vfmax > ctx->Vevne) {
      if (S.idelevne == 0) {
        S.idelevne = 1;
        S.tevneb = t;
        S.tevne = tevneinc;
      } else if (n.seis) {
        S.tevne += tevneinc;
      }
    } else {
      S.idelevne = 0;
    }
    return n;
  };
  // fault rows of d, v and tau after the step of record slot q, in pinned memory
  auto fetch_fault = [&](int q) -> int {
    if (int rc = d2h(ctx, ctx->rec_fault_host + (int64_t)q * 3 * NX, ctx->rec_fault + (int64_t)q * 3 * NX, 3 * NX))
      return rc;
    return EQ_OK;
  };
  auto put = [&](double* M, int64_t cap, int64_t col, const std::vector<double>& src) {
    if (M && col < cap) std::memcpy(M + col * NX, src.data(), sizeof(double) * NX);
  };
  // bookkeeping of one finished step whose record is in slot q; fault: its
  // fault rows on the host, or null (fetched here when needed)
  auto book = [&](const eq_step_info& info, int q, const double* fault) -> int {
    const int64_t k = R->n_steps;
    if (k < R->cap_steps) {
      if (R->time_) R->time_[k] = info.t;
      if (R->Vfmax) R->Vfmax[k] = info.Vfmax;
      if (R->isolver) R->isolver[k] = info.isolver_used;
      if (R->pcg_iters) { R->pcg_iters[2 * k] = info.pcg_iters[0]; R->pcg_iters[2 * k + 1] = info.pcg_iters[1]; }
    }
    R->n_steps++;
    R->total_pcg_iters += info.pcg_iters[0] + info.pcg_iters[1];
    R->nr_noconv += info.nr_noconv;
    const double t = info.t, Vfmax = info.Vfmax;
    BookScalars S{tvsx, tevneb, tevne, idelevne, slipstart};
    const Needs nd = needs(S, t, Vfmax);
    tvsx = S.tvsx;
    tevneb = S.tevneb;
    tevne = S.tevne;
    idelevne = S.idelevne;
    slipstart = S.slipstart;
    const bool need_start = nd.start, need_end = nd.end, need_is = nd.is, need_seis = nd.seis;
    if (nd.any() && !fault) {
      if (int rc = fetch_fault(q)) return rc;
      if (int rc = spin_sync(ctx)) return rc;
      fault = ctx->rec_fault_host + (int64_t)q * 3 * NX;
    }
    const double* dF = fault;
    const double* vF = fault ? fault + NX : nullptr;
    const double* tau = fault ? fault + 2 * NX : nullptr;
    if (need_start) {
      for (int64_t j = 0; j < NX; j++) delfref[j] = 2 * dF[j] + Vpl * t;
      const int64_t e = R->n_start;
      if (e < R->cap_events) {
        if (R->tStart) R->tStart[e] = t;
        for (int64_t j = 0; j < NX; j++) tmp[j] = (tau[j] + ctx->tauo[j]) / 1e6;
        put(R->taubefore, R->cap_events, e, tmp);
        if (R->hypo) R->hypo[e] = ctx->FltX[ctx->st_host[q].hypo];
      }
      R->n_start++;
    }
    if (need_end) {
      const int64_t e = R->n_end;
      if (e < R->cap_events) {
        for (int64_t j = 0; j < NX; j++) tmp[j] = 2 * dF[j] + Vpl * t - delfref[j];
        put(R->delfafter, R->cap_events, e, tmp);
        for (int64_t j = 0; j < NX; j++) tmp[j] = (tau[j] + ctx->tauo[j]) / 1e6;
        put(R->tauafter, R->cap_events, e, tmp);
        if (R->tEnd) R->tEnd[e] = t;
      }
      R->n_end++;
    }
    if (need_is) {
      const int64_t e = R->n_is;
      if (e < R->cap_is) {
        for (int64_t j = 0; j < NX; j++) tmp[j] = 2 * dF[j] + Vpl * t;
        put(R->is_slip, R->cap_is, e, tmp);
        for (int64_t j = 0; j < NX; j++) tmp[j] = 2 * vF[j] + Vpl;
        put(R->is_slipvel, R->cap_is, e, tmp);
        for (int64_t j = 0; j < NX; j++) tmp[j] = (tau[j] + ctx->tauo[j]) / 1e6;
        put(R->is_stress, R->cap_is, e, tmp);
      }
      R->n_is++;
    }
    if (Vfmax > ctx->Vevne) {
      if (need_seis) {
        const int64_t e = R->n_seis;
        if (e < R->cap_seis) {
          for (int64_t j = 0; j < NX; j++) tmp[j] = 2 * dF[j] + Vpl * t;
          put(R->seismic_slip, R->cap_seis, e, tmp);
          for (int64_t j = 0; j < NX; j++) tmp[j] = 2 * vF[j] + Vpl;
          put(R->seismic_slipvel, R->cap_seis, e, tmp);
          for (int64_t j = 0; j < NX; j++) tmp[j] = (tau[j] + ctx->tauo[j]) / 1e6;
          put(R->seismic_stress, R->cap_seis, e, tmp);
        }
        R->n_seis++;
      }
    }
    if (info.it % 10 == 0 && nseis > 0) {
      const int64_t e = R->n_rec;
      if (e < R->cap_rec) {
        const double* smp = ctx->rec_seis_host + (int64_t)q * 3 * nseis;
        for (size_t r = 0; r < nseis; r++) {
          if (R->dSeis) R->dSeis[r * R->cap_rec + e] = smp[3 * r];
          if (R->vSeis) R->vSeis[r * R->cap_rec + e] = smp[3 * r + 1];
          if (R->aSeis) R->aSeis[r * R->cap_rec + e] = smp[3 * r + 2];
        }
      }
      R->n_rec++;
    }
    return EQ_OK;
  };
  using clk = std::chrono::steady_clock;
  while (ctx->t < total_time && (max_steps <= 0 || R->n_steps < max_steps)) {
    const auto t0 = clk::now();
    if (ctx->isolver == 2) {
      // dynamic regime: a batch of steps, then the bookkeeping of each
      int m = 0;
      StepStatus expect{};
      const int64_t more = max_steps > 0 ? max_steps - R->n_steps : INT64_MAX;
      int rc = dyn_batch(ctx, total_time, more, &m, &expect);
      if (rc) {
        R->wall_dyn += std::chrono::duration<double>(clk::now() - t0).count();
        return rc;
      }
      // the records this batch's bookkeeping reads come back with the step
      // tail's status, in one synchronisation
      bool fetched[eq_ctx::kBatch] = {};
      {
        BookScalars S{tvsx, tevneb, tevne, idelevne, slipstart};
        double tt = ctx->t;
        for (int q = 0; q < m; q++) {
          tt += ctx->dt;
          const double vmax = q == m - 1 ? expect.Vfmax : ctx->st_host[q].Vfmax;
          if (needs(S, tt, vmax).any()) {
            if (int rc2 = fetch_fault(q)) return rc2;
            fetched[q] = true;
          }
        }
      }
      rc = spin_sync(ctx);
      R->wall_dyn += std::chrono::duration<double>(clk::now() - t0).count();
      if (rc) return rc;
      if (int rc2 = dyn_tail_check(ctx, m, expect)) return rc2;
      for (int q = 0; q < m; q++) {
        const StepStatus& st = ctx->st_host[q];
        eq_step_info info{};
        ctx->it += 1;
        ctx->t += ctx->dt;
        info.dt_used = ctx->dt;
        info.isolver_used = 2;
        info.t = ctx->t;
        info.it = ctx->it;
        info.Vfmax = st.Vfmax;
        info.dt_next = st.dt_next;
        info.isolver_next = st.isolver_next;
        info.nr_noconv = st.nr_noconv;
        ctx->nr_noconv_total += st.nr_noconv;
        R->n_dyn++;
        if (q == m - 1) {
          ctx->isolver = st.isolver_next;
          ctx->dt = st.dt_next;
        }
        if (st.nr_flag) {
          char buf[128];
          snprintf(buf, sizeof buf, "NR SEARCH FAILED! (fault location %d)", st.nr_flag);
          ctx->err = buf;
          return EQ_NR_NAN;
        }
              if (int rc2 = book(info, q, fetched[q] ? ctx->rec_fault_host + (int64_t)q * 3 * NX : nullptr)) return rc2;
      }
    } else {
      eq_step_info info;
      // (eq_step advances t by this dt, so this is the loop's test after it)
      ctx->run_more = ctx->t + ctx->dt < total_time && (max_steps <= 0 || R->n_steps + 1 < max_steps);
      int rc = eq_step(ctx, &info);
      ctx->run_more = false;
      const double el = std::chrono::duration<double>(clk::now() - t0).count();
      if (info.isolver_used == 1) { R->n_qs++; R->wall_qs += el; } else { R->n_dyn++; R->wall_dyn += el; }
      if (rc) return rc;
      if (int rc2 = book(info, 0, nullptr)) return rc2;
    }
  }
  return EQ_OK;
}

// ===========================================================================
// measurement hook
// ===========================================================================
template <typename T>
static void time_body(eq_ctx* ctx, int which) {
  CgWork<T>& w = work<T>(ctx);
  const Geom& g = ctx->g;
  switch (which) {
    case EQ_TIME_KU:
      launch_ku<T>(g, w.x, mu_of<T>(ctx), w.q, 1, 0, ctx->s);
      break;
    case EQ_TIME_CG_DIR:
      launch_cg_dir<T>(g, w.c, w.u[0], mu_of<T>(ctx), w.u[1], w.q, ctx->part_rho, tile_count(g), ctx->sc,
                       ctx->part_uq, ctx->s);
      break;
    case EQ_TIME_VCYCLE:
      vcycle<T>(ctx, w.r, w.c);
      break;
    case EQ_TIME_CHEB: {
      // one middle Chebyshev step: window d, reads r, x, mu, dinv; writes r, d', x
      ChebStep<T> st;
      st.src = w.chd[0];
      st.rin = w.chr[0];
      st.rout = w.chr[0];
      st.c1 = T(0.5);
      st.c2 = T(1e-3);
      st.xin = w.chx;
      st.xout = w.chx;
      st.dout = w.chd[1];
      st.mode = 0;
      launch_cheb<T>(g, st, mu_of<T>(ctx), dinv_of<T>(ctx), ctx->part_rr, ctx->npart, ctx->sc, ctx->part_rho,
                     ctx->s);
      break;
    }
    case EQ_TIME_SWEEP: {
      // one whole-tile sweep from x = 0 (pre-smoothing) at the hierarchy's degree
      ChebSweep<T> op = sweep_params<T>(ctx, 0);
      op.b = w.r;
      op.xout = w.chx;
      launch_cheb_sweep<T>(g, op, ctx->s);
      break;
    }
    case EQ_TIME_SWEEP_POST: {
      // one whole-tile sweep from chx + P0 coarse (post-smoothing), output c and <c, r>; in FP32 as
      // the V-cycle of an F32PC solve runs it: c widened to FP64, the partial with the FP64 residual
      ChebSweep<T> op = sweep_params<T>(ctx, 1);
      op.b = w.r;
      op.src = w.chx;
      op.coarse = lvecs<T>(ctx->lev[1]).xp;
      op.xout = w.c;
      op.rdot = w.r;
      if (sizeof(T) == 4) {
        op.xout64 = ctx->w64.c;
        op.rdot64 = ctx->w64.r;
      }
      launch_cheb_sweep<T>(g, op, ctx->s);
      break;
    }
    case EQ_TIME_RESTRICT:
      launch_restrict<T>(g, w.chx, w.r, mu_of<T>(ctx), lvecs<T>(ctx->lev[1]).b, ctx->sc, ctx->s);
      break;
    case EQ_TIME_CHEB2: {
      // two middle Chebyshev steps in one launch: window d, reads r, x, mu, dinv; writes r'', d'', x
      ChebStep<T> st[2];
      for (int m = 0; m < 2; m++) {
        st[m].src = w.chd[m];
        st[m].rin = w.chr[m];
        st[m].rout = w.chr[m + 1];
        st[m].c1 = T(0.5);
        st[m].c2 = T(1e-3);
        st[m].xin = w.chx;
        st[m].xout = w.chx;
        st[m].dout = w.chd[m + 1];
        st[m].mode = 0;
      }
      launch_cheb_pair<T>(g, st[0], st[1], mu_of<T>(ctx), dinv_of<T>(ctx), ctx->part_rr, ctx->npart, ctx->sc,
                          ctx->part_rho, ctx->s);
      break;
    }
    default:
      cg_iteration<T>(ctx, w.x, ctx->have_amg ? EQ_PC_AMG : EQ_PC_JACOBI);
      break;
  }
}

extern "C" int eq_time_kernel(eq_ctx* ctx, int which, int prec, int reps, double* avg_ms) {
  if (!ctx || !avg_ms || reps < 1) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (!ctx->have_material) return fail(ctx, EQ_ENOTREADY, "eq_set_material first");
  if (which == EQ_TIME_VCYCLE && !ctx->have_amg) return fail(ctx, EQ_ENOTREADY, "eq_amg_setup first");
  if ((which == EQ_TIME_SWEEP || which == EQ_TIME_SWEEP_POST || which == EQ_TIME_RESTRICT) &&
      (!ctx->have_amg || !ctx->pmg || !ctx->l0_cheb))
    return fail(ctx, EQ_ENOTREADY, "a p-multigrid hierarchy with Chebyshev smoothing first");
  if (which == EQ_TIME_NEWMARK && (!ctx->have_state || !ctx->have_params))
    return fail(ctx, EQ_ENOTREADY, "eq_state_init first");
  hipStream_t s = ctx->s;
  if (which == EQ_TIME_COPY) {
    const int64_t n = (int64_t)512 << 17;  // 512 MiB of doubles
    double *src = nullptr, *dst = nullptr;
    if (hipMalloc(&src, n * sizeof(double)) != hipSuccess) return fail(ctx, EQ_NOMEM, "copy buffers");
    if (hipMalloc(&dst, n * sizeof(double)) != hipSuccess) {
      (void)hipFree(src);
      return fail(ctx, EQ_NOMEM, "copy buffers");
    }
    launch_fill<double>(n, src, 1.0, s);
    launch_convert<double, double>(n, src, dst, s);  // warm-up
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0;
    bool ok = hipEventCreate(&e0) == hipSuccess;
    ok = ok && hipEventCreate(&e1) == hipSuccess;
    ok = ok && hipEventRecord(e0, s) == hipSuccess;
    for (int k = 0; ok && k < reps; k++) launch_convert<double, double>(n, src, dst, s);
    ok = ok && hipEventRecord(e1, s) == hipSuccess;
    ok = ok && hipEventSynchronize(e1) == hipSuccess;
    ok = ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (!ok) return fail(ctx, EQ_HIP_ERR, "copy timing");
    *avg_ms = (double)ms / reps;
    return sync_check(ctx);
  }
  // well-defined inputs: unit vectors, a CG state that never reports convergence
  if (prec == EQ_PREC_FP32) {
    launch_fill<float>(ctx->N, ctx->w32.x, 1.0f, s); launch_fill<float>(ctx->N, ctx->w32.r, 1.0f, s);
    launch_fill<float>(ctx->N, ctx->w32.c, 1.0f, s); launch_fill<float>(ctx->N, ctx->w32.u[0], 0.0f, s);
    launch_fill<double>(ctx->N, ctx->w64.r, 1.0, s);   // the FP64 residual of the F32PC post-sweep
  } else {
    launch_fill<double>(ctx->N, ctx->w64.x, 1.0, s); launch_fill<double>(ctx->N, ctx->w64.r, 1.0, s);
    launch_fill<double>(ctx->N, ctx->w64.c, 1.0, s); launch_fill<double>(ctx->N, ctx->w64.u[0], 0.0, s);
  }
  for (int k = 0; k < 3; k++) {
    if (prec == EQ_PREC_FP32) {
      launch_fill<float>(ctx->N, ctx->w32.chd[k], 1.0f, s);
      launch_fill<float>(ctx->N, ctx->w32.chr[k], 1.0f, s);
    } else {
      launch_fill<double>(ctx->N, ctx->w64.chd[k], 1.0, s);
      launch_fill<double>(ctx->N, ctx->w64.chr[k], 1.0, s);
    }
  }
  if (prec == EQ_PREC_FP32) launch_fill<float>(ctx->N, ctx->w32.chx, 1.0f, s);
  else launch_fill<double>(ctx->N, ctx->w64.chx, 1.0, s);
  CgScal h{};
  h.rho_prev = 1.0; h.reltol = -1.0; h.maxiter = INT_MAX;
  HIPCHK(ctx, hipMemcpyAsync(ctx->sc, &h, sizeof(CgScal), hipMemcpyHostToDevice, s));
  launch_fill<double>(ctx->npart, ctx->part_rr, 1.0, s);
  launch_fill<double>(ctx->npart, ctx->part_rho, 1.0, s);
  launch_fill<double>(ctx->npart, ctx->part_uq, 1.0, s);
  auto body = [&]() {
    if (which == EQ_TIME_NEWMARK) {
      // the same step from the same state each time, outputs in scratch
      double* w = ctx->api_nodal;
      launch_newmark(ctx->g, ctx->fp, ctx->fs_api, ctx->mu, ctx->M, w, ctx->v, ctx->a, w + ctx->N, w + 2 * ctx->N,
                     w + 3 * ctx->N, ctx->dt0, 0.5 * ctx->dt0, 0.5 * ctx->dt0 * ctx->dt0, ctx->flags, NewmarkRecord{},
                     s);
    } else if (prec == EQ_PREC_FP32) {
      time_body<float>(ctx, which);
    } else {
      time_body<double>(ctx, which);
    }
  };
  if (which == EQ_TIME_NEWMARK) {
    ctx->flags_dirty = true;
    // the fault routine works on the API scratch copy of the fault state
    HIPCHK(ctx, hipMemcpyAsync(ctx->fs_api.psi, ctx->fs.psi, sizeof(double) * 11 * ctx->NX,
                               hipMemcpyDeviceToDevice, s));
    if (!ctx->api_nodal) {
      ctx->api_nodal = ctx->alloc<double>(6 * ctx->N);
      NEED(ctx->api_nodal);
    }
    launch_newmark_predict(ctx->N, ctx->d, ctx->v, ctx->a, ctx->dt0, 0.5 * ctx->dt0 * ctx->dt0, ctx->api_nodal, s);
  }
  body();  // warm-up
  hipEvent_t e0, e1;
  HIPCHK(ctx, hipEventCreate(&e0));
  HIPCHK(ctx, hipEventCreate(&e1));
  HIPCHK(ctx, hipEventRecord(e0, s));
  for (int k = 0; k < reps; k++) body();
  HIPCHK(ctx, hipEventRecord(e1, s));
  HIPCHK(ctx, hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = (double)ms / reps;
  return sync_check(ctx);
}

// ===========================================================================
// learned initial guess
// ===========================================================================
#define LEARN_READY(ctx)                                                                  \
  do {                                                                                    \
    if (!(ctx)->learn.S) return fail((ctx), EQ_ENOTREADY, "eq_learn_init first");         \
  } while (0)

extern "C" int eq_learn_init(eq_ctx* ctx, int64_t capacity, int channels, int hidden, int nbasis, uint64_t seed) {
  if (!ctx || capacity < 1 || channels < 1 || hidden < 1) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  if (!ctx->have_params || !ctx->have_material) return fail(ctx, EQ_ENOTREADY, "material and params first");
  std::string err;
  int rc = learn_init(&ctx->learn, ctx->nI, ctx->NX, ctx->NX, (int)capacity, channels, hidden, nbasis, seed, &err);
  if (rc) return fail(ctx, rc == 7 ? EQ_NOMEM : EQ_EINVAL, err);
  if (!ctx->gp) {
    ctx->gp = ctx->alloc<double>(ctx->N);
    ctx->gq = ctx->alloc<double>(ctx->N);
    NEED(ctx->gp); NEED(ctx->gq);
  }
  ctx->learn_on = false;
  ctx->learn_record_on = 0;
  return EQ_OK;
}

extern "C" int eq_learn_record(eq_ctx* ctx, int on) {
  if (!ctx) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  LEARN_READY(ctx);
  ctx->learn_record_on = on ? 1 : 0;
  return EQ_OK;
}

extern "C" int eq_learn_info(const eq_ctx* ctx, int64_t* nsnap, int64_t* nparams, int* trained, int64_t* nlearned) {
  if (!ctx) return EQ_EINVAL;
  if (nsnap) *nsnap = ctx->learn.nsnap;
  if (nparams) *nparams = ctx->learn.d.nparams;
  if (trained) *trained = ctx->learn_on ? 1 : 0;
  if (nlearned) *nlearned = ctx->n_learned_solves;
  return EQ_OK;
}

extern "C" int eq_learn_gram(eq_ctx* ctx, double* G) {
  if (!ctx || !G) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  LEARN_READY(ctx);
  std::string err;
  if (int rc = learn_gram(&ctx->learn, G, ctx->s, &err)) return fail(ctx, rc == 7 ? EQ_NOMEM : EQ_HIP_ERR, err);
  return EQ_OK;
}

// G = Phi^T K_II Phi (one K_II product per basis vector), symmetrised, then
// inverted by Cholesky on the host for the Galerkin start.
static int galerkin_setup(eq_ctx* ctx) {
  LearnState& L = ctx->learn;
  const int NB = L.d.NB;
  std::vector<double> G((size_t)NB * NB), Ginv((size_t)NB * NB, 0.0);
  for (int k = 0; k < NB; k++) {
    learn_basis_column(&L, k, ctx->gp, ctx->s);
    launch_ku<double>(ctx->g, ctx->gp, ctx->mu, ctx->gq, 1, 0, ctx->s);
    learn_project(&L, ctx->gq, 0, ctx->s);
    if (learn_get_proj(&L, G.data() + (size_t)k * NB, ctx->s)) return fail(ctx, EQ_HIP_ERR, "Galerkin set-up");
  }
  for (int i = 0; i < NB; i++)
    for (int j = 0; j < i; j++) {
      const double a = 0.5 * (G[(size_t)i * NB + j] + G[(size_t)j * NB + i]);
      G[(size_t)i * NB + j] = G[(size_t)j * NB + i] = a;
    }
  // Cholesky G = C C^T (lower C in place), then Ginv column by column
  std::vector<double> C = G;
  for (int j = 0; j < NB; j++) {
    double d = C[(size_t)j * NB + j];
    for (int k = 0; k < j; k++) d -= C[(size_t)j * NB + k] * C[(size_t)j * NB + k];
    if (!(d > 0.0)) return fail(ctx, EQ_EINVAL, "Phi^T K_II Phi is not positive definite");
    d = std::sqrt(d);
    C[(size_t)j * NB + j] = d;
    for (int i = j + 1; i < NB; i++) {
      double a = C[(size_t)i * NB + j];
      for (int k = 0; k < j; k++) a -= C[(size_t)i * NB + k] * C[(size_t)j * NB + k];
      C[(size_t)i * NB + j] = a / d;
    }
  }
  std::vector<double> y(NB);
  for (int col = 0; col < NB; col++) {
    for (int i = 0; i < NB; i++) {   // C y = e_col
      double a = (i == col) ? 1.0 : 0.0;
      for (int k = 0; k < i; k++) a -= C[(size_t)i * NB + k] * y[k];
      y[i] = a / C[(size_t)i * NB + i];
    }
    for (int i = NB - 1; i >= 0; i--) {   // C^T z = y
      double a = y[i];
      for (int k = i + 1; k < NB; k++) a -= C[(size_t)k * NB + i] * Ginv[(size_t)k * NB + col];
      Ginv[(size_t)i * NB + col] = a / C[(size_t)i * NB + i];
    }
  }
  std::string err;
  if (learn_set_ginv(&L, Ginv.data(), ctx->s, &err)) return fail(ctx, EQ_HIP_ERR, err);
  return EQ_OK;
}

extern "C" int eq_learn_set_basis(eq_ctx* ctx, const double* V, const double* sigma, const int32_t* use) {
  if (!ctx || !V || !sigma) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  LEARN_READY(ctx);
  std::string err;
  if (int rc = learn_set_basis(&ctx->learn, V, sigma, use, ctx->s, &err))
    return fail(ctx, rc == 7 ? EQ_NOMEM : EQ_EINVAL, err);
  return galerkin_setup(ctx);
}

extern "C" int eq_learn_galerkin(eq_ctx* ctx, const double* x, const double* b, double* x_out) {
  if (!ctx || !x || !b || !x_out) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  LEARN_READY(ctx);
  if (!ctx->learn.gal_ready) return fail(ctx, EQ_ENOTREADY, "eq_learn_set_basis first");
  double* dx = ctx->gp;
  double* db = ctx->gq;
  if (int rc = upload_ni(ctx, x, dx)) return rc;
  if (int rc = upload_ni(ctx, b, db)) return rc;
  galerkin_correction(ctx, dx, db);
  return download_ni(ctx, dx, x_out);
}

extern "C" int eq_learn_norms(eq_ctx* ctx, double* out) {
  if (!ctx || !out) return EQ_EINVAL;
  LEARN_READY(ctx);
  if (learn_norms(&ctx->learn, out, ctx->s)) return fail(ctx, EQ_HIP_ERR, "norm copy failed");
  return EQ_OK;
}

extern "C" int eq_learn_train(eq_ctx* ctx, int epochs, int batch, double lr, uint64_t seed, double* hist) {
  if (!ctx || epochs < 0 || batch < 1) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  LEARN_READY(ctx);
  std::string err;
  if (int rc = learn_train(&ctx->learn, epochs, batch, lr, seed, hist, ctx->s, &err))
    return fail(ctx, rc == 7 ? EQ_NOMEM : EQ_EINVAL, err);
  ctx->learn_on = true;
  return EQ_OK;
}

extern "C" int eq_learn_loss_grad(eq_ctx* ctx, const float* params, int nsamp, const int32_t* idx, int use_bf16,
                                  double* loss, float* grad) {
  if (!ctx || !idx || !loss) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  LEARN_READY(ctx);
  std::string err;
  if (int rc = learn_loss_grad(&ctx->learn, params, nsamp, idx, use_bf16, loss, grad, ctx->s, &err))
    return fail(ctx, rc == 7 ? EQ_NOMEM : EQ_EINVAL, err);
  return EQ_OK;
}

extern "C" int eq_learn_get(eq_ctx* ctx, float* params, float* Phi, float* X, float* Y) {
  if (!ctx) return EQ_EINVAL;
  LEARN_READY(ctx);
  LearnState& L = ctx->learn;
  if (params) HIPCHK(ctx, hipMemcpyAsync(params, L.w32, sizeof(float) * L.d.nparams, hipMemcpyDeviceToHost, ctx->s));
  if (Phi) HIPCHK(ctx, hipMemcpyAsync(Phi, L.Phi, sizeof(float) * L.nI * L.d.NB, hipMemcpyDeviceToHost, ctx->s));
  if (X) HIPCHK(ctx, hipMemcpyAsync(X, L.X, sizeof(float) * (int64_t)L.nsnap * L.d.L0, hipMemcpyDeviceToHost, ctx->s));
  if (Y) HIPCHK(ctx, hipMemcpyAsync(Y, L.Y, sizeof(float) * (int64_t)L.nsnap * L.d.NB, hipMemcpyDeviceToHost, ctx->s));
  return sync_check(ctx);
}

extern "C" int eq_learn_snapshots(eq_ctx* ctx, float* S) {
  if (!ctx || !S) return EQ_EINVAL;
  LEARN_READY(ctx);
  LearnState& L = ctx->learn;
  HIPCHK(ctx, hipMemcpyAsync(S, L.S, sizeof(float) * L.nI * (int64_t)L.nsnap, hipMemcpyDeviceToHost, ctx->s));
  return sync_check(ctx);
}

extern "C" int eq_learn_predict(eq_ctx* ctx, const double* dF, double* coef) {
  if (!ctx || !dF || !coef) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  LEARN_READY(ctx);
  std::string err;
  if (learn_predict(&ctx->learn, dF, coef, 0, ctx->s, &err) != 0) return fail(ctx, EQ_HIP_ERR, err);
  return EQ_OK;
}

extern "C" int eq_learn_set_linear(eq_ctx* ctx, const float* A, const float* b) {
  if (!ctx || !A || !b) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  LEARN_READY(ctx);
  if (!ctx->learn.nbasis_ready) return fail(ctx, EQ_ENOTREADY, "linear predictor: no POD basis (eq_learn_set_basis)");
  std::string err;
  if (learn_set_linear(&ctx->learn, A, b, ctx->s, &err) != 0) return fail(ctx, EQ_NOMEM, err);
  return EQ_OK;
}

extern "C" int eq_learn_predict_linear(eq_ctx* ctx, const double* dF, double* coef) {
  if (!ctx || !dF || !coef) return EQ_EINVAL;
  ctx->qs_r_valid = false;
  LEARN_READY(ctx);
  if (!ctx->learn.lin_ready) return fail(ctx, EQ_ENOTREADY, "eq_learn_set_linear first");
  std::string err;
  if (learn_predict(&ctx->learn, dF, coef, 1, ctx->s, &err) != 0) return fail(ctx, EQ_HIP_ERR, err);
  return EQ_OK;
}

extern "C" int eq_launch_timing(eq_ctx* ctx, int on) {
  if (!ctx) return EQ_EINVAL;
  auto& L = ctx->lt;
  if (on) {
    if (L.ev.empty()) {
      L.ev.resize(2 * L.kSlots, nullptr);
      for (hipEvent_t& e : L.ev) HIPCHK(ctx, hipEventCreate(&e));
    }
    L.used = 0;
    for (int k = 0; k < EQ_TK_COUNT; k++) { L.count[k] = 0; L.ms[k] = 0.0; }
    L.on = true;
    return EQ_OK;
  }
  const int rc = lt_harvest(ctx);
  L.on = false;
  return rc;
}

extern "C" int eq_launch_timing_get(eq_ctx* ctx, int64_t* counts, double* total_ms) {
  if (!ctx || !counts || !total_ms) return EQ_EINVAL;
  if (int rc = lt_harvest(ctx)) return rc;
  for (int k = 0; k < EQ_TK_COUNT; k++) {
    counts[k] = ctx->lt.count[k];
    total_ms[k] = ctx->lt.ms[k];
  }
  return EQ_OK;
}

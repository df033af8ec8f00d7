// gfx950 kernels of libeqcycle_hip.
//
// Matrix-free stiffness.  For nodal rigidity mu (W_e[i,j] = w_i w_j mu, the
// form mat_trap / mat_gauss build, src/MaterialProperties.jl:77-128) the
// element operator of K_element (src/Kassemble.jl:31-63) summed over the mesh
// splits into 1-D operators along grid lines:
//   (K u)(I,J) = sum_{segments s on row J containing I} sum_p D[p][i] gx_p (D u)_p
//              + sum_{segments t on column I containing J} sum_q D[q][j] gy_q (D u)_q
// with gx = c1 * wline(J) * w_p * mu, gy = c2 * wline(I) * w_q * mu, where
// wline(L) sums the GLL weight of line L over the elements containing it.
// A workgroup owns a 64 x 16 node tile, stages a (64+5) x (16+5) window of
// the input and of mu in LDS (4-node halo below / left, 1 above / right) and
// evaluates every owned node by gather, so no atomics are needed and results
// are deterministic.  The same tile kernel carries fused epilogues: CG
// direction update + A u + <u, A u>, AMG level-0 Jacobi smoothing + residual,
// and the whole Newmark step with the fault Newton search.
//
// Reductions: every reducing kernel writes one FP64 partial per workgroup; the
// consumer kernel re-reduces the partials in a fixed order in each workgroup,
// so scalars never round-trip through the host and results are reproducible.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cmath>
#include <type_traits>

#include "eq_kernels.h"

namespace eq {

// --------------------------------------------------------------------------
// reductions (blockDim == BLOCK == 256, wave64)
// --------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// blockDim.x is a multiple of 64, at most 1024
__device__ double block_sum(double v) {
  __shared__ double sh[16];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  const int nw = (int)(blockDim.x >> 6);
  double r;
  if (nw == 4) {
    r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  } else {
    r = sh[0];
    for (int k = 1; k < nw; k++) r += sh[k];
  }
  return r;
}

__device__ double reduce_parts(const double* p, int n) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
  return block_sum(s);
}

// reduce_parts in two halves (the same sum, bit for bit), so that a kernel
// can issue its own loads between the partial loads and the block sum: the
// barriers of block_sum wait for LDS only, so those loads stay in flight.
__device__ __forceinline__ double parts_load(const double* p, int n) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
  return s;
}
__device__ __forceinline__ double parts_total(double s) { return block_sum(s); }

// --------------------------------------------------------------------------
// geometry helpers
// --------------------------------------------------------------------------
__device__ __forceinline__ double wline(const Geom& g, int L, int nl) {
  const int r = L & 3;
  if (r != 0) return g.w[r];
  return (L == 0 || L == nl - 1) ? g.w[0] : g.w[0] + g.w[4];
}

// One 5-node segment: all five outputs r[i] = sum_p D[p][i] w_p m_p sum_k D[p][k] u_k,
// then select output li.  u and m are strided views (stride ST).
template <typename T, int ST>
__device__ __forceinline__ T seg_eval(const Geom& g, const T* u, const T* m, int li) {
  T t[5];
#pragma unroll
  for (int p = 0; p < 5; p++) {
    T s = 0;
#pragma unroll
    for (int k = 0; k < 5; k++) s += (T)g.D[p * 5 + k] * u[k * ST];
    t[p] = s * ((T)g.w[p] * m[p * ST]);
  }
  T r0 = 0, r1 = 0, r2 = 0, r3 = 0, r4 = 0;
#pragma unroll
  for (int p = 0; p < 5; p++) {
    r0 += (T)g.D[p * 5 + 0] * t[p];
    r1 += (T)g.D[p * 5 + 1] * t[p];
    r2 += (T)g.D[p * 5 + 2] * t[p];
    r3 += (T)g.D[p * 5 + 3] * t[p];
    r4 += (T)g.D[p * 5 + 4] * t[p];
  }
  T r = r0;
  r = (li == 1) ? r1 : r;
  r = (li == 2) ? r2 : r;
  r = (li == 3) ? r3 : r;
  r = (li == 4) ? r4 : r;
  return r;
}

// (K u) at node (I, J) from a window with row stride WS whose (lc, lr)
// entry is node (I, J).
template <typename T, int WS>
__device__ __forceinline__ T ku_window(const Geom& g, const T* su, const T* sm, int lc, int lr, int I, int J) {
  T ax = 0, ay = 0;
  {
    const int i = I & 3, s = I >> 2;
    const int row = lr * WS;
    if (i != 0 || s < g.nelx)   // segment s, local index i
      if (s < g.nelx) ax += seg_eval<T, 1>(g, su + row + lc - i, sm + row + lc - i, i);
    if (i == 0 && s > 0)        // segment s-1, local index 4
      ax += seg_eval<T, 1>(g, su + row + lc - 4, sm + row + lc - 4, 4);
  }
  {
    const int j = J & 3, t = J >> 2;
    if (t < g.nely) ay += seg_eval<T, WS>(g, su + (lr - j) * WS + lc, sm + (lr - j) * WS + lc, j);
    if (j == 0 && t > 0) ay += seg_eval<T, WS>(g, su + (lr - 4) * WS + lc, sm + (lr - 4) * WS + lc, 4);
  }
  return (T)(g.c1 * wline(g, J, g.NY)) * ax + (T)(g.c2 * wline(g, I, g.NX)) * ay;
}

// Same evaluation straight from global memory (used for single fault rows).
template <typename T>
__device__ T ku_global(const Geom& g, const T* u, const T* mu, int I, int J) {
  const int64_t n = (int64_t)J * g.NX + I;
  T ax = 0, ay = 0;
  {
    const int i = I & 3, s = I >> 2;
    if (s < g.nelx) ax += seg_eval<T, 1>(g, u + n - i, mu + n - i, i);
    if (i == 0 && s > 0) ax += seg_eval<T, 1>(g, u + n - 4, mu + n - 4, 4);
  }
  {
    const int j = J & 3, t = J >> 2;
    const int NX = g.NX;
    if (t < g.nely) {
      const T* up = u + n - (int64_t)j * NX;
      const T* mp = mu + n - (int64_t)j * NX;
      T tt[5];
#pragma unroll
      for (int q = 0; q < 5; q++) {
        T s = 0;
#pragma unroll
        for (int l = 0; l < 5; l++) s += (T)g.D[q * 5 + l] * up[(int64_t)l * NX];
        tt[q] = s * ((T)g.w[q] * mp[(int64_t)q * NX]);
      }
      T r = 0;
      for (int q = 0; q < 5; q++) r += (T)g.D[q * 5 + j] * tt[q];
      ay += r;
    }
    if (j == 0 && t > 0) {
      const T* up = u + n - 4 * (int64_t)NX;
      const T* mp = mu + n - 4 * (int64_t)NX;
      T tt[5];
#pragma unroll
      for (int q = 0; q < 5; q++) {
        T s = 0;
#pragma unroll
        for (int l = 0; l < 5; l++) s += (T)g.D[q * 5 + l] * up[(int64_t)l * NX];
        tt[q] = s * ((T)g.w[q] * mp[(int64_t)q * NX]);
      }
      T r = 0;
      for (int q = 0; q < 5; q++) r += (T)g.D[q * 5 + 4] * tt[q];
      ay += r;
    }
  }
  return (T)(g.c1 * wline(g, J, g.NY)) * ax + (T)(g.c2 * wline(g, I, g.NX)) * ay;
}

// --------------------------------------------------------------------------
// generic tile kernel
// --------------------------------------------------------------------------
constexpr int WIN_I = TILE_I + 5;
constexpr int WIN_J = TILE_J + 5;

struct PreNone {};

// Segment outputs o[0..4] of one 5-node segment (strided view, stride ST).
template <typename T, int ST>
__device__ __forceinline__ void seg_all(const Geom& g, const T* u, const T* m, T o[5]) {
  T t[5];
#pragma unroll
  for (int p = 0; p < 5; p++) {
    T s = 0;
#pragma unroll
    for (int k = 0; k < 5; k++) s += (T)g.D[p * 5 + k] * u[k * ST];
    t[p] = s * ((T)g.w[p] * m[p * ST]);
  }
#pragma unroll
  for (int i = 0; i < 5; i++) {
    T r = 0;
#pragma unroll
    for (int p = 0; p < 5; p++) r += (T)g.D[p * 5 + i] * t[p];
    o[i] = r;
  }
}

// Output 4 only (the shared end node of the previous segment).
template <typename T, int ST>
__device__ __forceinline__ T seg_last(const Geom& g, const T* u, const T* m) {
  T r = 0;
#pragma unroll
  for (int p = 0; p < 5; p++) {
    T s = 0;
#pragma unroll
    for (int k = 0; k < 5; k++) s += (T)g.D[p * 5 + k] * u[k * ST];
    r += (T)g.D[p * 5 + 4] * (s * ((T)g.w[p] * m[p * ST]));
  }
  return r;
}

// XCD-aware tile order.  Consecutive workgroups go to the 8 XCDs in turn;
// workgroup b takes tile xcd_tile(b), which gives each XCD a contiguous run of
// tiles in row-major order, so the halos of neighbouring tiles meet in one L2.
// A bijection of 0 .. ntiles-1 (EQ_XCD_ORDER=0: the identity).
#ifndef EQ_XCD_ORDER
#define EQ_XCD_ORDER 1
#endif
__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
  if (!EQ_XCD_ORDER) return b;
  constexpr int kXcd = 8;
  const int base = ntiles / kXcd, rem = ntiles % kXcd;
  const int x = b % kXcd, k = b / kXcd;
  return x * base + min(x, rem) + k;
}

// Outputs i0 .. i0+NPT-1 of the 5-node segment at u, m (stride ST; has_seg:
// the segment exists), plus, when i0 == 0 and has_prev, the previous
// segment's output 4 added to the first (the shared node).  Each output is
// formed exactly as seg_all forms it.
template <typename T, int ST, int NPT>
__device__ __forceinline__ void seg_part(const Geom& g, const T* u, const T* m, int i0, bool has_seg, bool has_prev,
                                         T o[NPT]) {
  T a[5] = {0, 0, 0, 0, 0};
  if (has_seg) seg_all<T, ST>(g, u, m, a);
  if (NPT == 4) {
#pragma unroll
    for (int j = 0; j < 4; j++) o[j] = a[j];
  } else if (NPT == 2) {
    o[0] = i0 ? a[2] : a[0];
    o[1] = i0 ? a[3] : a[1];
  } else {
    T r = a[0];
    r = (i0 == 1) ? a[1] : r;
    r = (i0 == 2) ? a[2] : r;
    r = (i0 == 3) ? a[3] : r;
    o[0] = r;
  }
  if (i0 == 0 && has_prev) o[0] += seg_last<T, ST>(g, u - 4 * ST, m - 4 * ST);
}

// Tile kernel.  Phase 1 issues every global load of the block at once: the
// (64+5) x (16+5) window of the input vector (op.load) and of mu, held in
// registers, and the epilogue operands of the thread's TILE_NPT owned nodes
// (op.fetch); only then are the window values stored to LDS.  Phase 2: each
// thread takes TILE_NPT consecutive nodes of one row and accumulates their
// x-direction terms into an LDS tile.  Phase 3: each thread takes TILE_NPT
// consecutive nodes of one column, adds the y-direction terms and runs the
// epilogue.  With 4 nodes per thread every 5-node segment is evaluated once
// per direction (plus the shared end node of the previous segment); with
// fewer, a segment is evaluated by each of its threads, which buys more
// waves per tile.
static_assert(TILE_J * (TILE_I / TILE_NPT) == TBLOCK, "phase 2: one row strip per thread");
static_assert(TILE_I * (TILE_J / TILE_NPT) == TBLOCK, "phase 3: one column strip per thread");
static_assert(TILE_NPT == 1 || TILE_NPT == 2 || TILE_NPT == 4, "strips lie inside one segment");
static_assert(TBLOCK % 64 == 0 && TBLOCK <= 1024, "whole waves");
constexpr int WIN_LOADS = (WIN_J * WIN_I + TBLOCK - 1) / TBLOCK;

// Op::kLatePrepare (when declared true): the window and epilogue loads are
// issued before op.prepare, whose own reads then overlap them; load and fetch
// must not depend on Pre.  Otherwise prepare runs first and a block it stops
// issues no loads.
template <class Op, class = void>
struct LatePrepare : std::false_type {};
template <class Op>
struct LatePrepare<Op, std::void_t<decltype(Op::kLatePrepare)>> : std::bool_constant<Op::kLatePrepare> {};

// Op::kWindowAdd (when declared true): after the window loads, op.window_add
// adds a term of its own to each loaded value (before the values reach LDS).
template <class Op, class = void>
struct WindowAdd : std::false_type {};
template <class Op>
struct WindowAdd<Op, std::void_t<decltype(Op::kWindowAdd)>> : std::bool_constant<Op::kWindowAdd> {};
// Op::kWindowPost (when declared true): after a late op.prepare, op.window_post
// (pre, ...) may still change the loaded window values (before LDS).
template <class Op, class = void>
struct WindowPost : std::false_type {};
template <class Op>
struct WindowPost<Op, std::void_t<decltype(Op::kWindowPost)>> : std::bool_constant<Op::kWindowPost> {};

template <typename T, class Op>
__device__ __forceinline__ void tile_body(const Geom& g, const Op& op, int bx, int by, int ntx, int nty) {
  constexpr int NPT = TILE_NPT;
  constexpr bool late = LatePrepare<Op>::value;
  __shared__ T su[WIN_J * WIN_I];
  __shared__ T sm[WIN_J * WIN_I];
  __shared__ T sx[TILE_J * TILE_I];
  typename Op::Pre pre{};
  if (!late && !op.prepare(pre)) return;   // block-uniform
  const int I0 = bx * TILE_I, J0 = by * TILE_J;
  // phase-3 ownership of this thread: column c, rows Js .. Js+NPT-1
  const int c3 = threadIdx.x % TILE_I, q3 = threadIdx.x / TILE_I;
  const int I3 = I0 + c3, Js3 = J0 + NPT * q3;
  typename Op::Ext ext[NPT];
#pragma unroll
  for (int j = 0; j < NPT; j++)
    if (I3 < g.NX && Js3 + j < g.NY) op.fetch(pre, (int64_t)(Js3 + j) * g.NX + I3, I3, Js3 + j, ext[j]);
  {
    T ur[WIN_LOADS], mr[WIN_LOADS];
#pragma unroll
    for (int t = 0; t < WIN_LOADS; t++) {
      const int k = threadIdx.x + t * TBLOCK;
      const int r = k / WIN_I, c = k - r * WIN_I;
      const int I = I0 - 4 + c, J = J0 - 4 + r;
      ur[t] = 0;
      mr[t] = 0;
      if (k < WIN_J * WIN_I && I >= 0 && I < g.NX && J >= 0 && J < g.NY) {
        const int64_t n = (int64_t)J * g.NX + I;
        ur[t] = op.load(pre, n, I, J);
        mr[t] = op.mu[n];
      }
    }
    if constexpr (WindowAdd<Op>::value) op.window_add(g, I0, J0, ur);
    if (late && !op.prepare(pre)) return;   // block-uniform
    if constexpr (WindowPost<Op>::value) op.window_post(pre, g, I0, J0, ur);
#pragma unroll
    for (int t = 0; t < WIN_LOADS; t++) {
      const int k = threadIdx.x + t * TBLOCK;
      if (k < WIN_J * WIN_I) {
        su[k] = ur[t];
        sm[k] = mr[t];
      }
    }
  }
  __syncthreads();
  {
    // x-direction: thread -> (row r, strip q of NPT nodes); TILE_I / NPT strips per row
    constexpr int NSTRIP = TILE_I / NPT;
    const int r = threadIdx.x / NSTRIP, q = threadIdx.x - r * NSTRIP;
    const int J = J0 + r, Is = I0 + NPT * q;
    if (J < g.NY && Is < g.NX) {
      const int sg = Is >> 2, i0 = Is & 3;
      const int lr = r + 4, lc = NPT * q + 4 - i0;   // window column of the segment start
      T o[NPT];
      seg_part<T, 1, NPT>(g, su + lr * WIN_I + lc, sm + lr * WIN_I + lc, i0, sg < g.nelx, sg > 0, o);
      const T cx = (T)(g.c1 * wline(g, J, g.NY));
#pragma unroll
      for (int i = 0; i < NPT; i++) sx[r * TILE_I + NPT * q + i] = cx * o[i];
    }
  }
  __syncthreads();
  double acc0 = 0.0, acc1 = 0.0;
  if (I3 < g.NX && Js3 < g.NY) {
    // y-direction + epilogue
    const int t = Js3 >> 2, j0 = Js3 & 3;
    const int lc = c3 + 4, lr = NPT * q3 + 4;       // window row of Js3
    const int ls = lr - j0;                          // window row of the segment start
    T o[NPT];
    seg_part<T, WIN_I, NPT>(g, su + ls * WIN_I + lc, sm + ls * WIN_I + lc, j0, t < g.nely, t > 0, o);
    const T cy = (T)(g.c2 * wline(g, I3, g.NX));
#pragma unroll
    for (int j = 0; j < NPT; j++) {
      const int J = Js3 + j;
      if (J < g.NY) {
        const T ku = sx[(NPT * q3 + j) * TILE_I + c3] + cy * o[j];
        op.epilogue(pre, (int64_t)J * g.NX + I3, I3, J, ku, su[(lr + j) * WIN_I + lc], ext[j], acc0, acc1);
      }
    }
  }
  op.block_end(acc0, acc1);
  if (Op::kParts >= 1) {
    const double s = block_sum(acc0);
    if (threadIdx.x == 0) op.part0[by * ntx + bx] = s;
  }
  if (Op::kParts >= 2) {
    const double s = block_sum(acc1);
    if (threadIdx.x == 0) op.part1[by * ntx + bx] = s;
  }
  if (bx == 0 && by == 0) op.zero_tail(ntx * nty);
}

template <typename T, class Op>
__global__ __launch_bounds__(TBLOCK) void tile_kernel(Geom g, Op op) {
  const int t = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  tile_body<T, Op>(g, op, t % gridDim.x, t / gridDim.x, gridDim.x, gridDim.y);
}

// Epilogue operands with nothing to prefetch.
struct ExtNone {};

// y = sign K u  (interior: fault row of u read as 0, of y written as 0)
template <typename T>
struct OpApply {
  using Pre = PreNone;
  static constexpr int kParts = 0;
  __device__ void zero_tail(int) const {}
  __device__ void block_end(double, double) const {}
  const T* u;
  const T* mu;
  T* y;
  int interior;
  T sign;
  double* part0;
  double* part1;
  __device__ bool prepare(Pre&) const { return true; }
  using Ext = ExtNone;
  __device__ void fetch(const Pre&, int64_t, int, int, Ext&) const {}
  __device__ T load(const Pre&, int64_t n, int, int J) const { return (interior && J == 0) ? T(0) : u[n]; }
  __device__ void epilogue(const Pre&, int64_t n, int, int J, T ku, T, const Ext&, double&, double&) const {
    y[n] = (interior && J == 0) ? T(0) : sign * ku;
  }
};

// r = b - K_II x, partials of r.r and b.b
template <typename T>
struct OpResidual {
  using Pre = PreNone;
  static constexpr int kParts = 2;
  // partial arrays have npart entries; entries past the grid are cleared so that
  // consumers may sum all npart of them
  __device__ void block_end(double, double) const {}
  int npart;
  __device__ void zero_tail(int nb) const {
    for (int k = nb + threadIdx.x; k < npart; k += blockDim.x) {
      part0[k] = 0.0;
      part1[k] = 0.0;
    }
  }
  const T* x;
  const T* mu;
  const T* b;
  T* r;
  double* part0;
  double* part1;
  float* r32;        // optional FP32 copy of r (for an FP32 V-cycle)
  __device__ bool prepare(Pre&) const { return true; }
  struct Ext { T b; };
  __device__ void fetch(const Pre&, int64_t n, int, int J, Ext& e) const { e.b = J == 0 ? T(0) : b[n]; }
  __device__ T load(const Pre&, int64_t n, int, int J) const { return J == 0 ? T(0) : x[n]; }
  __device__ void epilogue(const Pre&, int64_t n, int, int J, T ku, T, const Ext& e, double& a0, double& a1) const {
    const T bv = e.b;
    const T rv = J == 0 ? T(0) : bv - ku;
    r[n] = rv;
    if (r32) r32[n] = (float)rv;
    a0 += (double)rv * (double)rv;
    a1 += (double)bv * (double)bv;
  }
};

// Galerkin start and its residual in one launch: the window loads x = d +
// Phi c off the fault row (the FP32 sum of expand_add16_kernel, in the same
// order, so x is that kernel's bit for bit), then r = b - K_II x with the partials of
// OpResidual; x goes to x_out at every owned node (the fault row as d).  Phi:
// the FP16 basis of the interior (node n = NX + i) in groups of 8 vectors,
// [ceil(nb / 8)][nI][8].  x_out must not alias d: neighbouring tiles still
// load d on their windows.
constexpr int kGalGroups = 8;   // groups of 8 basis vectors: up to 64
// kFull: nb = 8 kNG (no basis index to test: the coefficients load in blocks
// and the sums follow the loads as they arrive; fewer groups hold fewer
// registers)
template <bool kFull, int kNG = kGalGroups>
struct OpGalResidual {
  using Pre = PreNone;
  static constexpr int kParts = 2;
  static constexpr bool kWindowAdd = true;
  __device__ void block_end(double, double) const {}
  int npart;
  __device__ void zero_tail(int nb_) const {
    for (int k = nb_ + threadIdx.x; k < npart; k += blockDim.x) {
      part0[k] = 0.0;
      part1[k] = 0.0;
    }
  }
  const double* d;
  const half8* phi;
  int64_t nI;
  int nb;
  const double* c;
  const double* mu;
  const double* b;
  double* r;
  float* r32;
  double* x_out;
  double* part0;
  double* part1;
  const int* gate;   // or null; with it the launch does nothing unless *gate
  __device__ bool prepare(Pre&) const { return !gate || *gate; }
  struct Ext { double b, d0; };
  __device__ void fetch(const Pre&, int64_t n, int, int J, Ext& e) const {
    e.b = J == 0 ? 0.0 : b[n];
    e.d0 = J == 0 ? d[n] : 0.0;
  }
  __device__ double load(const Pre&, int64_t n, int, int J) const { return J == 0 ? 0.0 : d[n]; }
  template <int NL>
  __device__ void window_add(const Geom& g, int I0, int J0, double (&ur)[NL]) const {
    int64_t idx[NL];
    bool ok[NL];
    float acc[NL];
#pragma unroll
    for (int t = 0; t < NL; t++) {
      const int k = threadIdx.x + t * TBLOCK;
      const int rr = k / WIN_I, cc = k - rr * WIN_I;
      const int I = I0 - 4 + cc, J = J0 - 4 + rr;
      ok[t] = k < WIN_J * WIN_I && I >= 0 && I < g.NX && J >= 1 && J < g.NY;
      idx[t] = ok[t] ? (int64_t)(J - 1) * g.NX + I : 0;   // off the interior: node 0's row, sum discarded
      acc[t] = 0.0f;
    }
    // every group's loads first (up to 8 x NL 16-byte loads in flight), then
    // the sums in basis order; no branch per node, so the NL sums interleave
    constexpr int NG = kNG;
    const int ng = kFull ? NG : (nb + 7) / 8;
    half8 v[NG][NL];
#pragma unroll
    for (int gi = 0; gi < NG; gi++)
#pragma unroll
      for (int t = 0; t < NL; t++) v[gi][t] = gi < ng ? phi[(int64_t)gi * nI + idx[t]] : half8{};
#pragma unroll
    for (int gi = 0; gi < NG; gi++) {
#pragma unroll
      for (int j = 0; j < 8; j++) {
        if (kFull || 8 * gi + j < nb) {
          const float ck = (float)c[8 * gi + j];
#pragma unroll
          for (int t = 0; t < NL; t++) acc[t] = __builtin_fmaf((float)v[gi][t][j], ck, acc[t]);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < NL; t++) ur[t] += ok[t] ? (double)acc[t] : 0.0;
  }
  __device__ void epilogue(const Pre&, int64_t n, int, int J, double ku, double xn, const Ext& e, double& a0,
                           double& a1) const {
    x_out[n] = J == 0 ? e.d0 : xn;
    const double bv = e.b;
    const double rv = J == 0 ? 0.0 : bv - ku;
    r[n] = rv;
    if (r32) r32[n] = (float)rv;
    a0 += rv * rv;
    a1 += bv * bv;
  }
};

// beta = rho / rho_prev; u_new = c + beta u_old (u_new = c in the first
// iteration, u_old is not read); q = K_II u_new; partial u_new.q
template <typename T>
struct OpCgDir {
  struct Pre { T beta; bool first; };
  static constexpr int kParts = 1;
  // the window loads c before beta is known; after the first iteration
  // window_post adds beta u_old (c + beta u_old, the same rounding)
  static constexpr bool kLatePrepare = true;
  static constexpr bool kWindowPost = true;
  __device__ void zero_tail(int) const {}
  __device__ void block_end(double, double) const {}
  const T* c;
  const T* u_old;
  const T* mu;
  T* u_new;
  T* q;
  const double* part_rho;
  int nrho;
  const CgScal* sc;
  double* part0;
  double* part1;
  __device__ bool prepare(Pre& p) const {
    if (sc->done) return false;
    const double rho = reduce_parts(part_rho, nrho);
    p.beta = (T)(rho / sc->rho_prev);
    p.first = sc->iter == 0;
    return true;
  }
  using Ext = ExtNone;
  __device__ void fetch(const Pre&, int64_t, int, int, Ext&) const {}
  __device__ T load(const Pre&, int64_t n, int, int J) const { return J == 0 ? T(0) : c[n]; }
  template <int NL>
  __device__ void window_post(const Pre& p, const Geom& g, int I0, int J0, T (&ur)[NL]) const {
    if (p.first) return;   // block-uniform
#pragma unroll
    for (int t = 0; t < NL; t++) {
      const int k = threadIdx.x + t * TBLOCK;
      const int rr = k / WIN_I, cc = k - rr * WIN_I;
      const int I = I0 - 4 + cc, J = J0 - 4 + rr;
      if (k < WIN_J * WIN_I && I >= 0 && I < g.NX && J >= 1 && J < g.NY)
        ur[t] = ur[t] + p.beta * u_old[(int64_t)J * g.NX + I];
    }
  }
  __device__ void epilogue(const Pre&, int64_t n, int, int J, T ku, T uc, const Ext&, double& a0, double&) const {
    const T qv = J == 0 ? T(0) : ku;
    u_new[n] = uc;
    q[n] = qv;
    a0 += (double)uc * (double)qv;
  }
};

// Convergence test (cg! checks ||r|| <= tol ||b|| before each iteration),
// then x0 = omega D^-1 r and res = r - K_II x0.
template <typename T>
struct OpL0Down {
  using Pre = PreNone;
  static constexpr int kParts = 0;
  __device__ void zero_tail(int) const {}
  __device__ void block_end(double, double) const {}
  const T* r;
  const T* mu;
  const T* dinv;
  T omega;
  T* x0;
  T* res;
  const double* part_rr;
  int nrr;
  CgScal* sc;
  double* part0;
  double* part1;
  __device__ bool prepare(Pre&) const {
    if (sc->done) return false;
    const double rr = reduce_parts(part_rr, nrr);
    const double resid = sqrt(rr);
    const bool fin = (resid <= sc->reltol) || (sc->iter >= sc->maxiter);
    if (threadIdx.x == 0 && blockIdx.x == 0 && blockIdx.y == 0) {
      sc->resid = resid;
      if (fin) sc->done = 1;
    }
    return !fin;
  }
  struct Ext { T r; };
  __device__ void fetch(const Pre&, int64_t n, int, int J, Ext& e) const { e.r = J == 0 ? T(0) : r[n]; }
  __device__ T load(const Pre&, int64_t n, int, int J) const { return J == 0 ? T(0) : omega * dinv[n] * r[n]; }
  __device__ void epilogue(const Pre&, int64_t n, int, int J, T ku, T uc, const Ext& e, double&, double&) const {
    x0[n] = uc;
    res[n] = J == 0 ? T(0) : e.r - ku;
  }
};

// c = x0 + omega D^-1 (r - K_II x0), partial c.r
template <typename T>
struct OpL0Post {
  using Pre = PreNone;
  static constexpr int kParts = 1;
  __device__ void zero_tail(int) const {}
  __device__ void block_end(double, double) const {}
  const T* x0;
  const T* r;
  const T* mu;
  const T* dinv;
  T omega;
  T* c;
  const CgScal* sc;
  double* part0;
  double* part1;
  __device__ bool prepare(Pre&) const { return !sc->done; }
  struct Ext { T r, dinv; };
  __device__ void fetch(const Pre&, int64_t n, int, int J, Ext& e) const {
    e.r = r[n];
    e.dinv = J == 0 ? T(0) : dinv[n];
  }
  __device__ T load(const Pre&, int64_t n, int, int J) const { return J == 0 ? T(0) : x0[n]; }
  __device__ void epilogue(const Pre&, int64_t n, int, int J, T ku, T uc, const Ext& e, double& a0, double&) const {
    const T cv = J == 0 ? T(0) : uc + omega * e.dinv * (e.r - ku);
    c[n] = cv;
    a0 += (double)cv * (double)e.r;
  }
};

// Chebyshev (Jacobi-preconditioned) smoothing step on the matrix-free level:
//   w = window vector (d buffer, s0 dinv b for the first pre-smoothing step,
//       or x for the first post-smoothing step)
//   r = rin - K_II w
//   mode 0: d_new = c1 w + c2 dinv r, x_new = xsrc + d_new; write r, d_new, x_new
//   mode 1: write r only (residual after pre-smoothing; x = w if write_x_center)
//   mode 2: x_new as in mode 0 is the smoother output c; partial <c, rdot>
//           (with xout64/rdot64: c widened to FP64 into xout64, partial
//           <c, rdot64>)
// The first kernel of a V-cycle carries the cg! convergence test.
// Window value of a level-0 Chebyshev step's source at node n = (I, J):
// src_mode 0: src[n]; 1: s0 D^-1 b (the first iterate, formed on the fly);
// 2: src[n] + P0 coarse (bilinear interpolation of the element-vertex
// correction).  The fault row is 0.
template <typename T>
__device__ __forceinline__ T cheb_source(const T* src, const T* b, const T* dinv, T s0, int src_mode,
                                         const T* coarse, int nelx, int nely, const T* gx, int64_t n, int I,
                                         int J) {
  if (J == 0) return T(0);
  if (src_mode == 1) return s0 * dinv[n] * b[n];
  if (src_mode == 2) {
    const int ex = min(I >> 2, nelx - 1), i = I - 4 * ex;
    const int ey = min(J >> 2, nely - 1), j = J - 4 * ey;
    const T ax1 = T(0.5) * (T(1) + gx[i]), ax0 = T(1) - ax1;
    const T ay1 = T(0.5) * (T(1) + gx[j]), ay0 = T(1) - ay1;
    const int nvx = nelx + 1;
    T v = T(0);
    if (ey > 0) {
      const T* c0 = coarse + (int64_t)(ey - 1) * nvx + ex;
      v += ay0 * (ax0 * c0[0] + ax1 * c0[1]);
    }
    const T* c1 = coarse + (int64_t)ey * nvx + ex;
    v += ay1 * (ax0 * c1[0] + ax1 * c1[1]);
    return src[n] + v;
  }
  return src[n];
}

template <typename T>
struct OpCheb {
  using Pre = PreNone;
  static constexpr int kParts = 1;
  __device__ void zero_tail(int) const {}
  __device__ void block_end(double, double) const {}
  const T* src;
  const T* b;
  const T* mu;
  const T* dinv;
  T s0;
  int src_mode;      // 0: src[n], 1: s0 dinv[n] b[n]
  const T* rin;
  T* rout;
  T c1, c2;
  const T* xin;
  int xsrc_center;   // 1: x source is the window value itself
  T* dout;
  T* xout;
  int mode;
  int write_x_center;
  const T* rdot;
  double* xout64;
  const double* rdot64;
  int check;
  const double* part_rr;
  int nrr;
  CgScal* sc;
  double* part0;
  double* part1;
  __device__ bool prepare(Pre&) const {
    if (sc->done) return false;
    if (!check) return true;
    const double rr = reduce_parts(part_rr, nrr);
    const double resid = sqrt(rr);
    const bool fin = (resid <= sc->reltol) || (sc->iter >= sc->maxiter);
    if (threadIdx.x == 0 && blockIdx.x == 0 && blockIdx.y == 0) {
      sc->resid = resid;
      if (fin) sc->done = 1;
    }
    return !fin;
  }
  const T* coarse;    // src_mode 2: window = src + P0 coarse (bilinear vertex interpolation)
  int nelx, nely;
  T gx[5];
  __device__ T load(const Pre&, int64_t n, int I, int J) const {
    return cheb_source(src, b, dinv, s0, src_mode, coarse, nelx, nely, gx, n, I, J);
  }
  struct Ext { T rin, dinv, xin; double rdot; };
  __device__ void fetch(const Pre&, int64_t n, int, int J, Ext& e) const {
    if (J == 0) return;
    e.rin = rin[n];
    if (mode != 1) {
      e.dinv = dinv[n];
      if (!xsrc_center) e.xin = xin[n];
      if (mode == 2) e.rdot = rdot64 ? rdot64[n] : (double)rdot[n];
    }
  }
  __device__ void epilogue(const Pre&, int64_t n, int, int J, T ku, T wc, const Ext& e, double& a0, double&) const {
    if (J == 0) {
      if (mode == 0) { rout[n] = T(0); dout[n] = T(0); xout[n] = T(0); }
      else if (mode == 1) { rout[n] = T(0); if (write_x_center) xout[n] = T(0); }
      else if (xout64) xout64[n] = 0.0;
      else xout[n] = T(0);
      return;
    }
    const T r = e.rin - ku;
    if (mode == 1) {
      rout[n] = r;
      if (write_x_center) xout[n] = wc;
      return;
    }
    const T dn = c1 * wc + c2 * e.dinv * r;
    const T xn = (xsrc_center ? wc : e.xin) + dn;
    if (mode == 0) {
      xout[n] = xn;
      rout[n] = r;
      dout[n] = dn;
    } else {
      if (xout64) xout64[n] = (double)xn;
      else xout[n] = xn;
      a0 += (double)xn * e.rdot;
    }
  }
};

// Learned-guess safeguard: q = K_II p, partials <p, r0> and <p, q>.
struct OpGuessDot {
  using Pre = PreNone;
  static constexpr int kParts = 2;
  __device__ void zero_tail(int) const {}
  __device__ void block_end(double, double) const {}
  const double* p;
  const double* mu;
  const double* r0;
  double* q;
  double* part0;
  double* part1;
  __device__ bool prepare(Pre&) const { return true; }
  struct Ext { double r0; };
  __device__ void fetch(const Pre&, int64_t n, int, int, Ext& e) const { e.r0 = r0[n]; }
  __device__ double load(const Pre&, int64_t n, int, int J) const { return J == 0 ? 0.0 : p[n]; }
  __device__ void epilogue(const Pre&, int64_t n, int, int J, double ku, double pc, const Ext& e, double& a0,
                           double& a1) const {
    const double qv = J == 0 ? 0.0 : ku;
    q[n] = qv;
    a0 += pc * e.r0;
    a1 += pc * qv;
  }
};

// Order of Julia's maximum/findmax over the fault velocities (Vfmax and the
// hypocentre, src/main.jl:348,367): NaN is the largest value, and among equal
// values (NaN included) the lowest index wins.  True when (o, oi) replaces
// the running (m, mi).
template <typename I>
__device__ __forceinline__ bool findmax_before(double o, I oi, double m, I mi) {
  if (isnan(o)) return !isnan(m) || oi < mi;
  if (isnan(m)) return false;
  return o > m || (o == m && oi < mi);
}

// One launch, through hipExtLaunchKernelGGL when events are given
template <typename F, typename... Args>
static void launch_timed(F kernel, dim3 grid, dim3 block, hipStream_t s, LaunchEvents le, Args... args) {
  if (le.start)
    hipExtLaunchKernelGGL(kernel, grid, block, 0, s, le.start, le.stop, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}

// --------------------------------------------------------------------------
// fault physics (FP64, libm-equivalent device exp/log)
// --------------------------------------------------------------------------
// IDS! (src/otherFunctions_serial.jl:5-32)
__device__ double d_ids(double xLf, double Vo, double psi, double dt, double Vf, double cnd, int IDstate) {
  if (IDstate == 1) return psi + dt * ((Vo / xLf) * exp(-psi) - fabs(Vf) / xLf);
  const double VdtL = fabs(Vf) * dt / xLf;
  if (VdtL < cnd) return log(exp(psi - VdtL) + Vo * dt / xLf - 0.5 * Vo * fabs(Vf) * dt * dt / (xLf * xLf));
  return log(exp(psi - VdtL) + (Vo / fabs(Vf)) * (1 - exp(-VdtL)));
}

// IDS2! (src/otherFunctions_serial.jl:35-59)
__device__ double d_ids2(double xLf, double Vo, double psi, double psi1, double dt, double Vf, double Vf1,
                         int IDstate) {
  if (IDstate == 1)
    return psi + 0.5 * dt * ((Vo / xLf) * exp(-psi) - fabs(Vf) / xLf + (Vo / xLf) * exp(-psi1) - fabs(Vf1) / xLf);
  const double VdtL = 0.5 * fabs(Vf1 + Vf) * dt / xLf;
  if (VdtL < 1e-6)
    return log(exp(psi - VdtL) + Vo * dt / xLf - 0.5 * Vo * 0.5 * fabs(Vf1 + Vf) * dt * dt / (xLf * xLf));
  return log(exp(psi - VdtL) + (Vo / (0.5 * fabs(Vf + Vf1))) * (1 - exp(-VdtL)));
}

// NRsearch! (src/NRsearch_serial.jl:47-98): Newton on
//   g(tau) = Z Vfree - Z Vf(tau) + tauo - tau,  Vf = Vo (e^{h+fa} - e^{h-fa}),
// stopping when |delta| <= 0.001 a sigma; returns k (negative on the
// |delta| > 1e10 / k == 1000 exit, which returns the current iterate: Vf at
// the last point evaluated, tau after its step).
//
// The reference iterates in BigFloat (:9,51-54), where e^{h+fa} never
// overflows.  At a fast-slipping node the first step from tau = 0 lands far
// above the root; there, with M = h + |fa|, e^M dwarfs every other term, the
// step is c = -sign(fa) Vo / (fact Vo/(a sigma)) to far below FP64 rounding,
// and |fa| falls by 1 per step, for hundreds of steps.  While
// M >= max(h + 40, ln(A / (1e-20 Z Vo))), A = |Z Vfree + tauo - tau| + 1000|c|
// (a bound of the numerator's other term over any run of steps), the
// neglected terms stay below 1e-20 of the kept ones, so the run of steps down
// to that bound is taken at once (k += n, tau += n c: the same iteration count
// and, to rounding, the same iterate).  The other steps run as written, at M
// below about 65, far from the FP64 range of exp.
// Vo (exp(h + fa) - exp(h - fa)) and vfc (exp(h + fa) + exp(h - fa)) of the
// search.  For |fa| >= 20 the smaller exponential is below 2^-54 of the larger
// (e^-40 = 4.2e-18), so adding or subtracting it cannot change the rounded
// FP64 sum: it is not evaluated, and the results are the two-exponential
// form's bit for bit.
__device__ __forceinline__ void nr_rate(double help, double fa, double Vo, double vfc, double& Vf, double& Vfp) {
  if (fa >= 20.0) {
    const double e = exp(help + fa);
    Vf = Vo * e;
    Vfp = vfc * e;
  } else if (fa <= -20.0) {
    const double e = exp(help - fa);
    Vf = Vo * (-e);
    Vfp = vfc * e;
  } else {
    const double help1 = exp(help + fa), help2 = exp(help - fa);
    Vf = Vo * (help1 - help2);
    Vfp = vfc * (help1 + help2);
  }
}

__device__ int d_nrsearch(double fo, double Vo, double cca, double ccb, double Seff, double tau, double tauo,
                          double psi, double FltZ, double FltVfree, double& Vf_out, double& tau_out) {
  const double Vw = 1e10;
  const double fact = 1. + (Vo / Vw) * exp(-psi);
  const double eps = 0.001 * cca * Seff;
  const double help = -(fo + ccb * psi) / cca;
  // fa = fact tau / (sigma a) as tau times a loop constant: one division
  // fewer in the iteration's dependency chain (the iterates agree with the
  // divided form to rounding)
  const double kfa = fact / (Seff * cca);
  const double vfc = fact * (Vo / (cca * Seff));   // Vf' = vfc (help1 + help2)
  // |c| and a lower bound of the jump threshold Ta below (|FltZ FltVfree +
  // tauo - tau| >= 0), less a margin far above log's rounding: while M is
  // below it, M >= Ta is false and Ta is not evaluated
  const double cabs = Vo / vfc;
  const double Ta_lo = log(1000.0 * cabs / (1e-20 * FltZ * Vo)) - 1e-6;
  int k = 0;
  double delta = INFINITY, Vf = 0.0;
  while (fabs(delta) > eps) {
    const double fa = kfa * tau;
    const double M = fmax(help + fa, help - fa);
    if (isfinite(M) && M >= help + 40.0 && !(M < Ta_lo)) {
      const double c = copysign(cabs, -fa);
      const double Ta = log((fabs(FltZ * FltVfree + tauo - tau) + 1000.0 * fabs(c)) / (1e-20 * FltZ * Vo));
      if (M >= Ta) {   // false for a NaN Ta
        // the steps at M, M - dM, ... down to max(h + 40, Ta); one step if it
        // ends the loop
        double nd = 1.0;
        if (fabs(c) > eps && fabs(c) <= 1e10) nd = floor((M - fmax(help + 40.0, Ta)) / fabs(kfa * c)) + 1.0;
        if (nd > 1000 - k) nd = 1000 - k;
        const int n = (int)nd;
        const double tau_last = tau + (n - 1) * c;
        tau = tau + n * c;
        k += n;
        delta = c;
        if (fabs(delta) > 1e10 || k == 1000) {
          double Vfp;
          nr_rate(help, kfa * tau_last, Vo, vfc, Vf_out, Vfp);
          tau_out = tau;
          return -k;
        }
        continue;
      }
    }
    double Vfprime;
    nr_rate(help, fa, Vo, vfc, Vf, Vfprime);
    delta = (FltZ * FltVfree - FltZ * Vf + tauo - tau) / (1 + FltZ * Vfprime);
    tau = tau + delta;
    k = k + 1;
    if (fabs(delta) > 1e10 || k == 1000) {
      Vf_out = Vf;
      tau_out = tau;
      return -k;
    }
  }
  double Vfp;
  nr_rate(help, kfa * tau, Vo, vfc, Vf_out, Vfp);
  tau_out = tau;
  return k;
}

// FBC! body for one node j (src/NRsearch_serial.jl:17-43)
__device__ void d_fbc_node(const FaultDev& fp, const FaultState& fs, int j, double psi, double Vfj, double Vfree,
                           double dt, int* flags, double& psi2_out, double& tau2_out) {
  double psi1 = d_ids(fp.xLf[j], fp.Vo[j], psi, dt, Vfj, 1e-5, fp.IDstate);
  double Vf1, tau1, Vf2, tau2;
  const int k1 = d_nrsearch(fp.fo[j], fp.Vo[j], fp.cca[j], fp.ccb[j], fp.Seff[j], 0.0, fp.tauo[j], psi1, fp.FltZ[j],
                            Vfree, Vf1, tau1);
  if (Vfj > 1e10 || isnan(Vfj) || isnan(tau1)) atomicMin(&flags[0], j + 1);
  const double psi2 = d_ids2(fp.xLf[j], fp.Vo[j], psi, psi1, dt, Vfj, Vf1, fp.IDstate);
  const int k2 = d_nrsearch(fp.fo[j], fp.Vo[j], fp.cca[j], fp.ccb[j], fp.Seff[j], tau1, fp.tauo[j], psi2, fp.FltZ[j],
                            Vfree, Vf2, tau2);
  if (k1 < 0 || k2 < 0) atomicAdd(&flags[1], (k1 < 0) + (k2 < 0));
  fs.psi1[j] = psi1;
  fs.Vf1[j] = Vf1;
  fs.tau1[j] = tau1;
  fs.psi2[j] = psi2;
  fs.Vf2[j] = Vf2;
  fs.tau2[j] = tau2;
  psi2_out = psi2;
  tau2_out = tau2;
}

// Newmark step (src/main.jl:297-344) as a tile epilogue: -K d of the
// predicted displacement P (loaded as the window), absorbing edges, FBC!,
// corrector, and the next step's predictor.  Reads P, v, a and writes
// P2, v2, a2 (double-buffered state; P is the step's new d).
//
// In a batch of dynamic steps the kernel also keeps the step record that the
// step tail would write: the bottom tile row leaves the maximum of the new
// fault velocity per tile (value, first index) in vpart_out, the fault-row
// threads copy d, v, tau of their node into cap_fault, and the owners of the
// receiver nodes copy d, v, a into cap_seis.  The next step's kernel reads
// the previous record (vprev, nrprev) in prepare() and stops, in every block
// alike, when that step ended the dynamic regime (Vfmax < 2e-3, the switch
// rule of the step tail), failed the FBC! guard or did not run; a stopped
// kernel marks its own record as not run.
// Newmark predictor d + dt v + dt^2/2 a (src/main.jl:299-300), one
// expression for every place that forms it, so that all agree bit for bit
__device__ __forceinline__ double newmark_predict(double d, double v, double a, double dt, double half_dt_sq) {
  return d + dt * v + half_dt_sq * a;
}

struct OpNewmark {
  struct Pre { int seis; };
  static constexpr int kParts = 0;
  static constexpr bool kLatePrepare = true;   // the stop test reads the previous step's record
  __device__ void zero_tail(int) const {}
  FaultDev fp;
  FaultState fs;
  const double* mu;
  const double* M;
  // P: this step's predicted displacement d + dt v + dt^2/2 a (the step's
  // new d); v, a: the previous step's.  Out: v2, a2 and P2, the next step's
  // predictor formed from the new state as the previous step formed P.
  const double *P, *v, *a;
  double *P2, *v2, *a2;
  double dt, half_dt, half_dt_sq;
  int* flags;               // FBC! guard / non-convergence counters of this step
  int NY;
  double* part0;
  double* part1;
  // batch record (all null outside a batch)
  const double* vprev;      // [2 * nbx] previous step's (max, index) per bottom tile
  const int* nrprev;        // previous step's flags
  double* vpart_out;        // [2 * nbx] this step's
  double* cap_fault;        // [3 * nflt] d, v, tau of the fault row
  const int64_t* seis_idx;
  const int* seis_tile;     // [nseis] tile (row-major over the tile grid) holding each receiver
  int nseis;
  double* cap_seis;         // [3 * nseis]
  int nbx;
  // false: the previous step of the batch asked to stop (regime switch, FBC!
  // guard, non-finite or not run); block-uniform
  __device__ bool go() const {
    if (!vprev) return true;
    // (max, lowest index) over the fault records, one record per lane,
    // reduced across the wave; a record marked -1 means the previous step
    // did not run
    const int lane = threadIdx.x & 63;
    double m = -INFINITY, mi = 2147483647.0;
    int notrun = 0;
    for (int b = lane; b < nbx; b += 64) {
      const double o = vprev[2 * b], oi = vprev[2 * b + 1];
      if (oi < 0.0) notrun = 1;
      else if (findmax_before(o, oi, m, mi)) { m = o; mi = oi; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(m, off, 64), oi = __shfl_xor(mi, off, 64);
      if (findmax_before(o, oi, m, mi)) { m = o; mi = oi; }
    }
    const double Vfmax = 2 * m + fp.Vpl;
    // a NaN Vfmax keeps the dynamic regime (NaN < 2e-3 is false), as in the
    // reference; the step's own FBC! guard then stops on the NaN slip rate
    const bool stop = __any(notrun) || nrprev[0] != INT_MAX || Vfmax < 2e-3;
    if (stop && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) vpart_out[1] = -1.0;
    return !stop;
  }
  // tile blocks sit at blockIdx.y >= 1 (row 0 of the grid runs the fault row),
  // in the XCD-aware order of xcd_tile
  __device__ bool prepare(Pre& p) const {
    p.seis = 0;
    if (!go()) return false;
    if (cap_seis) {
      const int tile = xcd_tile((blockIdx.y - 1) * gridDim.x + blockIdx.x, gridDim.x * (gridDim.y - 1));
      for (int q = 0; q < nseis; q++)
        if (seis_tile[q] == tile) p.seis |= 1 << q;
    }
    return true;
  }
  struct Ext { double v, a, M; };
  __device__ void fetch(const Pre&, int64_t n, int, int, Ext& e) const {
    e.v = v[n];
    e.a = a[n];
    e.M = M[n];
  }
  __device__ double load(const Pre&, int64_t n, int, int) const { return P[n]; }
  // the fault row belongs to the fault blocks
  __device__ void epilogue(const Pre& p, int64_t n, int I, int J, double ku, double dn, const Ext& e, double& a0,
                           double& a1) const {
    if (J != 0) node(p, n, I, J, ku, dn, e, a0, a1);
  }
  // corrector of one node (src/main.jl:302-338): K d, absorbing edges, FBC!
  // on the fault row, a = f / M, v += dt/2 a
  __device__ void node(const Pre& p, int64_t n, int I, int J, double ku, double dn, const Ext& e, double& a0,
                       double& a1) const {
    const double vo = e.v;
    const double vp = vo + half_dt * e.a;
    double at = -ku;
    const bool fault = (J == 0);
    const bool deep = fault && (I < fp.j0);
    if (deep) at = 0.0;
    if (I == 0) at = at - fp.BcLC[J] * vp;
    if (J == NY - 1) at = at - fp.BcTC[I] * vp;
    double tau = 0.0;
    if (fault) {
      const int j = I;
      const double Mj = e.M;
      const double Vfree = 2 * vp + 2 * half_dt * at / Mj;
      const double Vfj = 2 * vo + fp.Vpl;
      fs.FltVfree[j] = Vfree;
      fs.Vf[j] = Vfj;
      double psi_new, tau2;
      if (!deep) {
        d_fbc_node(fp, fs, j, fs.psi[j], Vfj, Vfree, dt, flags, psi_new, tau2);
      } else {
        psi_new = fs.psi2[j];
        tau2 = fs.tau2[j];
      }
      tau = deep ? 0.0 : tau2 - fp.tauo[j];
      fs.tau[j] = tau;
      fs.psi[j] = psi_new;
      at = at - fp.FltB[j] * tau;
    }
    double an = at / e.M;
    double vn = vp + half_dt * an;
    if (deep) {
      vn = 0.0;
      an = 0.0;
    }
    P2[n] = newmark_predict(dn, vn, an, dt, half_dt_sq);
    v2[n] = vn;
    a2[n] = an;
    if (fault) {
      a0 = vn;
      a1 = (double)I;
      if (cap_fault) {
        cap_fault[I] = dn;
        cap_fault[fp.nflt + I] = vn;
        cap_fault[2 * fp.nflt + I] = tau;
      }
    }
    if (p.seis) {
      for (int q = 0; q < nseis; q++)
        if ((p.seis >> q & 1) && seis_idx[q] == n) {
          cap_seis[3 * q] = dn;
          cap_seis[3 * q + 1] = vn;
          cap_seis[3 * q + 2] = an;
        }
    }
  }
  __device__ void block_end(double, double) const {}
};

// K u at fault node (I, 0) of the predictor field P, read from global memory: the row-0 segment(s) along x and the column
// segment of rows 0..4, each value formed as the tile kernel forms it
// (seg_all / seg_last outputs, x term rounded before the y term is added).
__device__ __forceinline__ double newmark_ku_fault(const Geom& g, const OpNewmark& op, int I) {
  const OpNewmark::Pre pre{};
  double u[5], m[5];
  double ax = 0.0;
  const int i = I & 3, s = I >> 2;
  if (s < g.nelx) {
#pragma unroll
    for (int k = 0; k < 5; k++) {
      u[k] = op.load(pre, 4 * s + k, 4 * s + k, 0);
      m[k] = op.mu[4 * s + k];
    }
    ax = seg_eval<double, 1>(g, u, m, i);
  }
  if (i == 0 && s > 0) {
#pragma unroll
    for (int k = 0; k < 5; k++) {
      u[k] = op.load(pre, 4 * s - 4 + k, 4 * s - 4 + k, 0);
      m[k] = op.mu[4 * s - 4 + k];
    }
    ax += seg_last<double, 1>(g, u, m);
  }
  double ay = 0.0;
  if (g.nely > 0) {
#pragma unroll
    for (int k = 0; k < 5; k++) {
      const int64_t n = (int64_t)k * g.NX + I;
      u[k] = op.load(pre, n, I, k);
      m[k] = op.mu[n];
    }
    double o[5];
    seg_all<double, 1>(g, u, m, o);
    ay = o[0];
  }
  const double sx = __dmul_rn(g.c1 * wline(g, 0, g.NY), ax);
  return sx + (g.c2 * wline(g, I, g.NX)) * ay;
}

// Whole Newmark step.  Grid row 0: fault blocks, one fault node per thread,
// dispatched first so that the FBC! Newton searches run beside the bulk
// instead of after it; they also write the per-TILE_I fault-velocity records.
// Rows 1.. : the tiles of the mesh (tile_body with OpNewmark, fault row left
// to the fault blocks).
__global__ __launch_bounds__(TBLOCK) void newmark_kernel(Geom g, OpNewmark op) {
  if (blockIdx.y > 0) {
    const int t = xcd_tile((blockIdx.y - 1) * gridDim.x + blockIdx.x, gridDim.x * (gridDim.y - 1));
    tile_body<double, OpNewmark>(g, op, t % gridDim.x, t / gridDim.x, gridDim.x, gridDim.y - 1);
    return;
  }
  if (blockIdx.x * TBLOCK >= g.NX) return;   // block-uniform
  if (!op.go()) return;                      // block-uniform
  const int I = blockIdx.x * TBLOCK + threadIdx.x;
  double a0 = 0.0, a1 = 0.0;
  if (I < g.NX) {
    OpNewmark::Pre pre{};
    pre.seis = op.cap_seis ? (int)((1u << op.nseis) - 1u) : 0;
    OpNewmark::Ext e;
    op.fetch(pre, I, I, 0, e);
    const double dn = op.load(pre, I, I, 0);
    const double ku = newmark_ku_fault(g, op, I);
    op.node(pre, I, I, 0, ku, dn, e, a0, a1);
  }
  if (!op.vpart_out) return;
  // (max, first index) of the new fault velocity per TILE_I fault nodes, in
  // findmax order as the step tail; a tile with no fault node leaves
  // (-inf, INT_MAX)
  const bool valid = I < op.fp.nflt;
  double m = valid ? a0 : -INFINITY, mi = valid ? a1 : 2147483647.0;
#pragma unroll
  for (int off = TILE_I / 2; off > 0; off >>= 1) {
    const double o = __shfl_down(m, off, TILE_I), oi = __shfl_down(mi, off, TILE_I);
    if (findmax_before(o, oi, m, mi)) { m = o; mi = oi; }
  }
  const int b = I / TILE_I;
  if (I % TILE_I == 0 && b < op.nbx) {
    op.vpart_out[2 * b] = m;
    op.vpart_out[2 * b + 1] = mi;
  }
}

// --------------------------------------------------------------------------
// element-wise kernels
// --------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(BLOCK) void cg_update_kernel(int64_t n, T* x, T* r, const T* u, const T* q,
                                                          const double* part_rho, int nrho, const double* part_uq,
                                                          int nuq, CgScal* sc, double* part_rr, int npart,
                                                          float* r32) {
  // the partials of rho and u.q, then this thread's nodes, all in flight before
  // alpha is formed
  const double prho = parts_load(part_rho, nrho);
  const double puq = parts_load(part_uq, nuq);
  const int done = sc->done;
  const int64_t base = (int64_t)blockIdx.x * BLOCK * 4;
  T xv[4], uv[4], rv[4], qv[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int64_t i = base + k * BLOCK + threadIdx.x;
    xv[k] = uv[k] = rv[k] = qv[k] = T(0);
    if (i < n) {
      xv[k] = x[i];
      uv[k] = u[i];
      rv[k] = r[i];
      qv[k] = q[i];
    }
  }
  if (done) return;
  // the residual partials are consumed over npart entries: clear the tail
  if (blockIdx.x == 0)
    for (int i = gridDim.x + threadIdx.x; i < npart; i += BLOCK) part_rr[i] = 0.0;
  const double rho = parts_total(prho);
  const double uq = parts_total(puq);
  const T alpha = (T)(rho / uq);
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int64_t i = base + k * BLOCK + threadIdx.x;
    if (i < n) {
      x[i] = xv[k] + alpha * uv[k];
      const T rn = rv[k] - alpha * qv[k];
      r[i] = rn;
      if (r32) r32[i] = (float)rn;
      acc += (double)rn * (double)rn;
    }
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) part_rr[blockIdx.x] = s;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->rho_prev = rho;
    sc->iter = sc->iter + 1;
  }
}

// Jacobi preconditioner with the cg! convergence test: c = D^-1 r, partial c.r
template <typename T>
__global__ __launch_bounds__(BLOCK) void jacobi_pc_kernel(int64_t n, const T* dinv, const T* r, T* c,
                                                          const double* part_rr, int nrr, CgScal* sc,
                                                          double* part_rho) {
  if (sc->done) return;
  const double rr = reduce_parts(part_rr, nrr);
  const double resid = sqrt(rr);
  const bool fin = (resid <= sc->reltol) || (sc->iter >= sc->maxiter);
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    sc->resid = resid;
    if (fin) sc->done = 1;
  }
  if (fin) return;
  double acc = 0.0;
  const int64_t base = (int64_t)blockIdx.x * BLOCK * 4;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int64_t i = base + k * BLOCK + threadIdx.x;
    if (i < n) {
      const T cv = dinv[i] * r[i];
      c[i] = cv;
      acc += (double)cv * (double)r[i];
    }
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) part_rho[blockIdx.x] = s;
}

__global__ void cg_init_scal_kernel(CgScal* sc, const double* part_rr, const double* part_bb, int nparts, double tol,
                                    int maxiter) {
  const double rr = reduce_parts(part_rr, nparts);
  const double bb = reduce_parts(part_bb, nparts);
  if (threadIdx.x == 0) {
    sc->rho_prev = 1.0;
    sc->resid = sqrt(rr);
    sc->bnorm = sqrt(bb);
    sc->reltol = tol * sqrt(bb);
    sc->iter = 0;
    sc->maxiter = maxiter;
    sc->done = (sqrt(rr) <= tol * sqrt(bb) || maxiter <= 0) ? 1 : 0;
  }
}

// With a mirror (pinned host memory) the scalars also go there, so the host
// reads them after its synchronisation with no copy on the stream.
__global__ void cg_check_kernel(CgScal* sc, const double* part_rr, int nparts, CgScal* mirror) {
  if (!sc->done) {   // block-uniform
    const double rr = reduce_parts(part_rr, nparts);
    if (threadIdx.x == 0) {
      const double resid = sqrt(rr);
      sc->resid = resid;
      if (resid <= sc->reltol || sc->iter >= sc->maxiter) sc->done = 1;
    }
  }
  if (mirror && threadIdx.x == 0) *mirror = *sc;
}

// --------------------------------------------------------------------------
// CSR kernels for AMG levels >= 1 (LPR lanes per row)
// --------------------------------------------------------------------------
template <typename T, int LPR, int MODE>
__global__ __launch_bounds__(BLOCK) void csr_kernel(int64_t n, const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ col, const T* __restrict__ val,
                                                    const T* x, const T* b, const T* dinv, T omega, T* y, T* xout,
                                                    const CgScal* sc) {
  if (sc && sc->done) return;
  const int64_t gid = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const int64_t row = gid / LPR;
  const int lane = (int)(gid % LPR);
  double acc = 0.0;
  if (row < n) {
    const int q1 = rp[row + 1];
    int q = rp[row] + lane;
    // four entries at a time: their column and value loads, then the
    // gathers, then the sums in entry order (the gathers wait on the columns)
    for (; q + 3 * LPR < q1; q += 4 * LPR) {
      int c[4];
      T v[4], xv[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        c[u] = col[q + u * LPR];
        v[u] = val[q + u * LPR];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) xv[u] = (MODE == CSR_DOWN) ? omega * dinv[c[u]] * b[c[u]] : x[c[u]];
#pragma unroll
      for (int u = 0; u < 4; u++) acc += (double)v[u] * (double)xv[u];
    }
    for (; q < q1; q += LPR) {
      const int c = col[q];
      const T xv = (MODE == CSR_DOWN) ? omega * dinv[c] * b[c] : x[c];
      acc += (double)val[q] * (double)xv;
    }
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, LPR);
  if (row < n && lane == 0) {
    if (MODE == CSR_SPMV) {
      y[row] = (T)acc;
    } else if (MODE == CSR_DOWN) {
      xout[row] = omega * dinv[row] * b[row];
      y[row] = b[row] - (T)acc;
    } else if (MODE == CSR_PROLONG_ADD) {
      y[row] = y[row] + (T)acc;
    } else {  // CSR_POST: y = x + w dinv (b - A x)
      y[row] = x[row] + omega * dinv[row] * (b[row] - (T)acc);
    }
  }
}

// y = U b + V e, LPR lanes per row
template <typename T, int LPR>
__global__ __launch_bounds__(BLOCK) void csr_up_kernel(int64_t n, const int32_t* __restrict__ urp,
                                                       const int32_t* __restrict__ ucol, const T* __restrict__ uval,
                                                       const int32_t* __restrict__ vrp,
                                                       const int32_t* __restrict__ vcol, const T* __restrict__ vval,
                                                       const T* b, const T* e, T* y, const CgScal* sc) {
  if (sc && sc->done) return;
  const int64_t gid = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const int64_t row = gid / LPR;
  const int lane = (int)(gid % LPR);
  double acc = 0.0;
  if (row < n) {
    const int u1 = urp[row + 1];
    for (int q = urp[row] + lane; q < u1; q += LPR) acc += (double)uval[q] * (double)b[ucol[q]];
    const int v1 = vrp[row + 1];
    for (int q = vrp[row] + lane; q < v1; q += LPR) acc += (double)vval[q] * (double)e[vcol[q]];
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, LPR);
  if (row < n && lane == 0) y[row] = (T)acc;
}

// x = A_c^-1 b, one wavefront per row of the row-major inverse
template <typename T>
__global__ __launch_bounds__(BLOCK) void dense_kernel(int64_t n, const T* inv, const T* b, T* x,
                                                      const CgScal* sc) {
  if (sc && sc->done) return;
  const int64_t row = (int64_t)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  if (row < n) {
    // four independent partial sums keep four loads per lane in flight
    const T* a = inv + row * n;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int64_t j = lane;
    for (; j + 192 < n; j += 256) {
      s0 += (double)a[j] * (double)b[j];
      s1 += (double)a[j + 64] * (double)b[j + 64];
      s2 += (double)a[j + 128] * (double)b[j + 128];
      s3 += (double)a[j + 192] * (double)b[j + 192];
    }
    for (; j < n; j += 64) s0 += (double)a[j] * (double)b[j];
    s = (s0 + s1) + (s2 + s3);
  }
  s = wave_sum(s);
  if (row < n && lane == 0) x[row] = (T)s;
}


// --------------------------------------------------------------------------
// quasi-static step pieces (src/main.jl:226-295)
// --------------------------------------------------------------------------
// pass 1: dPre = d everywhere; fault: Vf0 = Vf = 2v + Vpl.  Both passes:
// F = dPre + v dt on the fault, d[fault] = F, b rows 1..4 = -(K F) (only the
// bottom element row couples to the fault).  db (optional, [4 NX]) receives
// b_new - b_old on rows 1..4 (row J at db[(J - 1) NX + I]).
// The fault-row part at fault node I (first pass: after dPre[I] = d[I]).
__device__ __forceinline__ void qs_begin_fault_node(const Geom& g, const FaultDev& fp, const FaultState& fs,
                                                    const double* mu, double* d, const double* dPre,
                                                    const double* v, double* b, double dt, int first_pass,
                                                    double* db, double theta, double* bold, int I) {
  const double vf = v[I];
  if (first_pass) {
    const double V0 = 2 * vf + fp.Vpl;
    fs.Vf0[I] = V0;
    fs.Vf[I] = V0;
  }
  const double F = (first_pass ? d[I] : dPre[I]) + vf * dt;
  // every load before the first store (the stores could alias them)
  double muq[5], bm[4], bo[4];
#pragma unroll
  for (int q = 0; q < 5; q++) muq[q] = mu[(int64_t)q * g.NX + I];
#pragma unroll
  for (int J = 1; J <= 4; J++) {
    bm[J - 1] = b[(int64_t)J * g.NX + I];
    bo[J - 1] = (db && theta != 0.0) ? bold[(int64_t)(J - 1) * g.NX + I] : 0.0;
  }
  d[I] = F;
  const double cy = g.c2 * wline(g, I, g.NX);
#pragma unroll
  for (int J = 1; J <= 4; J++) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 5; q++) s += g.D[q * 5 + J] * (g.w[q] * muq[q]) * g.D[q * 5 + 0];
    const int64_t m = (int64_t)J * g.NX + I;
    const int64_t r4 = (int64_t)(J - 1) * g.NX + I;
    const double bn = -(cy * s) * F;
    const double bv = bm[J - 1];
    // change of b since the solve the start comes from; with the extrapolation,
    // less theta times the change the step before (the start then solves that
    // extrapolated b off rows 1..4 already)
    if (db) db[r4] = theta != 0.0 ? (bn - bv) - theta * (bv - bo[J - 1]) : bn - bv;
    if (bold) bold[r4] = bv;
    b[m] = bn;
  }
}

__global__ __launch_bounds__(BLOCK) void qs_begin_kernel(Geom g, FaultDev fp, FaultState fs, const double* mu,
                                                         double* d, double* dPre, const double* v, double* b,
                                                         double dt, int first_pass, double* db, double theta,
                                                         double* bold, const StepStatus* st_next, double dt_prev) {
  if (st_next) {   // block-uniform
    if (!st_next->go) return;
    dt = st_next->dt_next;
    if (theta != 0.0) theta = dt / dt_prev;
  }
  const int64_t N = (int64_t)g.NX * g.NY;
  const int64_t n = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (first_pass) {
    if (n >= N) return;
    const double dn = d[n];
    // extrapolated start off the fault: d + theta (d - dPre), the last step's increment
    if (theta != 0.0 && n >= g.NX) d[n] = dn + theta * (dn - dPre[n]);
    dPre[n] = dn;
  }
  if (n >= g.NX) return;
  qs_begin_fault_node(g, fp, fs, mu, d, dPre, v, b, dt, first_pass, db, theta, bold, (int)n);
}

// tau1 = -(K d)[fault]/FltB (0 on creeping nodes), slrFunc!, Vf and fault v update
// save (or null): [psi | Vf | fault row of d], each nflt long, as they were
// before this update (the state a redone second pass restores)
// Fault update of a quasi-static pass at node j: tau1 = -(K d)[fault] / FltB, slrFunc!'s
// psi1 and Vf1, the average Vf and v on the fault (src/main.jl:262-282); with save,
// psi, Vf and d of the node first.
__device__ void qs_fault_node(const Geom& g, const FaultDev& fp, const FaultState& fs, const double* mu,
                              const double* d, double* v, double dt, double* save, int j) {
  if (save) {
    save[j] = fs.psi[j];
    save[fp.nflt + j] = fs.Vf[j];
    save[2 * fp.nflt + j] = d[j];
  }
  const bool deep = j < fp.j0;
  const double a = deep ? 0.0 : ku_global<double>(g, d, mu, j, 0);
  const double tau1 = -a / fp.FltB[j];
  fs.tau1[j] = tau1;
  double Vf1;
  if (!deep) {
    const double psi1 = d_ids(fp.xLf[j], fp.Vo[j], fs.psi[j], dt, fs.Vf[j], 1e-6, fp.IDstate);
    fs.psi1[j] = psi1;
    const double tauAB = tau1 + fp.tauo[j];
    const double fa = tauAB / (fp.Seff[j] * fp.cca[j]);
    const double help = -(fp.fo[j] + fp.ccb[j] * psi1) / fp.cca[j];
    Vf1 = fp.Vo[j] * (exp(help + fa) - exp(help - fa));
  } else {
    Vf1 = fp.Vpl;
  }
  fs.Vf1[j] = Vf1;
  const double Vf = (fs.Vf0[j] + Vf1) / 2;
  fs.Vf[j] = Vf;
  if (v) v[j] = 0.5 * (Vf - fp.Vpl);
}

__global__ __launch_bounds__(BLOCK) void qs_fault_kernel(Geom g, FaultDev fp, FaultState fs, const double* mu,
                                                         const double* d, double* v, double dt, double* save) {
  const int j = blockIdx.x * BLOCK + threadIdx.x;
  if (j >= fp.nflt) return;
  qs_fault_node(g, fp, fs, mu, d, v, dt, save, j);
}

// The fault row of the step finish at node j (src/main.jl:285-293): psi, tau, v on
// the fault; creeping nodes slide at Vpl with d = v = 0 there.
__device__ __forceinline__ void qs_finish_fault_node(const FaultDev& fp, const FaultState& fs, double* d, double* v,
                                                     int j) {
  const bool deep = j < fp.j0;
  fs.psi[j] = fs.psi1[j];
  fs.tau[j] = deep ? 0.0 : fs.tau1[j];
  if (deep) fs.Vf1[j] = fp.Vpl;
  double vv = 0.5 * (fs.Vf1[j] - fp.Vpl);
  if (deep) {
    vv = 0.0;
    d[j] = 0.0;
  }
  v[j] = vv;
}

// Vfmax (src/main.jl:348), regime switch (src/main.jl:439-443), dtevol! (src/dtevol.jl:5-44)
__device__ double dtevol_dev(const double* XiLf, const double* v, bool slip_rate_from_v, double Vpl, int j0, int n,
                             double dt, double dtmin, int isolver) {
  if (isolver != 1) return dtmin;
  const double dtmax = 50 * 24 * 60 * 60;
  double dtnx = dtmax;
  for (int i = j0 + threadIdx.x; i < n; i += BLOCK) {
    const double Vc = slip_rate_from_v ? 2 * v[i] + Vpl : v[i];
    if (fabs(Vc) * dtmax > XiLf[i]) {
      const double dtcell = XiLf[i] / fabs(Vc);
      if (dtcell < dtnx) dtnx = dtcell;
    }
  }
  __shared__ double sh[BLOCK];
  sh[threadIdx.x] = dtnx;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  dtnx = sh[0];
  __syncthreads();
  if (dtmin > dtnx) dtnx = dtmin;
  if (dtnx > 1.2 * dt) dtnx = 1.2 * dt;
  return dtnx;
}

// The step tail over the fault with a block of NT threads.
// qs_dPre (the quasi-static end, which runs beside the interior finish): the
// receivers off the fault take v = (d - dPre) / dt and a = 0, the values the
// finish writes there, and on the fault the block's own v and a = 0.
template <int NT>
__device__ void step_tail_body(const FaultDev& fp, const double* d, const double* v, const double* a,
                               const double* tau, double dt, double dtmin, int isolver, StepStatus* st, int* flags,
                               const StepCapture& cap, const double* qs_dPre = nullptr, int64_t NX = 0,
                               const CgScal* sc_go = nullptr) {
  // one pass over the fault: the maximum of v (first index on ties) and the
  // quasi-static dtevol! candidate min XiLf/|Vf| over the cells that qualify,
  // reduced across each wave by shuffles and across the four waves in LDS
  // (both reductions are order-independent)
  const double dtmax = 50 * 24 * 60 * 60;
  double m = -INFINITY, dtnx = dtmax;
  int mi = INT_MAX;
  for (int j = threadIdx.x; j < fp.nflt; j += NT) {
    const double vj = v[j];
    if (findmax_before(vj, j, m, mi)) { m = vj; mi = j; }
    if (j >= fp.j0) {
      const double Vc = 2 * vj + fp.Vpl;
      if (fabs(Vc) * dtmax > fp.XiLf[j]) {
        const double dtcell = fp.XiLf[j] / fabs(Vc);
        if (dtcell < dtnx) dtnx = dtcell;
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double om = __shfl_down(m, off, 64);
    const int oi = __shfl_down(mi, off, 64);
    if (findmax_before(om, oi, m, mi)) { m = om; mi = oi; }
    dtnx = fmin(dtnx, __shfl_down(dtnx, off, 64));
  }
  __shared__ double sm[NT / 64], sdt[NT / 64];
  __shared__ int si[NT / 64];
  if ((threadIdx.x & 63) == 0) {
    sm[threadIdx.x >> 6] = m;
    si[threadIdx.x >> 6] = mi;
    sdt[threadIdx.x >> 6] = dtnx;
  }
  __syncthreads();
  m = sm[0];
  mi = si[0];
  dtnx = sdt[0];
  for (int w = 1; w < NT / 64; w++) {
    if (findmax_before(sm[w], si[w], m, mi)) { m = sm[w]; mi = si[w]; }
    dtnx = fmin(dtnx, sdt[w]);
  }
  const double Vfmax = 2 * m + fp.Vpl;
  const int next = ((isolver == 1 && Vfmax < 5e-3) || (isolver == 2 && Vfmax < 2e-3)) ? 1 : 2;
  double dtn = dtmin;
  if (next == 1) {
    // dtevol! (src/dtevol.jl:5-44)
    if (dtmin > dtnx) dtnx = dtmin;
    if (dtnx > 1.2 * dt) dtnx = 1.2 * dt;
    dtn = dtnx;
  }
  if (cap.fault) {
    for (int j = threadIdx.x; j < fp.nflt; j += NT) {
      cap.fault[j] = d[j];
      cap.fault[fp.nflt + j] = v[j];
      cap.fault[2 * fp.nflt + j] = tau[j];
    }
  }
  if (cap.seis) {
    for (int q = threadIdx.x; q < cap.nseis; q += NT) {
      const int64_t n = cap.seis_idx[q];
      const double dn = d[n];
      const double vn = (qs_dPre && n >= NX) ? (dn - qs_dPre[n]) / dt : v[n];
      const double an = qs_dPre ? 0.0 : a[n];
      cap.seis[3 * q] = dn;
      cap.seis[3 * q + 1] = vn;
      cap.seis[3 * q + 2] = an;
      if (cap.seis_mirror) {
        cap.seis_mirror[3 * q] = dn;
        cap.seis_mirror[3 * q + 1] = vn;
        cap.seis_mirror[3 * q + 2] = an;
      }
    }
  }
  if (threadIdx.x == 0) {
    StepStatus o{};
    o.Vfmax = Vfmax;
    o.dt_next = dtn;
    o.isolver_next = next;
    o.hypo = mi;
    o.nr_flag = (flags[0] == INT_MAX) ? 0 : flags[0];
    o.nr_noconv = flags[1];
    o.go = (next == 1 && o.nr_flag == 0 && (!sc_go || sc_go->done)) ? 1 : 0;
    *st = o;
    if (cap.st_mirror) *cap.st_mirror = o;
    flags[0] = INT_MAX;
    flags[1] = 0;
  }
}

__global__ __launch_bounds__(BLOCK) void step_tail_kernel(FaultDev fp, const double* d, const double* v,
                                                          const double* a, const double* tau, double dt,
                                                          double dtmin, int isolver, StepStatus* st, int* flags,
                                                          StepCapture cap) {
  step_tail_body<BLOCK>(fp, d, v, a, tau, dt, dtmin, isolver, st, flags, cap);
}

// The end of a quasi-static step.  Block 0 (QE_BLOCK threads, fault nodes
// strided): with sc, the second solve's convergence test (cg_check_kernel's,
// mirror included); the second pass's fault update (with save: psi, Vf and d of
// each node first); the fault row of the step finish; the step tail, which reads
// every fault node.  Blocks 1.. : the interior part of the finish (v = (d -
// dPre) / dt off the fault, a = 0), which block 0 neither reads nor writes (its
// receiver capture forms those values itself).
constexpr int QE_BLOCK = 1024;
__global__ __launch_bounds__(QE_BLOCK) void qs_end_kernel(Geom g, FaultDev fp, FaultState fs, const double* mu,
                                                          double* d, const double* dPre, double* v, double* a,
                                                          double dt, double dtmin, double* save, CgScal* sc,
                                                          const double* part_rr, int nparts, CgScal* sc_mirror,
                                                          StepStatus* st, int* flags, StepCapture cap) {
  if (blockIdx.x > 0) {
    const int64_t N = (int64_t)g.NX * g.NY;
    const int64_t n = (int64_t)(blockIdx.x - 1) * QE_BLOCK + threadIdx.x;
    if (n >= N) return;
    if (n >= g.NX) v[n] = (d[n] - dPre[n]) / dt;
    a[n] = 0.0;
    return;
  }
  if (sc) {
    if (!sc->done) {   // block-uniform
      const double rr = reduce_parts(part_rr, nparts);
      if (threadIdx.x == 0) {
        const double resid = sqrt(rr);
        sc->resid = resid;
        if (resid <= sc->reltol || sc->iter >= sc->maxiter) sc->done = 1;
      }
    }
    if (sc_mirror && threadIdx.x == 0) *sc_mirror = *sc;
  }
  for (int j = threadIdx.x; j < fp.nflt; j += QE_BLOCK) qs_fault_node(g, fp, fs, mu, d, nullptr, dt, save, j);
  __syncthreads();   // every tau1 read d's fault row before creeping nodes reset it
  for (int j = threadIdx.x; j < fp.nflt; j += QE_BLOCK) qs_finish_fault_node(fp, fs, d, v, j);
  __syncthreads();
  step_tail_body<QE_BLOCK>(fp, d, v, a, fs.tau, dt, dtmin, 1, st, flags, cap, dPre, g.NX, sc);
}

// Between the passes of a quasi-static step, in one block: the first pass's
// fault update (qs_fault_node at every fault node), then, once every node's
// K d has read the fault row, the second pass's begin on the fault row
// (qs_begin_fault_node: F = dPre + v dt, b rows 1..4, db).  With sc, first
// the first solve's convergence test (cg_check_kernel's, mirror included);
// while that solve has not converged the kernel changes nothing else.
__global__ __launch_bounds__(QE_BLOCK) void qs_mid_kernel(Geom g, FaultDev fp, FaultState fs, const double* mu,
                                                          double* d, const double* dPre, double* v, double* b,
                                                          double dt, double* db, CgScal* sc, const double* part_rr,
                                                          int nparts, CgScal* sc_mirror) {
  if (sc) {
    __shared__ int done;
    if (!sc->done) {   // block-uniform
      const double rr = reduce_parts(part_rr, nparts);
      if (threadIdx.x == 0) {
        const double resid = sqrt(rr);
        sc->resid = resid;
        if (resid <= sc->reltol || sc->iter >= sc->maxiter) sc->done = 1;
      }
    }
    if (threadIdx.x == 0) {
      done = sc->done;
      if (sc_mirror) *sc_mirror = *sc;
    }
    __syncthreads();
    if (!done) return;
  }
  for (int j = threadIdx.x; j < fp.nflt; j += QE_BLOCK) qs_fault_node(g, fp, fs, mu, d, v, dt, nullptr, j);
  __syncthreads();
  for (int I = threadIdx.x; I < g.NX; I += QE_BLOCK)
    qs_begin_fault_node(g, fp, fs, mu, d, dPre, v, b, dt, 0, db, 0.0, nullptr, I);
}

// --------------------------------------------------------------------------
// stand-alone fault kernels
// --------------------------------------------------------------------------
__global__ void ids_kernel(int64_t n, const double* xLf, const double* Vo, const double* psi, double dt,
                           const double* Vf, double cnd, int IDstate, double* psi1) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) psi1[i] = d_ids(xLf[i], Vo[i], psi[i], dt, Vf[i], cnd, IDstate);
}
__global__ void ids2_kernel(int64_t n, const double* xLf, const double* Vo, const double* psi, const double* psi1,
                            double dt, const double* Vf, const double* Vf1, int IDstate, double* psi2) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) psi2[i] = d_ids2(xLf[i], Vo[i], psi[i], psi1[i], dt, Vf[i], Vf1[i], IDstate);
}
__global__ void nrsearch_kernel(int64_t n, const double* fo, const double* Vo, const double* cca, const double* ccb,
                                const double* Seff, const double* tau, const double* tauo, const double* psi,
                                const double* FltZ, const double* FltVfree, double* Vf, double* tau_out, int* k) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  double vf, t;
  k[i] = d_nrsearch(fo[i], Vo[i], cca[i], ccb[i], Seff[i], tau[i], tauo[i], psi[i], FltZ[i], FltVfree[i], vf, t);
  Vf[i] = vf;
  tau_out[i] = t;
}
__global__ void fbc_kernel(FaultDev fp, FaultState fs, const double* FltVfree, double dt, int* flags) {
  const int j = fp.j0 + blockIdx.x * BLOCK + threadIdx.x;
  if (j >= fp.nflt) return;
  double p2, t2;
  d_fbc_node(fp, fs, j, fs.psi[j], fs.Vf[j], FltVfree[j], dt, flags, p2, t2);
}
__global__ void slr_kernel(FaultDev fp, FaultState fs, double dt) {
  const int j = fp.j0 + blockIdx.x * BLOCK + threadIdx.x;
  if (j >= fp.nflt) return;
  const double psi1 = d_ids(fp.xLf[j], fp.Vo[j], fs.psi[j], dt, fs.Vf[j], 1e-6, fp.IDstate);
  fs.psi1[j] = psi1;
  const double tauAB = fs.tau1[j] + fp.tauo[j];
  const double fa = tauAB / (fp.Seff[j] * fp.cca[j]);
  const double help = -(fp.fo[j] + fp.ccb[j] * psi1) / fp.cca[j];
  fs.Vf1[j] = fp.Vo[j] * (exp(help + fa) - exp(help - fa));
}
__global__ __launch_bounds__(BLOCK) void dtevol_kernel(const double* XiLf, const double* Vf, int j0, int n,
                                                       double dt, double dtmin, int isolver, double* out) {
  const double r = dtevol_dev(XiLf, Vf, false, 0.0, j0, n, dt, dtmin, isolver);
  if (threadIdx.x == 0) *out = r;
}

// --------------------------------------------------------------------------
// misc
// --------------------------------------------------------------------------
__global__ void diag_kernel(Geom g, const double* mu, double* diag) {
  const int64_t N = (int64_t)g.NX * g.NY;
  const int64_t n = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (n >= N) return;
  const int I = (int)(n % g.NX), J = (int)(n / g.NX);
  double ax = 0, ay = 0;
  const int i = I & 3, s = I >> 2;
  for (int ss = (i == 0 ? s - 1 : s); ss <= s; ss++) {
    if (ss < 0 || ss >= g.nelx) continue;
    const int li = I - 4 * ss;
    for (int p = 0; p < 5; p++)
      ax += g.D[p * 5 + li] * g.D[p * 5 + li] * g.w[p] * mu[(int64_t)J * g.NX + 4 * ss + p];
  }
  const int j = J & 3, t = J >> 2;
  for (int tt = (j == 0 ? t - 1 : t); tt <= t; tt++) {
    if (tt < 0 || tt >= g.nely) continue;
    const int lj = J - 4 * tt;
    for (int q = 0; q < 5; q++)
      ay += g.D[q * 5 + lj] * g.D[q * 5 + lj] * g.w[q] * mu[(int64_t)(4 * tt + q) * g.NX + I];
  }
  diag[n] = g.c1 * wline(g, J, g.NY) * ax + g.c2 * wline(g, I, g.NX) * ay;
}

// element_computation! (src/PCG.jl:71-86): a_e = c1 H (W.(H' u_e)) + c2 (W.(u_e H)) H'
__global__ void element_ku_kernel(Geom g, const double* u, const double* mu, double* a_elem) {
  const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const int64_t Nel = (int64_t)g.nelx * g.nely;
  if (e >= Nel) return;
  const int ex = (int)(e % g.nelx), ey = (int)(e / g.nelx);
  double ue[5][5], W[5][5];
  for (int j = 0; j < 5; j++)
    for (int i = 0; i < 5; i++) {
      const int64_t n = (int64_t)(4 * ey + j) * g.NX + 4 * ex + i;
      ue[i][j] = u[n];
      W[i][j] = g.w[i] * g.w[j] * mu[n];
    }
  for (int j = 0; j < 5; j++)
    for (int i = 0; i < 5; i++) {
      double sx = 0, sy = 0;
      for (int p = 0; p < 5; p++) {
        double tx = 0, ty = 0;
        for (int k = 0; k < 5; k++) {
          tx += g.D[p * 5 + k] * ue[k][j];
          ty += g.D[p * 5 + k] * ue[i][k];
        }
        sx += g.D[p * 5 + i] * W[p][j] * tx;
        sy += g.D[p * 5 + j] * W[i][p] * ty;
      }
      a_elem[e * 25 + j * 5 + i] = g.c1 * sx + g.c2 * sy;
    }
}

template <typename T, typename U>
__global__ void convert_kernel(int64_t n, const T* x, U* y) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) y[i] = (U)x[i];
}
template <typename T>
__global__ void fill_kernel(int64_t n, T* y, T v) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) y[i] = v;
}
template <typename T>
__global__ void fill_strided_kernel(int64_t n, T* y, int stride, T v) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) y[i * stride] = v;
}
// y[k n + i] = x_k[i] (k = 0, 1, 2), or the reverse
__global__ void copy3_kernel(int64_t n, double* x0, double* x1, double* x2, double* y, int gather) {
  const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (e >= 3 * n) return;
  const int64_t k = e / n, i = e - k * n;
  double* x = k == 0 ? x0 : (k == 1 ? x1 : x2);
  if (gather) y[e] = x[i];
  else x[i] = y[e];
}
// x += gamma p, r -= gamma q with gamma = <p, r0> / <p, K p> (0 if not positive
// definite in floating point); partial ||r||^2.  The energy-norm error of the
// guess is then never larger than that of x before the correction.
__global__ __launch_bounds__(BLOCK) void guess_update_kernel(int64_t n, double* x, double* r, const double* p,
                                                             const double* q, const double* part_pr,
                                                             const double* part_pq, int nparts, double* part_rr,
                                                             int npart_out) {
  const double pr = reduce_parts(part_pr, nparts);
  const double pq = reduce_parts(part_pq, nparts);
  const double gamma = (pq > 0.0) ? pr / pq : 0.0;
  if (blockIdx.x == 0)
    for (int i = gridDim.x + threadIdx.x; i < npart_out; i += BLOCK) part_rr[i] = 0.0;
  double acc = 0.0;
  const int64_t base = (int64_t)blockIdx.x * BLOCK * 4;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int64_t i = base + k * BLOCK + threadIdx.x;
    if (i < n) {
      x[i] += gamma * p[i];
      const double rv = r[i] - gamma * q[i];
      r[i] = rv;
      acc += rv * rv;
    }
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) part_rr[blockIdx.x] = s;
}

// FP32 preconditioner output back to FP64 with the partial <c, r> in FP64
__global__ __launch_bounds__(BLOCK) void widen_dot_kernel(int64_t n, const float* c32, const double* r, double* c,
                                                          const CgScal* sc, double* part) {
  if (sc->done) return;
  double acc = 0.0;
  const int64_t base = (int64_t)blockIdx.x * BLOCK * 4;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int64_t i = base + k * BLOCK + threadIdx.x;
    if (i < n) {
      const double cv = (double)c32[i];
      c[i] = cv;
      acc += cv * r[i];
    }
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void axpy_kernel(int64_t n, double alpha, const double* x, double* y) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) y[i] += alpha * x[i];
}
__global__ void scale_mask_kernel(int64_t n, int nx, const double* x, double sign, double* y) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) y[i] = (i < nx) ? 0.0 : sign * x[i];
}

// --------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------
static inline int nblk(int64_t n) { return (int)((n + BLOCK - 1) / BLOCK); }

template <typename T>
void launch_ku(const Geom& g, const T* u, const T* mu, T* y, int interior, int negate, hipStream_t s) {
  OpApply<T> op{u, mu, y, interior, negate ? T(-1) : T(1), nullptr, nullptr};
  hipLaunchKernelGGL((tile_kernel<T, OpApply<T>>), tile_grid(g), dim3(TBLOCK), 0, s, g, op);
}

void launch_diag(const Geom& g, const double* mu, double* diag, hipStream_t s) {
  const int64_t N = (int64_t)g.NX * g.NY;
  hipLaunchKernelGGL(diag_kernel, dim3(nblk(N)), dim3(BLOCK), 0, s, g, mu, diag);
}

void launch_element_ku(const Geom& g, const double* u, const double* mu, double* a_elem, hipStream_t s) {
  const int64_t Nel = (int64_t)g.nelx * g.nely;
  hipLaunchKernelGGL(element_ku_kernel, dim3(nblk(Nel)), dim3(BLOCK), 0, s, g, u, mu, a_elem);
}

template <typename T>
void launch_cg_residual(const Geom& g, const T* x, const T* mu, const T* b, T* r, double* part_rr, double* part_bb,
                        int npart, hipStream_t s, float* r32) {
  OpResidual<T> op{npart, x, mu, b, r, part_rr, part_bb, r32};
  hipLaunchKernelGGL((tile_kernel<T, OpResidual<T>>), tile_grid(g), dim3(TBLOCK), 0, s, g, op);
}

void launch_gal_residual(const Geom& g, const double* d, const half8* phi16, int64_t nI, int nb, const double* c,
                         const double* mu, const double* b, double* r, float* r32, double* x_out, double* part_rr,
                         double* part_bb, int npart, hipStream_t s, const int* gate) {
#define EQ_GAL_FULL(NG)                                                                                    \
  {                                                                                                        \
    OpGalResidual<true, NG> op{npart, d, phi16, nI, nb, c, mu, b, r, r32, x_out, part_rr, part_bb, gate};  \
    hipLaunchKernelGGL((tile_kernel<double, OpGalResidual<true, NG>>), tile_grid(g), dim3(TBLOCK), 0, s, g, op); \
  }
  if (nb == 64) EQ_GAL_FULL(8)
  else if (nb == 48) EQ_GAL_FULL(6)
  else if (nb == 40) EQ_GAL_FULL(5)
  else if (nb == 32) EQ_GAL_FULL(4)
  else if (nb == 24) EQ_GAL_FULL(3)
  else if (nb == 16) EQ_GAL_FULL(2)
  else if (nb == 8) EQ_GAL_FULL(1)
#undef EQ_GAL_FULL
  else {
    OpGalResidual<false> op{npart, d, phi16, nI, nb, c, mu, b, r, r32, x_out, part_rr, part_bb, gate};
    hipLaunchKernelGGL((tile_kernel<double, OpGalResidual<false>>), tile_grid(g), dim3(TBLOCK), 0, s, g, op);
  }
}

void launch_cg_init_scal(CgScal* sc, const double* part_rr, const double* part_bb, int nparts, double tol,
                         int maxiter, hipStream_t s) {
  hipLaunchKernelGGL(cg_init_scal_kernel, dim3(1), dim3(BLOCK), 0, s, sc, part_rr, part_bb, nparts, tol, maxiter);
}

void launch_cg_check(CgScal* sc, const double* part_rr, int nparts, hipStream_t s, CgScal* mirror) {
  hipLaunchKernelGGL(cg_check_kernel, dim3(1), dim3(BLOCK), 0, s, sc, part_rr, nparts, mirror);
}

template <typename T>
void launch_cg_dir(const Geom& g, const T* c, const T* u_old, const T* mu, T* u_new, T* q, const double* part_rho,
                   int nrho, const CgScal* sc, double* part_uq, hipStream_t s, LaunchEvents le) {
  OpCgDir<T> op{c, u_old, mu, u_new, q, part_rho, nrho, sc, part_uq, nullptr};
  launch_timed(tile_kernel<T, OpCgDir<T>>, tile_grid(g), dim3(TBLOCK), s, le, g, op);
}

template <typename T>
void launch_cg_update(int64_t n, T* x, T* r, const T* u, const T* q, const double* part_rho, int nrho,
                      const double* part_uq, int nuq, CgScal* sc, double* part_rr, int npart, float* r32,
                      hipStream_t s, LaunchEvents le) {
  launch_timed(cg_update_kernel<T>, dim3(vec_blocks(n)), dim3(BLOCK), s, le, n, x, r, u, q, part_rho, nrho,
               part_uq, nuq, sc, part_rr, npart, r32);
}

template <typename T>
void launch_jacobi_pc(int64_t n, const T* dinv, const T* r, T* c, const double* part_rr, int nrr, CgScal* sc,
                      double* part_rho, hipStream_t s) {
  hipLaunchKernelGGL((jacobi_pc_kernel<T>), dim3(vec_blocks(n)), dim3(BLOCK), 0, s, n, dinv, r, c, part_rr, nrr, sc,
                     part_rho);
}

template <typename T>
void launch_l0_down(const Geom& g, const T* r, const T* mu, const T* dinv, T omega, T* x0, T* res,
                    const double* part_rr, int nrr, CgScal* sc, hipStream_t s) {
  OpL0Down<T> op{r, mu, dinv, omega, x0, res, part_rr, nrr, sc, nullptr, nullptr};
  hipLaunchKernelGGL((tile_kernel<T, OpL0Down<T>>), tile_grid(g), dim3(TBLOCK), 0, s, g, op);
}

template <typename T>
void launch_l0_post(const Geom& g, const T* x0, const T* r, const T* mu, const T* dinv, T omega, T* c,
                    double* part_rho, const CgScal* sc, hipStream_t s) {
  OpL0Post<T> op{x0, r, mu, dinv, omega, c, sc, part_rho, nullptr};
  hipLaunchKernelGGL((tile_kernel<T, OpL0Post<T>>), tile_grid(g), dim3(TBLOCK), 0, s, g, op);
}


template <typename T>
void launch_cheb(const Geom& g, const ChebStep<T>& c, const T* mu, const T* dinv, const double* part_rr, int nrr,
                 CgScal* sc, double* part_rho, hipStream_t s, LaunchEvents le) {
  OpCheb<T> op{c.src, c.b, mu, dinv, c.s0, c.src_mode, c.rin, c.rout, c.c1, c.c2, c.xin, c.xsrc_center, c.dout,
               c.xout, c.mode, c.write_x_center, c.rdot, c.xout64, c.rdot64, c.check, part_rr, nrr, sc, part_rho,
               nullptr,
               c.coarse, g.nelx, g.nely, {}};
  for (int p = 0; p < 5; p++) op.gx[p] = (T)g.x[p];
  launch_timed(tile_kernel<T, OpCheb<T>>, tile_grid(g), dim3(TBLOCK), s, le, g, op);
}

template <typename T>
__global__ __launch_bounds__(BLOCK) void ku_rows_kernel(Geom g, const T* u, const T* mu, T* y, int nrows) {
  const int64_t k = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (k >= (int64_t)nrows * g.NX) return;
  const int J = (int)(k / g.NX), I = (int)(k - (int64_t)J * g.NX);
  y[k] = ku_global<T>(g, u, mu, I, J);
}

template <typename T>
void launch_ku_rows(const Geom& g, const T* u, const T* mu, T* y, int nrows, hipStream_t s) {
  nrows = nrows < g.NY ? nrows : g.NY;
  const int64_t n = (int64_t)nrows * g.NX;
  hipLaunchKernelGGL((ku_rows_kernel<T>), dim3((unsigned)((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, s, g, u, mu, y,
                     nrows);
}

// --------------------------------------------------------------------------
// Two consecutive level-0 Chebyshev steps in one launch
// --------------------------------------------------------------------------
// The second step needs the first step's new direction d1 on the tile and on
// the ring of nodes its stencil reaches: one element beyond the tile along the
// owned rows (4 columns before, 1 after) and along the owned columns (4 rows
// before, 1 after).  The first step is therefore evaluated on the tile and
// that ring, from a source window one element wider again (columns I0-8 ..
// I0+TILE_I+4, rows J0-8 .. J0+TILE_J+4); d1 stays in LDS, r1 and x1 of the
// owned nodes in registers, and the second step runs on the tile from there.
// HBM traffic per node is that of one step (source, mu, r, D^-1 and x in;
// the second step's outputs out); the ring and the wider window are re-read
// from L2.  Each node value is formed by the same operations in the same
// order as in two OpCheb launches, so the results are those of two launches.
constexpr int FW_I = TILE_I + 13;
constexpr int FW_J = TILE_J + 13;
constexpr int FB_I = TILE_I + 5;               // box of the first step: columns I0-4 .. I0+TILE_I
constexpr int FB_J = TILE_J + 5;               // rows J0-4 .. J0+TILE_J
constexpr int FW_LOADS = (FW_J * FW_I + PBLOCK - 1) / PBLOCK;
constexpr int FB_STRIPS_I = (FB_I + 3) / 4;    // 4-node row strips per box row
constexpr int FB_STRIPS_J = (FB_J + 3) / 4;    // 4-node column strips per box column
constexpr int F_RING = 2 * TILE_I + 5 * (TILE_J / 4);   // column strips of the ring
static_assert(F_RING <= PBLOCK, "one ring strip per thread");
static_assert(FB_STRIPS_J == TILE_J / 4 + 2, "ring rows: one strip below, one above");

template <typename T>
struct ChebPair {
  // first step (OpCheb, mode 0)
  const T* src;
  const T* b;
  T s0;
  int src_mode;
  const T* coarse;
  int nelx, nely;
  T gx[5];
  const T* rin;
  T c1a, c2a;
  const T* xin;
  int xsrc_center;
  T* x1out;          // second step in mode 1: the first step's x goes here
  // second step (OpCheb with the first step's outputs as inputs)
  T c1b, c2b;
  int mode;
  T* rout;
  T* dout;
  T* xout;
  const T* rdot;
  double* xout64;          // mode 2 widened output and FP64 dot operand (see ChebStep)
  const double* rdot64;
  // both
  const T* mu;
  const T* dinv;
  int check;
  const double* part_rr;
  int nrr;
  CgScal* sc;
  double* part0;
};

template <typename T>
__global__ __launch_bounds__(PBLOCK) void cheb_pair_kernel(Geom g, ChebPair<T> op) {
  __shared__ T sw[FW_J * FW_I];    // first step's source
  __shared__ T smw[FW_J * FW_I];   // mu
  __shared__ T sbx[FB_J * FB_I];   // x-direction terms: first step on the box, then second step on the tile
  __shared__ T sd1[FB_J * FW_I];   // d1 on the box, window column numbering
  if (op.sc->done) return;         // block-uniform, as OpCheb::prepare
  if (op.check) {
    const double rr = reduce_parts(op.part_rr, op.nrr);
    const double resid = sqrt(rr);
    const bool fin = (resid <= op.sc->reltol) || (op.sc->iter >= op.sc->maxiter);
    if (threadIdx.x == 0 && blockIdx.x == 0 && blockIdx.y == 0) {
      op.sc->resid = resid;
      if (fin) op.sc->done = 1;
    }
    if (fin) return;
  }
  const int tile = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int I0 = (tile % gridDim.x) * TILE_I, J0 = (tile / gridDim.x) * TILE_J;
  const int tid = threadIdx.x;
  const int NX = g.NX, NY = g.NY;
  // owned column strip (both steps): column c3, rows Js3 .. Js3+3
  const int c3 = tid % TILE_I, q3 = tid / TILE_I;
  const int I3 = I0 + c3, Js3 = J0 + 4 * q3;
  // ring column strip of the first step (box column hc, box strip hk), if any
  int hc = -1, hk = 0;
  if (tid < 2 * TILE_I) {
    hc = 4 + tid % TILE_I;
    hk = tid < TILE_I ? 0 : FB_STRIPS_J - 1;
  } else if (tid < F_RING) {
    const int h = tid - 2 * TILE_I, col = h / (TILE_J / 4);
    hc = col < 4 ? col : FB_I - 1;
    hk = 1 + h % (TILE_J / 4);
  }
  // every global load of the block first
  T o_r[4], o_di[4], o_x[4], h_r[4], h_di[4];
  double o_rd[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    o_r[j] = o_di[j] = o_x[j] = h_r[j] = h_di[j] = T(0);
    o_rd[j] = 0.0;
    const int J = Js3 + j;
    if (I3 < NX && J > 0 && J < NY) {
      const int64_t n = (int64_t)J * NX + I3;
      o_r[j] = op.rin[n];
      o_di[j] = op.dinv[n];
      if (!op.xsrc_center) o_x[j] = op.xin[n];
      if (op.mode == 2) o_rd[j] = op.rdot64 ? op.rdot64[n] : (double)op.rdot[n];
    }
  }
  if (hc >= 0) {
    const int I = I0 - 4 + hc;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int J = J0 - 4 + 4 * hk + j;
      if (4 * hk + j < FB_J && I >= 0 && I < NX && J > 0 && J < NY) {
        const int64_t n = (int64_t)J * NX + I;
        h_r[j] = op.rin[n];
        h_di[j] = op.dinv[n];
      }
    }
  }
  {
    T ur[FW_LOADS], mr[FW_LOADS];
#pragma unroll
    for (int t = 0; t < FW_LOADS; t++) {
      const int k = tid + t * PBLOCK;
      const int r = k / FW_I, c = k - r * FW_I;
      const int I = I0 - 8 + c, J = J0 - 8 + r;
      ur[t] = 0;
      mr[t] = 0;
      if (k < FW_J * FW_I && I >= 0 && I < NX && J >= 0 && J < NY) {
        const int64_t n = (int64_t)J * NX + I;
        ur[t] = cheb_source(op.src, op.b, op.dinv, op.s0, op.src_mode, op.coarse, op.nelx, op.nely, op.gx, n, I,
                            J);
        mr[t] = op.mu[n];
      }
    }
#pragma unroll
    for (int t = 0; t < FW_LOADS; t++) {
      const int k = tid + t * PBLOCK;
      if (k < FW_J * FW_I) {
        sw[k] = ur[t];
        smw[k] = mr[t];
      }
    }
  }
  __syncthreads();
  // first step, x direction, on every box row
  for (int sidx = tid; sidx < FB_J * FB_STRIPS_I; sidx += PBLOCK) {
    const int rb = sidx / FB_STRIPS_I, k = sidx - rb * FB_STRIPS_I;
    const int J = J0 - 4 + rb, Is = I0 - 4 + 4 * k;
    if (J < 0 || J >= NY || Is < 0 || Is >= NX) continue;
    const int sg = Is >> 2;
    const int wr = rb + 4, wc = 4 * k + 4;
    T o[5] = {0, 0, 0, 0, 0};
    if (sg < g.nelx) seg_all<T, 1>(g, sw + wr * FW_I + wc, smw + wr * FW_I + wc, o);
    if (sg > 0) o[0] += seg_last<T, 1>(g, sw + wr * FW_I + wc - 4, smw + wr * FW_I + wc - 4);
    const T cx = (T)(g.c1 * wline(g, J, NY));
#pragma unroll
    for (int i = 0; i < 4; i++)
      if (4 * k + i < FB_I) sbx[rb * FB_I + 4 * k + i] = cx * o[i];
  }
  __syncthreads();
  // first step, y direction and update, on one box column strip: r1 = r - K d,
  // d1 = c1 d + c2 D^-1 r1 (0 on the fault row and outside the mesh)
  T r1[4], x1[4];
  auto step1 = [&](int cb, int kk, const T* rr, const T* di, bool owned) {
    const int I = I0 - 4 + cb, Js = J0 - 4 + 4 * kk;
    const int wc = cb + 4, wr = 4 * kk + 4;
    const bool inI = I >= 0 && I < NX;
    T o[5] = {0, 0, 0, 0, 0};
    if (inI && Js >= 0 && Js < NY) {
      const int t = Js >> 2;
      if (t < g.nely) seg_all<T, FW_I>(g, sw + wr * FW_I + wc, smw + wr * FW_I + wc, o);
      if (t > 0) o[0] += seg_last<T, FW_I>(g, sw + (wr - 4) * FW_I + wc, smw + (wr - 4) * FW_I + wc);
    }
    const T cy = inI ? (T)(g.c2 * wline(g, I, NX)) : T(0);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int rb = 4 * kk + j;
      if (rb >= FB_J) break;
      const int J = Js + j;
      T dn = T(0);
      if (owned) {
        r1[j] = T(0);
        x1[j] = T(0);
      }
      if (inI && J > 0 && J < NY) {
        const T ku = sbx[rb * FB_I + cb] + cy * o[j];
        const T wcv = sw[(rb + 4) * FW_I + wc];
        const T r = rr[j] - ku;
        dn = op.c1a * wcv + op.c2a * di[j] * r;
        if (owned) {
          r1[j] = r;
          x1[j] = (op.xsrc_center ? wcv : o_x[j]) + dn;
        }
      }
      sd1[rb * FW_I + wc] = dn;
    }
  };
  step1(c3 + 4, q3 + 1, o_r, o_di, true);
  if (hc >= 0) step1(hc, hk, h_r, h_di, false);
  __syncthreads();
  // second step, x direction, on the tile
  {
    constexpr int NSTRIP = TILE_I / 4;
    const int r = tid / NSTRIP, q = tid - r * NSTRIP;
    const int J = J0 + r, Is = I0 + 4 * q;
    if (J < NY && Is < NX) {
      const int sg = Is >> 2;
      const int dr = r + 4, wc = 4 * q + 8;
      T o[5] = {0, 0, 0, 0, 0};
      if (sg < g.nelx) seg_all<T, 1>(g, sd1 + dr * FW_I + wc, smw + (dr + 4) * FW_I + wc, o);
      if (sg > 0) o[0] += seg_last<T, 1>(g, sd1 + dr * FW_I + wc - 4, smw + (dr + 4) * FW_I + wc - 4);
      const T cx = (T)(g.c1 * wline(g, J, NY));
#pragma unroll
      for (int i = 0; i < 4; i++) sbx[r * TILE_I + 4 * q + i] = cx * o[i];
    }
  }
  __syncthreads();
  // second step, y direction and the OpCheb epilogue of its mode
  double acc0 = 0.0;
  if (I3 < NX && Js3 < NY) {
    const int t = Js3 >> 2;
    const int wc = c3 + 8, dr = 4 * q3 + 4;
    T o[5] = {0, 0, 0, 0, 0};
    if (t < g.nely) seg_all<T, FW_I>(g, sd1 + dr * FW_I + wc, smw + (dr + 4) * FW_I + wc, o);
    if (t > 0) o[0] += seg_last<T, FW_I>(g, sd1 + (dr - 4) * FW_I + wc, smw + dr * FW_I + wc);
    const T cy = (T)(g.c2 * wline(g, I3, NX));
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int J = Js3 + j;
      if (J >= NY) break;
      const int64_t n = (int64_t)J * NX + I3;
      if (J == 0) {
        if (op.mode == 0) { op.rout[n] = T(0); op.dout[n] = T(0); op.xout[n] = T(0); }
        else if (op.mode == 1) { op.rout[n] = T(0); op.x1out[n] = T(0); }
        else if (op.xout64) op.xout64[n] = 0.0;
        else op.xout[n] = T(0);
        continue;
      }
      const T ku = sbx[(4 * q3 + j) * TILE_I + c3] + cy * o[j];
      const T wcv = sd1[(dr + j) * FW_I + wc];
      const T r = r1[j] - ku;
      if (op.mode == 1) {
        op.rout[n] = r;
        op.x1out[n] = x1[j];
        continue;
      }
      const T dn = op.c1b * wcv + op.c2b * o_di[j] * r;
      const T xn = x1[j] + dn;
      if (op.mode == 0) {
        op.xout[n] = xn;
        op.rout[n] = r;
        op.dout[n] = dn;
      } else {
        if (op.xout64) op.xout64[n] = (double)xn;
        else op.xout[n] = xn;
        acc0 += (double)xn * o_rd[j];
      }
    }
  }
  const double s = block_sum(acc0);
  if (threadIdx.x == 0) op.part0[tile] = s;
}

template <typename T>
void launch_cheb_pair(const Geom& g, const ChebStep<T>& a, const ChebStep<T>& b, const T* mu, const T* dinv,
                      const double* part_rr, int nrr, CgScal* sc, double* part_rho, hipStream_t s,
                      LaunchEvents le) {
  ChebPair<T> op{};
  op.src = a.src;
  op.b = a.b;
  op.s0 = a.s0;
  op.src_mode = a.src_mode;
  op.coarse = a.coarse;
  op.nelx = g.nelx;
  op.nely = g.nely;
  for (int p = 0; p < 5; p++) op.gx[p] = (T)g.x[p];
  op.rin = a.rin;
  op.c1a = a.c1;
  op.c2a = a.c2;
  op.xin = a.xin;
  op.xsrc_center = a.xsrc_center;
  op.x1out = a.xout;
  op.c1b = b.c1;
  op.c2b = b.c2;
  op.mode = b.mode;
  op.rout = b.rout;
  op.dout = b.dout;
  op.xout = b.xout;
  op.rdot = b.rdot;
  op.xout64 = b.xout64;
  op.rdot64 = b.rdot64;
  op.mu = mu;
  op.dinv = dinv;
  op.check = a.check;
  op.part_rr = part_rr;
  op.nrr = nrr;
  op.sc = sc;
  op.part0 = part_rho;
  launch_timed(cheb_pair_kernel<T>, tile_grid(g), dim3(PBLOCK), s, le, g, op);
}

// --------------------------------------------------------------------------
// A whole level-0 Chebyshev sweep per tile in one launch (restricted additive
// Schwarz with a ring of SW_RING nodes)
// --------------------------------------------------------------------------
// The block of a TILE_I x TILE_J tile works on the window W: the tile and a
// ring of SW_RING = 2 nodes around it (68 x 20 nodes).  The sweep's residual is
// restricted to W, its k - 1 products with K_II are those of K_II[W, W]
// (directions vanish outside W and on the fault row), and only the tile's own
// nodes are written:
//   x <- x + R_tile p(D_W^-1 K_W) D_W^-1 (b - K_II x)|_W
// with p the degree-k Chebyshev polynomial of the global sweep.  The stencil of
// a node of W reaches the elements that hold it, so the LDS arrays cover the
// element-aligned box around W: columns I0-4 .. I0+TILE_I+4 and rows
// J0-4 .. J0+TILE_J+4 (SW_I x SW_J); W starts at box index SW_OFF.
//   pre  (post = 0): from x = 0: r = b, d0 = s0 D^-1 b, x = d0, then for
//        m = 1 .. k-1: r -= K_W d, d = c1[m] d + c2[m] D^-1 r, x += d.
//   post (post = 1): v = src (+ P0 coarse, formed in the box load) on the whole
//        box; r = b - K_II v on W, d0 = s0 D^-1 r, x = v + d0, then the same
//        k - 1 local steps; output c = x and the tile's partial <c, rdot>.
// Each node of W keeps r, x, d and D^-1 in registers of the thread that owns
// its column strip; d goes through LDS (two buffers) for the stencils.
constexpr int SW_RING = 2;
constexpr int SW_OFF = 4 - SW_RING;              // box index of W's first row / column
constexpr int SW_WI = TILE_I + 2 * SW_RING;      // window columns
constexpr int SW_WJ = TILE_J + 2 * SW_RING;      // window rows
constexpr int SW_I = TILE_I + 9;                 // box columns
constexpr int SW_J = TILE_J + 9;                 // box rows
constexpr int SW_BOX = SW_I * SW_J;
constexpr int SW_BLOCK = 448;
constexpr int SW_XS = TILE_I / 4 + 2;            // x-strips (element segments) per window row
constexpr int SW_YS = TILE_J / 4 + 2;            // y-strips per window column
constexpr int SW_BOX_LOADS = (SW_BOX + SW_BLOCK - 1) / SW_BLOCK;
constexpr int SW_SXS = SW_WI + 1;                // row stride of the x-term array (odd: no LDS bank conflicts)
constexpr int SW_VI = TILE_I / 4 + 4;            // element vertices I0/4 - 1 .. I0/4 + TILE_I/4 + 2: those
constexpr int SW_VJ = TILE_J / 4 + 4;            // of the box's elements
static_assert(SW_VI * SW_VJ <= SW_BLOCK, "one vertex value per thread");
static_assert(SW_WJ * SW_XS <= SW_BLOCK && SW_WI * SW_YS <= SW_BLOCK, "one strip per thread and pass");
static_assert(TILE_I % 4 == 0 && TILE_J % 4 == 0, "element-aligned tiles");

// Segment outputs o[0..4] of one 5-node segment (strided view, stride ST)
// with the quadrature weight folded into wm (wm_p = w_p mu_p):
// o_i = sum_p D[p][i] wm_p sum_k D[p][k] u_k.
template <typename T, int ST>
__device__ __forceinline__ void seg_all_wm(const Geom& g, const T* u, const T* wm, T o[5]) {
  T t[5];
#pragma unroll
  for (int p = 0; p < 5; p++) {
    T s = 0;
#pragma unroll
    for (int k = 0; k < 5; k++) s += (T)g.D[p * 5 + k] * u[k * ST];
    t[p] = s * wm[p * ST];
  }
#pragma unroll
  for (int i = 0; i < 5; i++) {
    T r = 0;
#pragma unroll
    for (int p = 0; p < 5; p++) r += (T)g.D[p * 5 + i] * t[p];
    o[i] = r;
  }
}

// The same segment as one symmetric 5 x 5 operator, o = M u with
// M = s D^T diag(wm) D (s: a row factor folded in), stored as its 15 upper
// entries (row-major, i <= k): M = sum_p (s wm_p) dd[p], dd[p] the products
// D[p][i] D[p][k].
template <int ST>
__device__ __forceinline__ void seg_matrix(const float (&dd)[5][15], const float* wm, bool on, float s,
                                           float m[15]) {
  float w[5];
#pragma unroll
  for (int p = 0; p < 5; p++) w[p] = on ? s * wm[p * ST] : 0.0f;
#pragma unroll
  for (int e = 0; e < 15; e++) {
    float a = dd[0][e] * w[0];
#pragma unroll
    for (int p = 1; p < 5; p++) a += dd[p][e] * w[p];
    m[e] = a;
  }
}
template <typename T, int ST>
__device__ __forceinline__ void seg_apply(const T* u, const T m[15], T o[5]) {
  T v[5];
#pragma unroll
  for (int k = 0; k < 5; k++) v[k] = u[k * ST];
  constexpr int row[5] = {0, 5, 9, 12, 14};   // index of m(i, i)
#pragma unroll
  for (int i = 0; i < 5; i++) {
    T r = 0;
#pragma unroll
    for (int k = 0; k < 5; k++) r += m[i <= k ? row[i] + k - i : row[k] + i - k] * v[k];
    o[i] = r;
  }
}

// One product with K per step in two phases.  Phase A: every x-strip thread
// (a window row, an element segment along x) and every y-strip thread (a
// window column, a segment along y) evaluates its segment once: outputs 0..3
// go to the segment's first four nodes (x: the LDS term array; y: registers),
// output 4, the segment's end node, to a carry slot that the next strip's
// node 0 adds in phase B.  Phase B: the owner of each node of W forms K d
// there and updates r, d and x.
template <typename T>
__global__ __launch_bounds__(SW_BLOCK) void cheb_sweep_kernel(Geom g, ChebSweep<T> op) {
  __shared__ T sd[2][SW_BOX];        // directions (post: sd[0] first holds v)
  __shared__ T swx[SW_BOX];          // w_(I mod 4) mu: quadrature weight along x times mu
  __shared__ T swy[SW_BOX];          // w_(J mod 4) mu
  __shared__ T sx[SW_WJ * SW_SXS];   // x terms (times c1 wline(J)) on W
  __shared__ T cxr[SW_XS * SW_WJ];   // x carry of strip k in row r (k * SW_WJ + r): its segment's output 4,
                                     // times c1 wline(J)
  __shared__ T cyr[SW_YS * SW_WI];   // y carry of strip k: its segment's output 4
  __shared__ T scv[SW_VJ * SW_VI];   // post: the coarse (element-vertex) correction around the box
  // the convergence test's partials (and with init ||b||^2's) load first and are
  // summed below, while the tile's own loads are in flight
  const bool test = op.init || op.check;
  const double prr = test ? parts_load(op.part_rr, op.nrr) : 0.0;
  const double pbb = op.init ? parts_load(op.part_bb, op.nrr) : 0.0;
  const int done_in = op.init ? 0 : op.sc->done;
  const int tile = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int I0 = (tile % gridDim.x) * TILE_I, J0 = (tile / gridDim.x) * TILE_J;
  const int tid = threadIdx.x;
  const int NX = g.NX, NY = g.NY;
  // y-strip of this thread (and owner of its nodes): window column yc, box rows 4 yk .. 4 yk + 3
  const int yc = tid % SW_WI, yk = tid / SW_WI;
  const int Iy = I0 - SW_RING + yc, bcy = yc + SW_OFF;
  const bool ycol = tid < SW_WI * SW_YS && Iy >= 0 && Iy < NX;
  const int ytg = (J0 >> 2) - 1 + yk;   // element segment along y: box rows 4 yk .. 4 yk + 4
  const bool yseg = ytg >= 0 && ytg < g.nely;
  const T cy = ycol ? (T)(g.c2 * wline(g, Iy, NX)) : T(0);
  const int xs_own = bcy >> 2;          // x-strip holding this thread's column
  const bool xi0 = (bcy & 3) == 0;      // the column is a segment's node 0 (adds the previous strip's carry)
  // x-strip of this thread: window row xwr, box columns 4 xk .. 4 xk + 4 (rows vary fastest:
  // consecutive threads read LDS rows apart, an odd stride)
  const bool xtask = tid < SW_WJ * SW_XS;
  const int xk = tid / SW_WJ, xwr = tid - xk * SW_WJ;
  const int xbr = xwr + SW_OFF, xJ = J0 - 4 + xbr;
  const bool xrow = xtask && xJ >= 1 && xJ < NY;
  const int xsg = (I0 >> 2) - 1 + xk;
  const bool xseg = xsg >= 0 && xsg < g.nelx;
  const T cx = xrow ? (T)(g.c1 * wline(g, xJ, NY)) : T(0);
  T r[4], x[4], d[4], di[4];
  double rd[4];
  bool act[4], own[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int br = 4 * yk + j, J = J0 - 4 + br;
    act[j] = ycol && br >= SW_OFF && br < SW_OFF + SW_WJ && J >= 1 && J < NY;
    own[j] = ycol && br >= 4 && br < 4 + TILE_J && bcy >= 4 && bcy < 4 + TILE_I && J < NY;
    r[j] = x[j] = d[j] = di[j] = T(0);
    rd[j] = 0.0;
    if (act[j]) {
      const int64_t n = (int64_t)J * NX + Iy;
      r[j] = op.b[n];
      di[j] = op.dinv[n];
      if (op.post && own[j]) rd[j] = op.rdot64 ? op.rdot64[n] : (double)op.rdot[n];
    }
  }
  {
    // post with a coarse correction: the values at the vertices of the box's elements (vertex
    // columns I0/4 - 1 .. I0/4 + TILE_I/4 + 2, rows likewise; 0 outside the mesh and on the
    // fault row) go to LDS first
    const bool pcoarse = op.post && op.coarse;
    if (pcoarse && tid < SW_VI * SW_VJ) {
      const int a = tid % SW_VI, c = tid / SW_VI;
      const int vx = (I0 >> 2) - 1 + a, vy = (J0 >> 2) - 1 + c;
      T v = T(0);
      if (vx >= 0 && vx <= op.nelx && vy >= 1 && vy <= op.nely) v = op.coarse[(int64_t)(vy - 1) * (op.nelx + 1) + vx];
      scv[tid] = v;
    }
    T ub[SW_BOX_LOADS], mb[SW_BOX_LOADS];
#pragma unroll
    for (int t = 0; t < SW_BOX_LOADS; t++) {
      const int k = tid + t * SW_BLOCK;
      const int br = k / SW_I, bc = k - br * SW_I;
      const int I = I0 - 4 + bc, J = J0 - 4 + br;
      ub[t] = mb[t] = T(0);
      if (k < SW_BOX && I >= 0 && I < NX && J >= 0 && J < NY) {
        const int64_t n = (int64_t)J * NX + I;
        mb[t] = op.mu[n];
        if (op.post && J > 0) ub[t] = op.src[n];
      }
    }
    // the test (block-uniform)
    if (done_in) return;               // as OpCheb::prepare
    if (op.init) {
      // a solve's first kernel: block 0 writes the scalars (nothing else of this launch reads them)
      const double resid = sqrt(parts_total(prr)), bnorm = sqrt(parts_total(pbb));
      const bool fin = resid <= op.tol * bnorm || op.maxiter <= 0;
      if (threadIdx.x == 0 && blockIdx.x == 0 && blockIdx.y == 0) {
        CgScal h{};
        h.rho_prev = 1.0;
        h.resid = resid;
        h.bnorm = bnorm;
        h.reltol = op.tol * bnorm;
        h.iter = 0;
        h.maxiter = op.maxiter;
        h.done = fin ? 1 : 0;
        *op.sc = h;
      }
      if (fin) return;
    } else if (op.check) {
      const double resid = sqrt(parts_total(prr));
      const bool fin = (resid <= op.sc->reltol) || (op.sc->iter >= op.sc->maxiter);
      if (threadIdx.x == 0 && blockIdx.x == 0 && blockIdx.y == 0) {
        op.sc->resid = resid;
        if (fin) op.sc->done = 1;
      }
      if (fin) return;
    }
    if (pcoarse) __syncthreads();
#pragma unroll
    for (int t = 0; t < SW_BOX_LOADS; t++) {
      const int k = tid + t * SW_BLOCK;
      if (k < SW_BOX) {
        const int br = k / SW_I, bc = k - br * SW_I;   // I0, J0 are multiples of 4: local index = box index mod 4
        const int I = I0 - 4 + bc, J = J0 - 4 + br;
        T v = ub[t];
        if (pcoarse && I >= 0 && I < NX && J > 0 && J < NY) {
          // bilinear interpolation of the vertex values, as cheb_source forms it
          const int ex = min(I >> 2, op.nelx - 1), i = I - 4 * ex;
          const int ey = min(J >> 2, op.nely - 1), j = J - 4 * ey;
          const T ax1 = T(0.5) * (T(1) + op.gx[i]), ax0 = T(1) - ax1;
          const T ay1 = T(0.5) * (T(1) + op.gx[j]), ay0 = T(1) - ay1;
          const T* c0 = scv + (ey - (J0 >> 2) + 1) * SW_VI + ex - (I0 >> 2) + 1;   // vertex (ex, ey)
          T p = T(0);
          if (ey > 0) p += ay0 * (ax0 * c0[0] + ax1 * c0[1]);
          p += ay1 * (ax0 * c0[SW_VI] + ax1 * c0[SW_VI + 1]);
          v = v + p;
        }
        swx[k] = (T)g.w[bc & 3] * mb[t];
        swy[k] = (T)g.w[br & 3] * mb[t];
        sd[0][k] = v;
        sd[1][k] = T(0);
      }
    }
  }
  __syncthreads();
  // FP32: the thread's x and y segments (the same in every step) as 5 x 5
  // operators in registers, the x one times cx: a segment then costs 25
  // multiply-adds instead of 55 and reads no weights from LDS
  constexpr bool kSegM = std::is_same<T, float>::value;
  T mx[15], my[15];
  if constexpr (kSegM) {
    const bool xon = xrow && xseg, yon = ycol && yseg;
    seg_matrix<1>(op.dd, swx + (xon ? xbr * SW_I + 4 * xk : 0), xon, cx, mx);
    seg_matrix<SW_I>(op.dd, swy + (yon ? 4 * yk * SW_I + bcy : 0), yon, 1.0f, my);
  }
  // phase A on direction su: x terms and carries to LDS, this thread's y outputs to oy
  T oy[5];
  auto phase_a = [&](const T* su) {
    if (xrow) {
      T o[5] = {0, 0, 0, 0, 0};
      if constexpr (kSegM) {
        if (xseg) seg_apply<T, 1>(su + xbr * SW_I + 4 * xk, mx, o);
      } else {
        if (xseg) seg_all_wm<T, 1>(g, su + xbr * SW_I + 4 * xk, swx + xbr * SW_I + 4 * xk, o);
#pragma unroll
        for (int i = 0; i < 5; i++) o[i] = cx * o[i];
      }
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int bc = 4 * xk + i;
        if (bc >= SW_OFF && bc < SW_OFF + SW_WI) sx[xwr * SW_SXS + bc - SW_OFF] = o[i];
      }
      cxr[xk * SW_WJ + xwr] = o[4];
    }
#pragma unroll
    for (int j = 0; j < 5; j++) oy[j] = T(0);
    if (ycol) {
      if constexpr (kSegM) {
        if (yseg) seg_apply<T, SW_I>(su + 4 * yk * SW_I + bcy, my, oy);
      } else {
        if (yseg) seg_all_wm<T, SW_I>(g, su + 4 * yk * SW_I + bcy, swy + 4 * yk * SW_I + bcy, oy);
      }
      cyr[yk * SW_WI + yc] = oy[4];
    }
  };
  // phase B: (K su) at this thread's active nodes
  auto phase_b = [&](T ku[4]) {
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (act[j]) {
        const int wr = 4 * yk + j - SW_OFF;
        T X = sx[wr * SW_SXS + yc];
        if (xi0 && xs_own > 0) X += cxr[(xs_own - 1) * SW_WJ + wr];
        T Y = oy[j];
        if (j == 0 && yk > 0) Y += cyr[(yk - 1) * SW_WI + yc];
        ku[j] = X + cy * Y;
      }
  };
  int cur = 0;
  if (!op.post) {
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (act[j]) {
        d[j] = op.s0 * di[j] * r[j];
        x[j] = d[j];
        sd[0][(4 * yk + j) * SW_I + bcy] = d[j];
      }
    __syncthreads();
  } else {
    phase_a(sd[0]);
    __syncthreads();
    T ku[4];
    phase_b(ku);
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (act[j]) {
        const int k = (4 * yk + j) * SW_I + bcy;
        r[j] = r[j] - ku[j];
        d[j] = op.s0 * di[j] * r[j];
        x[j] = sd[0][k] + d[j];
        sd[1][k] = d[j];
      }
    __syncthreads();
    cur = 1;
  }
#pragma unroll
  for (int m = 1; m < kSweepMaxDeg; m++) {   // unrolled: c1[m], c2[m] at constant offsets
    if (m >= op.degree) break;
    T* nxt = sd[1 - cur];
    phase_a(sd[cur]);
    if (op.post && m == 1)   // sd[0] held v on the whole box: clear it before it takes directions
      for (int k = tid; k < SW_BOX; k += SW_BLOCK) nxt[k] = T(0);
    __syncthreads();
    T ku[4];
    phase_b(ku);
    const T c1 = op.c1[m], c2 = op.c2[m];
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (act[j]) {
        r[j] = r[j] - ku[j];
        d[j] = c1 * d[j] + c2 * di[j] * r[j];
        x[j] = x[j] + d[j];
        nxt[(4 * yk + j) * SW_I + bcy] = d[j];
      }
    __syncthreads();
    cur = 1 - cur;
  }
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++)
    if (own[j]) {
      const int J = J0 - 4 + 4 * yk + j;
      const int64_t n = (int64_t)J * NX + Iy;   // x = 0 on the fault row
      if (!op.post) {
        op.xout[n] = x[j];
      } else {
        if (op.xout64) op.xout64[n] = (double)x[j];
        else op.xout[n] = x[j];
        acc += (double)x[j] * rd[j];
      }
    }
  if (op.post) {
    const double s = block_sum(acc);
    if (threadIdx.x == 0) op.part0[tile] = s;
  }
}

template <typename T>
void launch_cheb_sweep(const Geom& g, const ChebSweep<T>& op, hipStream_t s, LaunchEvents le) {
  launch_timed(cheb_sweep_kernel<T>, tile_grid(g), dim3(SW_BLOCK), s, le, g, op);
}

// --------------------------------------------------------------------------
// Residual after pre-smoothing, restricted to the element vertices (P0^T)
// --------------------------------------------------------------------------
// b1 = P0^T (b - K_II x) without forming the residual in HBM.  The tile owns
// the vertices (vx, vy) with 4 vx in [I0, I0 + TILE_I) and 4 vy in [J0, J0 +
// TILE_J), vy >= 1; the bilinear hat of a vertex covers the nodes within 3 of
// it in each direction (weight 1 at the vertex, 0.5 (1 -+ xi) at the GLL nodes
// of the elements on either side).  So the residual is formed on the tile and
// the 3 columns / rows before it, RS_I x RS_J nodes, all of which the tile
// kernel's window (x and mu on columns I0-4 .. I0+TILE_I, rows J0-4 ..
// J0+TILE_J) covers with their elements.  The fault row J = 0 is not an
// unknown (r = 0 there).
constexpr int RS_I = TILE_I + 3;    // residual columns I0-3 .. I0+TILE_I-1 (window columns 1 ..)
constexpr int RS_J = TILE_J + 3;
constexpr int RS_XS = TILE_I / 4 + 1;   // x-strips (segments at window columns 4k .. 4k+4) reaching them
constexpr int RS_YS = TILE_J / 4 + 1;
constexpr int RS_BLOCK = 256;
constexpr int RS_LOADS = (WIN_J * WIN_I + RS_BLOCK - 1) / RS_BLOCK;
static_assert(RS_J * RS_XS <= 2 * RS_BLOCK && RS_I * RS_YS <= 2 * RS_BLOCK, "two strips per thread and pass");
static_assert((TILE_I / 4) * (TILE_J / 4) <= RS_BLOCK, "one vertex per thread");

template <typename T>
__global__ __launch_bounds__(RS_BLOCK) void restrict_kernel(Geom g, const T* __restrict__ x, const T* __restrict__ b,
                                                            const T* __restrict__ mu, T* __restrict__ b1,
                                                            const CgScal* sc) {
  __shared__ T su[WIN_J * WIN_I];
  __shared__ T sm[WIN_J * WIN_I];
  __shared__ T sr[RS_J * RS_I];   // x-direction terms, then the residual
  if (sc->done) return;
  const int tile = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int I0 = (tile % gridDim.x) * TILE_I, J0 = (tile / gridDim.x) * TILE_J;
  const int tid = threadIdx.x, NX = g.NX, NY = g.NY;
  // b on the residual region: thread -> column c of RS_I, rows q, q + RS_BLOCK / RS_I ...
  constexpr int NB = (RS_I * RS_J + RS_BLOCK - 1) / RS_BLOCK;
  T br_[NB];
#pragma unroll
  for (int t = 0; t < NB; t++) {
    const int k = tid + t * RS_BLOCK;
    const int rr = k / RS_I, cc = k - rr * RS_I;
    const int I = I0 - 3 + cc, J = J0 - 3 + rr;
    br_[t] = T(0);
    if (k < RS_I * RS_J && I >= 0 && I < NX && J >= 1 && J < NY) br_[t] = b[(int64_t)J * NX + I];
  }
  {
    T ur[RS_LOADS], mr[RS_LOADS];
#pragma unroll
    for (int t = 0; t < RS_LOADS; t++) {
      const int k = tid + t * RS_BLOCK;
      const int rr = k / WIN_I, cc = k - rr * WIN_I;
      const int I = I0 - 4 + cc, J = J0 - 4 + rr;
      ur[t] = mr[t] = T(0);
      if (k < WIN_J * WIN_I && I >= 0 && I < NX && J >= 0 && J < NY) {
        const int64_t n = (int64_t)J * NX + I;
        ur[t] = J > 0 ? x[n] : T(0);
        mr[t] = mu[n];
      }
    }
#pragma unroll
    for (int t = 0; t < RS_LOADS; t++) {
      const int k = tid + t * RS_BLOCK;
      if (k < WIN_J * WIN_I) {
        su[k] = ur[t];
        sm[k] = mr[t];
      }
    }
  }
  __syncthreads();
  // x direction on window rows 1 .. RS_J, strips k: segment at window columns 4k .. 4k+4
  for (int task = tid; task < RS_J * RS_XS; task += RS_BLOCK) {
    const int rr = task / RS_XS, k = task - rr * RS_XS;
    const int wr = rr + 1, J = J0 - 4 + wr;
    if (J < 1 || J >= NY) continue;
    const int sg = (I0 >> 2) - 1 + k;
    const T* up = su + wr * WIN_I + 4 * k;
    const T* mp = sm + wr * WIN_I + 4 * k;
    T o[5] = {0, 0, 0, 0, 0};
    if (sg >= 0 && sg < g.nelx) seg_all<T, 1>(g, up, mp, o);
    if (k > 0 && sg > 0 && sg - 1 < g.nelx) o[0] += seg_last<T, 1>(g, up - 4, mp - 4);
    const T cx = (T)(g.c1 * wline(g, J, NY));
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int wc = 4 * k + i;   // residual column wc - 1
      if (wc >= 1 && wc <= RS_I) sr[rr * RS_I + wc - 1] = cx * o[i];
    }
  }
  __syncthreads();
  // y direction on window columns 1 .. RS_I; r = b - K x into sr
  for (int task = tid; task < RS_I * RS_YS; task += RS_BLOCK) {
    const int cc = task % RS_I, k = task / RS_I;
    const int wc = cc + 1, I = I0 - 4 + wc;
    if (I < 0 || I >= NX) continue;
    const int tg = (J0 >> 2) - 1 + k;
    const T* up = su + 4 * k * WIN_I + wc;
    const T* mp = sm + 4 * k * WIN_I + wc;
    T o[5] = {0, 0, 0, 0, 0};
    if (tg >= 0 && tg < g.nely) seg_all<T, WIN_I>(g, up, mp, o);
    if (k > 0 && tg > 0 && tg - 1 < g.nely) o[0] += seg_last<T, WIN_I>(g, up - 4 * WIN_I, mp - 4 * WIN_I);
    const T cy = (T)(g.c2 * wline(g, I, NX));
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int wr = 4 * k + j, rr = wr - 1, J = J0 - 4 + wr;
      if (wr >= 1 && wr <= RS_J && J >= 1 && J < NY) sr[rr * RS_I + cc] += cy * o[j];
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < NB; t++) {
    const int k = tid + t * RS_BLOCK;
    if (k < RS_I * RS_J) {
      const int rr = k / RS_I, cc = k - rr * RS_I;
      const int I = I0 - 3 + cc, J = J0 - 3 + rr;
      sr[k] = (I >= 0 && I < NX && J >= 1 && J < NY) ? br_[t] - sr[k] : T(0);
    }
  }
  __syncthreads();
  // one owned vertex per thread: (a, c) at residual column 4a + 3, row 4c + 3
  constexpr int NVX = TILE_I / 4;
  if (tid < NVX * (TILE_J / 4)) {
    const int a = tid % NVX, c = tid / NVX;
    const int vx = (I0 >> 2) + a, vy = (J0 >> 2) + c;
    if (vy >= 1 && vx <= g.nelx && vy <= g.nely) {
      T w[7];
      w[3] = T(1);
#pragma unroll
      for (int o = 1; o <= 3; o++) {
        w[3 - o] = (T)(0.5 * (1.0 + g.x[4 - o]));   // node 4 - o of the element before the vertex
        w[3 + o] = (T)(0.5 * (1.0 - g.x[o]));       // node o of the element after it
      }
      T s = T(0);
#pragma unroll
      for (int j = 0; j < 7; j++) {
        const T* row = sr + (4 * c + j) * RS_I + 4 * a;
        T t = T(0);
#pragma unroll
        for (int i = 0; i < 7; i++) t += w[i] * row[i];
        s += w[j] * t;
      }
      b1[(int64_t)(vy - 1) * (g.nelx + 1) + vx] = s;
    }
  }
}

template <typename T>
void launch_restrict(const Geom& g, const T* x, const T* b, const T* mu, T* b1, const CgScal* sc, hipStream_t s) {
  hipLaunchKernelGGL((restrict_kernel<T>), tile_grid(g), dim3(RS_BLOCK), 0, s, g, x, b, mu, b1, sc);
}

template <typename T, int MODE>
static void csr_dispatch(const DevCsr& M, const T* val, const T* x, const T* b, const T* dinv, T omega, T* y, T* xout,
                         const CgScal* sc, hipStream_t s) {
  const int64_t threads = M.n * (int64_t)M.lpr;
  const dim3 grid((unsigned)((threads + BLOCK - 1) / BLOCK));
  switch (M.lpr) {
    case 1:
      hipLaunchKernelGGL((csr_kernel<T, 1, MODE>), grid, dim3(BLOCK), 0, s, M.n, M.rowptr, M.col, val, x, b, dinv,
                         omega, y, xout, sc);
      break;
    case 2:
      hipLaunchKernelGGL((csr_kernel<T, 2, MODE>), grid, dim3(BLOCK), 0, s, M.n, M.rowptr, M.col, val, x, b, dinv,
                         omega, y, xout, sc);
      break;
    case 4:
      hipLaunchKernelGGL((csr_kernel<T, 4, MODE>), grid, dim3(BLOCK), 0, s, M.n, M.rowptr, M.col, val, x, b, dinv,
                         omega, y, xout, sc);
      break;
    case 8:
      hipLaunchKernelGGL((csr_kernel<T, 8, MODE>), grid, dim3(BLOCK), 0, s, M.n, M.rowptr, M.col, val, x, b, dinv,
                         omega, y, xout, sc);
      break;
    default:
      hipLaunchKernelGGL((csr_kernel<T, 16, MODE>), grid, dim3(BLOCK), 0, s, M.n, M.rowptr, M.col, val, x, b, dinv,
                         omega, y, xout, sc);
      break;
  }
}

template <typename T>
static const T* csr_vals(const DevCsr& M);
template <>
const double* csr_vals<double>(const DevCsr& M) { return M.val; }
template <>
const float* csr_vals<float>(const DevCsr& M) { return M.val32; }

template <typename T>
void launch_csr(CsrMode mode, const DevCsr& M, const T* x, const T* b, const T* dinv, T omega, T* y, T* xout,
                const CgScal* sc, hipStream_t s) {
  if (M.n == 0) return;
  const T* val = csr_vals<T>(M);
  switch (mode) {
    case CSR_SPMV: csr_dispatch<T, CSR_SPMV>(M, val, x, b, dinv, omega, y, xout, sc, s); break;
    case CSR_DOWN: csr_dispatch<T, CSR_DOWN>(M, val, x, b, dinv, omega, y, xout, sc, s); break;
    case CSR_PROLONG_ADD: csr_dispatch<T, CSR_PROLONG_ADD>(M, val, x, b, dinv, omega, y, xout, sc, s); break;
    default: csr_dispatch<T, CSR_POST>(M, val, x, b, dinv, omega, y, xout, sc, s); break;
  }
}

template <typename T>
void launch_csr_up(const DevCsr& U, const DevCsr& V, const T* b, const T* e, T* y, const CgScal* sc, hipStream_t s) {
  const T* uval = csr_vals<T>(U);
  const T* vval = csr_vals<T>(V);
  const int lpr = std::max(U.lpr, V.lpr);
  const int64_t threads = U.n * (int64_t)lpr;
  const dim3 grid((unsigned)((threads + BLOCK - 1) / BLOCK));
#define EQ_UP(L)                                                                                                   \
  hipLaunchKernelGGL((csr_up_kernel<T, L>), grid, dim3(BLOCK), 0, s, U.n, U.rowptr, U.col, uval, V.rowptr, V.col, \
                     vval, b, e, y, sc)
  switch (lpr) {
    case 1: EQ_UP(1); break;
    case 2: EQ_UP(2); break;
    case 4: EQ_UP(4); break;
    case 8: EQ_UP(8); break;
    default: EQ_UP(16); break;
  }
#undef EQ_UP
}

template <typename T>
void launch_dense(int64_t n, const T* inv, const T* b, T* x, const CgScal* sc, hipStream_t s) {
  const int64_t blocks = (n + BLOCK / 64 - 1) / (BLOCK / 64);
  hipLaunchKernelGGL((dense_kernel<T>), dim3((unsigned)blocks), dim3(BLOCK), 0, s, n, inv, b, x, sc);
}

__global__ void newmark_predict_kernel(int64_t n, const double* d, const double* v, const double* a, double dt,
                                       double half_dt_sq, double* P) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) P[i] = newmark_predict(d[i], v[i], a[i], dt, half_dt_sq);
}
__global__ void permute_kernel(int64_t n, const int32_t* perm, const double* src, double* dst, int to_structured) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (to_structured)
    dst[perm[k]] = src[k];
  else
    dst[k] = src[perm[k]];
}

void launch_permute(int64_t n, const int32_t* perm, const double* src, double* dst, int to_structured,
                    hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(permute_kernel, dim3(nblk(n)), dim3(BLOCK), 0, s, n, perm, src, dst, to_structured);
}

void launch_newmark_predict(int64_t n, const double* d, const double* v, const double* a, double dt,
                            double half_dt_sq, double* P, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(newmark_predict_kernel, dim3(nblk(n)), dim3(BLOCK), 0, s, n, d, v, a, dt, half_dt_sq, P);
}

void launch_newmark(const Geom& g, const FaultDev& fp, FaultState fs, const double* mu, const double* M,
                    const double* P, const double* v, const double* a, double* P2, double* v2, double* a2,
                    double dt, double half_dt, double half_dt_sq, int* flags, const NewmarkRecord& rec,
                    hipStream_t s, LaunchEvents le) {
  OpNewmark op{fp, fs, mu, M, P, v, a, P2, v2, a2, dt, half_dt, half_dt_sq, flags, g.NY, nullptr, nullptr,
               rec.vprev, rec.nrprev, rec.vpart_out, rec.cap_fault, rec.seis_idx, rec.seis_tile, rec.nseis,
               rec.cap_seis, (int)tile_grid(g).x};
  const dim3 tg = tile_grid(g);
  launch_timed(newmark_kernel, dim3(tg.x, tg.y + 1), dim3(TBLOCK), s, le, g, op);
}

void launch_qs_begin(const Geom& g, const FaultDev& fp, FaultState fs, const double* mu, double* d, double* dPre,
                     const double* v, double* b, double dt, int first_pass, hipStream_t s, double* db, double theta,
                     double* bold, const StepStatus* st_next, double dt_prev) {
  const int64_t N = first_pass ? (int64_t)g.NX * g.NY : g.NX;
  hipLaunchKernelGGL(qs_begin_kernel, dim3(nblk(N)), dim3(BLOCK), 0, s, g, fp, fs, mu, d, dPre, v, b, dt,
                     first_pass, db, theta, bold, st_next, dt_prev);
}

void launch_qs_fault(const Geom& g, const FaultDev& fp, FaultState fs, const double* mu, const double* d, double* v,
                     double dt, hipStream_t s, double* save) {
  hipLaunchKernelGGL(qs_fault_kernel, dim3(nblk(fp.nflt)), dim3(BLOCK), 0, s, g, fp, fs, mu, d, v, dt, save);
}

void launch_qs_end(const Geom& g, const FaultDev& fp, FaultState fs, const double* mu, double* d, const double* dPre,
                   double* v, double* a, double dt, double dtmin, double* save, CgScal* sc, const double* part_rr,
                   int nparts, CgScal* sc_mirror, StepStatus* st, int* flags, const StepCapture& cap, hipStream_t s) {
  const int64_t N = (int64_t)g.NX * g.NY;
  const int nb = (int)(1 + (N + QE_BLOCK - 1) / QE_BLOCK);
  hipLaunchKernelGGL(qs_end_kernel, dim3(nb), dim3(QE_BLOCK), 0, s, g, fp, fs, mu, d, dPre, v, a, dt, dtmin, save,
                     sc, part_rr, nparts, sc_mirror, st, flags, cap);
}

void launch_qs_mid(const Geom& g, const FaultDev& fp, FaultState fs, const double* mu, double* d, const double* dPre,
                   double* v, double* b, double dt, double* db, hipStream_t s, CgScal* sc, const double* part_rr,
                   int nparts, CgScal* sc_mirror) {
  hipLaunchKernelGGL(qs_mid_kernel, dim3(1), dim3(QE_BLOCK), 0, s, g, fp, fs, mu, d, dPre, v, b, dt, db, sc, part_rr,
                     nparts, sc_mirror);
}

void launch_step_tail(const FaultDev& fp, const double* d, const double* v, const double* a, const double* tau,
                      double dt, double dtmin, int isolver, StepStatus* st, int* flags, const StepCapture& cap,
                      hipStream_t s) {
  hipLaunchKernelGGL(step_tail_kernel, dim3(1), dim3(BLOCK), 0, s, fp, d, v, a, tau, dt, dtmin, isolver, st, flags,
                     cap);
}

void launch_ids(int64_t n, const double* xLf, const double* Vo, const double* psi, double dt, const double* Vf,
                double cnd, int IDstate, double* psi1, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(ids_kernel, dim3(nblk(n)), dim3(BLOCK), 0, s, n, xLf, Vo, psi, dt, Vf, cnd, IDstate, psi1);
}
void launch_ids2(int64_t n, const double* xLf, const double* Vo, const double* psi, const double* psi1, double dt,
                 const double* Vf, const double* Vf1, int IDstate, double* psi2, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(ids2_kernel, dim3(nblk(n)), dim3(BLOCK), 0, s, n, xLf, Vo, psi, psi1, dt, Vf, Vf1, IDstate,
                     psi2);
}
void launch_nrsearch(int64_t n, const double* fo, const double* Vo, const double* cca, const double* ccb,
                     const double* Seff, const double* tau, const double* tauo, const double* psi,
                     const double* FltZ, const double* FltVfree, double* Vf, double* tau_out, int* k,
                     hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(nrsearch_kernel, dim3(nblk(n)), dim3(BLOCK), 0, s, n, fo, Vo, cca, ccb, Seff, tau, tauo, psi,
                     FltZ, FltVfree, Vf, tau_out, k);
}
void launch_fbc(const FaultDev& fp, FaultState fs, const double* FltVfree, double dt, int* flags, hipStream_t s) {
  const int n = fp.nflt - fp.j0;
  if (n <= 0) return;
  hipLaunchKernelGGL(fbc_kernel, dim3(nblk(n)), dim3(BLOCK), 0, s, fp, fs, FltVfree, dt, flags);
}
void launch_slr(const FaultDev& fp, FaultState fs, double dt, hipStream_t s) {
  const int n = fp.nflt - fp.j0;
  if (n <= 0) return;
  hipLaunchKernelGGL(slr_kernel, dim3(nblk(n)), dim3(BLOCK), 0, s, fp, fs, dt);
}
void launch_dtevol(const double* XiLf, const double* Vf, int j0, int n, double dt, double dtmin, int isolver,
                   double* out, hipStream_t s) {
  hipLaunchKernelGGL(dtevol_kernel, dim3(1), dim3(BLOCK), 0, s, XiLf, Vf, j0, n, dt, dtmin, isolver, out);
}

template <typename T, typename U>
void launch_convert(int64_t n, const T* x, U* y, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((convert_kernel<T, U>), dim3(nblk(n)), dim3(BLOCK), 0, s, n, x, y);
}
template <typename T>
void launch_fill(int64_t n, T* y, T val, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((fill_kernel<T>), dim3(nblk(n)), dim3(BLOCK), 0, s, n, y, val);
}
template <typename T>
void launch_fill_strided(int64_t n, T* y, int stride, T val, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((fill_strided_kernel<T>), dim3(nblk(n)), dim3(BLOCK), 0, s, n, y, stride, val);
}
void launch_copy3(int64_t n, double* x0, double* x1, double* x2, double* y, bool gather, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(copy3_kernel, dim3(nblk(3 * n)), dim3(BLOCK), 0, s, n, x0, x1, x2, y, gather ? 1 : 0);
}
void launch_guess_dot(const Geom& g, const double* p, const double* mu, const double* r0, double* q,
                      double* part_pr, double* part_pq, hipStream_t s) {
  OpGuessDot op{p, mu, r0, q, part_pr, part_pq};
  hipLaunchKernelGGL((tile_kernel<double, OpGuessDot>), tile_grid(g), dim3(TBLOCK), 0, s, g, op);
}
void launch_guess_update(int64_t n, double* x, double* r, const double* p, const double* q, const double* part_pr,
                         const double* part_pq, int nparts, double* part_rr, int npart_out, hipStream_t s) {
  hipLaunchKernelGGL(guess_update_kernel, dim3(vec_blocks(n)), dim3(BLOCK), 0, s, n, x, r, p, q, part_pr, part_pq,
                     nparts, part_rr, npart_out);
}
void launch_widen_dot(int64_t n, const float* c32, const double* r, double* c, const CgScal* sc, double* part,
                      hipStream_t s) {
  hipLaunchKernelGGL(widen_dot_kernel, dim3(vec_blocks(n)), dim3(BLOCK), 0, s, n, c32, r, c, sc, part);
}
void launch_axpy(int64_t n, double alpha, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(axpy_kernel, dim3(nblk(n)), dim3(BLOCK), 0, s, n, alpha, x, y);
}
void launch_scale_mask(int64_t n, int nx, const double* x, double sign, double* y, hipStream_t s) {
  hipLaunchKernelGGL(scale_mask_kernel, dim3(nblk(n)), dim3(BLOCK), 0, s, n, nx, x, sign, y);
}

// explicit instantiations
#define EQ_INST(T)                                                                                                  \
  template void launch_ku<T>(const Geom&, const T*, const T*, T*, int, int, hipStream_t);                          \
  template void launch_cg_residual<T>(const Geom&, const T*, const T*, const T*, T*, double*, double*, int,         \
                                      hipStream_t, float*);                                                         \
  template void launch_cg_dir<T>(const Geom&, const T*, const T*, const T*, T*, T*, const double*, int,            \
                                 const CgScal*, double*, hipStream_t, LaunchEvents);                                \
  template void launch_cg_update<T>(int64_t, T*, T*, const T*, const T*, const double*, int, const double*, int,    \
                                    CgScal*, double*, int, float*, hipStream_t, LaunchEvents);                      \
  template void launch_jacobi_pc<T>(int64_t, const T*, const T*, T*, const double*, int, CgScal*, double*,          \
                                    hipStream_t);                                                                   \
  template void launch_l0_down<T>(const Geom&, const T*, const T*, const T*, T, T*, T*, const double*, int,         \
                                  CgScal*, hipStream_t);                                                            \
  template void launch_l0_post<T>(const Geom&, const T*, const T*, const T*, const T*, T, T*, double*,             \
                                  const CgScal*, hipStream_t);                                                      \
  template void launch_csr<T>(CsrMode, const DevCsr&, const T*, const T*, const T*, T, T*, T*, const CgScal*,       \
                              hipStream_t);                                                                         \
  template void launch_dense<T>(int64_t, const T*, const T*, T*, const CgScal*, hipStream_t);                      \
  template void launch_csr_up<T>(const DevCsr&, const DevCsr&, const T*, const T*, T*, const CgScal*, hipStream_t); \
  template void launch_ku_rows<T>(const Geom&, const T*, const T*, T*, int, hipStream_t);                         \
  template void launch_cheb_pair<T>(const Geom&, const ChebStep<T>&, const ChebStep<T>&, const T*, const T*,       \
                                    const double*, int, CgScal*, double*, hipStream_t, LaunchEvents);             \
  template void launch_cheb<T>(const Geom&, const ChebStep<T>&, const T*, const T*, const double*, int, CgScal*,   \
                               double*, hipStream_t, LaunchEvents);                                                \
  template void launch_cheb_sweep<T>(const Geom&, const ChebSweep<T>&, hipStream_t, LaunchEvents);            \
  template void launch_restrict<T>(const Geom&, const T*, const T*, const T*, T*, const CgScal*, hipStream_t);    \
  template void launch_fill<T>(int64_t, T*, T, hipStream_t);
EQ_INST(double)
EQ_INST(float)
template void launch_convert<double, float>(int64_t, const double*, float*, hipStream_t);
template void launch_convert<float, double>(int64_t, const float*, double*, hipStream_t);
template void launch_convert<double, double>(int64_t, const double*, double*, hipStream_t);
template void launch_fill<int>(int64_t, int*, int, hipStream_t);
template void launch_fill_strided<int>(int64_t, int*, int, int, hipStream_t);

}  // namespace eq

// Learned initial guess: snapshot store, POD basis, conv1d + MLP network with
// hand-written forward / backward / Adam kernels (BF16 weights, FP32 master
// weights and accumulation), single-workgroup inference.
//
// Network: input = fault slip increment dF (length L0) divided by |dF|;
// NL conv layers (C channels, kernel 5, stride 2, padding 2, ReLU); an MLP
// flat -> H (ReLU) -> NB.  The output approximates the POD coefficients of
// the interior displacement increment divided by |dF|: the quasi-static
// response is linear in dF, so the scale is factored out exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "eq_learn.h"

namespace eq {

namespace {
constexpr int LB = 256;     // block size of the element-wise kernels
constexpr int NBMAX = 64;   // max POD coefficients

__device__ __forceinline__ uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  const uint32_t r = 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
  return (uint16_t)((u + r) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }
__device__ __forceinline__ float ldw(const float* w, int64_t i) { return w[i]; }
__device__ __forceinline__ float ldw(const uint16_t* w, int64_t i) { return bf2f(w[i]); }

__device__ double lblock_sum(double v) {
  __shared__ double sh[LB / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0;
  for (int k = 0; k < LB / 64; k++) r += sh[k];
  return r;
}

inline int gblk(int64_t n) { return (int)((n + LB - 1) / LB); }

#define LCHK(call)                                       \
  do {                                                   \
    const hipError_t e_ = (call);                        \
    if (e_ != hipSuccess) {                              \
      if (err) *err = hipGetErrorString(e_);             \
      return 2;                                          \
    }                                                    \
  } while (0)

// ---- snapshots -------------------------------------------------------------
__global__ void record_kernel(const double* d, const double* dPre, int64_t nflt, int64_t off, int64_t nI, float* Fcol,
                              float* Scol) {
  const int64_t i = (int64_t)blockIdx.x * LB + threadIdx.x;
  if (i < nflt) Fcol[i] = (float)(d[i] - dPre[i]);
  if (i < nI) Scol[i] = (float)(d[off + i] - dPre[off + i]);
}

// |dF| and |F| (fault row of the new displacement) of one snapshot
__global__ __launch_bounds__(LB) void snap_norm_kernel(const double* d, const double* dPre, int64_t nflt,
                                                       double* out) {
  double a = 0.0, b = 0.0;
  for (int64_t i = threadIdx.x; i < nflt; i += LB) {
    const double df = d[i] - dPre[i];
    a += df * df;
    b += d[i] * d[i];
  }
  a = lblock_sum(a);
  b = lblock_sum(b);
  if (threadIdx.x == 0) {
    out[0] = sqrt(a);
    out[1] = sqrt(b);
  }
}

// inference: x = (d - base)[0:nflt] / |.|, scale stored in coef[NB]
__global__ __launch_bounds__(LB) void infer_input_kernel(const double* x, const double* base, int64_t nflt, float* a0,
                                                         float* coef, int NB) {
  __shared__ double s_n;
  double ss = 0.0;
  for (int64_t i = threadIdx.x; i < nflt; i += LB) {
    const double v = x[i] - base[i];
    ss += v * v;
  }
  ss = lblock_sum(ss);
  if (threadIdx.x == 0) {
    s_n = sqrt(ss);
    coef[NB] = (float)s_n;
  }
  __syncthreads();
  const double inv = s_n > 0 ? 1.0 / s_n : 0.0;
  for (int64_t i = threadIdx.x; i < nflt; i += LB) a0[i] = (float)((x[i] - base[i]) * inv);
}

// out[n] = act(bias[n] + W[n, :] . in), one wavefront per row; scale by
// *scale when given (the final layer multiplies back the input norm)
__global__ __launch_bounds__(LB) void gemv_rows_kernel(int N, int K, const float* in, const uint16_t* W,
                                                       const float* bias, float* out, int relu, const float* scale) {
  const int row = blockIdx.x * (LB / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  float acc = 0.f;
  if (row < N)
    for (int k = lane; k < K; k += 64) acc += bf2f(W[(int64_t)row * K + k]) * in[k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (row < N && lane == 0) {
    float v = acc + bias[row];
    if (relu) v = fmaxf(v, 0.f);
    if (scale) v *= *scale;
    out[row] = v;
  }
}

// G = S^T S for 64 x 64 tiles with bi <= bj; FP64 accumulation.
// Gram matrix G = S^T S of the snapshot columns (S: nI x m, column-major, FP32)
// on the matrix cores: V_MFMA_F64_16X16X4_F64 with the FP32 entries widened,
// so every product is exact and the sums run in FP64 as in an FP64 GEMM.  A
// workgroup (4 waves) owns a GM x GM tile of the upper triangle; each wave a
// GW x GW quarter as GB x GB blocks of 16 x 16.  Rows of S stream through LDS
// in chunks of GK, the next chunk's global loads in flight while the current
// one feeds the MFMAs.  Few output tiles (a Gram of the 64 basis vectors)
// split the rows over gridDim.z workgroups per tile, each writing its partial
// Gram at G + z m^2; gram_sum_kernel adds the partials in z order.
//   Operand layout of the 16x16x4 FP64 MFMA (lane l, 64 lanes):
//   A (16 x 4): A[l % 16][l / 16];  B (4 x 16): B[l / 16][l % 16];
//   D (16 x 16, four FP64 registers r): D[4 r + l / 16][l % 16] (accumulator
//   groups of one row, unlike the FP32 16x16 forms' groups of four).
typedef double gram_d4 __attribute__((ext_vector_type(4)));
constexpr int GM = 128;                      // output tile side
constexpr int GW = GM / 2, GB = GW / 16;     // a wave's quarter, in 16 x 16 blocks per side
constexpr int GK = 32;                       // rows of S per LDS chunk
constexpr int GLD = (GK * GM) / 256;         // loads per thread per operand and chunk
static_assert(GLD * 256 == GK * GM, "whole loads");

__global__ __launch_bounds__(256) void gram_kernel(const float* S, int64_t nI, int m, int64_t rows_per_split,
                                                   double* G) {
  const int bi = blockIdx.y, bj = blockIdx.x;
  if (bi > bj) return;   // block-uniform: the upper triangle only
  const int64_t rbeg = (int64_t)blockIdx.z * rows_per_split;
  const int64_t rend = std::min<int64_t>(nI, rbeg + rows_per_split);
  G += (int64_t)blockIdx.z * m * m;
  __shared__ float As[GK][GM + 1];
  __shared__ float Bs[GK][GM + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wi = (w >> 1) * GW, wj = (w & 1) * GW;   // the wave's quarter of the tile
  const int li = lane & 15, lk = lane >> 4;
  gram_d4 acc[GB][GB];
#pragma unroll
  for (int a = 0; a < GB; a++)
#pragma unroll
    for (int b = 0; b < GB; b++) acc[a][b] = gram_d4{0.0, 0.0, 0.0, 0.0};
  // chunk loads: element e = tid + 256 q is column e / GK, row e % GK of the
  // chunk (consecutive threads read consecutive rows of one column)
  float pa[GLD], pb[GLD];
  auto fetch = [&](int64_t r0) {
#pragma unroll
    for (int q = 0; q < GLD; q++) {
      const int e = tid + 256 * q, c = e / GK, rr = e - c * GK;
      const int64_t row = r0 + rr;
      const int ca = bi * GM + c, cb = bj * GM + c;
      pa[q] = (row < rend && ca < m) ? S[(int64_t)ca * nI + row] : 0.f;
      pb[q] = (row < rend && cb < m) ? S[(int64_t)cb * nI + row] : 0.f;
    }
  };
  fetch(rbeg);
  for (int64_t r0 = rbeg; r0 < rend; r0 += GK) {
#pragma unroll
    for (int q = 0; q < GLD; q++) {
      const int e = tid + 256 * q, c = e / GK, rr = e - c * GK;
      As[rr][c] = pa[q];
      Bs[rr][c] = pb[q];
    }
    __syncthreads();
    if (r0 + GK < rend) fetch(r0 + GK);
#pragma unroll
    for (int ks = 0; ks < GK / 4; ks++) {
      const int k = 4 * ks + lk;
      double a[GB], b[GB];
#pragma unroll
      for (int t = 0; t < GB; t++) {
        a[t] = (double)As[k][wi + 16 * t + li];
        b[t] = (double)Bs[k][wj + 16 * t + li];
      }
#pragma unroll
      for (int ta = 0; ta < GB; ta++)
#pragma unroll
        for (int tb = 0; tb < GB; tb++)
          acc[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ta], b[tb], acc[ta][tb], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int ta = 0; ta < GB; ta++)
#pragma unroll
    for (int tb = 0; tb < GB; tb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = bi * GM + wi + 16 * ta + 4 * r + lk, j = bj * GM + wj + 16 * tb + li;
        if (i < m && j < m) G[(int64_t)i * m + j] = acc[ta][tb][r];
      }
}

// G[i, j] = sum_z part[z][i, j] (z in order) on the upper triangle
__global__ void gram_sum_kernel(int m, int nz, const double* part, double* G) {
  const int64_t e = (int64_t)blockIdx.x * LB + threadIdx.x;
  if (e >= (int64_t)m * m) return;
  const int i = (int)(e / m), j = (int)(e - (int64_t)i * m);
  if (i > j) return;
  double s = 0.0;
  for (int z = 0; z < nz; z++) s += part[(int64_t)z * m * m + e];
  G[e] = s;
}

// Upper triangle of S^T S (S: nI x m column-major) into G (m x m, row-major,
// device); partials from the row split in scratch (nz m^2 doubles, nz from
// gram_splits) when more than one.
inline int gram_splits(int64_t nI, int m) {
  const int nt = (m + GM - 1) / GM, tiles = nt * (nt + 1) / 2;
  int nz = 1;
  while (tiles * nz < 512 && (nI + GK - 1) / GK >= 32 * (int64_t)nz * 2) nz *= 2;
  return nz;
}
hipError_t launch_gram(const float* S, int64_t nI, int m, double* G, double* part, hipStream_t s) {
  const int nt = (m + GM - 1) / GM, nz = gram_splits(nI, m);
  int64_t rps = (nI + nz - 1) / nz;
  rps = (rps + GK - 1) / GK * GK;
  hipLaunchKernelGGL(gram_kernel, dim3(nt, nt, nz), dim3(256), 0, s, S, nI, m, rps, nz > 1 ? part : G);
  if (nz > 1)
    hipLaunchKernelGGL(gram_sum_kernel, dim3(gblk((int64_t)m * m)), dim3(LB), 0, s, m, nz, part, G);
  return hipGetLastError();
}

// ---- network: training kernels --------------------------------------------
__global__ void gather_kernel(const float* X, int L0, const int32_t* idx, int B, float* out) {
  const int64_t e = (int64_t)blockIdx.x * LB + threadIdx.x;
  if (e >= (int64_t)B * L0) return;
  const int b = (int)(e / L0), i = (int)(e - (int64_t)b * L0);
  out[e] = X[(int64_t)idx[b] * L0 + i];
}

// ---- training GEMMs on the matrix cores --------------------------------------
// C[M x N] = A[M x K] B[K x N] with V_MFMA_F32_16X16X4_F32: FP32 operands (the
// BF16 weights widen exactly), FP32 accumulation, so the products are those of
// an FP32 GEMM.  The operands are functors (im2col gathers for the
// convolutions, transposed views for the linear layers) and the output goes
// through an epilogue functor.  A workgroup (4 waves) owns a 64 x 64 tile, each
// wave a 32 x 32 quarter as 2 x 2 blocks of 16 x 16; K streams through LDS in
// chunks of 16.  gridDim.z > 1 splits K: each split writes its FP32 partial
// tile to part + z M N and gemm_finish_kernel sums the splits in order (FP64)
// and applies the epilogue.
//   Operand layout of the 16x16x4 FP32 MFMA (lane l): A[l % 16][l / 16],
//   B[l / 16][l % 16], D (four registers r) D[4 (l / 16) + r][l % 16].
typedef float gemm_f4 __attribute__((ext_vector_type(4)));
constexpr int QM = 64, QK = 16;

template <class AOp, class BOp, class Ep>
__global__ __launch_bounds__(256) void mfma_gemm_kernel(int M, int N, int K, int kps, AOp a, BOp b, Ep ep,
                                                        float* part) {
  __shared__ float As[QK][QM + 4];
  __shared__ float Bs[QK][QM + 4];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wi = (w >> 1) * 32, wj = (w & 1) * 32, li = lane & 15, lk = lane >> 4;
  const int m0 = blockIdx.y * QM, n0 = blockIdx.x * QM;
  const int kbeg = blockIdx.z * kps, kend = min(K, kbeg + kps);
  gemm_f4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = gemm_f4{0.f, 0.f, 0.f, 0.f};
  // element e = tid + 256 q of a chunk: row / column e % 64, k = e / 64; the
  // next chunk's operands are fetched into registers while this one computes
  constexpr int QL = (QM * QK) / 256;
  float pa[QL], pb[QL];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < QL; q++) {
      const int e = tid + 256 * q, mm = e & 63, kk = e >> 6;
      const int k = k0 + kk;
      pa[q] = (k < kend && m0 + mm < M) ? a(m0 + mm, k) : 0.f;
      pb[q] = (k < kend && n0 + mm < N) ? b(k, n0 + mm) : 0.f;
    }
  };
  fetch(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += QK) {
#pragma unroll
    for (int q = 0; q < QL; q++) {
      const int e = tid + 256 * q, mm = e & 63, kk = e >> 6;
      As[kk][mm] = pa[q];
      Bs[kk][mm] = pb[q];
    }
    __syncthreads();
    if (k0 + QK < kend) fetch(k0 + QK);
#pragma unroll
    for (int ks = 0; ks < QK / 4; ks++) {
      const int k = 4 * ks + lk;
      const float a0 = As[k][wi + li], a1 = As[k][wi + 16 + li];
      const float b0 = Bs[k][wj + li], b1 = Bs[k][wj + 16 + li];
      acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int ta = 0; ta < 2; ta++)
#pragma unroll
    for (int tb = 0; tb < 2; tb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = m0 + wi + 16 * ta + 4 * lk + r, j = n0 + wj + 16 * tb + li;
        if (i < M && j < N) {
          if (part) part[((int64_t)blockIdx.z * M + i) * N + j] = acc[ta][tb][r];
          else ep(i, j, acc[ta][tb][r]);
        }
      }
}

// Few splits: a thread per output element adds them in order.  Many: one wave
// per element, lane z adding splits z, z + 64, ..., then a shuffle tree (fixed
// order) summing the lanes; both in FP64.
template <class Ep>
__global__ __launch_bounds__(LB) void gemm_sum_kernel(int M, int N, int nz, const float* part, Ep ep) {
  const int64_t e = (int64_t)blockIdx.x * LB + threadIdx.x;
  if (e >= (int64_t)M * N) return;
  double v = 0.0;
  for (int z = 0; z < nz; z++) v += part[(int64_t)z * M * N + e];
  ep((int)(e / N), (int)(e % N), (float)v);
}

template <class Ep>
__global__ __launch_bounds__(LB) void gemm_finish_kernel(int M, int N, int nz, const float* part, Ep ep) {
  const int64_t e = (int64_t)blockIdx.x * (LB / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (e >= (int64_t)M * N) return;   // wave-uniform
  double v = 0.0;
  for (int z = lane; z < nz; z += 64) v += part[(int64_t)z * M * N + e];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if (lane == 0) ep((int)(e / N), (int)(e % N), (float)v);
}

// Launch C = A B with the K split so that the grid has about 256 workgroups
// (each split at least 64 deep; the partials must fit part_cap floats).
template <class AOp, class BOp, class Ep>
void gemm(int M, int N, int K, AOp a, BOp b, Ep ep, float* part, int64_t part_cap, hipStream_t s) {
  const int tm = (M + QM - 1) / QM, tn = (N + QM - 1) / QM, tiles = tm * tn;
  int nz = 1;
  while (tiles * nz * 2 <= 256 && K >= 64 * nz * 2 && (int64_t)nz * 2 * M * N <= part_cap) nz *= 2;
  int kps = (K + nz - 1) / nz;
  kps = (kps + QK - 1) / QK * QK;
  nz = (K + kps - 1) / kps;
  if (nz <= 1) {
    hipLaunchKernelGGL((mfma_gemm_kernel<AOp, BOp, Ep>), dim3(tn, tm, 1), dim3(256), 0, s, M, N, K, K, a, b, ep,
                       (float*)nullptr);
    return;
  }
  hipLaunchKernelGGL((mfma_gemm_kernel<AOp, BOp, Ep>), dim3(tn, tm, nz), dim3(256), 0, s, M, N, K, kps, a, b, ep,
                     part);
  if (nz <= 8)
    hipLaunchKernelGGL((gemm_sum_kernel<Ep>), dim3(gblk((int64_t)M * N)), dim3(LB), 0, s, M, N, nz, part, ep);
  else
    hipLaunchKernelGGL((gemm_finish_kernel<Ep>), dim3((unsigned)(((int64_t)M * N + LB / 64 - 1) / (LB / 64))),
                       dim3(LB), 0, s, M, N, nz, part, ep);
}

// operands --------------------------------------------------------------------
struct RowMat {   // A(r, c) = p[r ld + c]
  const float* p;
  int ld;
  __device__ float operator()(int r, int c) const { return p[(int64_t)r * ld + c]; }
};
struct ColMat {   // A(r, c) = p[c ld + r] (the transpose of a row-major matrix)
  const float* p;
  int ld;
  __device__ float operator()(int r, int c) const { return p[(int64_t)c * ld + r]; }
};
template <typename WT>
struct WRow {     // weights, row-major [rows x ld]: W(r, c)
  const WT* p;
  int ld;
  __device__ float operator()(int r, int c) const { return ldw(p, (int64_t)r * ld + c); }
};
template <typename WT>
struct WCol {     // the transpose of row-major weights: W^T(r, c) = W(c, r)
  const WT* p;
  int ld;
  __device__ float operator()(int r, int c) const { return ldw(p, (int64_t)c * ld + r); }
};
struct RowMatOnes {   // row-major [rows x n] with a column of ones appended at c = n
  const float* p;
  int n;
  __device__ float operator()(int r, int c) const { return c < n ? p[(int64_t)r * n + c] : 1.f; }
};
// im2col of a conv layer's input x [B x Cin x Lin] for stride 2, padding 2,
// kernel 5: row (b, o), column (ci, k) -> x[b, ci, 2o + k - 2] (0 outside)
struct ConvIn {
  const float* x;
  int Cin, Lin, Lout;
  __device__ float operator()(int m, int kk) const {
    const int b = m / Lout, o = m - b * Lout, ci = kk / 5, k = kk - 5 * ci;
    const int i = 2 * o + k - 2;
    return (i >= 0 && i < Lin) ? x[((int64_t)b * Cin + ci) * Lin + i] : 0.f;
  }
};
// the same gather with rows and columns exchanged, plus a column of ones at
// (ci, k) = Cin * 5 (the bias gradient): B((b, o), (ci, k))
struct ConvInT {
  const float* x;
  int Cin, Lin, Lout;
  __device__ float operator()(int kk, int n) const {
    if (n == 5 * Cin) return 1.f;
    const int b = kk / Lout, o = kk - b * Lout, ci = n / 5, k = n - 5 * ci;
    const int i = 2 * o + k - 2;
    return (i >= 0 && i < Lin) ? x[((int64_t)b * Cin + ci) * Lin + i] : 0.f;
  }
};
// dout [B x Cout x Lout] as A(co, (b, o))
struct ConvGradT {
  const float* g;
  int Cout, Lout;
  __device__ float operator()(int co, int kk) const {
    const int b = kk / Lout, o = kk - b * Lout;
    return g[((int64_t)b * Cout + co) * Lout + o];
  }
};
// the transposed convolution's gather of dout: row (b, i), column (co, k) ->
// dout[b, co, (i + 2 - k) / 2] when that index is whole and in range
struct ConvGradIm {
  const float* g;
  int Cout, Lin, Lout;
  __device__ float operator()(int m, int kk) const {
    const int b = m / Lin, i = m - b * Lin, co = kk / 5, k = kk - 5 * co;
    const int t = i + 2 - k;
    return (t >= 0 && (t & 1) == 0 && (t >> 1) < Lout) ? g[((int64_t)b * Cout + co) * Lout + (t >> 1)] : 0.f;
  }
};
// conv weights W [Cout x Cin x 5] as B((co, k), ci) = W[co, ci, k]
template <typename WT>
struct ConvWT {
  const WT* p;
  int Cin;
  __device__ float operator()(int kk, int ci) const {
    const int co = kk / 5, k = kk - 5 * co;
    return ldw(p, ((int64_t)co * Cin + ci) * 5 + k);
  }
};

// epilogues -------------------------------------------------------------------
struct EpDense {   // out[i N + j] = f(v + bias[j]), f = ReLU or identity
  float* out;
  const float* bias;
  int N, relu;
  __device__ void operator()(int i, int j, float v) const {
    v += bias[j];
    out[(int64_t)i * N + j] = relu ? fmaxf(v, 0.f) : v;
  }
};
struct EpMasked {  // out[i N + j] = v where the layer input (a ReLU output) is positive
  float* out;
  const float* in;
  int N;
  __device__ void operator()(int i, int j, float v) const {
    const int64_t e = (int64_t)i * N + j;
    out[e] = in[e] > 0.f ? v : 0.f;
  }
};
struct EpGrad {    // weight gradient [M x N]; column N is the bias gradient
  float* dW;
  float* db;
  int N;
  __device__ void operator()(int i, int j, float v) const {
    if (j < N) dW[(int64_t)i * N + j] = v;
    else db[i] = v;
  }
};
struct EpConv {    // conv output [B x Cout x Lout] from row (b, o), column co, with bias and ReLU
  float* out;
  const float* bias;
  int Cout, Lout;
  __device__ void operator()(int m, int co, float v) const {
    const int b = m / Lout, o = m - b * Lout;
    out[((int64_t)b * Cout + co) * Lout + o] = fmaxf(v + bias[co], 0.f);
  }
};
struct EpConvGradIn {   // din [B x Cin x Lin] from row (b, i), column ci, masked by the input
  float* din;
  const float* in;
  int Cin, Lin;
  __device__ void operator()(int m, int ci, float v) const {
    const int b = m / Lin, i = m - b * Lin;
    const int64_t e = ((int64_t)b * Cin + ci) * Lin + i;
    din[e] = in[e] > 0.f ? v : 0.f;
  }
};
struct EpColOut {       // column-major output: out[j ld + i] = v
  float* out;
  int64_t ld;
  __device__ void operator()(int i, int j, float v) const { out[(int64_t)j * ld + i] = v; }
};
struct EpConvGradW {    // dW [Cout x Cin x 5] from row co, column (ci, k); column Cin * 5 is db
  float* dW;
  float* db;
  int n;
  __device__ void operator()(int co, int j, float v) const {
    if (j < n) dW[(int64_t)co * n + j] = v;
    else db[co] = v;
  }
};

template <typename WT>
__global__ void conv_fwd_kernel(int B, int Cin, int Lin, int Cout, int Lout, const float* in, const WT* W,
                                const float* bias, float* out) {
  const int64_t e = (int64_t)blockIdx.x * LB + threadIdx.x;
  if (e >= (int64_t)B * Cout * Lout) return;
  const int o = (int)(e % Lout);
  const int co = (int)((e / Lout) % Cout);
  const int b = (int)(e / ((int64_t)Lout * Cout));
  float acc = bias[co];
  for (int ci = 0; ci < Cin; ci++) {
    const float* x = in + ((int64_t)b * Cin + ci) * Lin;
    const int64_t wb = ((int64_t)co * Cin + ci) * 5;
#pragma unroll
    for (int k = 0; k < 5; k++) {
      const int i = 2 * o + k - 2;
      if (i >= 0 && i < Lin) acc += ldw(W, wb + k) * x[i];
    }
  }
  out[e] = fmaxf(acc, 0.f);
}

// mean squared error over B x NB outputs; dout = d loss / d out
__global__ __launch_bounds__(LB) void mse_kernel(int B, int NB, const float* out, const float* Y, const int32_t* idx,
                                                 float* dout, double* part) {
  const int64_t e = (int64_t)blockIdx.x * LB + threadIdx.x;
  double l = 0.0;
  if (e < (int64_t)B * NB) {
    const int b = (int)(e / NB), j = (int)(e - (int64_t)b * NB);
    const float diff = out[e] - Y[(int64_t)idx[b] * NB + j];
    const float inv = 1.f / (float)(B * NB);
    dout[e] = 2.f * diff * inv;
    l = (double)diff * (double)diff / (double)(B * NB);
  }
  const double s = lblock_sum(l);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void adam_kernel(int64_t n, float* w, uint16_t* w16, const float* g, float* m, float* v, float lr,
                            float b1, float b2, float c1, float c2, float eps) {
  const int64_t i = (int64_t)blockIdx.x * LB + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  w16[i] = f2bf(wi);
}

__global__ void to_bf16_kernel(int64_t n, const float* w, uint16_t* w16) {
  const int64_t i = (int64_t)blockIdx.x * LB + threadIdx.x;
  if (i < n) w16[i] = f2bf(w[i]);
}

// ---- inference: one workgroup per launch -----------------------------------
constexpr int IB = 1024;   // threads of an inference workgroup

// a0[i] = (x[i] - base[i]) / |x - base| over the fault; returns |x - base|
// (every thread) and leaves it in coef[NB]
__device__ double infer_input(const double* x, const double* base, int64_t nflt, float* a0, float* coef, int NB) {
  __shared__ double sred[IB / 64];
  __shared__ double s_n;
  double ss = 0.0;
  for (int64_t i = threadIdx.x; i < nflt; i += IB) {
    const double v = x[i] - base[i];
    ss += v * v;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) ss += __shfl_down(ss, off, 64);
  if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = ss;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < IB / 64; k++) t += sred[k];
    s_n = sqrt(t);
    coef[NB] = (float)s_n;
  }
  __syncthreads();
  const double nrm = s_n;
  const double inv = nrm > 0 ? 1.0 / nrm : 0.0;
  for (int64_t i = threadIdx.x; i < nflt; i += IB) a0[i] = (float)((x[i] - base[i]) * inv);
  __syncthreads();
  return nrm;
}

// out[row] = act(bias[row] + W[row, :] . in) (times scale when scaled), one
// wave per row, lanes striding the row, as a shuffle-reduced sum
template <typename WT>
__device__ void rows_gemv(int N, int K, const float* in, const WT* W, const float* bias, float* out, int relu,
                          bool scaled, float scale) {
  const int lane = threadIdx.x & 63;
  for (int row = threadIdx.x >> 6; row < N; row += IB / 64) {
    float acc = 0.f;
    for (int k = lane; k < K; k += 64) acc += ldw(W, (int64_t)row * K + k) * in[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) {
      float v = acc + bias[row];
      if (relu) v = fmaxf(v, 0.f);
      if (scaled) v *= scale;
      out[row] = v;
    }
  }
}

// The linear predictor in one launch: coef = |dF| (A dF/|dF| + b).
__global__ __launch_bounds__(IB) void lin_infer_kernel(int64_t nflt, int NB, const double* x, const double* base,
                                                       const float* A, const float* b, float* buf, float* coef) {
  const double nrm = infer_input(x, base, nflt, buf, coef, NB);
  rows_gemv(NB, (int)nflt, buf, A, b, coef, 0, true, (float)nrm);
}

// p[off + i] = sum_k Phi[k, i] coef[k]; p[0:off] = 0
__global__ __launch_bounds__(LB) void guess_kernel(int64_t nI, int64_t off, int NB, const float* Phi,
                                                   const float* coef, double* p) {
  __shared__ float c[NBMAX];
  if (threadIdx.x < NB) c[threadIdx.x] = coef[threadIdx.x];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * LB + threadIdx.x;
  if (i < off) p[i] = 0.0;
  if (i >= nI) return;
  float acc = 0.f;
  for (int k = 0; k < NB; k++) acc += Phi[(int64_t)k * nI + i] * c[k];
  p[off + i] = (double)acc;
}

// ---- Galerkin start on span(Phi) --------------------------------------------
// Projection s = Phi^T r reads the row-major copy Phi_rm [nI x NB]: lane k of
// a wave owns column k, so a row is one coalesced 256-byte load and no
// cross-lane reduction is needed; r[off + i] is wave-uniform.
constexpr int PROJ_RPW = 64;                  // rows per wave (many waves: bytes in flight)
constexpr int PROJ_ROWS = 4 * PROJ_RPW;       // rows per block (4 waves)
static_assert(LB == 4 * NBMAX, "projection kernels: four waves of NBMAX lanes");

__global__ __launch_bounds__(LB) void proj_dot_kernel(int64_t nI, int64_t off, int NB, const float* Phi_rm,
                                                      const double* r, double* part) {
  __shared__ double sw[4][NBMAX];
  __shared__ double sr[4][PROJ_RPW];
  const int w = threadIdx.x >> 6, k = threadIdx.x & 63;
  const int64_t i0 = (int64_t)blockIdx.x * PROJ_ROWS + (int64_t)w * PROJ_RPW;
  // the wave's 64 values of r, one coalesced load, then LDS broadcasts
  sr[w][k] = (i0 + k < nI) ? r[off + i0 + k] : 0.0;
  __syncthreads();
  double acc = 0.0;
  if (k < NB) {
    if (i0 + PROJ_RPW <= nI) {
#pragma unroll 16
      for (int u = 0; u < PROJ_RPW; u++) acc += (double)Phi_rm[(i0 + u) * NB + k] * sr[w][u];
    } else {
      for (int u = 0; i0 + u < nI; u++) acc += (double)Phi_rm[(i0 + u) * NB + k] * sr[w][u];
    }
  }
  sw[w][k] = acc;
  __syncthreads();
  if (threadIdx.x < NB)
    part[(int64_t)blockIdx.x * NB + threadIdx.x] =
        (sw[0][threadIdx.x] + sw[1][threadIdx.x]) + (sw[2][threadIdx.x] + sw[3][threadIdx.x]);
}

// s[k] = sum_b part[b, k], one block per column, fixed order
__global__ __launch_bounds__(LB) void proj_sum_kernel(int nblk, int NB, const double* part, double* s) {
  const int k = blockIdx.x;
  double a = 0.0;
  for (int b = threadIdx.x; b < nblk; b += LB) a += part[(int64_t)b * NB + k];
  a = lblock_sum(a);
  if (threadIdx.x == 0) s[k] = a;
}

// c = Ginv s (solve) or c = s: four lanes per row of Ginv, each with its
// sixteen loads issued before the products, then a fixed-order sum of the four
__global__ __launch_bounds__(LB) void proj_solve_kernel(int NB, const double* s, const double* Ginv, int solve,
                                                        double* c) {
  static_assert(LB == 4 * NBMAX, "four lanes per coefficient");
  __shared__ double sv[NBMAX];
  __shared__ double sp[LB];
  const int t = threadIdx.x >> 2, q = threadIdx.x & 3;
  const int j0 = q * (NBMAX / 4);
  double gr[NBMAX / 4];
#pragma unroll
  for (int u = 0; u < NBMAX / 4; u++) gr[u] = (solve && t < NB && j0 + u < NB) ? Ginv[t * NB + j0 + u] : 0.0;
  if (threadIdx.x < NBMAX) sv[threadIdx.x] = threadIdx.x < NB ? s[threadIdx.x] : 0.0;
  __syncthreads();
  double p = 0.0;
#pragma unroll
  for (int u = 0; u < NBMAX / 4; u++) p += gr[u] * sv[j0 + u];
  sp[threadIdx.x] = p;
  __syncthreads();
  if (threadIdx.x < NB) {
    const int r = threadIdx.x;
    c[r] = solve ? (sp[4 * r] + sp[4 * r + 1]) + (sp[4 * r + 2] + sp[4 * r + 3]) : sv[r];
  }
}

// T[k, i] = sum_j Ginv[k, j] Phi[i, j] over the first nrows interior rows
// (the operator that maps a vector living on those rows to its Galerkin
// coefficients)
__global__ __launch_bounds__(LB) void gal_rows_op_kernel(int64_t nrows, int NB, const double* Ginv,
                                                         const float* Phi_rm, double* T) {
  const int64_t e = (int64_t)blockIdx.x * LB + threadIdx.x;
  if (e >= nrows * NB) return;
  const int64_t i = e / NB;
  const int k = (int)(e - i * NB);
  double a = 0.0;
  for (int j = 0; j < NB; j++) a += Ginv[k * NB + j] * (double)Phi_rm[i * NB + j];
  T[(int64_t)k * nrows + i] = a;
}

// c[k] = sum_i T[k, i] v[i], one block per coefficient, fixed order: each
// thread sums i = tid, tid + LB, ... in turn, its loads of a 16 LB chunk
// issued first (a few thousand rows are one chunk)
__global__ __launch_bounds__(LB) void gal_coef_kernel(int64_t nrows, const double* T, const double* v, double* c) {
  const int k = blockIdx.x;
  const double* Tk = T + (int64_t)k * nrows;
  constexpr int U = 16;
  double a = 0.0;
  for (int64_t i0 = threadIdx.x; i0 < nrows; i0 += (int64_t)U * LB) {
    double tv[U], vv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = i0 + (int64_t)u * LB;
      tv[u] = i < nrows ? Tk[i] : 0.0;
      vv[u] = i < nrows ? v[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (i0 + (int64_t)u * LB < nrows) a += tv[u] * vv[u];
  }
  a = lblock_sum(a);
  if (threadIdx.x == 0) c[k] = a;
}

// FP16 basis in groups of 8 vectors: y[(g nI + i) 8 + j] = Phi[8 g + j, i]
// (zero past NB), so that one 16-byte load holds 8 coefficients of a node
__global__ void to_f16_blocked_kernel(int64_t nI, int NB, const float* Phi, _Float16* y) {
  const int64_t e = (int64_t)blockIdx.x * LB + threadIdx.x;
  if (e >= nI * 8 * phi16_groups(NB)) return;
  const int64_t g = e / (nI * 8), rem = e - g * nI * 8;
  const int64_t i = rem >> 3;
  const int k = (int)(8 * g + (rem & 7));
  y[e] = k < NB ? (_Float16)Phi[(int64_t)k * nI + i] : (_Float16)0.0f;
}

// Phi_rm[i, k] = Phi[k, i]
__global__ void phi_rowmajor_kernel(int64_t nI, int NB, const float* Phi, float* Phi_rm) {
  const int64_t e = (int64_t)blockIdx.x * LB + threadIdx.x;
  if (e >= nI * NB) return;
  const int64_t i = e / NB;
  const int k = (int)(e - i * NB);
  Phi_rm[e] = Phi[(int64_t)k * nI + i];
}

// x[off + i] += sum_k Phi[k, i] c[k]
__global__ __launch_bounds__(LB) void expand_add_kernel(int64_t nI, int64_t off, int NB, const float* Phi,
                                                        const double* c, double* x) {
  __shared__ double cs[NBMAX];
  if (threadIdx.x < NB) cs[threadIdx.x] = c[threadIdx.x];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * LB + threadIdx.x;
  if (i >= nI) return;
  double acc = 0.0;
#pragma unroll 16
  for (int k = 0; k < NB; k++) acc += (double)Phi[(int64_t)k * nI + i] * cs[k];
  x[off + i] += acc;
}

// The same from the blocked FP16 basis: one 16-byte load per 8 vectors, the
// sum in FP32 fused multiply-adds (Phi's FP16 rounding, 5e-4 of the
// increment, dwarfs the FP32 sum's), k = 0, 1, ... in order; the residual
// kernel that expands on its own windows (launch_gal_residual) sums alike
__global__ __launch_bounds__(LB) void expand_add16_kernel(int64_t nI, int64_t off, int NB, const half8* Phi8,
                                                          const double* c, double* x) {
  __shared__ float cs[NBMAX];
  if (threadIdx.x < NB) cs[threadIdx.x] = (float)c[threadIdx.x];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * LB + threadIdx.x;
  if (i >= nI) return;
  float acc = 0.0f;
  const int ng = phi16_groups(NB);
#pragma unroll 2
  for (int gi = 0; gi < ng; gi++) {
    const half8 v = Phi8[(int64_t)gi * nI + i];
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (8 * gi + j < NB) acc = __builtin_fmaf((float)v[j], cs[8 * gi + j], acc);
  }
  x[off + i] += (double)acc;
}

__global__ void basis_col_kernel(int64_t nI, int64_t off, const float* col, double* out) {
  const int64_t i = (int64_t)blockIdx.x * LB + threadIdx.x;
  if (i < off) out[i] = 0.0;
  if (i < nI) out[off + i] = (double)col[i];
}

template <typename T>
T* dalloc(int64_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, sizeof(T) * (size_t)std::max<int64_t>(n, 1)) != hipSuccess) return nullptr;
  return (T*)p;
}

uint64_t smix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double sm_uniform(uint64_t& s) { return (double)(smix(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

void NetDims::init(int l0, int c, int nl, int h, int nb) {
  L0 = l0; C = c; NL = nl; H = h; NB = nb;
  L[0] = l0;
  for (int l = 1; l <= NL; l++) L[l] = (L[l - 1] - 1) / 2 + 1;
  int64_t p = 0;
  for (int l = 1; l <= NL; l++) {
    const int cin = (l == 1) ? 1 : C;
    w_off[l - 1] = p; p += (int64_t)C * cin * 5;
    b_off[l - 1] = p; p += C;
  }
  w_off[NL] = p; p += (int64_t)H * flat();
  b_off[NL] = p; p += H;
  w_off[NL + 1] = p; p += (int64_t)NB * H;
  b_off[NL + 1] = p; p += NB;
  nparams = p;
}

int learn_init(LearnState* ls, int64_t nI, int64_t nflt, int64_t offset, int cap, int C, int H, int NB, uint64_t seed,
               std::string* err) {
  learn_free(ls);
  if (NB < 1 || NB > NBMAX) { *err = "nbasis must be 1..64"; return 1; }
  if (cap < NB) { *err = "snapshot capacity must be >= nbasis"; return 1; }
  ls->nI = nI; ls->nflt = nflt; ls->offset = offset; ls->cap = cap;
  ls->d.init((int)nflt, C, 4, H, NB);
  const NetDims& d = ls->d;
  ls->S = dalloc<float>(nI * (int64_t)cap);
  ls->F = dalloc<float>(nflt * (int64_t)cap);
  ls->Phi = dalloc<float>(nI * (int64_t)NB);
  ls->X = dalloc<float>(nflt * (int64_t)cap);
  ls->Y = dalloc<float>((int64_t)NB * cap);
  ls->w32 = dalloc<float>(d.nparams);
  ls->w16 = dalloc<uint16_t>(d.nparams);
  ls->grad = dalloc<float>(d.nparams);
  ls->adam_m = dalloc<float>(d.nparams);
  ls->adam_v = dalloc<float>(d.nparams);
  ls->inf_act = dalloc<float>(2 * (int64_t)std::max<int64_t>(nflt, (int64_t)C * d.L[1]) + 2 * H);
  ls->coef = dalloc<float>(NBMAX + 1);
  ls->norms = dalloc<double>(2 * (int64_t)cap);
  ls->gal_nblk = (int)((nI + PROJ_ROWS - 1) / PROJ_ROWS);
  ls->gal_part = dalloc<double>((int64_t)ls->gal_nblk * NB);
  ls->gal_c = dalloc<double>(2 * NBMAX);   // [0, NBMAX): coefficients, [NBMAX, 2 NBMAX): projections
  ls->Phi_rm = dalloc<float>(nI * (int64_t)NB);
  ls->Phi16 = dalloc<_Float16>(nI * (int64_t)phi16_groups(NB) * 8);
  ls->gal_Ginv = dalloc<double>((int64_t)NB * NB);
  if (!ls->gal_part || !ls->gal_c || !ls->gal_Ginv || !ls->Phi_rm || !ls->Phi16 || !ls->norms || !ls->S || !ls->F || !ls->Phi || !ls->X || !ls->Y || !ls->w32 || !ls->w16 || !ls->grad || !ls->adam_m ||
      !ls->adam_v || !ls->inf_act || !ls->coef) {
    *err = "device allocation failed (learned guess)";
    learn_free(ls);
    return 7;
  }
  // He-uniform weights, zero biases (splitmix64 stream from seed)
  std::vector<float> w(d.nparams, 0.f);
  uint64_t st = seed;
  auto fill = [&](int64_t off, int64_t n, double fan_in) {
    const double a = std::sqrt(6.0 / fan_in);
    for (int64_t i = 0; i < n; i++) w[off + i] = (float)((2.0 * sm_uniform(st) - 1.0) * a);
  };
  for (int l = 1; l <= d.NL; l++) {
    const int cin = (l == 1) ? 1 : d.C;
    fill(d.w_off[l - 1], (int64_t)d.C * cin * 5, cin * 5.0);
  }
  fill(d.w_off[d.NL], (int64_t)d.H * d.flat(), d.flat());
  fill(d.w_off[d.NL + 1], (int64_t)d.NB * d.H, d.H * 4.0);
  LCHK(hipMemcpy(ls->w32, w.data(), sizeof(float) * d.nparams, hipMemcpyHostToDevice));
  LCHK(hipMemset(ls->adam_m, 0, sizeof(float) * d.nparams));
  LCHK(hipMemset(ls->adam_v, 0, sizeof(float) * d.nparams));
  hipLaunchKernelGGL(to_bf16_kernel, dim3(gblk(d.nparams)), dim3(LB), 0, 0, d.nparams, ls->w32, ls->w16);
  ls->adam_t = 0;
  ls->nsnap = 0;
  ls->trained = 0;
  ls->nbasis_ready = 0;
  ls->gal_ready = 0;
  return hipDeviceSynchronize() == hipSuccess ? 0 : 2;
}

void learn_free(LearnState* ls) {
  void* ptrs[] = {ls->S, ls->F, ls->Phi, ls->X, ls->Y, ls->w32, ls->w16, ls->grad, ls->adam_m, ls->adam_v,
                  ls->act, ls->dact, ls->out, ls->idx, ls->loss_part, ls->inf_act, ls->coef, ls->norms,
                  ls->gal_part, ls->gal_c, ls->gal_Ginv, ls->Phi_rm, ls->gal_T, ls->lin_A, ls->lin_b, ls->Phi16,
                  ls->gemm_part};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  NetDims d = ls->d;
  *ls = LearnState();
  ls->d = d;
}

void learn_record(LearnState* ls, const double* d, const double* dPre, hipStream_t s) {
  if (ls->nsnap >= ls->cap) return;
  const int j = ls->nsnap++;
  const int64_t n = std::max(ls->nI, ls->nflt);
  hipLaunchKernelGGL(record_kernel, dim3(gblk(n)), dim3(LB), 0, s, d, dPre, ls->nflt, ls->offset, ls->nI,
                     ls->F + (int64_t)j * ls->nflt, ls->S + (int64_t)j * ls->nI);
  hipLaunchKernelGGL(snap_norm_kernel, dim3(1), dim3(LB), 0, s, d, dPre, ls->nflt, ls->norms + 2 * (int64_t)j);
}

int learn_norms(LearnState* ls, double* out, hipStream_t s) {
  if (hipMemcpyAsync(out, ls->norms, sizeof(double) * 2 * ls->nsnap, hipMemcpyDeviceToHost, s) != hipSuccess) return 2;
  return hipStreamSynchronize(s) == hipSuccess ? 0 : 2;
}

int learn_gram(LearnState* ls, double* G_host, hipStream_t s, std::string* err) {
  const int m = ls->nsnap;
  if (m < 1) { *err = "no snapshots"; return 6; }
  const int nz = gram_splits(ls->nI, m);
  double* G = dalloc<double>((int64_t)m * m * (nz > 1 ? nz + 1 : 1));
  if (!G) { *err = "device allocation failed"; return 7; }
  std::vector<double> h((int64_t)m * m);
  hipError_t e = launch_gram(ls->S, ls->nI, m, G, G + (int64_t)m * m, s);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), G, sizeof(double) * m * m, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(G);
  if (e != hipSuccess) { *err = hipGetErrorString(e); return 2; }
  for (int i = 0; i < m; i++)
    for (int j = 0; j < m; j++) G_host[(int64_t)i * m + j] = (i <= j) ? h[(int64_t)i * m + j] : h[(int64_t)j * m + i];
  return 0;
}

int learn_set_basis(LearnState* ls, const double* V, const double* sigma, const int32_t* use, hipStream_t s,
                    std::string* err) {
  const int m = ls->nsnap, NB = ls->d.NB, L0 = ls->d.L0;
  ls->gal_ready = 0;
  ls->used.clear();
  for (int j = 0; j < m; j++)
    if (!use || use[j]) ls->used.push_back(j);
  if ((int)ls->used.size() < NB) { *err = "fewer usable snapshots than basis vectors"; return 6; }
  std::vector<float> W((int64_t)m * NB);
  for (int j = 0; j < m; j++)
    for (int k = 0; k < NB; k++) W[(int64_t)j * NB + k] = sigma[k] > 0 ? (float)(V[(int64_t)j * NB + k] / sigma[k]) : 0.f;
  float* dW = dalloc<float>((int64_t)m * NB);
  if (!dW) { *err = "device allocation failed"; return 7; }
  LCHK(hipMemcpyAsync(dW, W.data(), sizeof(float) * W.size(), hipMemcpyHostToDevice, s));
  // Phi = S W on the matrix cores (FP32 GEMM, [nI x m] [m x NB])
  gemm((int)ls->nI, NB, m, ColMat{ls->S, (int)ls->nI}, RowMat{dW, NB}, EpColOut{ls->Phi, ls->nI}, nullptr, 0, s);
  // Cholesky-QR pass: S V / sigma loses orthogonality in FP32 along the small
  // singular values, so G = Phi^T Phi (FP64), G = L L^T, Phi <- Phi L^-T
  // (the row-major copy is the scratch output, then the pointers swap)
  std::vector<double> Winv((size_t)NB * NB, 0.0);
  bool reortho = false;
  {
    const int nz = gram_splits(ls->nI, NB);
    double* G = dalloc<double>((int64_t)NB * NB * (nz > 1 ? nz + 1 : 1));
    if (!G) { (void)hipFree(dW); *err = "device allocation failed"; return 7; }
    const hipError_t e0 = launch_gram(ls->Phi, ls->nI, NB, G, G + (int64_t)NB * NB, s);
    std::vector<double> h((size_t)NB * NB);
    const hipError_t e1 = e0 != hipSuccess ? e0 : hipMemcpyAsync(h.data(), G, sizeof(double) * NB * NB,
                                                                 hipMemcpyDeviceToHost, s);
    const hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFree(G);
    if (e1 != hipSuccess || e2 != hipSuccess) { (void)hipFree(dW); *err = "Gram of the basis failed"; return 2; }
    std::vector<double> L((size_t)NB * NB, 0.0);
    reortho = true;
    for (int j = 0; j < NB && reortho; j++) {
      double d = h[(size_t)j * NB + j];   // upper triangle holds G
      for (int k = 0; k < j; k++) d -= L[(size_t)j * NB + k] * L[(size_t)j * NB + k];
      if (!(d > 0.0)) { reortho = false; break; }
      d = std::sqrt(d);
      L[(size_t)j * NB + j] = d;
      for (int i = j + 1; i < NB; i++) {
        double a = h[(size_t)j * NB + i];
        for (int k = 0; k < j; k++) a -= L[(size_t)i * NB + k] * L[(size_t)j * NB + k];
        L[(size_t)i * NB + j] = a / d;
      }
    }
    if (reortho) {
      // Winv = L^-T (upper triangular): column c solves L^T w = e_c
      for (int c = 0; c < NB; c++)
        for (int i = NB - 1; i >= 0; i--) {
          double a = (i == c) ? 1.0 : 0.0;
          for (int k = i + 1; k < NB; k++) a -= L[(size_t)k * NB + i] * Winv[(size_t)k * NB + c];
          Winv[(size_t)i * NB + c] = a / L[(size_t)i * NB + i];
        }
      std::vector<float> Wf((size_t)NB * NB);
      for (size_t q = 0; q < Wf.size(); q++) Wf[q] = (float)Winv[q];
      LCHK(hipMemcpyAsync(dW, Wf.data(), sizeof(float) * Wf.size(), hipMemcpyHostToDevice, s));
      gemm((int)ls->nI, NB, NB, ColMat{ls->Phi, (int)ls->nI}, RowMat{dW, NB}, EpColOut{ls->Phi_rm, ls->nI}, nullptr, 0,
           s);
      std::swap(ls->Phi, ls->Phi_rm);
    }
  }
  hipLaunchKernelGGL(phi_rowmajor_kernel, dim3(gblk(ls->nI * NB)), dim3(LB), 0, s, ls->nI, NB, ls->Phi, ls->Phi_rm);
  // the FP16 copy the Galerkin expansion may read (eq_solver_opts.basis_fp16)
  hipLaunchKernelGGL(to_f16_blocked_kernel, dim3(gblk(ls->nI * 8 * phi16_groups(NB))), dim3(LB), 0, s, ls->nI, NB,
                     ls->Phi, ls->Phi16);
  // dataset: X_j = dF_j / |dF_j|, Y_j = Phi^T s_j / |dF_j| = Winv^T (sigma_k V_jk) / |dF_j|
  std::vector<float> F((int64_t)m * L0), X((int64_t)m * L0), Y((int64_t)m * NB);
  LCHK(hipMemcpyAsync(F.data(), ls->F, sizeof(float) * F.size(), hipMemcpyDeviceToHost, s));
  hipError_t e = hipStreamSynchronize(s);
  (void)hipFree(dW);
  if (e != hipSuccess) { *err = hipGetErrorString(e); return 2; }
  for (int j = 0; j < m; j++) {
    if (use && !use[j]) {
      for (int i = 0; i < L0; i++) X[(int64_t)j * L0 + i] = 0.f;
      for (int k = 0; k < NB; k++) Y[(int64_t)j * NB + k] = 0.f;
      continue;
    }
    double nrm = 0;
    for (int i = 0; i < L0; i++) nrm += (double)F[(int64_t)j * L0 + i] * F[(int64_t)j * L0 + i];
    nrm = std::sqrt(nrm);
    const double inv = nrm > 0 ? 1.0 / nrm : 0.0;
    for (int i = 0; i < L0; i++) X[(int64_t)j * L0 + i] = (float)(F[(int64_t)j * L0 + i] * inv);
    for (int k = 0; k < NB; k++) {
      double y = 0.0;
      if (reortho) {
        for (int i = 0; i <= k; i++) y += Winv[(size_t)i * NB + k] * sigma[i] * V[(int64_t)j * NB + i];
      } else {
        y = sigma[k] * V[(int64_t)j * NB + k];
      }
      Y[(int64_t)j * NB + k] = (float)(y * inv);
    }
  }
  LCHK(hipMemcpyAsync(ls->X, X.data(), sizeof(float) * X.size(), hipMemcpyHostToDevice, s));
  LCHK(hipMemcpyAsync(ls->Y, Y.data(), sizeof(float) * Y.size(), hipMemcpyHostToDevice, s));
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) { *err = hipGetErrorString(e); return 2; }
  ls->nbasis_ready = 1;
  return 0;
}

static int ensure_batch(LearnState* ls, int B) {
  if (ls->bmax >= B) return 0;
  for (void* p : {(void*)ls->act, (void*)ls->dact, (void*)ls->out, (void*)ls->idx, (void*)ls->loss_part,
                  (void*)ls->gemm_part})
    if (p) (void)hipFree(p);
  const NetDims& d = ls->d;
  int64_t off = 0;
  ls->act_off[0] = off; off += (int64_t)B * d.L0;
  for (int l = 1; l <= d.NL; l++) { ls->act_off[l] = off; off += (int64_t)B * d.C * d.L[l]; }
  ls->act_off[d.NL + 1] = off; off += (int64_t)B * d.H;
  ls->act_off[d.NL + 2] = off;
  ls->act = dalloc<float>(off);
  ls->dact = dalloc<float>(off);
  ls->out = dalloc<float>((int64_t)B * d.NB * 2);
  ls->idx = dalloc<int32_t>(B);
  ls->loss_part = dalloc<double>(gblk((int64_t)B * d.NB) + 1);
  ls->gemm_part_cap = (int64_t)1 << 21;   // split-K partials of the layer GEMMs (8 MB)
  ls->gemm_part = dalloc<float>(ls->gemm_part_cap);
  ls->bmax = B;
  return (ls->act && ls->dact && ls->out && ls->idx && ls->loss_part && ls->gemm_part) ? 0 : 7;
}

// forward + backward of one batch; fills grad, returns the loss partials in
// loss_part.  Every layer is a GEMM on the matrix cores (mfma_gemm_kernel):
// the convolutions through their im2col gathers, the weight gradients with a
// column of ones that gives the bias gradient.
template <typename WT>
static void fwd_bwd(LearnState* ls, const WT* W, const float* w32, int B, const int32_t* idx, double* loss_part,
                    hipStream_t s) {
  const NetDims& d = ls->d;
  float* A = ls->act;
  float* dA = ls->dact;
  float* out = ls->out;
  float* dout = ls->out + (int64_t)B * d.NB;
  float* part = ls->gemm_part;
  const int64_t cap = ls->gemm_part_cap;
  hipLaunchKernelGGL(gather_kernel, dim3(gblk((int64_t)B * d.L0)), dim3(LB), 0, s, ls->X, d.L0, idx, B,
                     A + ls->act_off[0]);
  for (int l = 1; l <= d.NL; l++) {
    const int cin = (l == 1) ? 1 : d.C, Lin = d.L[l - 1], Lout = d.L[l];
    gemm(B * Lout, d.C, 5 * cin, ConvIn{A + ls->act_off[l - 1], cin, Lin, Lout},
         WCol<WT>{W + d.w_off[l - 1], 5 * cin}, EpConv{A + ls->act_off[l], w32 + d.b_off[l - 1], d.C, Lout}, part,
         cap, s);
  }
  const int K = d.flat();
  gemm(B, d.H, K, RowMat{A + ls->act_off[d.NL], K}, WCol<WT>{W + d.w_off[d.NL], K},
       EpDense{A + ls->act_off[d.NL + 1], w32 + d.b_off[d.NL], d.H, 1}, part, cap, s);
  gemm(B, d.NB, d.H, RowMat{A + ls->act_off[d.NL + 1], d.H}, WCol<WT>{W + d.w_off[d.NL + 1], d.H},
       EpDense{out, w32 + d.b_off[d.NL + 1], d.NB, 0}, part, cap, s);
  hipLaunchKernelGGL(mse_kernel, dim3(gblk((int64_t)B * d.NB)), dim3(LB), 0, s, B, d.NB, out, ls->Y, idx, dout,
                     loss_part);
  // backward
  float* g = ls->grad;
  gemm(d.NB, d.H + 1, B, ColMat{dout, d.NB}, RowMatOnes{A + ls->act_off[d.NL + 1], d.H},
       EpGrad{g + d.w_off[d.NL + 1], g + d.b_off[d.NL + 1], d.H}, part, cap, s);
  gemm(B, d.H, d.NB, RowMat{dout, d.NB}, WRow<WT>{W + d.w_off[d.NL + 1], d.H},
       EpMasked{dA + ls->act_off[d.NL + 1], A + ls->act_off[d.NL + 1], d.H}, part, cap, s);
  gemm(d.H, K + 1, B, ColMat{dA + ls->act_off[d.NL + 1], d.H}, RowMatOnes{A + ls->act_off[d.NL], K},
       EpGrad{g + d.w_off[d.NL], g + d.b_off[d.NL], K}, part, cap, s);
  gemm(B, K, d.H, RowMat{dA + ls->act_off[d.NL + 1], d.H}, WRow<WT>{W + d.w_off[d.NL], K},
       EpMasked{dA + ls->act_off[d.NL], A + ls->act_off[d.NL], K}, part, cap, s);
  for (int l = d.NL; l >= 1; l--) {
    const int cin = (l == 1) ? 1 : d.C, Lin = d.L[l - 1], Lout = d.L[l];
    gemm(d.C, 5 * cin + 1, B * Lout, ConvGradT{dA + ls->act_off[l], d.C, Lout},
         ConvInT{A + ls->act_off[l - 1], cin, Lin, Lout}, EpConvGradW{g + d.w_off[l - 1], g + d.b_off[l - 1], 5 * cin},
         part, cap, s);
    if (l > 1)
      gemm(B * Lin, cin, 5 * d.C, ConvGradIm{dA + ls->act_off[l], d.C, Lin, Lout}, ConvWT<WT>{W + d.w_off[l - 1], cin},
           EpConvGradIn{dA + ls->act_off[l - 1], A + ls->act_off[l - 1], cin, Lin}, part, cap, s);
  }
}

int learn_train(LearnState* ls, int epochs, int batch, double lr, uint64_t seed, double* hist, hipStream_t s,
                std::string* err) {
  if (!ls->nbasis_ready) { *err = "POD basis not set"; return 6; }
  const int m = (int)ls->used.size();
  const int B = std::min(batch, m);
  if (B < 1) { *err = "empty dataset"; return 6; }
  if (ensure_batch(ls, B)) { *err = "device allocation failed"; return 7; }
  const NetDims& d = ls->d;
  const int nparts = gblk((int64_t)B * d.NB);
  const int nbatch = m / B;
  // every epoch's shuffled order up front, and every batch's loss partials
  // read back once at the end: the steps run with no host round trip
  std::vector<int32_t> perm((size_t)epochs * m);
  uint64_t st = seed;
  for (int ep = 0; ep < epochs; ep++) {
    int32_t* pe = perm.data() + (size_t)ep * m;
    for (int i = 0; i < m; i++) pe[i] = ls->used[i];
    for (int i = m - 1; i > 0; i--) std::swap(pe[i], pe[smix(st) % (uint64_t)(i + 1)]);
  }
  int32_t* perm_dev = dalloc<int32_t>((int64_t)epochs * m);
  double* parts_dev = dalloc<double>((int64_t)epochs * nbatch * nparts);
  if (!perm_dev || !parts_dev) {
    if (perm_dev) (void)hipFree(perm_dev);
    if (parts_dev) (void)hipFree(parts_dev);
    *err = "device allocation failed";
    return 7;
  }
  hipError_t e = hipMemcpyAsync(perm_dev, perm.data(), sizeof(int32_t) * perm.size(), hipMemcpyHostToDevice, s);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs && e == hipSuccess; ep++)
    for (int q = 0; q < nbatch; q++) {
      fwd_bwd<uint16_t>(ls, ls->w16, ls->w32, B, perm_dev + (int64_t)ep * m + (int64_t)q * B,
                        parts_dev + ((int64_t)ep * nbatch + q) * nparts, s);
      ls->adam_t++;
      const float c1 = 1.f - std::pow(b1, (float)ls->adam_t), c2 = 1.f - std::pow(b2, (float)ls->adam_t);
      hipLaunchKernelGGL(adam_kernel, dim3(gblk(d.nparams)), dim3(LB), 0, s, d.nparams, ls->w32, ls->w16, ls->grad,
                         ls->adam_m, ls->adam_v, (float)lr, b1, b2, c1, c2, eps);
    }
  std::vector<double> parts((size_t)epochs * nbatch * nparts);
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(parts.data(), parts_dev, sizeof(double) * parts.size(), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(perm_dev);
  (void)hipFree(parts_dev);
  if (e != hipSuccess) { *err = hipGetErrorString(e); return 2; }
  for (int ep = 0; ep < epochs; ep++) {
    double eloss = 0;
    for (int q = 0; q < nbatch; q++) {
      double l = 0;
      const double* pq = parts.data() + ((size_t)ep * nbatch + q) * nparts;
      for (int k = 0; k < nparts; k++) l += pq[k];
      eloss += l;
    }
    if (hist) hist[ep] = nbatch ? eloss / nbatch : 0.0;
  }
  ls->trained = 1;
  return 0;
}

int learn_loss_grad(LearnState* ls, const float* params, int nsamp, const int32_t* idx, int use_bf16, double* loss,
                    float* grad_host, hipStream_t s, std::string* err) {
  if (!ls->nbasis_ready) { *err = "POD basis not set"; return 6; }
  if (nsamp < 1 || nsamp > ls->nsnap) { *err = "bad sample count"; return 1; }
  if (ensure_batch(ls, nsamp)) { *err = "device allocation failed"; return 7; }
  const NetDims& d = ls->d;
  for (int i = 0; i < nsamp; i++)
    if (idx[i] < 0 || idx[i] >= ls->nsnap) { *err = "sample index out of range"; return 1; }
  if (params) {
    LCHK(hipMemcpyAsync(ls->w32, params, sizeof(float) * d.nparams, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(to_bf16_kernel, dim3(gblk(d.nparams)), dim3(LB), 0, s, d.nparams, ls->w32, ls->w16);
  }
  LCHK(hipMemcpyAsync(ls->idx, idx, sizeof(int32_t) * nsamp, hipMemcpyHostToDevice, s));
  if (use_bf16) fwd_bwd<uint16_t>(ls, ls->w16, ls->w32, nsamp, ls->idx, ls->loss_part, s);
  else fwd_bwd<float>(ls, ls->w32, ls->w32, nsamp, ls->idx, ls->loss_part, s);
  const int nparts = gblk((int64_t)nsamp * d.NB);
  std::vector<double> parts(nparts);
  LCHK(hipMemcpyAsync(parts.data(), ls->loss_part, sizeof(double) * nparts, hipMemcpyDeviceToHost, s));
  if (grad_host) LCHK(hipMemcpyAsync(grad_host, ls->grad, sizeof(float) * d.nparams, hipMemcpyDeviceToHost, s));
  const hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) { *err = hipGetErrorString(e); return 2; }
  double l = 0;
  for (double p : parts) l += p;
  *loss = l;
  return 0;
}

// Coefficients for the input x - base into ls->coef (coef[NB] = |x - base|).
// The linear predictor: one launch.  The network: per layer, input
// normalisation, conv layers (thread per output), wave-per-row GEMVs for the
// MLP -- 7 launches spread over the chip.  The network's whole forward in one
// workgroup (activations in global scratch between barriers) measured about
// 280 us per call on the 641-node fault, against tens of microseconds for
// the per-layer launches, so the launches stay.
static void infer(LearnState* ls, const double* x, const double* base, int linear, hipStream_t s) {
  const NetDims& d = ls->d;
  float* bufA = ls->inf_act;
  float* bufB = ls->inf_act + std::max<int64_t>(ls->nflt, (int64_t)d.C * d.L[1]) + d.H;
  if (linear) {
    hipLaunchKernelGGL(lin_infer_kernel, dim3(1), dim3(IB), 0, s, ls->nflt, d.NB, x, base, ls->lin_A, ls->lin_b,
                       bufA, ls->coef);
    return;
  }
  hipLaunchKernelGGL(infer_input_kernel, dim3(1), dim3(LB), 0, s, x, base, ls->nflt, bufA, ls->coef, d.NB);
  float* in = bufA;
  float* out = bufB;
  for (int l = 1; l <= d.NL; l++) {
    const int cin = (l == 1) ? 1 : d.C;
    hipLaunchKernelGGL((conv_fwd_kernel<uint16_t>), dim3(gblk((int64_t)d.C * d.L[l])), dim3(LB), 0, s, 1, cin,
                       d.L[l - 1], d.C, d.L[l], in, ls->w16 + d.w_off[l - 1], ls->w32 + d.b_off[l - 1], out);
    std::swap(in, out);
  }
  hipLaunchKernelGGL(gemv_rows_kernel, dim3((d.H + 3) / 4), dim3(LB), 0, s, d.H, d.flat(), in, ls->w16 + d.w_off[d.NL],
                     ls->w32 + d.b_off[d.NL], out, 1, (const float*)nullptr);
  hipLaunchKernelGGL(gemv_rows_kernel, dim3((d.NB + 3) / 4), dim3(LB), 0, s, d.NB, d.H, out,
                     ls->w16 + d.w_off[d.NL + 1], ls->w32 + d.b_off[d.NL + 1], ls->coef, 0,
                     (const float*)(ls->coef + d.NB));
}

static void expand_guess(LearnState* ls, double* p, hipStream_t s) {
  const int64_t n = std::max(ls->nI, ls->offset);
  hipLaunchKernelGGL(guess_kernel, dim3((unsigned)((n + LB - 1) / LB)), dim3(LB), 0, s, ls->nI, ls->offset, ls->d.NB,
                     ls->Phi, ls->coef, p);
}

void learn_guess(LearnState* ls, const double* x, const double* base, double* p, hipStream_t s) {
  infer(ls, x, base, 0, s);
  expand_guess(ls, p, s);
}

void learn_guess_linear(LearnState* ls, const double* x, const double* base, double* p, hipStream_t s) {
  infer(ls, x, base, 1, s);
  expand_guess(ls, p, s);
}

int learn_set_linear(LearnState* ls, const float* A_host, const float* b_host, hipStream_t s, std::string* err) {
  const int64_t nA = (int64_t)ls->d.NB * ls->nflt;
  if (!ls->lin_A) {
    ls->lin_A = dalloc<float>(nA);
    ls->lin_b = dalloc<float>(ls->d.NB);
    if (!ls->lin_A || !ls->lin_b) { *err = "device allocation failed (linear predictor)"; return 7; }
  }
  LCHK(hipMemcpyAsync(ls->lin_A, A_host, sizeof(float) * nA, hipMemcpyHostToDevice, s));
  LCHK(hipMemcpyAsync(ls->lin_b, b_host, sizeof(float) * ls->d.NB, hipMemcpyHostToDevice, s));
  LCHK(hipStreamSynchronize(s));
  ls->lin_ready = 1;
  return 0;
}

int learn_get_basis(LearnState* ls, float* Phi_host, hipStream_t s) {
  if (hipMemcpyAsync(Phi_host, ls->Phi, sizeof(float) * ls->nI * ls->d.NB, hipMemcpyDeviceToHost, s) != hipSuccess)
    return 2;
  return hipStreamSynchronize(s) == hipSuccess ? 0 : 2;
}

int learn_get_params(LearnState* ls, float* out, hipStream_t s) {
  if (hipMemcpyAsync(out, ls->w32, sizeof(float) * ls->d.nparams, hipMemcpyDeviceToHost, s) != hipSuccess) return 2;
  return hipStreamSynchronize(s) == hipSuccess ? 0 : 2;
}

int learn_predict(LearnState* ls, const double* dF_host, double* coef_host, int linear, hipStream_t s,
                  std::string* err) {
  double* x = dalloc<double>(ls->nflt);
  double* z = dalloc<double>(ls->nflt);
  if (!x || !z) { *err = "device allocation failed"; return 7; }
  LCHK(hipMemcpyAsync(x, dF_host, sizeof(double) * ls->nflt, hipMemcpyHostToDevice, s));
  LCHK(hipMemsetAsync(z, 0, sizeof(double) * ls->nflt, s));
  infer(ls, x, z, linear, s);
  std::vector<float> c(ls->d.NB);
  LCHK(hipMemcpyAsync(c.data(), ls->coef, sizeof(float) * ls->d.NB, hipMemcpyDeviceToHost, s));
  const hipError_t e = hipStreamSynchronize(s);
  (void)hipFree(x);
  (void)hipFree(z);
  if (e != hipSuccess) { *err = hipGetErrorString(e); return 2; }
  for (int k = 0; k < ls->d.NB; k++) coef_host[k] = c[k];
  return 0;
}

void learn_project(LearnState* ls, const double* r, int solve, hipStream_t s) {
  hipLaunchKernelGGL(proj_dot_kernel, dim3(ls->gal_nblk), dim3(LB), 0, s, ls->nI, ls->offset, ls->d.NB, ls->Phi_rm,
                     r, ls->gal_part);
  hipLaunchKernelGGL(proj_sum_kernel, dim3(ls->d.NB), dim3(LB), 0, s, ls->gal_nblk, ls->d.NB, ls->gal_part,
                     ls->gal_c + NBMAX);
  hipLaunchKernelGGL(proj_solve_kernel, dim3(1), dim3(LB), 0, s, ls->d.NB, ls->gal_c + NBMAX, ls->gal_Ginv, solve,
                     ls->gal_c);
}

void learn_project_rows(LearnState* ls, const double* v, int64_t nrows, hipStream_t s) {
  nrows = std::min<int64_t>(nrows, ls->nI);
  if (ls->gal_T && nrows == ls->gal_trows) {
    hipLaunchKernelGGL(gal_coef_kernel, dim3(ls->d.NB), dim3(LB), 0, s, nrows, ls->gal_T, v, ls->gal_c);
    return;
  }
  const int nblk = (int)((nrows + PROJ_ROWS - 1) / PROJ_ROWS);
  if (nblk == 0) return;
  hipLaunchKernelGGL(proj_dot_kernel, dim3(nblk), dim3(LB), 0, s, nrows, (int64_t)0, ls->d.NB, ls->Phi_rm, v,
                     ls->gal_part);
  hipLaunchKernelGGL(proj_sum_kernel, dim3(ls->d.NB), dim3(LB), 0, s, nblk, ls->d.NB, ls->gal_part,
                     ls->gal_c + NBMAX);
  hipLaunchKernelGGL(proj_solve_kernel, dim3(1), dim3(LB), 0, s, ls->d.NB, ls->gal_c + NBMAX, ls->gal_Ginv, 1,
                     ls->gal_c);
}

void learn_expand_add(LearnState* ls, double* x, hipStream_t s, bool fp16) {
  if (fp16 && ls->Phi16)
    hipLaunchKernelGGL(expand_add16_kernel, dim3(gblk(ls->nI)), dim3(LB), 0, s, ls->nI, ls->offset, ls->d.NB,
                       (const half8*)ls->Phi16, ls->gal_c, x);
  else
    hipLaunchKernelGGL(expand_add_kernel, dim3(gblk(ls->nI)), dim3(LB), 0, s, ls->nI, ls->offset, ls->d.NB,
                       ls->Phi, ls->gal_c, x);
}

void learn_basis_column(LearnState* ls, int k, double* out, hipStream_t s) {
  hipLaunchKernelGGL(basis_col_kernel, dim3(gblk(std::max(ls->nI, ls->offset))), dim3(LB), 0, s, ls->nI, ls->offset,
                     ls->Phi + (int64_t)k * ls->nI, out);
}

int learn_get_proj(LearnState* ls, double* host, hipStream_t s) {
  if (hipMemcpyAsync(host, ls->gal_c, sizeof(double) * ls->d.NB, hipMemcpyDeviceToHost, s) != hipSuccess) return 2;
  return hipStreamSynchronize(s) == hipSuccess ? 0 : 2;
}

int learn_set_ginv(LearnState* ls, const double* Ginv_host, hipStream_t s, std::string* err) {
  LCHK(hipMemcpyAsync(ls->gal_Ginv, Ginv_host, sizeof(double) * ls->d.NB * ls->d.NB, hipMemcpyHostToDevice, s));
  // Ginv Phi^T restricted to the rows next to the fault (rows 1..4 of the
  // grid: the first 4 nflt interior rows), where a change of b lives
  const int64_t nrows = std::min<int64_t>(4 * ls->nflt, ls->nI);
  if (!ls->gal_T) ls->gal_T = dalloc<double>(nrows * ls->d.NB);
  if (!ls->gal_T) { *err = "device allocation failed (Galerkin rows operator)"; return 2; }
  ls->gal_trows = nrows;
  hipLaunchKernelGGL(gal_rows_op_kernel, dim3((unsigned)((nrows * ls->d.NB + LB - 1) / LB)), dim3(LB), 0, s, nrows,
                     ls->d.NB, ls->gal_Ginv, ls->Phi_rm, ls->gal_T);
  LCHK(hipGetLastError());
  LCHK(hipStreamSynchronize(s));
  ls->gal_ready = 1;
  return 0;
}

}  // namespace eq

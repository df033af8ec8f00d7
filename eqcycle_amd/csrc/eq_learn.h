// Learned initial guess for the quasi-static solve: snapshot store, POD
// basis and the conv1d + MLP network (BF16 weights, FP32 master copy and
// accumulation).  Device code in eq_learn.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace eq {

// 8 FP16 basis coefficients of one node (one 16-byte load)
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
// groups of 8 basis vectors in the FP16 copy of the basis
__host__ __device__ inline int phi16_groups(int NB) { return (NB + 7) / 8; }

struct NetDims {
  int L0 = 0;        // input length (fault nodes)
  int C = 32;        // channels of every conv layer
  int NL = 4;        // conv layers (kernel 5, stride 2, padding 2, ReLU)
  int H = 256;       // hidden width of the MLP
  int NB = 64;       // outputs = POD coefficients
  int L[8] = {0};    // L[l]: length after conv layer l (L[0] = L0)
  // parameter offsets in the flat FP32 vector
  int64_t w_off[8], b_off[8];
  int64_t nparams = 0;
  int flat() const { return C * L[NL]; }
  void init(int l0, int c, int nl, int h, int nb);
};

struct LearnState {
  NetDims d;
  int64_t nI = 0, nflt = 0, offset = 0;   // interior count, fault count, N-vector offset of the interior
  int cap = 0, nsnap = 0, recording = 0;
  // snapshots (column-major): S [nI x cap] interior increments, F [nflt x cap] fault increments
  float* S = nullptr;
  float* F = nullptr;
  // basis (column-major [nI x NB]) and dataset
  float* Phi = nullptr;
  int nbasis_ready = 0;
  float* X = nullptr;        // [nsnap x L0] normalised inputs
  float* Y = nullptr;        // [nsnap x NB] normalised targets
  // parameters
  float* w32 = nullptr;      // FP32 master
  uint16_t* w16 = nullptr;   // BF16 copy used by forward / backward
  float* grad = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  int64_t adam_t = 0;
  int trained = 0;
  // training buffers (capacity bmax)
  int bmax = 0;
  float *act = nullptr, *dact = nullptr;   // all layer activations for a batch
  int64_t act_off[12];
  float* out = nullptr;
  int32_t* idx = nullptr;
  double* loss_part = nullptr;
  float* gemm_part = nullptr;  // split-K partials of the layer GEMMs
  int64_t gemm_part_cap = 0;
  // inference scratch
  float* inf_act = nullptr;
  float* coef = nullptr;     // [NB + 1]: coefficients, last = input scale
  double* norms = nullptr;   // [2 x cap]: |dF|, |F| per snapshot
  std::vector<int32_t> used; // snapshots used for the dataset
  // Galerkin start on span(Phi): c = (Phi^T K_II Phi)^-1 Phi^T r, x += Phi c
  double* gal_part = nullptr;   // [gal_nblk x NB] partial projections
  double* gal_c = nullptr;      // [NB] projections / coefficients
  double* gal_Ginv = nullptr;   // [NB x NB] row-major inverse of Phi^T K_II Phi
  float* Phi_rm = nullptr;      // [nI x NB] row-major copy of Phi (projection)
  _Float16* Phi16 = nullptr;    // FP16 copy of Phi for the expansion, in groups of 8 vectors:
                                // [phi16_groups(NB)][nI][8] (zero past NB)
  double* gal_T = nullptr;      // [NB x gal_trows] Ginv Phi^T on the first gal_trows interior rows
  int64_t gal_trows = 0;
  int gal_nblk = 0;
  int gal_ready = 0;
  // linear least-squares predictor of the same coefficients from the same
  // input: coef = |dF| (A dF/|dF| + b)
  float* lin_A = nullptr;       // [NB x L0] row-major
  float* lin_b = nullptr;       // [NB]
  int lin_ready = 0;
};

int learn_init(LearnState* ls, int64_t nI, int64_t nflt, int64_t offset, int cap, int C, int H, int NB,
               uint64_t seed, std::string* err);
void learn_free(LearnState* ls);
// append one (dF, dd) pair: dF = d[0:nflt] - dPre[0:nflt], dd = d[off:] - dPre[off:]
void learn_record(LearnState* ls, const double* d, const double* dPre, hipStream_t s);
// G = S^T S (FP64, row-major nsnap x nsnap)
int learn_gram(LearnState* ls, double* G_host, hipStream_t s, std::string* err);
// Phi = S V diag(1/sigma); dataset X, Y from the snapshots and sigma V
int learn_set_basis(LearnState* ls, const double* V, const double* sigma, const int32_t* use, hipStream_t s,
                    std::string* err);
int learn_norms(LearnState* ls, double* out, hipStream_t s);
// Adam training on the dataset; loss per epoch in hist (may be null)
int learn_train(LearnState* ls, int epochs, int batch, double lr, uint64_t seed, double* hist, hipStream_t s,
                std::string* err);
// loss and gradient for given parameters on samples idx (use_bf16 = 0: FP32 weights)
int learn_loss_grad(LearnState* ls, const float* params, int nsamp, const int32_t* idx, int use_bf16, double* loss,
                    float* grad_host, hipStream_t s, std::string* err);
// p[off:] = |dF| Phi net(dF / |dF|), p[0:off] = 0, with dF = x[0:nflt] - base[0:nflt]
void learn_guess(LearnState* ls, const double* x, const double* base, double* p, hipStream_t s);
// the same with the linear predictor's coefficients
void learn_guess_linear(LearnState* ls, const double* x, const double* base, double* p, hipStream_t s);
// upload the linear predictor (A [NB x L0] row-major, b [NB])
int learn_set_linear(LearnState* ls, const float* A_host, const float* b_host, hipStream_t s, std::string* err);
int learn_get_basis(LearnState* ls, float* Phi_host, hipStream_t s);
int learn_get_params(LearnState* ls, float* out, hipStream_t s);
// coefficients for one input (test entry point); linear: the linear predictor's
int learn_predict(LearnState* ls, const double* dF_host, double* coef_host, int linear, hipStream_t s,
                  std::string* err);

// Galerkin start.  learn_project: gal_c = Phi^T r[off:] (solve = 0) or
// gal_c = Ginv Phi^T r[off:] (solve = 1), FP64 accumulation in a fixed order.
void learn_project(LearnState* ls, const double* r, int solve, hipStream_t s);
// gal_c = Ginv Phi[0:nrows, :]^T v: the projection of a vector that is zero
// beyond the first nrows interior rows (v holds just those rows)
void learn_project_rows(LearnState* ls, const double* v, int64_t nrows, hipStream_t s);
// x[off:] += Phi gal_c
// fp16: read the FP16 copy of the basis (when set)
void learn_expand_add(LearnState* ls, double* x, hipStream_t s, bool fp16 = false);
// out = column k of Phi as an N-vector (out[0:off] = 0)
void learn_basis_column(LearnState* ls, int k, double* out, hipStream_t s);
int learn_get_proj(LearnState* ls, double* host, hipStream_t s);
// upload Ginv (row-major NB x NB) and mark the Galerkin start ready
int learn_set_ginv(LearnState* ls, const double* Ginv_host, hipStream_t s, std::string* err);

}  // namespace eq

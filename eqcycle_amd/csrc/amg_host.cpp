// Host-side set-up for the GPU preconditioner and operators.
//
// * assemble_kii: K_II (stiffness rows/columns of the non-fault nodes) as CSR
//   in the structured interior order, from nodal rigidity.  Same operator as
//   stiffness_assembly / K_element (src/Kassemble.jl:3-63) restricted to
//   FltNI (src/main.jl:195).
// * amg_build: Ruge-Stueben hierarchy of the kind ruge_stuben(kni) builds
//   (src/main.jl:200): classical strength with threshold theta, RS C/F
//   splitting (Ruge & Stueben 1987, bucketed max-lambda selection), direct
//   interpolation with separate positive / negative weight sums, R = P^T,
//   Galerkin coarse operators, dense inverse on the coarsest level.
//   AlgebraicMultigrid.jl is not part of the reference repository and its
//   version is unpinned; this follows the published algorithm.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>

#include "eq_internal.h"

namespace eq {


// Sum of GLL weights of the elements that contain grid line L (of nl lines).
static inline double wline(const Geom& g, int L, int nl) {
  int r = L & 3;
  if (r != 0) return g.w[r];
  return (L == 0 || L == nl - 1) ? g.w[0] : g.w[0] + g.w[4];
}

void assemble_kii(const Geom& g, const double* mu, HostCsr* K) {
  const int NX = g.NX, NY = g.NY;
  const int64_t nI = (int64_t)NX * (NY - 1);
  K->n = K->m = nI;
  K->rowptr.assign(nI + 1, 0);
  K->col.clear();
  K->val.clear();
  K->col.reserve(nI * 13);
  K->val.reserve(nI * 13);
  // gx / gy: line coefficient times the GLL weight of local index p in its segment
  auto gx = [&](int I, int J, int p) { return g.c1 * wline(g, J, NY) * g.w[p] * mu[(int64_t)J * NX + I]; };
  auto gy = [&](int I, int J, int q) { return g.c2 * wline(g, I, NX) * g.w[q] * mu[(int64_t)J * NX + I]; };
  int64_t cols[32];
  double vals[32];
  for (int J = 1; J < NY; J++) {
    for (int I = 0; I < NX; I++) {
      int nc = 0;
      auto add = [&](int64_t c, double v) {
        for (int q = 0; q < nc; q++)
          if (cols[q] == c) { vals[q] += v; return; }
        cols[nc] = c; vals[nc] = v; nc++;
      };
      // x segments containing I
      int s_lo = (I & 3) == 0 ? (I >> 2) - 1 : (I >> 2);
      int s_hi = (I >> 2);
      for (int s = s_lo; s <= s_hi; s++) {
        if (s < 0 || s >= g.nelx) continue;
        int i = I - 4 * s;
        for (int k = 0; k < 5; k++) {
          double v = 0;
          for (int p = 0; p < 5; p++) v += g.D[p * 5 + i] * g.D[p * 5 + k] * gx(4 * s + p, J, p);
          add((int64_t)J * NX + 4 * s + k, v);
        }
      }
      int t_lo = (J & 3) == 0 ? (J >> 2) - 1 : (J >> 2);
      int t_hi = (J >> 2);
      for (int t = t_lo; t <= t_hi; t++) {
        if (t < 0 || t >= g.nely) continue;
        int j = J - 4 * t;
        for (int l = 0; l < 5; l++) {
          double v = 0;
          for (int q = 0; q < 5; q++) v += g.D[q * 5 + j] * g.D[q * 5 + l] * gy(I, 4 * t + q, q);
          int Jc = 4 * t + l;
          if (Jc == 0) continue;  // fault column: not an unknown
          add((int64_t)Jc * NX + I, v);
        }
      }
      // sort by column, drop exact zeros, shift to the reduced index
      int ord[32];
      for (int q = 0; q < nc; q++) ord[q] = q;
      std::sort(ord, ord + nc, [&](int a, int b) { return cols[a] < cols[b]; });
      for (int q = 0; q < nc; q++) {
        double v = vals[ord[q]];
        if (v == 0.0) continue;
        K->col.push_back((int32_t)(cols[ord[q]] - NX));
        K->val.push_back(v);
      }
      int64_t r = (int64_t)(J - 1) * NX + I;
      K->rowptr[r + 1] = (int64_t)K->col.size();
    }
  }
}

void csr_transpose(const HostCsr& A, HostCsr* T) {
  T->n = A.m;
  T->m = A.n;
  T->rowptr.assign(A.m + 1, 0);
  int64_t nnz = A.nnz();
  T->col.resize(nnz);
  T->val.resize(nnz);
  for (int64_t q = 0; q < nnz; q++) T->rowptr[A.col[q] + 1]++;
  for (int64_t i = 0; i < A.m; i++) T->rowptr[i + 1] += T->rowptr[i];
  std::vector<int64_t> pos(T->rowptr.begin(), T->rowptr.end() - 1);
  for (int64_t i = 0; i < A.n; i++)
    for (int64_t q = A.rowptr[i]; q < A.rowptr[i + 1]; q++) {
      int64_t c = A.col[q];
      T->col[pos[c]] = (int32_t)i;
      T->val[pos[c]] = A.val[q];
      pos[c]++;
    }
}

static void csr_matmul(const HostCsr& A, const HostCsr& B, HostCsr* C) {
  const int64_t n = A.n, m = B.m;
  C->n = n;
  C->m = m;
  C->rowptr.assign(n + 1, 0);
  C->col.clear();
  C->val.clear();
  std::vector<int64_t> mark(m, -1);
  std::vector<double> acc(m, 0.0);
  std::vector<int32_t> cur;
  for (int64_t i = 0; i < n; i++) {
    cur.clear();
    for (int64_t q = A.rowptr[i]; q < A.rowptr[i + 1]; q++) {
      int64_t k = A.col[q];
      double a = A.val[q];
      for (int64_t s = B.rowptr[k]; s < B.rowptr[k + 1]; s++) {
        int32_t j = B.col[s];
        if (mark[j] != i) { mark[j] = i; cur.push_back(j); acc[j] = a * B.val[s]; }
        else acc[j] += a * B.val[s];
      }
    }
    std::sort(cur.begin(), cur.end());
    for (int32_t j : cur) { C->col.push_back(j); C->val.push_back(acc[j]); }
    C->rowptr[i + 1] = (int64_t)C->col.size();
  }
}

// C = alpha A + beta B (same shape), sorted union of the row patterns
static void csr_add(double alpha, const HostCsr& A, double beta, const HostCsr& B, HostCsr* C) {
  C->n = A.n;
  C->m = A.m;
  C->rowptr.assign(A.n + 1, 0);
  C->col.clear();
  C->val.clear();
  for (int64_t i = 0; i < A.n; i++) {
    int64_t p = A.rowptr[i], q = B.rowptr[i];
    const int64_t pe = A.rowptr[i + 1], qe = B.rowptr[i + 1];
    while (p < pe || q < qe) {
      const int32_t ca = p < pe ? A.col[p] : INT32_MAX, cb = q < qe ? B.col[q] : INT32_MAX;
      if (ca == cb) { C->col.push_back(ca); C->val.push_back(alpha * A.val[p++] + beta * B.val[q++]); }
      else if (ca < cb) { C->col.push_back(ca); C->val.push_back(alpha * A.val[p++]); }
      else { C->col.push_back(cb); C->val.push_back(beta * B.val[q++]); }
    }
    C->rowptr[i + 1] = (int64_t)C->col.size();
  }
}

// A <- diag(r) A diag(c) (either may be null)
static void csr_scale(HostCsr* A, const double* r, const double* c) {
  for (int64_t i = 0; i < A->n; i++)
    for (int64_t q = A->rowptr[i]; q < A->rowptr[i + 1]; q++) {
      if (r) A->val[q] *= r[i];
      if (c) A->val[q] *= c[A->col[q]];
    }
}

// sort the column indices (and values) of every row
static void csr_sort_rows(HostCsr* A) {
  std::vector<std::pair<int32_t, double>> row;
  for (int64_t i = 0; i < A->n; i++) {
    row.clear();
    for (int64_t q = A->rowptr[i]; q < A->rowptr[i + 1]; q++) row.emplace_back(A->col[q], A->val[q]);
    std::sort(row.begin(), row.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    int64_t q = A->rowptr[i];
    for (const auto& e : row) { A->col[q] = e.first; A->val[q] = e.second; q++; }
  }
}

void condense_level(const HostLevel& Lin, double omega, HostCsr* M, HostCsr* U, HostCsr* V) {
  HostLevel L = Lin;
  csr_sort_rows(&L.A);
  csr_sort_rows(&L.P);
  csr_sort_rows(&L.R);
  const double* dinv = L.dinv.data();
  // M = R - omega (R A) D^-1
  HostCsr RA;
  csr_matmul(L.R, L.A, &RA);
  csr_scale(&RA, nullptr, dinv);
  csr_add(1.0, L.R, -omega, RA, M);
  // U = 2 omega D^-1 - omega^2 D^-1 A D^-1
  HostCsr DAD = L.A, Dg;
  csr_scale(&DAD, dinv, dinv);
  Dg.n = Dg.m = L.A.n;
  Dg.rowptr.resize(L.A.n + 1);
  Dg.col.resize(L.A.n);
  Dg.val.resize(L.A.n);
  for (int64_t i = 0; i < L.A.n; i++) {
    Dg.rowptr[i] = i;
    Dg.col[i] = (int32_t)i;
    Dg.val[i] = dinv[i];
  }
  Dg.rowptr[L.A.n] = L.A.n;
  csr_add(2.0 * omega, Dg, -omega * omega, DAD, U);
  // V = P - omega D^-1 (A P)
  HostCsr AP;
  csr_matmul(L.A, L.P, &AP);
  csr_scale(&AP, dinv, nullptr);
  csr_add(1.0, L.P, -omega, AP, V);
}

void collapse_coarse(const std::vector<HostCsr>& M, const std::vector<HostCsr>& U, const std::vector<HostCsr>& V,
                     const std::vector<int64_t>& col_off, int64_t ncols, HostCsr* S, HostCsr* C) {
  const size_t k = M.size();   // condensed levels 1 .. k
  // down: T_1 = M_1, T_j = M_j T_(j-1), stacked
  std::vector<HostCsr> T(k);
  T[0] = M[0];
  for (size_t j = 1; j < k; j++) csr_matmul(M[j], T[j - 1], &T[j]);
  S->n = 0;
  S->m = M[0].m;
  S->rowptr.assign(1, 0);
  S->col.clear();
  S->val.clear();
  for (const HostCsr& B : T) {
    for (int64_t i = 0; i < B.n; i++) {
      for (int64_t q = B.rowptr[i]; q < B.rowptr[i + 1]; q++) {
        S->col.push_back(B.col[q]);
        S->val.push_back(B.val[q]);
      }
      S->rowptr.push_back((int64_t)S->col.size());
    }
    S->n += B.n;
  }
  // up: column blocks Q_j U_(j+1) (j = 0 .. k-1) and Q_k, Q_0 = I, Q_j = Q_(j-1) V_j
  std::vector<HostCsr> blk(k + 1);
  HostCsr Q;
  for (size_t j = 0; j < k; j++) {
    if (j == 0) {
      blk[0] = U[0];
      Q = V[0];
    } else {
      csr_matmul(Q, U[j], &blk[j]);
      HostCsr Qn;
      csr_matmul(Q, V[j], &Qn);
      Q = std::move(Qn);
    }
  }
  blk[k] = std::move(Q);
  const int64_t n = U[0].n;
  C->n = n;
  C->m = ncols;
  C->rowptr.assign(n + 1, 0);
  C->col.clear();
  C->val.clear();
  for (int64_t i = 0; i < n; i++) {
    for (size_t j = 0; j <= k; j++)
      for (int64_t q = blk[j].rowptr[i]; q < blk[j].rowptr[i + 1]; q++) {
        C->col.push_back((int32_t)(col_off[j] + blk[j].col[q]));
        C->val.push_back(blk[j].val[q]);
      }
    C->rowptr[i + 1] = (int64_t)C->col.size();
  }
}

// Strong connections of row i: |a_ij| >= theta max_{k!=i} |a_ik| (abs) or
// -a_ij >= theta max_{k!=i} (-a_ik) (signed).  The diagonal is kept, with A's
// values stored at the strong positions.
static void strength(const HostCsr& A, double theta, bool use_abs, HostCsr* S) {
  S->n = S->m = A.n;
  S->rowptr.assign(A.n + 1, 0);
  S->col.clear();
  S->val.clear();
  for (int64_t i = 0; i < A.n; i++) {
    double mx = 0.0;
    for (int64_t q = A.rowptr[i]; q < A.rowptr[i + 1]; q++)
      if (A.col[q] != i) mx = std::max(mx, use_abs ? std::fabs(A.val[q]) : -A.val[q]);
    double thr = theta * mx;
    for (int64_t q = A.rowptr[i]; q < A.rowptr[i + 1]; q++) {
      double a = A.val[q];
      bool keep = (A.col[q] == i) ||
                  (a != 0.0 && (use_abs ? std::fabs(a) >= thr : (mx > 0 && -a >= thr)));
      if (keep) { S->col.push_back(A.col[q]); S->val.push_back(a); }
    }
    S->rowptr[i + 1] = (int64_t)S->col.size();
  }
}

enum { U_PT = 0, C_PT = 1, F_PT = 2 };

// Bucketed RS first pass: repeatedly take the undecided point with the most
// strong dependents as C, make its dependents F, and raise the weight of the
// points those new F points depend on.
static void rs_split(const HostCsr& S, const HostCsr& T, std::vector<int>& split) {
  const int64_t n = S.n;
  std::vector<int64_t> lam(n), bucket_start(n + 1, 0), bucket_cnt(n + 1, 0), at(n), pos(n);
  for (int64_t i = 0; i < n; i++) { lam[i] = T.rowptr[i + 1] - T.rowptr[i]; bucket_cnt[lam[i]]++; }
  int64_t run = 0;
  for (int64_t b = 0; b < n; b++) { bucket_start[b] = run; run += bucket_cnt[b]; bucket_cnt[b] = 0; }
  for (int64_t i = 0; i < n; i++) {
    int64_t b = lam[i], p = bucket_start[b] + bucket_cnt[b]++;
    at[p] = i;
    pos[i] = p;
  }
  split.assign(n, U_PT);
  for (int64_t i = 0; i < n; i++)
    if (lam[i] == 0 || (lam[i] == 1 && T.col[T.rowptr[i]] == i)) split[i] = F_PT;
  auto swap_slots = [&](int64_t node, int64_t newp) {
    int64_t oldp = pos[node], other = at[newp];
    at[oldp] = other;
    pos[other] = oldp;
    at[newp] = node;
    pos[node] = newp;
  };
  for (int64_t top = n - 1; top >= 0; top--) {
    int64_t i = at[top];
    bucket_cnt[lam[i]]--;
    if (split[i] == F_PT) continue;
    split[i] = C_PT;
    for (int64_t q = T.rowptr[i]; q < T.rowptr[i + 1]; q++) {
      int64_t j = T.col[q];
      if (split[j] != U_PT) continue;
      split[j] = F_PT;
      for (int64_t r = S.rowptr[j]; r < S.rowptr[j + 1]; r++) {
        int64_t k = S.col[r];
        if (split[k] != U_PT || lam[k] >= n - 1) continue;
        int64_t b = lam[k], newp = bucket_start[b] + bucket_cnt[b] - 1;
        swap_slots(k, newp);
        bucket_cnt[b]--;
        bucket_cnt[b + 1]++;
        bucket_start[b + 1] = newp;
        lam[k]++;
      }
    }
    for (int64_t q = S.rowptr[i]; q < S.rowptr[i + 1]; q++) {
      int64_t j = S.col[q];
      if (split[j] != U_PT || lam[j] == 0) continue;
      int64_t b = lam[j], newp = bucket_start[b];
      swap_slots(j, newp);
      bucket_cnt[b]--;
      bucket_cnt[b - 1]++;
      bucket_start[b]++;
      lam[j]--;
    }
  }
}

static void direct_interpolation(const HostCsr& A, const HostCsr& S, const std::vector<int>& split, HostCsr* P) {
  const int64_t n = A.n;
  std::vector<int64_t> cidx(n, -1);
  int64_t nc = 0;
  for (int64_t i = 0; i < n; i++)
    if (split[i] == C_PT) cidx[i] = nc++;
  P->n = n;
  P->m = nc;
  P->rowptr.assign(n + 1, 0);
  P->col.clear();
  P->val.clear();
  for (int64_t i = 0; i < n; i++) {
    if (split[i] == C_PT) {
      P->col.push_back((int32_t)cidx[i]);
      P->val.push_back(1.0);
    } else {
      double spos = 0, sneg = 0, apos = 0, aneg = 0, diag = 0;
      for (int64_t q = S.rowptr[i]; q < S.rowptr[i + 1]; q++) {
        int64_t j = S.col[q];
        if (j != i && split[j] == C_PT) (S.val[q] < 0 ? sneg : spos) += S.val[q];
      }
      for (int64_t q = A.rowptr[i]; q < A.rowptr[i + 1]; q++) {
        if (A.col[q] == i) diag += A.val[q];
        else (A.val[q] < 0 ? aneg : apos) += A.val[q];
      }
      double alpha = sneg != 0 ? aneg / sneg : 0.0;
      double beta = spos != 0 ? apos / spos : 0.0;
      if (spos == 0) { diag += apos; beta = 0; }
      double cneg = -alpha / diag, cpos = -beta / diag;
      for (int64_t q = S.rowptr[i]; q < S.rowptr[i + 1]; q++) {
        int64_t j = S.col[q];
        if (j != i && split[j] == C_PT) {
          P->col.push_back((int32_t)cidx[j]);
          P->val.push_back(S.val[q] < 0 ? cneg * S.val[q] : cpos * S.val[q]);
        }
      }
    }
    P->rowptr[i + 1] = (int64_t)P->col.size();
  }
}

static bool dense_inverse(int64_t n, std::vector<double> a, std::vector<double>* inv) {
  inv->assign(n * n, 0.0);
  for (int64_t i = 0; i < n; i++) (*inv)[i * n + i] = 1.0;
  for (int64_t c = 0; c < n; c++) {
    int64_t p = c;
    for (int64_t r = c + 1; r < n; r++)
      if (std::fabs(a[r * n + c]) > std::fabs(a[p * n + c])) p = r;
    if (a[p * n + c] == 0.0) return false;
    if (p != c)
      for (int64_t j = 0; j < n; j++) {
        std::swap(a[c * n + j], a[p * n + j]);
        std::swap((*inv)[c * n + j], (*inv)[p * n + j]);
      }
    double d = a[c * n + c];
    for (int64_t j = 0; j < n; j++) { a[c * n + j] /= d; (*inv)[c * n + j] /= d; }
    for (int64_t r = 0; r < n; r++) {
      if (r == c) continue;
      double f = a[r * n + c];
      if (f == 0.0) continue;
      for (int64_t j = 0; j < n; j++) {
        a[r * n + j] -= f * a[c * n + j];
        (*inv)[r * n + j] -= f * (*inv)[c * n + j];
      }
    }
  }
  return true;
}

double estimate_lambda_max_dinvA(const HostCsr& A, const std::vector<double>& dinv, int iters) {
  const int64_t n = A.n;
  if (n == 0) return 1.0;
  std::vector<double> x(n), y(n);
  // deterministic start vector
  uint64_t s = 0x2545F4914F6CDD1DULL;
  for (int64_t i = 0; i < n; i++) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    x[i] = 0.5 + (double)(s >> 11) * (1.0 / 9007199254740992.0);
  }
  double lam = 1.0;
  for (int it = 0; it < iters; it++) {
    double nx = 0;
    for (int64_t i = 0; i < n; i++) nx += x[i] * x[i];
    nx = std::sqrt(nx);
    for (int64_t i = 0; i < n; i++) x[i] /= nx;
    for (int64_t i = 0; i < n; i++) {
      double acc = 0;
      for (int64_t q = A.rowptr[i]; q < A.rowptr[i + 1]; q++) acc += A.val[q] * x[A.col[q]];
      y[i] = dinv[i] * acc;
    }
    double num = 0;
    for (int64_t i = 0; i < n; i++) num += x[i] * y[i];
    lam = num;
    x.swap(y);
  }
  return lam;
}

int amg_build(const HostCsr& A0, const AmgOptions& opt, HostHierarchy* H, std::string* err,
              std::vector<std::vector<int>>* splits) {
  H->lev.clear();
  if (splits) splits->clear();
  HostLevel L0;
  L0.A = A0;
  const double sgn = opt.build_on_negated ? -1.0 : 1.0;
  if (sgn < 0)
    for (auto& v : L0.A.val) v = -v;
  H->lev.push_back(std::move(L0));
  while ((int)H->lev.size() < opt.max_levels && H->lev.back().A.n > opt.max_coarse) {
    HostLevel& cur = H->lev.back();
    HostCsr S, T;
    strength(cur.A, opt.theta, opt.strength_abs != 0, &S);
    csr_transpose(S, &T);
    std::vector<int> split;
    rs_split(S, T, split);
    direct_interpolation(cur.A, S, split, &cur.P);
    if (cur.P.m == 0 || cur.P.m == cur.A.n) {
      cur.P = HostCsr();
      break;
    }
    if (splits) {
      splits->emplace_back(split.size());
      for (size_t i = 0; i < split.size(); i++) splits->back()[i] = split[i] == C_PT;
    }
    csr_transpose(cur.P, &cur.R);
    HostCsr RA, Ac;
    csr_matmul(cur.R, cur.A, &RA);
    csr_matmul(RA, cur.P, &Ac);
    HostLevel next;
    next.A = std::move(Ac);
    H->lev.push_back(std::move(next));
  }
  // back to the sign of the operator the GPU applies (K_II, SPD)
  if (sgn < 0)
    for (auto& L : H->lev)
      for (auto& v : L.A.val) v = -v;
  return finish_hierarchy(H, nullptr, err);
}

// D^-1, lambda_max(D^-1 A) (or the caller's smoothing weights) and the dense
// coarsest inverse of a hierarchy whose operators and prolongations are set.
int finish_hierarchy(HostHierarchy* H, const double* omega, std::string* err) {
  for (size_t l = 0; l < H->lev.size(); l++) {
    HostLevel& L = H->lev[l];
    L.dinv.assign(L.A.n, 0.0);
    for (int64_t i = 0; i < L.A.n; i++)
      for (int64_t q = L.A.rowptr[i]; q < L.A.rowptr[i + 1]; q++)
        if (L.A.col[q] == i) L.dinv[i] = 1.0 / L.A.val[q];
    for (int64_t i = 0; i < L.A.n; i++)
      if (!(L.dinv[i] > 0) || !std::isfinite(L.dinv[i])) {
        if (err) *err = "AMG: level " + std::to_string(l) + " has a zero or negative diagonal entry";
        return EQ_EINVAL;
      }
    // omega = 4 / (3 lambda_max) is the Jacobi weight the V-cycle uses
    L.lambda_max = omega ? 4.0 / (3.0 * omega[l]) : estimate_lambda_max_dinvA(L.A, L.dinv, 30);
  }
  const HostCsr& Ac = H->lev.back().A;
  H->ncoarse = Ac.n;
  std::vector<double> dense(Ac.n * Ac.n, 0.0);
  for (int64_t i = 0; i < Ac.n; i++)
    for (int64_t q = Ac.rowptr[i]; q < Ac.rowptr[i + 1]; q++) dense[i * Ac.n + Ac.col[q]] += Ac.val[q];
  if (!dense_inverse(Ac.n, dense, &H->coarse_inv)) {
    if (err) *err = "AMG: singular coarsest operator";
    return EQ_EINVAL;
  }
  return EQ_OK;
}

}  // namespace eq

namespace eq {

// p-multigrid first step: bilinear (Q1) functions on the element vertices
// embedded in the degree-4 GLL space.  Rows: K_II unknowns (structured
// interior order), columns: vertices (ex, ey) with ey >= 1 (vertices on the
// fault row carry the Dirichlet data and are not unknowns).
void build_pmg_p0(const Geom& g, HostCsr* P) {
  const int NX = g.NX, NY = g.NY;
  const int nvx = g.nelx + 1;
  const int64_t nI = (int64_t)NX * (NY - 1);
  P->n = nI;
  P->m = (int64_t)nvx * g.nely;
  P->rowptr.assign(nI + 1, 0);
  P->col.clear();
  P->val.clear();
  for (int J = 1; J < NY; J++) {
    const int ey = std::min(J / 4, g.nely - 1), j = J - 4 * ey;
    const double wy[2] = {0.5 * (1.0 - g.x[j]), 0.5 * (1.0 + g.x[j])};
    for (int I = 0; I < NX; I++) {
      const int ex = std::min(I / 4, g.nelx - 1), i = I - 4 * ex;
      const double wx[2] = {0.5 * (1.0 - g.x[i]), 0.5 * (1.0 + g.x[i])};
      int32_t cc[4];
      double vv[4];
      int nc = 0;
      for (int b = 0; b < 2; b++)
        for (int a = 0; a < 2; a++) {
          const double w = wx[a] * wy[b];
          const int vy = ey + b;
          if (vy == 0 || w == 0.0) continue;
          cc[nc] = (int32_t)((int64_t)(vy - 1) * nvx + ex + a);
          vv[nc] = w;
          nc++;
        }
      // columns ascending
      for (int p = 1; p < nc; p++)
        for (int q = p; q > 0 && cc[q] < cc[q - 1]; q--) {
          std::swap(cc[q], cc[q - 1]);
          std::swap(vv[q], vv[q - 1]);
        }
      for (int p = 0; p < nc; p++) {
        P->col.push_back(cc[p]);
        P->val.push_back(vv[p]);
      }
      P->rowptr[(int64_t)(J - 1) * NX + I + 1] = (int64_t)P->col.size();
    }
  }
}

// Geometric semi-coarsening of a structured vertex grid (nx columns, ny
// unknown rows above the Dirichlet fault row, index row * nx + column) in one
// direction: every other line is kept (and the last), a dropped line takes the
// mean of its two neighbours (linear interpolation on a uniform grid); in y
// the Dirichlet row below the first unknown row counts as a kept line of zeros.
static void semi_coarsen_p(int nx, int ny, bool in_x, HostCsr* P, int* cnx, int* cny) {
  std::vector<int> keep;
  const int nl = in_x ? nx : ny;
  // kept lines: x columns 0, 2, 4, ... and nx - 1; y rows (1-based above the
  // fault) 2, 4, ... and ny, i.e. 0-based unknown rows 1, 3, ... and ny - 1
  std::vector<int> pos(nl, -1);
  for (int l = 0; l < nl; l++) {
    const bool kept = in_x ? (l % 2 == 0 || l == nl - 1) : (l % 2 == 1 || l == nl - 1);
    if (kept) { pos[l] = (int)keep.size(); keep.push_back(l); }
  }
  const int nk = (int)keep.size();
  *cnx = in_x ? nk : nx;
  *cny = in_x ? ny : nk;
  P->n = (int64_t)nx * ny;
  P->m = (int64_t)(*cnx) * (*cny);
  P->rowptr.assign(P->n + 1, 0);
  P->col.clear();
  P->val.clear();
  for (int yy = 0; yy < ny; yy++)
    for (int xx = 0; xx < nx; xx++) {
      const int l = in_x ? xx : yy;
      auto coarse = [&](int lc) {
        return in_x ? (int64_t)yy * (*cnx) + pos[lc] : (int64_t)pos[lc] * (*cnx) + xx;
      };
      if (pos[l] >= 0) {
        P->col.push_back((int32_t)coarse(l));
        P->val.push_back(1.0);
      } else {
        // neighbours l - 1 and l + 1; below y row 0 lies the Dirichlet row
        if (l - 1 >= 0) { P->col.push_back((int32_t)coarse(l - 1)); P->val.push_back(0.5); }
        if (l + 1 < nl) { P->col.push_back((int32_t)coarse(l + 1)); P->val.push_back(0.5); }
      }
      P->rowptr[(int64_t)yy * nx + xx + 1] = (int64_t)P->col.size();
    }
}

int pmg_build(const Geom& g, const HostCsr& Kii, const AmgOptions& opt, HostHierarchy* H, std::string* err) {
  HostLevel L0;
  L0.A = Kii;
  build_pmg_p0(g, &L0.P);
  csr_transpose(L0.P, &L0.R);
  HostCsr RA, A1;
  csr_matmul(L0.R, Kii, &RA);
  csr_matmul(RA, L0.P, &A1);
  H->lev.clear();
  L0.dinv.assign(Kii.n, 0.0);
  for (int64_t i = 0; i < Kii.n; i++)
    for (int64_t q = Kii.rowptr[i]; q < Kii.rowptr[i + 1]; q++)
      if (Kii.col[q] == i) L0.dinv[i] = 1.0 / Kii.val[q];
  L0.lambda_max = estimate_lambda_max_dinvA(Kii, L0.dinv, 40);
  H->lev.push_back(std::move(L0));
  // Elongated elements: the Q1 operator is anisotropic (and has positive
  // off-diagonal entries), which defeats point smoothing with RS coarsening;
  // semi-coarsen the vertex grid geometrically in the finer direction until
  // its spacing is within a factor 1.6 of the other, then hand over to RS.
  int nx = g.nelx + 1, ny = g.nely;
  double hx = 1.0 / g.c1, hy = 1.0;   // element aspect dx / dy = dye / dxe scaled
  HostCsr Acur = std::move(A1);
  while (Acur.n > opt.max_coarse && (int)H->lev.size() < opt.max_levels - 1) {
    const bool in_x = hx * 1.6 <= hy, in_y = hy * 1.6 <= hx;
    if (!in_x && !in_y) break;
    if ((in_x ? nx : ny) < 3) break;
    HostLevel L;
    L.A = std::move(Acur);
    int cnx, cny;
    semi_coarsen_p(nx, ny, in_x, &L.P, &cnx, &cny);
    csr_transpose(L.P, &L.R);
    HostCsr RA2, Ac;
    csr_matmul(L.R, L.A, &RA2);
    csr_matmul(RA2, L.P, &Ac);
    L.dinv.assign(L.A.n, 0.0);
    for (int64_t i = 0; i < L.A.n; i++)
      for (int64_t q = L.A.rowptr[i]; q < L.A.rowptr[i + 1]; q++)
        if (L.A.col[q] == i) L.dinv[i] = 1.0 / L.A.val[q];
    L.lambda_max = estimate_lambda_max_dinvA(L.A, L.dinv, 30);
    H->lev.push_back(std::move(L));
    Acur = std::move(Ac);
    if (in_x) hx *= 2; else hy *= 2;
    nx = cnx;
    ny = cny;
  }
  AmgOptions o = opt;
  o.max_levels = std::max(1, opt.max_levels - (int)H->lev.size());
  HostHierarchy sub;
  if (int rc = amg_build(Acur, o, &sub, err)) return rc;
  for (auto& L : sub.lev) H->lev.push_back(std::move(L));
  H->coarse_inv = std::move(sub.coarse_inv);
  H->ncoarse = sub.ncoarse;
  return EQ_OK;
}

}  // namespace eq

// ---------------------------------------------------------------------------
// Host-side Ruge-Stueben set-up exported for inspection (no GPU involved):
// the hierarchy eq_amg_setup builds with its "rs" choice, from a caller's
// matrix.
// ---------------------------------------------------------------------------
struct eq_rs {
  eq::HostHierarchy H;
  std::vector<std::vector<int>> split;
};

extern "C" int eq_rs_build(const eq_csr* A, double theta, int max_levels, int64_t max_coarse, int negate,
                           eq_rs** out) {
  if (!A || !out || !A->rowptr || A->n <= 0 || A->m != A->n || !(theta >= 0) || max_levels < 1) return EQ_EINVAL;
  *out = nullptr;
  const int64_t n = A->n;
  if (A->rowptr[0] != 0) return EQ_EINVAL;
  for (int64_t i = 0; i < n; i++)
    if (A->rowptr[i + 1] < A->rowptr[i]) return EQ_EINVAL;
  const int64_t nnz = A->rowptr[n];
  if (nnz > 0 && (!A->col || !A->val)) return EQ_EINVAL;
  eq::HostCsr C;
  C.n = C.m = n;
  C.rowptr.assign(A->rowptr, A->rowptr + n + 1);
  C.col.resize(nnz);
  C.val.assign(A->val, A->val + nnz);
  for (int64_t q = 0; q < nnz; q++) {
    if (A->col[q] < 0 || A->col[q] >= n) return EQ_EINVAL;
    C.col[q] = (int32_t)A->col[q];
  }
  eq::AmgOptions o;
  o.theta = theta;
  o.max_levels = max_levels;
  o.max_coarse = max_coarse;
  o.build_on_negated = negate != 0;
  auto* h = new (std::nothrow) eq_rs;
  if (!h) return EQ_NOMEM;
  std::string err;
  if (int rc = eq::amg_build(C, o, &h->H, &err, &h->split)) {
    delete h;
    return rc;
  }
  *out = h;
  return EQ_OK;
}

extern "C" int eq_rs_info(const eq_rs* h, int* nlev, int64_t* n, int64_t* nnz_a, int64_t* nnz_p, int cap) {
  if (!h || !nlev) return EQ_EINVAL;
  *nlev = (int)h->H.lev.size();
  for (int l = 0; l < *nlev && l < cap; l++) {
    const eq::HostLevel& L = h->H.lev[l];
    if (n) n[l] = L.A.n;
    if (nnz_a) nnz_a[l] = L.A.nnz();
    if (nnz_p) nnz_p[l] = (l + 1 < *nlev) ? L.P.nnz() : 0;
  }
  return EQ_OK;
}

extern "C" int eq_rs_level(const eq_rs* h, int l, int which, int64_t* rowptr, int64_t* col, double* val) {
  if (!h || l < 0 || l >= (int)h->H.lev.size() || (which != 0 && which != 1)) return EQ_EINVAL;
  if (which == 1 && l + 1 >= (int)h->H.lev.size()) return EQ_EINVAL;
  const eq::HostCsr& M = which == 0 ? h->H.lev[l].A : h->H.lev[l].P;
  for (int64_t i = 0; i <= M.n; i++) rowptr[i] = M.rowptr[i];
  for (int64_t q = 0; q < M.nnz(); q++) {
    col[q] = M.col[q];
    val[q] = M.val[q];
  }
  return EQ_OK;
}

extern "C" int eq_rs_split(const eq_rs* h, int l, int* is_c) {
  if (!h || !is_c || l < 0 || l >= (int)h->split.size()) return EQ_EINVAL;
  std::copy(h->split[l].begin(), h->split[l].end(), is_c);
  return EQ_OK;
}

extern "C" void eq_rs_free(eq_rs* h) { delete h; }

/*
 * eqoracle.c -- CPU FP64 restatement of the eqcycle time-stepping arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the reference that the HIP library's
 * results are compared against (and the CPU baseline that bench.py times).  It is never
 * linked into, loaded by or called from the product library.
 *
 * Every function names the reference file:line it restates.  Index arrays are
 * 0-based here; the Python wrapper converts the reference's 1-based arrays.
 * Sparse matrices are CSR with int64 indices.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int64_t n, m;       /* rows, cols */
    int64_t *rowptr;    /* n+1 */
    int64_t *col;       /* nnz */
    double *val;        /* nnz */
} orc_csr;

static int64_t nnz_of(const orc_csr *A) { return A->rowptr[A->n]; }

void orc_csr_free(orc_csr *A) {
    if (!A) return;
    free(A->rowptr); free(A->col); free(A->val);
    A->rowptr = NULL; A->col = NULL; A->val = NULL; A->n = A->m = 0;
}

static int cmp_i64(const void *a, const void *b) {
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

/* ------------------------------------------------------------------------ */
/* Stiffness assembly: K_element (src/Kassemble.jl:31-63) summed per element  */
/* as FEMSparse.assemble_local_matrix! does (src/Kassemble.jl:20-29).         */
/* iglob: [Nel][5 j][5 i] 0-based; W: [Nel][5 j][5 i]; H column-major:       */
/* H[i + 5 p] = h'_i(xi_p).                                                   */
/* ------------------------------------------------------------------------ */
static void k_element(const double *W, const double *H, double dxe, double dye, double *Ke) {
    const double dx_dxi = 0.5 * dxe, dy_deta = 0.5 * dye, jac = dx_dxi * dy_deta;
    const double c_eta = jac / (dy_deta * dy_deta), c_xi = jac / (dx_dxi * dx_dxi);
#define WW(a, b) W[(b) * 5 + (a)]   /* ww[a,b] = W[a + 5 b] */
#define HH(a, b) H[(b) * 5 + (a)]
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++)
            for (int k = 0; k < 5; k++)
                for (int l = 0; l < 5; l++) {
                    double term1 = 0.0, term2 = 0.0;
                    for (int p = 0; p < 5; p++) {
                        double dik = (i == k) ? 1.0 : 0.0, djl = (j == l) ? 1.0 : 0.0;
                        term1 += dik * WW(k, p) * c_eta * HH(j, p) * HH(l, p);
                        term2 += djl * WW(p, j) * c_xi * HH(i, p) * HH(k, p);
                    }
                    /* Ke[(i,j),(k,l)] with (i,j) -> i + 5 j */
                    Ke[(i + 5 * j) * 25 + (k + 5 * l)] = term1 + term2;
                }
#undef WW
#undef HH
}

int orc_assemble_K(int64_t nel, int64_t nglob, const int64_t *iglob, const double *W,
                   const double *H, double dxe, double dye, orc_csr *K) {
    /* node -> element incidence (each node lies in at most 4 elements) */
    int64_t *ecount = calloc(nglob + 1, sizeof(int64_t));
    if (!ecount) return -1;
    for (int64_t e = 0; e < nel; e++)
        for (int a = 0; a < 25; a++) ecount[iglob[e * 25 + a] + 1]++;
    for (int64_t i = 0; i < nglob; i++) ecount[i + 1] += ecount[i];
    int64_t *elist = malloc(sizeof(int64_t) * ecount[nglob]);
    int64_t *eloc = malloc(sizeof(int64_t) * ecount[nglob]);
    int64_t *fill = calloc(nglob, sizeof(int64_t));
    double *Ke = malloc(sizeof(double) * 625 * (size_t)nel);
    if (!elist || !eloc || !fill || !Ke) return -1;
    for (int64_t e = 0; e < nel; e++) {
        for (int a = 0; a < 25; a++) {
            int64_t g = iglob[e * 25 + a];
            elist[ecount[g] + fill[g]] = e;
            eloc[ecount[g] + fill[g]] = a;
            fill[g]++;
        }
        k_element(W + e * 25, H, dxe, dye, Ke + (size_t)e * 625);
    }
    /* row pattern: union of the element node sets */
    int64_t *rowptr = calloc(nglob + 1, sizeof(int64_t));
    int64_t buf[128];
    for (int64_t r = 0; r < nglob; r++) {
        int nb = 0;
        for (int64_t q = ecount[r]; q < ecount[r + 1]; q++)
            for (int b = 0; b < 25; b++) buf[nb++] = iglob[elist[q] * 25 + b];
        qsort(buf, nb, sizeof(int64_t), cmp_i64);
        int u = 0;
        for (int t = 0; t < nb; t++) if (t == 0 || buf[t] != buf[t - 1]) u++;
        rowptr[r + 1] = rowptr[r] + u;
    }
    int64_t *col = malloc(sizeof(int64_t) * rowptr[nglob]);
    double *val = calloc(rowptr[nglob], sizeof(double));
    for (int64_t r = 0; r < nglob; r++) {
        int nb = 0;
        for (int64_t q = ecount[r]; q < ecount[r + 1]; q++)
            for (int b = 0; b < 25; b++) buf[nb++] = iglob[elist[q] * 25 + b];
        qsort(buf, nb, sizeof(int64_t), cmp_i64);
        int64_t o = rowptr[r];
        for (int t = 0; t < nb; t++) if (t == 0 || buf[t] != buf[t - 1]) col[o++] = buf[t];
        /* sum element contributions in element order */
        for (int64_t q = ecount[r]; q < ecount[r + 1]; q++) {
            int64_t e = elist[q];
            int a = (int)eloc[q];
            for (int b = 0; b < 25; b++) {
                int64_t c = iglob[e * 25 + b];
                /* binary search c in the row */
                int64_t lo = rowptr[r], hi = rowptr[r + 1] - 1;
                while (lo < hi) { int64_t mid = (lo + hi) / 2; if (col[mid] < c) lo = mid + 1; else hi = mid; }
                val[lo] += Ke[(size_t)e * 625 + a * 25 + b];
            }
        }
    }
    /* dropzeros! (src/Kassemble.jl:16) */
    int64_t o = 0, start = 0;
    for (int64_t r = 0; r < nglob; r++) {
        int64_t end = rowptr[r + 1];
        for (int64_t q = start; q < end; q++)
            if (val[q] != 0.0) { col[o] = col[q]; val[o] = val[q]; o++; }
        start = end;
        rowptr[r + 1] = o;
    }
    free(ecount); free(elist); free(eloc); free(fill); free(Ke);
    K->n = K->m = nglob; K->rowptr = rowptr; K->col = col; K->val = val;
    return 0;
}

/* Threads of the OpenMP loops below (SpMV, the coarse dense product, CG's  */
/* dots and updates).  One thread, the default, is the serial order; more   */
/* threads change only the summation order of the dot products.  The        */
/* Gauss-Seidel sweeps stay sequential, as in the reference's smoother.     */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

#define ORC_PAR_MIN 16384   /* shorter loops run on one thread */

void orc_spmv(const orc_csr *A, const double *x, double *y) {
#pragma omp parallel for schedule(static) if (A->n >= ORC_PAR_MIN)
    for (int64_t i = 0; i < A->n; i++) {
        double s = 0.0;
        for (int64_t q = A->rowptr[i]; q < A->rowptr[i + 1]; q++) s += A->val[q] * x[A->col[q]];
        y[i] = s;
    }
}

/* B = scale * A[idx, idx] with idx sorted ascending (kni = -K[FltNI,FltNI], src/main.jl:195) */
int orc_csr_submatrix(const orc_csr *A, const int64_t *idx, int64_t nidx, double scale, orc_csr *B) {
    int64_t *map = malloc(sizeof(int64_t) * A->m);
    if (!map) return -1;
    for (int64_t i = 0; i < A->m; i++) map[i] = -1;
    for (int64_t k = 0; k < nidx; k++) map[idx[k]] = k;
    int64_t *rp = calloc(nidx + 1, sizeof(int64_t));
    for (int64_t k = 0; k < nidx; k++) {
        int64_t r = idx[k], c = 0;
        for (int64_t q = A->rowptr[r]; q < A->rowptr[r + 1]; q++) if (map[A->col[q]] >= 0) c++;
        rp[k + 1] = rp[k] + c;
    }
    int64_t *cl = malloc(sizeof(int64_t) * rp[nidx]);
    double *vl = malloc(sizeof(double) * rp[nidx]);
    for (int64_t k = 0; k < nidx; k++) {
        int64_t r = idx[k], o = rp[k];
        for (int64_t q = A->rowptr[r]; q < A->rowptr[r + 1]; q++) {
            int64_t m = map[A->col[q]];
            if (m >= 0) { cl[o] = m; vl[o] = scale * A->val[q]; o++; }
        }
    }
    free(map);
    B->n = B->m = nidx; B->rowptr = rp; B->col = cl; B->val = vl;
    return 0;
}

static int csr_transpose(const orc_csr *A, orc_csr *T) {
    int64_t nnz = nnz_of(A);
    int64_t *rp = calloc(A->m + 1, sizeof(int64_t));
    int64_t *cl = malloc(sizeof(int64_t) * (nnz ? nnz : 1));
    double *vl = malloc(sizeof(double) * (nnz ? nnz : 1));
    if (!rp || !cl || !vl) return -1;
    for (int64_t q = 0; q < nnz; q++) rp[A->col[q] + 1]++;
    for (int64_t i = 0; i < A->m; i++) rp[i + 1] += rp[i];
    int64_t *pos = malloc(sizeof(int64_t) * (A->m + 1));
    memcpy(pos, rp, sizeof(int64_t) * (A->m + 1));
    for (int64_t i = 0; i < A->n; i++)
        for (int64_t q = A->rowptr[i]; q < A->rowptr[i + 1]; q++) {
            int64_t c = A->col[q];
            cl[pos[c]] = i; vl[pos[c]] = A->val[q]; pos[c]++;
        }
    free(pos);
    T->n = A->m; T->m = A->n; T->rowptr = rp; T->col = cl; T->val = vl;
    return 0;
}

/* C = A * B (Gustavson, columns sorted per row) */
static int csr_matmul(const orc_csr *A, const orc_csr *B, orc_csr *C) {
    int64_t n = A->n, m = B->m;
    int64_t *mark = malloc(sizeof(int64_t) * (m ? m : 1));
    double *acc = calloc(m ? m : 1, sizeof(double));
    int64_t *rp = calloc(n + 1, sizeof(int64_t));
    if (!mark || !acc || !rp) return -1;
    for (int64_t j = 0; j < m; j++) mark[j] = -1;
    for (int64_t i = 0; i < n; i++) {
        int64_t c = 0;
        for (int64_t q = A->rowptr[i]; q < A->rowptr[i + 1]; q++) {
            int64_t k = A->col[q];
            for (int64_t s = B->rowptr[k]; s < B->rowptr[k + 1]; s++) {
                int64_t j = B->col[s];
                if (mark[j] != i) { mark[j] = i; c++; }
            }
        }
        rp[i + 1] = rp[i] + c;
    }
    int64_t *cl = malloc(sizeof(int64_t) * (rp[n] ? rp[n] : 1));
    double *vl = malloc(sizeof(double) * (rp[n] ? rp[n] : 1));
    for (int64_t j = 0; j < m; j++) mark[j] = -1;
    for (int64_t i = 0; i < n; i++) {
        int64_t o = rp[i];
        for (int64_t q = A->rowptr[i]; q < A->rowptr[i + 1]; q++) {
            int64_t k = A->col[q];
            double a = A->val[q];
            for (int64_t s = B->rowptr[k]; s < B->rowptr[k + 1]; s++) {
                int64_t j = B->col[s];
                if (mark[j] != i) { mark[j] = i; cl[o++] = j; acc[j] = a * B->val[s]; }
                else acc[j] += a * B->val[s];
            }
        }
        qsort(cl + rp[i], o - rp[i], sizeof(int64_t), cmp_i64);
        for (int64_t t = rp[i]; t < o; t++) vl[t] = acc[cl[t]];
    }
    free(mark); free(acc);
    C->n = n; C->m = m; C->rowptr = rp; C->col = cl; C->val = vl;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Ruge-Stueben AMG as AlgebraicMultigrid.jl's ruge_stuben(A) defaults       */
/* (called at src/main.jl:200): Classical(0.25) strength on |a_ij|, RS       */
/* C/F splitting, direct interpolation, R = P', Galerkin R*A*P, max_levels   */
/* 10, max_coarse 10, symmetric Gauss-Seidel pre/post smoothing, pinv        */
/* coarse solve.  The package is not vendored in the reference and its       */
/* version is unpinned; this restates its published algorithm.               */
/* ------------------------------------------------------------------------ */
#define ORC_MAXLEV 32
typedef struct {
    int nlev;                    /* number of levels incl. coarsest */
    orc_csr A[ORC_MAXLEV];       /* A[0] = fine operator (a copy) */
    orc_csr P[ORC_MAXLEV];
    orc_csr R[ORC_MAXLEV];
    double *coarse_inv;          /* dense inverse of A[nlev-1], row-major */
    int64_t ncoarse;
    double *work[ORC_MAXLEV][3]; /* res, coarse_b, coarse_x */
    int smoother;                /* 0 = symmetric GS */
    int *split[ORC_MAXLEV];      /* C_NODE / F_NODE of each level with a coarser one */
} orc_amg;

/* strength: S_ij kept when |a_ij| >= theta * max_{k != i} |a_ik|; diagonal kept */
static int classical_strength(const orc_csr *A, double theta, orc_csr *S) {
    int64_t n = A->n;
    int64_t *rp = calloc(n + 1, sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) {
        double mx = 0.0;
        for (int64_t q = A->rowptr[i]; q < A->rowptr[i + 1]; q++)
            if (A->col[q] != i) mx = fmax(mx, fabs(A->val[q]));
        double thr = theta * mx;
        int64_t c = 0;
        for (int64_t q = A->rowptr[i]; q < A->rowptr[i + 1]; q++)
            if (A->col[q] == i || (fabs(A->val[q]) >= thr && A->val[q] != 0.0)) c++;
        rp[i + 1] = rp[i] + c;
    }
    int64_t *cl = malloc(sizeof(int64_t) * rp[n]);
    double *vl = malloc(sizeof(double) * rp[n]);
    for (int64_t i = 0; i < n; i++) {
        double mx = 0.0;
        for (int64_t q = A->rowptr[i]; q < A->rowptr[i + 1]; q++)
            if (A->col[q] != i) mx = fmax(mx, fabs(A->val[q]));
        double thr = theta * mx;
        int64_t o = rp[i];
        for (int64_t q = A->rowptr[i]; q < A->rowptr[i + 1]; q++)
            if (A->col[q] == i || (fabs(A->val[q]) >= thr && A->val[q] != 0.0)) {
                cl[o] = A->col[q]; vl[o] = A->val[q]; o++;   /* A's values at strong positions */
            }
    }
    S->n = S->m = n; S->rowptr = rp; S->col = cl; S->val = vl;
    return 0;
}

enum { U_NODE = 0, C_NODE = 1, F_NODE = 2 };

static void rs_split(const orc_csr *S, const orc_csr *T, int *split) {
    int64_t n = S->n;
    int64_t *lambda = malloc(sizeof(int64_t) * n);
    int64_t *iptr = calloc(n + 1, sizeof(int64_t));
    int64_t *icnt = calloc(n + 1, sizeof(int64_t));
    int64_t *i2n = malloc(sizeof(int64_t) * n);
    int64_t *n2i = malloc(sizeof(int64_t) * n);
    for (int64_t i = 0; i < n; i++) { lambda[i] = T->rowptr[i + 1] - T->rowptr[i]; icnt[lambda[i]]++; }
    for (int64_t i = 0, cum = 0; i < n; i++) { iptr[i] = cum; cum += icnt[i]; icnt[i] = 0; }
    for (int64_t i = 0; i < n; i++) {
        int64_t li = lambda[i], idx = iptr[li] + icnt[li];
        i2n[idx] = i; n2i[i] = idx; icnt[li]++;
    }
    for (int64_t i = 0; i < n; i++) {
        split[i] = U_NODE;
        if (lambda[i] == 0 || (lambda[i] == 1 && T->col[T->rowptr[i]] == i)) split[i] = F_NODE;
    }
    for (int64_t top = n - 1; top >= 0; top--) {
        int64_t i = i2n[top], li = lambda[i];
        icnt[li]--;
        if (split[i] == F_NODE) continue;
        split[i] = C_NODE;
        for (int64_t jj = T->rowptr[i]; jj < T->rowptr[i + 1]; jj++) {
            int64_t j = T->col[jj];
            if (split[j] != U_NODE) continue;
            split[j] = F_NODE;
            for (int64_t kk = S->rowptr[j]; kk < S->rowptr[j + 1]; kk++) {
                int64_t k = S->col[kk];
                if (split[k] != U_NODE) continue;
                if (lambda[k] >= n - 1) continue;
                int64_t lk = lambda[k], oldp = n2i[k], newp = iptr[lk] + icnt[lk] - 1;
                n2i[i2n[newp]] = oldp; i2n[oldp] = i2n[newp];
                n2i[k] = newp; i2n[newp] = k;
                icnt[lk]--; icnt[lk + 1]++; iptr[lk + 1] = newp;
                lambda[k]++;
            }
        }
        for (int64_t jj = S->rowptr[i]; jj < S->rowptr[i + 1]; jj++) {
            int64_t j = S->col[jj];
            if (split[j] != U_NODE) continue;
            if (lambda[j] == 0) continue;
            int64_t lj = lambda[j], oldp = n2i[j], newp = iptr[lj];
            n2i[i2n[newp]] = oldp; i2n[oldp] = i2n[newp];
            n2i[j] = newp; i2n[newp] = j;
            icnt[lj]--; icnt[lj - 1]++; iptr[lj]++;
            lambda[j]--;
        }
    }
    free(lambda); free(iptr); free(icnt); free(i2n); free(n2i);
}

static int direct_interp(const orc_csr *A, const orc_csr *S, const int *split, orc_csr *P) {
    int64_t n = A->n;
    int64_t *cidx = malloc(sizeof(int64_t) * n);
    int64_t nc = 0;
    for (int64_t i = 0; i < n; i++) cidx[i] = (split[i] == C_NODE) ? nc++ : -1;
    int64_t *rp = calloc(n + 1, sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) {
        int64_t c = 0;
        if (split[i] == C_NODE) c = 1;
        else
            for (int64_t q = S->rowptr[i]; q < S->rowptr[i + 1]; q++)
                if (split[S->col[q]] == C_NODE && S->col[q] != i) c++;
        rp[i + 1] = rp[i] + c;
    }
    int64_t *cl = malloc(sizeof(int64_t) * (rp[n] ? rp[n] : 1));
    double *vl = malloc(sizeof(double) * (rp[n] ? rp[n] : 1));
    for (int64_t i = 0; i < n; i++) {
        int64_t o = rp[i];
        if (split[i] == C_NODE) { cl[o] = cidx[i]; vl[o] = 1.0; continue; }
        double spos = 0, sneg = 0, apos = 0, aneg = 0, diag = 0;
        for (int64_t q = S->rowptr[i]; q < S->rowptr[i + 1]; q++) {
            int64_t j = S->col[q];
            if (split[j] == C_NODE && j != i) { if (S->val[q] < 0) sneg += S->val[q]; else spos += S->val[q]; }
        }
        for (int64_t q = A->rowptr[i]; q < A->rowptr[i + 1]; q++) {
            int64_t j = A->col[q];
            if (j == i) diag += A->val[q];
            else if (A->val[q] < 0) aneg += A->val[q];
            else apos += A->val[q];
        }
        double alpha = (sneg != 0) ? aneg / sneg : 0.0;
        double beta = (spos != 0) ? apos / spos : 0.0;
        if (spos == 0) { diag += apos; beta = 0; }
        double negc = -alpha / diag, posc = -beta / diag;
        for (int64_t q = S->rowptr[i]; q < S->rowptr[i + 1]; q++) {
            int64_t j = S->col[q];
            if (split[j] == C_NODE && j != i) {
                cl[o] = cidx[j];
                vl[o] = (S->val[q] < 0) ? negc * S->val[q] : posc * S->val[q];
                o++;
            }
        }
    }
    free(cidx);
    P->n = n; P->m = nc; P->rowptr = rp; P->col = cl; P->val = vl;
    return 0;
}

static int csr_copy(const orc_csr *A, orc_csr *B) {
    int64_t nnz = nnz_of(A);
    B->n = A->n; B->m = A->m;
    B->rowptr = malloc(sizeof(int64_t) * (A->n + 1));
    B->col = malloc(sizeof(int64_t) * (nnz ? nnz : 1));
    B->val = malloc(sizeof(double) * (nnz ? nnz : 1));
    if (!B->rowptr || !B->col || !B->val) return -1;
    memcpy(B->rowptr, A->rowptr, sizeof(int64_t) * (A->n + 1));
    memcpy(B->col, A->col, sizeof(int64_t) * nnz);
    memcpy(B->val, A->val, sizeof(double) * nnz);
    return 0;
}

/* dense inverse by Gauss-Jordan with partial pivoting (pinv of a nonsingular matrix) */
static int dense_inverse(int64_t n, double *a, double *inv) {
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++) inv[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int64_t c = 0; c < n; c++) {
        int64_t p = c;
        for (int64_t r = c + 1; r < n; r++) if (fabs(a[r * n + c]) > fabs(a[p * n + c])) p = r;
        if (a[p * n + c] == 0.0) return -1;
        if (p != c)
            for (int64_t j = 0; j < n; j++) {
                double t = a[c * n + j]; a[c * n + j] = a[p * n + j]; a[p * n + j] = t;
                t = inv[c * n + j]; inv[c * n + j] = inv[p * n + j]; inv[p * n + j] = t;
            }
        double d = a[c * n + c];
        for (int64_t j = 0; j < n; j++) { a[c * n + j] /= d; inv[c * n + j] /= d; }
        for (int64_t r = 0; r < n; r++) {
            if (r == c) continue;
            double f = a[r * n + c];
            if (f == 0.0) continue;
            for (int64_t j = 0; j < n; j++) { a[r * n + j] -= f * a[c * n + j]; inv[r * n + j] -= f * inv[c * n + j]; }
        }
    }
    return 0;
}

orc_amg *orc_amg_build(const orc_csr *A0, double theta, int max_levels, int64_t max_coarse) {
    orc_amg *ml = calloc(1, sizeof(orc_amg));
    if (!ml) return NULL;
    csr_copy(A0, &ml->A[0]);
    int lev = 0;
    /* while length(levels)+1 < max_levels && size(A,1) > max_coarse */
    while (lev + 1 < max_levels && ml->A[lev].n > max_coarse && lev + 1 < ORC_MAXLEV) {
        orc_csr S, T;
        classical_strength(&ml->A[lev], theta, &S);
        csr_transpose(&S, &T);
        int *split = malloc(sizeof(int) * ml->A[lev].n);
        rs_split(&S, &T, split);
        direct_interp(&ml->A[lev], &S, split, &ml->P[lev]);
        orc_csr_free(&S); orc_csr_free(&T);
        if (ml->P[lev].m == 0 || ml->P[lev].m == ml->A[lev].n) { orc_csr_free(&ml->P[lev]); free(split); break; }
        ml->split[lev] = split;
        csr_transpose(&ml->P[lev], &ml->R[lev]);
        orc_csr RA;
        csr_matmul(&ml->R[lev], &ml->A[lev], &RA);
        csr_matmul(&RA, &ml->P[lev], &ml->A[lev + 1]);
        orc_csr_free(&RA);
        lev++;
    }
    ml->nlev = lev + 1;
    int64_t nc = ml->A[lev].n;
    ml->ncoarse = nc;
    double *dense = calloc(nc * nc, sizeof(double));
    ml->coarse_inv = malloc(sizeof(double) * nc * nc);
    for (int64_t i = 0; i < nc; i++)
        for (int64_t q = ml->A[lev].rowptr[i]; q < ml->A[lev].rowptr[i + 1]; q++)
            dense[i * nc + ml->A[lev].col[q]] += ml->A[lev].val[q];
    if (dense_inverse(nc, dense, ml->coarse_inv) != 0) memset(ml->coarse_inv, 0, sizeof(double) * nc * nc);
    free(dense);
    for (int l = 0; l < ml->nlev; l++) {
        int64_t n = ml->A[l].n, ncn = (l + 1 < ml->nlev) ? ml->A[l + 1].n : 1;
        ml->work[l][0] = calloc(n, sizeof(double));
        ml->work[l][1] = calloc(ncn, sizeof(double));
        ml->work[l][2] = calloc(ncn, sizeof(double));
    }
    return ml;
}

void orc_amg_free(orc_amg *ml) {
    if (!ml) return;
    for (int l = 0; l < ml->nlev; l++) {
        orc_csr_free(&ml->A[l]); orc_csr_free(&ml->P[l]); orc_csr_free(&ml->R[l]);
        free(ml->split[l]);
        for (int k = 0; k < 3; k++) free(ml->work[l][k]);
    }
    free(ml->coarse_inv);
    free(ml);
}

int orc_amg_nlev(const orc_amg *ml) { return ml->nlev; }
int64_t orc_amg_level_n(const orc_amg *ml, int l) { return ml->A[l].n; }
int64_t orc_amg_level_nnz(const orc_amg *ml, int l) { return nnz_of(&ml->A[l]); }
int64_t orc_amg_level_pnnz(const orc_amg *ml, int l) { return (l + 1 < ml->nlev) ? nnz_of(&ml->P[l]) : 0; }
void orc_amg_get_split(const orc_amg *ml, int l, int *is_c) {
    for (int64_t i = 0; i < ml->A[l].n; i++) is_c[i] = ml->split[l][i] == C_NODE;
}
/* copy level l's P (or A) out: which 0 = A, 1 = P */
void orc_amg_get_csr(const orc_amg *ml, int l, int which, int64_t *rowptr, int64_t *col, double *val) {
    const orc_csr *M = which ? &ml->P[l] : &ml->A[l];
    memcpy(rowptr, M->rowptr, sizeof(int64_t) * (M->n + 1));
    memcpy(col, M->col, sizeof(int64_t) * nnz_of(M));
    memcpy(val, M->val, sizeof(double) * nnz_of(M));
}

/* gs!(A, b, x, start, step, stop): row i of a symmetric A */
static void gs_sweep(const orc_csr *A, const double *b, double *x, int forward) {
    int64_t n = A->n;
    for (int64_t t = 0; t < n; t++) {
        int64_t i = forward ? t : n - 1 - t;
        double rsum = 0.0, d = 0.0;
        for (int64_t q = A->rowptr[i]; q < A->rowptr[i + 1]; q++) {
            if (A->col[q] == i) d = A->val[q];
            else rsum += A->val[q] * x[A->col[q]];
        }
        if (d != 0.0) x[i] = (b[i] - rsum) / d;
    }
}

static void vcycle(orc_amg *ml, int lvl, double *x, const double *b) {
    const orc_csr *A = &ml->A[lvl];
    gs_sweep(A, b, x, 1); gs_sweep(A, b, x, 0);
    double *res = ml->work[lvl][0], *cb = ml->work[lvl][1], *cx = ml->work[lvl][2];
    orc_spmv(A, x, res);
    for (int64_t i = 0; i < A->n; i++) res[i] = b[i] - res[i];
    orc_spmv(&ml->R[lvl], res, cb);
    int64_t nc = ml->A[lvl + 1].n;
    memset(cx, 0, sizeof(double) * nc);
    if (lvl + 1 == ml->nlev - 1) {
#pragma omp parallel for schedule(static) if (nc >= 256)
        for (int64_t i = 0; i < nc; i++) {
            double s = 0.0;
            for (int64_t j = 0; j < nc; j++) s += ml->coarse_inv[i * nc + j] * cb[j];
            cx[i] = s;
        }
    } else {
        vcycle(ml, lvl + 1, cx, cb);
    }
    orc_spmv(&ml->P[lvl], cx, res);
    for (int64_t i = 0; i < A->n; i++) x[i] += res[i];
    gs_sweep(A, b, x, 1); gs_sweep(A, b, x, 0);
}

/* ldiv!(x, aspreconditioner(ml), b): one V-cycle from x = 0 */
void orc_amg_apply(orc_amg *ml, const double *b, double *x) {
    memset(x, 0, sizeof(double) * ml->A[0].n);
    if (ml->nlev == 1) {
        int64_t nc = ml->ncoarse;
        for (int64_t i = 0; i < nc; i++) {
            double s = 0.0;
            for (int64_t j = 0; j < nc; j++) s += ml->coarse_inv[i * nc + j] * b[j];
            x[i] = s;
        }
        return;
    }
    vcycle(ml, 0, x, b);
}

/* IterativeSolvers.cg!(x, A, b; Pl, tol, maxiter) (src/main.jl:256):       */
/* r = b - A x, stop when ||r||_2 <= tol ||b||_2 or after maxiter updates.    */
static double dot(int64_t n, const double *a, const double *b) {
    double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s) if (n >= ORC_PAR_MIN)
    for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

int orc_cg(const orc_csr *A, double *x, const double *b, orc_amg *Pl, double tol, int64_t maxiter,
           int64_t *iters_out, double *relres_out) {
    int64_t n = A->n;
    double *r = malloc(sizeof(double) * n), *c = malloc(sizeof(double) * n), *u = calloc(n, sizeof(double));
    orc_spmv(A, x, c);
#pragma omp parallel for schedule(static) if (n >= ORC_PAR_MIN)
    for (int64_t i = 0; i < n; i++) r[i] = b[i] - c[i];
    double residual = sqrt(dot(n, r, r));
    double nb = sqrt(dot(n, b, b));
    double reltol = nb * tol;
    double rho = 1.0;
    int64_t it = 0;
    while (!(it >= maxiter || residual <= reltol)) {
        if (Pl) orc_amg_apply(Pl, r, c); else memcpy(c, r, sizeof(double) * n);
        double rho_prev = rho;
        rho = dot(n, c, r);
        double beta = rho / rho_prev;
#pragma omp parallel for schedule(static) if (n >= ORC_PAR_MIN)
        for (int64_t i = 0; i < n; i++) u[i] = c[i] + beta * u[i];
        orc_spmv(A, u, c);
        double alpha = rho / dot(n, u, c);
#pragma omp parallel for schedule(static) if (n >= ORC_PAR_MIN)
        for (int64_t i = 0; i < n; i++) { x[i] += alpha * u[i]; r[i] -= alpha * c[i]; }
        residual = sqrt(dot(n, r, r));
        it++;
    }
    if (iters_out) *iters_out = it;
    if (relres_out) *relres_out = (nb > 0) ? residual / nb : residual;
    free(r); free(c); free(u);
    return (residual <= reltol) ? 0 : 1;
}

/* PCG! (src/PCG.jl:8-67): Jacobi-preconditioned CG on K_II, rhs = -(K F)[FltNI] */
int orc_pcg_jacobi(const orc_csr *Kii, const double *diag, double *x, const double *Fnew, double tol,
                   int64_t maxit, int64_t *iters_out) {
    int64_t n = Kii->n;
    double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
    double *p = malloc(sizeof(double) * n), *a = malloc(sizeof(double) * n);
    orc_spmv(Kii, x, a);
    for (int64_t i = 0; i < n; i++) { r[i] = Fnew[i] - a[i]; z[i] = r[i] / diag[i]; p[i] = z[i]; }
    double nF = sqrt(dot(n, Fnew, Fnew));
    int status = 1;
    int64_t it;
    for (it = 1; it <= maxit; it++) {
        orc_spmv(Kii, p, a);
        double zr = dot(n, z, r);
        double alpha = zr / dot(n, p, a);
        for (int64_t i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * a[i]; z[i] = r[i] / diag[i]; }
        double beta = dot(n, z, r) / zr;
        for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        double rel = sqrt(dot(n, r, r)) / nF;
        if (rel < tol) { status = 0; break; }
        if (it == maxit || rel > 1e10) { status = 1; break; }
    }
    if (iters_out) *iters_out = it;
    free(r); free(z); free(p); free(a);
    return status;
}

/* ------------------------------------------------------------------------ */
/* Fault physics (src/otherFunctions_serial.jl, src/NRsearch_serial.jl,      */
/* src/dtevol.jl), FP64 throughout.                                           */
/* ------------------------------------------------------------------------ */

/* IDS! (src/otherFunctions_serial.jl:5-32), IDstate 1 and 2 */
double orc_ids(double xLf, double Vo, double psi, double dt, double Vf, double cnd, int IDstate) {
    if (IDstate == 1) return psi + dt * ((Vo / xLf) * exp(-psi) - fabs(Vf) / xLf);
    double VdtL = fabs(Vf) * dt / xLf;
    if (VdtL < cnd)
        return log(exp(psi - VdtL) + Vo * dt / xLf - 0.5 * Vo * fabs(Vf) * dt * dt / (xLf * xLf));
    return log(exp(psi - VdtL) + (Vo / fabs(Vf)) * (1 - exp(-VdtL)));
}

/* IDS2! (src/otherFunctions_serial.jl:35-59) */
double orc_ids2(double xLf, double Vo, double psi, double psi1, double dt, double Vf, double Vf1, int IDstate) {
    if (IDstate == 1)
        return psi + 0.5 * dt * ((Vo / xLf) * exp(-psi) - fabs(Vf) / xLf + (Vo / xLf) * exp(-psi1) - fabs(Vf1) / xLf);
    double VdtL = 0.5 * fabs(Vf1 + Vf) * dt / xLf;
    if (VdtL < 1e-6)
        return log(exp(psi - VdtL) + Vo * dt / xLf - 0.5 * Vo * 0.5 * fabs(Vf1 + Vf) * dt * dt / (xLf * xLf));
    return log(exp(psi - VdtL) + (Vo / (0.5 * fabs(Vf + Vf1))) * (1 - exp(-VdtL)));
}

/* NRsearch! (src/NRsearch_serial.jl:47-98) in FP64; returns the iteration count, */
/* negative on the |delta| > 1e10 / k == 1000 exit, which returns the current    */
/* iterate (Vf at the last point evaluated, tau after its step).                  */
/*                                                                                */
/* The reference keeps fa, help1, help2, delta and tau in BigFloat (:9,51-54), so */
/* its iterates never overflow.  At a fast-slipping node the first step from      */
/* tau = 0 lands far above the root.  There e^M, M = help + |fa|, dwarfs every    */
/* other term of the step: delta equals c = -sign(fa) Vo / (fact Vo/(a sigma))    */
/* to far below FP64 rounding, and |fa| falls by |c| fact/(sigma a) = 1 per step, */
/* for hundreds of steps.  While M >= max(help + 40, ln(A / (1e-20 Z Vo))), with   */
/* A = |Z Vfree + tauo - tau| + 1000 |c| bounding the numerator's other term over */
/* any run of steps, the neglected terms stay below 1e-20 of the kept ones        */
/* (e^{-2|fa|} <= e^{-80} and A e^{-M} <= 1e-20 Z Vo), so that run of steps is    */
/* taken at once: k += n, tau += n c.  Every other step runs as written, at M     */
/* below about 65, where FP64 exp is far from overflow.                           */
int orc_nrsearch(double fo, double Vo, double cca, double ccb, double Seff, double tau, double tauo,
                 double psi, double FltZ, double FltVfree, double *Vf_out, double *tau_out) {
    const double Vw = 1e10;
    double fact = 1. + (Vo / Vw) * exp(-psi);
    double eps = 0.001 * cca * Seff;
    double help = -(fo + ccb * psi) / cca;
    double vfc = fact * (Vo / (cca * Seff));   /* Vf' = vfc (help1 + help2) */
    int k = 0;
    double delta = INFINITY, Vf = 0.0;
    while (fabs(delta) > eps) {
        double fa = fact * tau / (Seff * cca);
        double M = fmax(help + fa, help - fa);
        if (isfinite(M) && M >= help + 40.0) {
            double c = copysign(Vo / vfc, -fa);
            double Ta = log((fabs(FltZ * FltVfree + tauo - tau) + 1000.0 * fabs(c)) / (1e-20 * FltZ * Vo));
            if (M >= Ta) {   /* false for a NaN Ta */
                /* the steps at M, M - dM, ... down to max(help + 40, Ta); one step
                   if it ends the loop */
                double Tlin = fmax(help + 40.0, Ta);
                double nd = 1.0;
                if (fabs(c) > eps && fabs(c) <= 1e10) nd = floor((M - Tlin) / fabs(fact * c / (Seff * cca))) + 1.0;
                if (nd > 1000 - k) nd = 1000 - k;
                int n = (int)nd;
                double tau_last = tau + (n - 1) * c;
                tau = tau + n * c;
                k += n;
                delta = c;
                if (fabs(delta) > 1e10 || k == 1000) {
                    double fl = fact * tau_last / (Seff * cca);
                    *Vf_out = Vo * (exp(help + fl) - exp(help - fl));
                    *tau_out = tau;
                    return -k;
                }
                continue;
            }
        }
        double help1 = exp(help + fa), help2 = exp(help - fa);
        Vf = Vo * (help1 - help2);
        double Vfprime = vfc * (help1 + help2);
        delta = (FltZ * FltVfree - FltZ * Vf + tauo - tau) / (1 + FltZ * Vfprime);
        tau = tau + delta;
        k = k + 1;
        if (fabs(delta) > 1e10 || k == 1000) { *Vf_out = Vf; *tau_out = tau; return -k; }
    }
    double fa = fact * tau / (Seff * cca);
    Vf = Vo * (exp(help + fa) - exp(help - fa));
    *Vf_out = Vf; *tau_out = tau;
    return k;
}

/* FBC! (src/NRsearch_serial.jl:6-45); j0 = NFBC-1 (0-based).  Returns 0, or */
/* 1 + j for the first node whose guard (Vf > 1e10, NaN) fired.              */
int orc_fbc(int IDstate, const double *fo, const double *Vo, const double *xLf, const double *cca,
            const double *ccb, const double *Seff, const double *tauo, const double *FltZ, int64_t j0,
            int64_t n, double *psi1, double *Vf1, double *tau1, double *psi2, double *Vf2, double *tau2,
            const double *psi, const double *Vf, const double *FltVfree, double dt, int64_t *nr_noconv) {
    int64_t noconv = 0;
    for (int64_t j = j0; j < n; j++) {
        psi1[j] = orc_ids(xLf[j], Vo[j], psi[j], dt, Vf[j], 1e-5, IDstate);
        int k = orc_nrsearch(fo[j], Vo[j], cca[j], ccb[j], Seff[j], 0.0, tauo[j], psi1[j], FltZ[j], FltVfree[j],
                             &Vf1[j], &tau1[j]);
        if (k < 0) noconv++;
        if (Vf[j] > 1e10 || isnan(Vf[j]) || isnan(tau1[j])) { if (nr_noconv) *nr_noconv = noconv; return (int)(1 + j); }
        psi2[j] = orc_ids2(xLf[j], Vo[j], psi[j], psi1[j], dt, Vf[j], Vf1[j], IDstate);
        k = orc_nrsearch(fo[j], Vo[j], cca[j], ccb[j], Seff[j], tau1[j], tauo[j], psi2[j], FltZ[j], FltVfree[j],
                         &Vf2[j], &tau2[j]);
        if (k < 0) noconv++;
    }
    if (nr_noconv) *nr_noconv = noconv;
    return 0;
}

/* slrFunc! (src/otherFunctions_serial.jl:62-83) */
void orc_slr(int IDstate, const double *fo, const double *Vo, const double *xLf, const double *cca,
             const double *ccb, const double *Seff, const double *tauo, int64_t j0, int64_t n,
             const double *psi, double *psi1, const double *Vf, double *Vf1, const double *tau1, double dt) {
    for (int64_t j = j0; j < n; j++) {
        psi1[j] = orc_ids(xLf[j], Vo[j], psi[j], dt, Vf[j], 1e-6, IDstate);
        double tauAB = tau1[j] + tauo[j];
        double fa = tauAB / (Seff[j] * cca[j]);
        double help = -(fo[j] + ccb[j] * psi1[j]) / cca[j];
        Vf1[j] = Vo[j] * (exp(help + fa) - exp(help - fa));
    }
}

/* dtevol! (src/dtevol.jl:5-44) */
double orc_dtevol(double dt, double dtmin, const double *XiLf, int64_t n, int64_t j0, const double *Vf, int isolver) {
    const double dtmax = 50 * 24 * 60 * 60;
    const double dtincf = 1.2;
    if (isolver == 1) {
        double dtnx = dtmax;
        for (int64_t i = j0; i < n; i++)
            if (fabs(Vf[i]) * dtmax > XiLf[i]) {
                double dtcell = XiLf[i] / fabs(Vf[i]);
                if (dtcell < dtnx) dtnx = dtcell;
            }
        if (dtmin > dtnx) dtnx = dtmin;
        if (dtnx > dtincf * dt) dtnx = dtincf * dt;
        return dtnx;
    }
    return dtmin;
}

/* ------------------------------------------------------------------------ */
/* SplitMix64 (for fixtures): n outputs from seed                             */
/* ------------------------------------------------------------------------ */
void orc_splitmix64(uint64_t seed, int64_t n, uint64_t *out) {
    for (int64_t k = 1; k <= n; k++) {
        uint64_t z = seed + (uint64_t)k * 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        out[k - 1] = z ^ (z >> 31);
    }
}

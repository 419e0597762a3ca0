/* CPU FP64 restatement of the reference forward solve: incompressible SIMPLE
 * with stored Rhie-Chow fluxes and the SST-2003 k-omega closure.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP engine). Every function
 * cites the reference code it restates; the floating-point operations are
 * written in the reference's order so that, built with the same compiler
 * flags (-O2, no contraction), results can be compared bit for bit with the
 * reference compiled from its own sources. The mesh is 2D (all face normals
 * have zero z), so the reference's z terms, which are exact zeros, are
 * omitted.
 */
#include "rans_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* std::max / std::min semantics: (a < b) ? b : a and (b < a) ? b : a. */
static double smax(double a, double b) { return (a < b) ? b : a; }
static double smin(double a, double b) { return (b < a) ? b : a; }

/* ------------------------------------------------------------------------ */
/* LDU linear system (core/include/LinearSolver.hpp:20-116)                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n, nif;
    const int *own, *nbr;
    double *diag, *upper, *lower, *source;
} lsys_t;

static void ls_zero(lsys_t *A) {
    memset(A->diag, 0, sizeof(double) * A->n);
    memset(A->source, 0, sizeof(double) * A->n);
    memset(A->upper, 0, sizeof(double) * A->nif);
    memset(A->lower, 0, sizeof(double) * A->nif);
}

/* LinearSystem::matvec (LinearSolver.hpp:99-108) */
static void ls_matvec(const lsys_t *A, const double *x, double *y) {
    for (int i = 0; i < A->n; ++i) y[i] = A->diag[i] * x[i];
    for (int f = 0; f < A->nif; ++f) {
        y[A->own[f]] += A->upper[f] * x[A->nbr[f]];
        y[A->nbr[f]] += A->lower[f] * x[A->own[f]];
    }
}

/* LinearSystem::residual (LinearSolver.hpp:111-115) */
static void ls_residual(const lsys_t *A, const double *x, double *r) {
    ls_matvec(A, x, r);
    for (int i = 0; i < A->n; ++i) r[i] = A->source[i] - r[i];
}

/* LinearSystem::validateSolve (LinearSolver.hpp:63-89): 0 when admitted. */
static int ls_admit(const lsys_t *A, const double *x) {
    for (int f = 0; f < A->nif; ++f)
        if (A->own[f] < 0 || A->own[f] >= A->n || A->nbr[f] < 0 || A->nbr[f] >= A->n) return 1;
    for (int i = 0; i < A->n; ++i)
        if (!isfinite(A->diag[i]) || !isfinite(A->source[i])) return 2;
    for (int f = 0; f < A->nif; ++f)
        if (!isfinite(A->upper[f]) || !isfinite(A->lower[f])) return 2;
    for (int i = 0; i < A->n; ++i)
        if (!isfinite(x[i])) return 3;
    return 0;
}

static double vdot(int n, const double *a, const double *b) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
static double vnorm(int n, const double *a) { return sqrt(vdot(n, a, a)); }

/* linalg::jacobiPrecond (LinearSolver.hpp:173-177) */
static void jacobi(const lsys_t *A, const double *r, double *z) {
    for (int i = 0; i < A->n; ++i) z[i] = (fabs(A->diag[i]) > 1e-30) ? r[i] / A->diag[i] : r[i];
}

typedef struct {
    int iterations;
    double final_res, initial_res;
    int converged;
} sres_t;

typedef struct {
    double *r, *z, *p, *Ap, *rh, *v, *s, *t, *ph, *sh;
} kwork_t;

/* PCGSolver::solve (LinearSolver.hpp:191-244) */
static sres_t pcg_solve(const lsys_t *A, double *x, int maxIter, double tol, kwork_t *w) {
    const int n = A->n;
    double *r = w->r, *z = w->z, *p = w->p, *Ap = w->Ap;
    sres_t res = {0, 0.0, 0.0, 0};
    ls_residual(A, x, r);
    double r0 = vnorm(n, r);
    res.initial_res = r0;
    if (r0 < 1e-30) {
        res.converged = 1;
        return res;
    }
    jacobi(A, r, z);
    memcpy(p, z, sizeof(double) * n);
    double rz = vdot(n, r, z);
    for (int it = 0; it < maxIter; ++it) {
        ls_matvec(A, p, Ap);
        double pAp = vdot(n, p, Ap);
        if (fabs(pAp) < 1e-30) break;
        double alpha = rz / pAp;
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
        }
        double rn = vnorm(n, r);
        res.iterations = it + 1;
        res.final_res = rn / r0;
        if (res.final_res < tol) {
            res.converged = 1;
            return res;
        }
        jacobi(A, r, z);
        double rzNew = vdot(n, r, z);
        double beta = rzNew / (rz + 1e-30);
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        rz = rzNew;
    }
    return res;
}

/* BiCGSTABSolver::solve (LinearSolver.hpp:259-343) */
static sres_t bicgstab_solve(const lsys_t *A, double *x, int maxIter, double tol, kwork_t *w) {
    const int n = A->n;
    double *r = w->r, *rh = w->rh, *p = w->p, *v = w->v, *s = w->s, *t = w->t, *ph = w->ph,
           *sh = w->sh;
    memset(p, 0, sizeof(double) * n);
    memset(v, 0, sizeof(double) * n);
    sres_t res = {0, 0.0, 0.0, 0};
    ls_residual(A, x, r);
    double r0 = vnorm(n, r);
    res.initial_res = r0;
    if (r0 < 1e-30) {
        res.converged = 1;
        return res;
    }
    memcpy(rh, r, sizeof(double) * n);
    double rho = 1, alpha = 1, omega = 1;
    for (int it = 0; it < maxIter; ++it) {
        double rhoN = vdot(n, rh, r);
        if (fabs(rhoN) < 1e-30) break;
        double beta = (rhoN / (rho + 1e-30)) * (alpha / (omega + 1e-30));
        for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        jacobi(A, p, ph);
        ls_matvec(A, ph, v);
        double rv = vdot(n, rh, v);
        if (fabs(rv) < 1e-30) break;
        alpha = rhoN / rv;
        for (int i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
        double sn = vnorm(n, s);
        if (sn / r0 < tol) {
            for (int i = 0; i < n; ++i) x[i] += alpha * ph[i];
            res.iterations = it + 1;
            res.final_res = sn / r0;
            res.initial_res = r0;
            res.converged = 1;
            return res;
        }
        jacobi(A, s, sh);
        ls_matvec(A, sh, t);
        double tt = vdot(n, t, t);
        omega = (tt > 1e-30) ? vdot(n, t, s) / tt : 0.0;
        for (int i = 0; i < n; ++i) x[i] += alpha * ph[i] + omega * sh[i];
        for (int i = 0; i < n; ++i) r[i] = s[i] - omega * t[i];
        double rn = vnorm(n, r);
        res.iterations = it + 1;
        res.final_res = rn / r0;
        if (res.final_res < tol) {
            res.converged = 1;
            return res;
        }
        if (fabs(omega) < 1e-30) break;
        rho = rhoN;
    }
    return res;
}

/* ------------------------------------------------------------------------ */
/* Aggregation AMG-PCG (LinearSolver.hpp:411-985)                           */
/* ------------------------------------------------------------------------ */
typedef struct {
    int nRows;
    int *rowPtr, *colIdx;
    double *values, *diag;
} csr_t;

typedef struct {
    csr_t A;
    int *aggMap, nCoarse;
    int *galerkinIdx, *diagPos;
    double *x, *b, *r;
} amg_level_t;

#define AMG_MAX_LEVELS 21
typedef struct {
    amg_level_t lev[AMG_MAX_LEVELS];
    int nlev;
    int *patternOwn, *patternNbr, patternN, patternNif;
    int *upperPos, *lowerPos;
    int mappable, needFullSetup, itersAfterSetup;
    double *lu;
    int *piv;
    int coarseDirect;
} amg_t;

static void csr_free(csr_t *c) {
    free(c->rowPtr);
    free(c->colIdx);
    free(c->values);
    free(c->diag);
    memset(c, 0, sizeof *c);
}

static void amg_clear_levels(amg_t *g) {
    for (int l = 0; l < g->nlev; ++l) {
        amg_level_t *L = &g->lev[l];
        csr_free(&L->A);
        free(L->aggMap);
        free(L->galerkinIdx);
        free(L->diagPos);
        free(L->x);
        free(L->b);
        free(L->r);
        memset(L, 0, sizeof *L);
    }
    g->nlev = 0;
}

static void amg_free(amg_t *g) {
    amg_clear_levels(g);
    free(g->patternOwn);
    free(g->patternNbr);
    free(g->upperPos);
    free(g->lowerPos);
    free(g->lu);
    free(g->piv);
    memset(g, 0, sizeof *g);
}

/* CSRMatrix::find (LinearSolver.hpp:464-468): lower_bound over a sorted row */
static int csr_find(const csr_t *A, int r, int c) {
    int lo = A->rowPtr[r], hi = A->rowPtr[r + 1];
    const int last = hi;
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if (A->colIdx[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    return (lo != last && A->colIdx[lo] == c) ? lo : -1;
}

/* lduToCSR (LinearSolver.hpp:472-532) */
static void ldu_to_csr(const lsys_t *A, csr_t *csr) {
    const int n = A->n;
    int *count = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) count[i] = 1;
    for (int f = 0; f < A->nif; ++f) {
        count[A->own[f]]++;
        count[A->nbr[f]]++;
    }
    csr->nRows = n;
    csr->rowPtr = (int *)calloc(n + 1, sizeof(int));
    for (int i = 0; i < n; ++i) csr->rowPtr[i + 1] = csr->rowPtr[i] + count[i];
    int nnz = csr->rowPtr[n];
    csr->colIdx = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    csr->values = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
    csr->diag = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    int *pos = count;
    for (int i = 0; i < n; ++i) pos[i] = csr->rowPtr[i];
    for (int i = 0; i < n; ++i) {
        csr->colIdx[pos[i]] = i;
        csr->values[pos[i]] = A->diag[i];
        csr->diag[i] = A->diag[i];
        pos[i]++;
    }
    for (int f = 0; f < A->nif; ++f) {
        int o = A->own[f], nb = A->nbr[f];
        csr->colIdx[pos[o]] = nb;
        csr->values[pos[o]] = A->upper[f];
        pos[o]++;
        csr->colIdx[pos[nb]] = o;
        csr->values[pos[nb]] = A->lower[f];
        pos[nb]++;
    }
    for (int i = 0; i < n; ++i) {
        int start = csr->rowPtr[i], end = csr->rowPtr[i + 1];
        for (int a = start + 1; a < end; ++a) {
            int key_c = csr->colIdx[a];
            double key_v = csr->values[a];
            int b = a - 1;
            while (b >= start && csr->colIdx[b] > key_c) {
                csr->colIdx[b + 1] = csr->colIdx[b];
                csr->values[b + 1] = csr->values[b];
                b--;
            }
            csr->colIdx[b + 1] = key_c;
            csr->values[b + 1] = key_v;
        }
    }
    free(count);
}

/* buildAggregates (LinearSolver.hpp:535-575) */
static int build_aggregates(const csr_t *A, int *map) {
    const int n = A->nRows;
    int nCoarse = 0;
    for (int i = 0; i < n; ++i) map[i] = -1;
    for (int i = 0; i < n; ++i) {
        if (map[i] >= 0) continue;
        int agg = nCoarse++;
        map[i] = agg;
        double aii = fabs(A->diag[i]);
        for (int j = A->rowPtr[i]; j < A->rowPtr[i + 1]; ++j) {
            int c = A->colIdx[j];
            if (c == i || map[c] >= 0) continue;
            double ajj = fabs(A->diag[c]);
            double strength = fabs(A->values[j]) / sqrt(aii * ajj + 1e-30);
            if (strength > 0.25) map[c] = agg;
        }
    }
    for (int i = 0; i < n; ++i) {
        if (map[i] >= 0) continue;
        double bestStr = -1.0;
        int bestAgg = 0;
        for (int j = A->rowPtr[i]; j < A->rowPtr[i + 1]; ++j) {
            int c = A->colIdx[j];
            if (c == i && map[c] < 0) continue;
            if (map[c] >= 0 && fabs(A->values[j]) > bestStr) {
                bestStr = fabs(A->values[j]);
                bestAgg = map[c];
            }
        }
        map[i] = bestAgg;
    }
    return nCoarse;
}

typedef struct {
    int col;
    double val;
} centry_t;

static int centry_cmp(const void *a, const void *b) {
    const centry_t *x = (const centry_t *)a, *y = (const centry_t *)b;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return (x->val < y->val) ? -1 : (x->val > y->val);
}

/* galerkinProduct (LinearSolver.hpp:578-618): A_c[I,J] = sum of A[i,j] over
 * the aggregates, summed in fine row/column order, rows sorted by column. */
static void galerkin_product(const csr_t *A, const int *map, int nCoarse, csr_t *Ac) {
    int *cnt = (int *)calloc(nCoarse, sizeof(int));
    int *cap = (int *)calloc(nCoarse, sizeof(int));
    centry_t **rows = (centry_t **)calloc(nCoarse, sizeof(centry_t *));
    for (int i = 0; i < A->nRows; ++i) {
        int I = map[i];
        for (int j = A->rowPtr[i]; j < A->rowPtr[i + 1]; ++j) {
            int J = map[A->colIdx[j]];
            int k = 0;
            for (; k < cnt[I]; ++k)
                if (rows[I][k].col == J) break;
            if (k == cnt[I]) {
                if (cnt[I] == cap[I]) {
                    cap[I] = cap[I] ? 2 * cap[I] : 8;
                    rows[I] = (centry_t *)realloc(rows[I], sizeof(centry_t) * cap[I]);
                }
                rows[I][k].col = J;
                rows[I][k].val = 0.0;
                cnt[I]++;
            }
            rows[I][k].val += A->values[j];
        }
    }
    int nnz = 0;
    for (int I = 0; I < nCoarse; ++I) nnz += cnt[I];
    Ac->nRows = nCoarse;
    Ac->rowPtr = (int *)calloc(nCoarse + 1, sizeof(int));
    Ac->diag = (double *)calloc(nCoarse > 0 ? nCoarse : 1, sizeof(double));
    Ac->colIdx = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    Ac->values = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
    int pos = 0;
    for (int I = 0; I < nCoarse; ++I) {
        Ac->rowPtr[I] = pos;
        if (cnt[I] > 1) qsort(rows[I], cnt[I], sizeof(centry_t), centry_cmp);
        for (int k = 0; k < cnt[I]; ++k) {
            Ac->colIdx[pos] = rows[I][k].col;
            Ac->values[pos] = rows[I][k].val;
            if (rows[I][k].col == I) Ac->diag[I] = rows[I][k].val;
            pos++;
        }
        free(rows[I]);
    }
    Ac->rowPtr[nCoarse] = pos;
    free(rows);
    free(cnt);
    free(cap);
}

/* AMGSolver::sumCoarse (LinearSolver.hpp:699-706) */
static void sum_coarse(const amg_level_t *fine, amg_level_t *coarse) {
    const int nnzc = coarse->A.rowPtr[coarse->A.nRows];
    for (int k = 0; k < nnzc; ++k) coarse->A.values[k] = 0.0;
    const int nnz = fine->A.rowPtr[fine->A.nRows];
    for (int k = 0; k < nnz; ++k) coarse->A.values[fine->galerkinIdx[k]] += fine->A.values[k];
    for (int I = 0; I < coarse->A.nRows; ++I)
        coarse->A.diag[I] = coarse->diagPos[I] >= 0 ? coarse->A.values[coarse->diagPos[I]] : 0.0;
}

static void compute_diag_pos(amg_level_t *L) {
    L->diagPos = (int *)malloc(sizeof(int) * (L->A.nRows > 0 ? L->A.nRows : 1));
    for (int i = 0; i < L->A.nRows; ++i) L->diagPos[i] = csr_find(&L->A, i, i);
}

/* AMGSolver::factorCoarsest (LinearSolver.hpp:708-746): dense partial-pivot LU */
static void factor_coarsest(amg_t *g) {
    const amg_level_t *L = &g->lev[g->nlev - 1];
    const int n = L->A.nRows;
    g->coarseDirect = n <= 400;
    if (!g->coarseDirect) return;
    free(g->lu);
    free(g->piv);
    g->lu = (double *)calloc((size_t)n * n > 0 ? (size_t)n * n : 1, sizeof(double));
    g->piv = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    double *lu = g->lu;
    for (int i = 0; i < n; ++i)
        for (int j = L->A.rowPtr[i]; j < L->A.rowPtr[i + 1]; ++j)
            lu[(size_t)i * n + L->A.colIdx[j]] += L->A.values[j];
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (fabs(lu[(size_t)i * n + k]) > fabs(lu[(size_t)p * n + k])) p = i;
        g->piv[k] = p;
        if (fabs(lu[(size_t)p * n + k]) < 1e-300) {
            g->coarseDirect = 0;
            return;
        }
        if (p != k)
            for (int j = 0; j < n; ++j) {
                double tmp = lu[(size_t)k * n + j];
                lu[(size_t)k * n + j] = lu[(size_t)p * n + j];
                lu[(size_t)p * n + j] = tmp;
            }
        const double inv = 1.0 / lu[(size_t)k * n + k];
        for (int i = k + 1; i < n; ++i) {
            double *lik = &lu[(size_t)i * n + k];
            *lik *= inv;
            if (*lik != 0.0)
                for (int j = k + 1; j < n; ++j) lu[(size_t)i * n + j] -= *lik * lu[(size_t)k * n + j];
        }
    }
}

/* CSRMatrix::gaussSeidelSweep / ...Backward (LinearSolver.hpp:437-461) */
static void gs_sweep(const csr_t *A, double *x, const double *b, int backward) {
    for (int ii = 0; ii < A->nRows; ++ii) {
        const int i = backward ? A->nRows - 1 - ii : ii;
        double s = b[i];
        for (int j = A->rowPtr[i]; j < A->rowPtr[i + 1]; ++j)
            if (A->colIdx[j] != i) s -= A->values[j] * x[A->colIdx[j]];
        if (fabs(A->diag[i]) > 1e-30) x[i] = s / A->diag[i];
    }
}

static void csr_residual(const csr_t *A, const double *x, const double *b, double *r) {
    for (int i = 0; i < A->nRows; ++i) {
        double s = 0.0;
        for (int j = A->rowPtr[i]; j < A->rowPtr[i + 1]; ++j) s += A->values[j] * x[A->colIdx[j]];
        r[i] = b[i] - s;
    }
}

/* AMGSolver::solveCoarsest (LinearSolver.hpp:748-768) */
static void solve_coarsest(amg_t *g, amg_level_t *L) {
    const int n = L->A.nRows;
    if (!g->coarseDirect) {
        for (int s = 0; s < 25; ++s) {
            gs_sweep(&L->A, L->x, L->b, 0);
            gs_sweep(&L->A, L->x, L->b, 1);
        }
        return;
    }
    double *y = L->x;
    memcpy(y, L->b, sizeof(double) * n);
    for (int k = 0; k < n; ++k) {
        double tmp = y[k];
        y[k] = y[g->piv[k]];
        y[g->piv[k]] = tmp;
    }
    const double *lu = g->lu;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) y[i] -= lu[(size_t)i * n + j] * y[j];
    for (int i = n - 1; i >= 0; --i) {
        for (int j = i + 1; j < n; ++j) y[i] -= lu[(size_t)i * n + j] * y[j];
        y[i] /= lu[(size_t)i * n + i];
    }
}

/* AMGSolver::fullSetup (LinearSolver.hpp:770-842) */
static void amg_full_setup(amg_t *g, const lsys_t *A) {
    amg_clear_levels(g);
    amg_level_t *L0 = &g->lev[0];
    ldu_to_csr(A, &L0->A);
    compute_diag_pos(L0);
    free(g->upperPos);
    free(g->lowerPos);
    g->upperPos = (int *)malloc(sizeof(int) * (A->nif > 0 ? A->nif : 1));
    g->lowerPos = (int *)malloc(sizeof(int) * (A->nif > 0 ? A->nif : 1));
    for (int f = 0; f < A->nif; ++f) g->upperPos[f] = g->lowerPos[f] = -1;
    g->mappable = 1;
    for (int i = 0; i < L0->A.nRows && g->mappable; ++i)
        for (int j = L0->A.rowPtr[i] + 1; j < L0->A.rowPtr[i + 1]; ++j)
            if (L0->A.colIdx[j] == L0->A.colIdx[j - 1]) {
                g->mappable = 0;
                break;
            }
    if (g->mappable)
        for (int f = 0; f < A->nif; ++f) {
            g->upperPos[f] = csr_find(&L0->A, A->own[f], A->nbr[f]);
            g->lowerPos[f] = csr_find(&L0->A, A->nbr[f], A->own[f]);
        }
    free(g->patternOwn);
    free(g->patternNbr);
    g->patternOwn = (int *)malloc(sizeof(int) * (A->nif > 0 ? A->nif : 1));
    g->patternNbr = (int *)malloc(sizeof(int) * (A->nif > 0 ? A->nif : 1));
    memcpy(g->patternOwn, A->own, sizeof(int) * A->nif);
    memcpy(g->patternNbr, A->nbr, sizeof(int) * A->nif);
    g->patternN = A->n;
    g->patternNif = A->nif;
    g->nlev = 1;
    for (int l = 0; l < AMG_MAX_LEVELS - 1; ++l) {
        amg_level_t *fine = &g->lev[l];
        if (fine->A.nRows <= 50) break;
        int *map = (int *)malloc(sizeof(int) * fine->A.nRows);
        int nc = build_aggregates(&fine->A, map);
        if (nc == 0 || nc >= fine->A.nRows) {
            free(map);
            break;
        }
        fine->aggMap = map;
        fine->nCoarse = nc;
        amg_level_t *coarse = &g->lev[l + 1];
        galerkin_product(&fine->A, fine->aggMap, nc, &coarse->A);
        compute_diag_pos(coarse);
        const int nnz = fine->A.rowPtr[fine->A.nRows];
        fine->galerkinIdx = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
        for (int i = 0; i < fine->A.nRows; ++i)
            for (int j = fine->A.rowPtr[i]; j < fine->A.rowPtr[i + 1]; ++j)
                fine->galerkinIdx[j] =
                    csr_find(&coarse->A, fine->aggMap[i], fine->aggMap[fine->A.colIdx[j]]);
        sum_coarse(fine, coarse);
        g->nlev = l + 2;
    }
    for (int l = 0; l < g->nlev; ++l) {
        int n = g->lev[l].A.nRows;
        g->lev[l].x = (double *)calloc(n > 0 ? n : 1, sizeof(double));
        g->lev[l].b = (double *)calloc(n > 0 ? n : 1, sizeof(double));
        g->lev[l].r = (double *)calloc(n > 0 ? n : 1, sizeof(double));
    }
    factor_coarsest(g);
    g->needFullSetup = 0;
    g->itersAfterSetup = -1;
}

/* AMGSolver::refreshValues (LinearSolver.hpp:844-851) */
static void amg_refresh(amg_t *g, const lsys_t *A) {
    amg_level_t *L0 = &g->lev[0];
    for (int i = 0; i < A->n; ++i) {
        L0->A.values[L0->diagPos[i]] = A->diag[i];
        L0->A.diag[i] = A->diag[i];
    }
    for (int f = 0; f < A->nif; ++f) {
        L0->A.values[g->upperPos[f]] = A->upper[f];
        L0->A.values[g->lowerPos[f]] = A->lower[f];
    }
    for (int l = 0; l + 1 < g->nlev; ++l) sum_coarse(&g->lev[l], &g->lev[l + 1]);
    factor_coarsest(g);
}

/* AMGSolver::vcycle (LinearSolver.hpp:853-886) */
static void amg_vcycle(amg_t *g, int level) {
    amg_level_t *L = &g->lev[level];
    const int n = L->A.nRows;
    if (level == g->nlev - 1) {
        solve_coarsest(g, L);
        return;
    }
    for (int s = 0; s < 2; ++s) gs_sweep(&L->A, L->x, L->b, 0);
    csr_residual(&L->A, L->x, L->b, L->r);
    amg_level_t *C = &g->lev[level + 1];
    memset(C->b, 0, sizeof(double) * C->A.nRows);
    memset(C->x, 0, sizeof(double) * C->A.nRows);
    for (int i = 0; i < n; ++i) C->b[L->aggMap[i]] += L->r[i];
    amg_vcycle(g, level + 1);
    for (int i = 0; i < n; ++i) L->x[i] += C->x[L->aggMap[i]];
    for (int s = 0; s < 2; ++s) gs_sweep(&L->A, L->x, L->b, 1);
}

static void amg_precondition(amg_t *g, const double *r, double *z) {
    amg_level_t *L0 = &g->lev[0];
    memcpy(L0->b, r, sizeof(double) * L0->A.nRows);
    memset(L0->x, 0, sizeof(double) * L0->A.nRows);
    amg_vcycle(g, 0);
    memcpy(z, L0->x, sizeof(double) * L0->A.nRows);
}

static int amg_same_pattern(const amg_t *g, const lsys_t *A) {
    return g->nlev > 0 && g->lev[0].A.nRows == A->n && g->patternNif == A->nif &&
           memcmp(g->patternOwn, A->own, sizeof(int) * A->nif) == 0 &&
           memcmp(g->patternNbr, A->nbr, sizeof(int) * A->nif) == 0;
}

/* AMGSolver::solve (LinearSolver.hpp:908-982), including the frozen-hierarchy
 * degradation rule of the IterTimer destructor (:924-940). */
static sres_t amg_solve(amg_t *g, const lsys_t *A, double *x, int maxIter, double tol,
                        kwork_t *w) {
    const int n = A->n;
    const int wasFull = g->needFullSetup || !g->mappable || !amg_same_pattern(g, A);
    if (wasFull) amg_full_setup(g, A);
    else amg_refresh(g, A);
    double *r = w->r, *z = w->z, *p = w->p, *Ap = w->Ap;
    sres_t res = {0, 0.0, 0.0, 0};
    ls_residual(A, x, r);
    double r0 = vnorm(n, r);
    res.initial_res = r0;
    if (r0 < 1e-30) {
        res.converged = 1;
        goto done;
    }
    amg_precondition(g, r, z);
    memcpy(p, z, sizeof(double) * n);
    double rz = vdot(n, r, z);
    for (int it = 0; it < maxIter; ++it) {
        ls_matvec(A, p, Ap);
        double pAp = vdot(n, p, Ap);
        if (fabs(pAp) < 1e-30) break;
        double alpha = rz / pAp;
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
        }
        double rn = vnorm(n, r);
        res.iterations = it + 1;
        res.final_res = rn / r0;
        if (res.final_res < tol) {
            res.converged = 1;
            goto done;
        }
        amg_precondition(g, r, z);
        double rzNew = vdot(n, r, z);
        double beta = rzNew / (rz + 1e-30);
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        rz = rzNew;
    }
done:
    if (wasFull || g->itersAfterSetup <= 0) g->itersAfterSetup = res.iterations;
    else if (!res.converged || res.iterations > 1.5 * g->itersAfterSetup + 2) g->needFullSetup = 1;
    return res;
}

/* Dispatch one linear solve by solver kind. Returns nonzero if the system is
 * not admitted (the reference throws std::invalid_argument there). */
static int linear_solve(int kind, amg_t *amg, const lsys_t *A, double *x, int maxIter, double tol,
                        kwork_t *w, sres_t *out) {
    if (ls_admit(A, x)) return 1;
    if (kind == OR_SOLVER_PCG) *out = pcg_solve(A, x, maxIter, tol, w);
    else if (kind == OR_SOLVER_BICGSTAB) *out = bicgstab_solve(A, x, maxIter, tol, w);
    else *out = amg_solve(amg, A, x, maxIter, tol, w);
    return 0;
}

static void kwork_alloc(kwork_t *w, int n) {
    double **slots[] = {&w->r, &w->z, &w->p, &w->Ap, &w->rh, &w->v, &w->s, &w->t, &w->ph, &w->sh};
    for (size_t i = 0; i < sizeof slots / sizeof slots[0]; ++i)
        *slots[i] = (double *)calloc(n > 0 ? n : 1, sizeof(double));
}
static void kwork_free(kwork_t *w) {
    double *slots[] = {w->r, w->z, w->p, w->Ap, w->rh, w->v, w->s, w->t, w->ph, w->sh};
    for (size_t i = 0; i < sizeof slots / sizeof slots[0]; ++i) free(slots[i]);
}

int or_linear_solve(int solver, int n, int nif, const int *own, const int *nbr, const double *diag,
                    const double *upper, const double *lower, const double *source, double *x,
                    int max_iter, double tol, int *iterations, double *initial_res,
                    double *final_res) {
    lsys_t A = {n, nif, own, nbr, (double *)diag, (double *)upper, (double *)lower,
                (double *)source};
    kwork_t w;
    kwork_alloc(&w, n);
    amg_t amg;
    memset(&amg, 0, sizeof amg);
    amg.needFullSetup = 1;
    amg.itersAfterSetup = -1;
    sres_t res = {0, 0, 0, 0};
    int rc = linear_solve(solver, &amg, &A, x, max_iter, tol, &w, &res);
    amg_free(&amg);
    kwork_free(&w);
    if (iterations) *iterations = res.iterations;
    if (initial_res) *initial_res = res.initial_res;
    if (final_res) *final_res = res.final_res;
    return rc ? -1 : res.converged;
}

/* ------------------------------------------------------------------------ */
/* SST-2003 pointwise closure (core/src/SSTModel.cpp:15-108)                */
/* Coefficient vector order: 0 sk1 1 sw1 2 b1 3 a1c 4 sk2 5 sw2 6 b2 7 a2c  */
/*                           8 betaStar 9 a1 10 kappa                       */
/* ------------------------------------------------------------------------ */
enum { C_SK1 = 0, C_SW1, C_B1, C_AL1, C_SK2, C_SW2, C_B2, C_AL2, C_BS, C_A1, C_KAPPA };

static double blend(double p1, double p2, double F1) { return F1 * p1 + (1.0 - F1) * p2; }

double or_sst_f1(const double c[11], double k, double omega, double y, double nu, double CDkw_pos) {
    double sqrtK = sqrt(smax(k, 0.0));
    double omSafe = smax(omega, 1e-20);
    double ySafe = smax(y, 1e-20);
    double y2s = ySafe * ySafe;
    double term1 = sqrtK / (c[C_BS] * omSafe * ySafe);
    double term2 = 500.0 * nu / (y2s * omSafe);
    double CDpos = smax(CDkw_pos, 1e-10);
    double term3 = 4.0 * c[C_SW2] * smax(k, 0.0) / (CDpos * y2s);
    double arg1 = smin(smax(term1, term2), term3);
    double a4 = arg1 * arg1 * arg1 * arg1;
    return tanh(a4);
}

double or_sst_f2(const double c[11], double k, double omega, double y, double nu) {
    double sqrtK = sqrt(smax(k, 0.0));
    double omSafe = smax(omega, 1e-20);
    double ySafe = smax(y, 1e-20);
    double y2 = ySafe * ySafe;
    double term1 = 2.0 * sqrtK / (c[C_BS] * omSafe * ySafe);
    double term2 = 500.0 * nu / (y2 * omSafe);
    double arg2 = smax(term1, term2);
    return tanh(arg2 * arg2);
}

double or_sst_nut(const double c[11], int variant, double k, double omega, double S, double F2) {
    double kSafe = smax(k, 0.0);
    if (variant == 1) return kSafe / smax(omega, 1e-20);
    double denom = smax(c[C_A1] * omega, S * F2);
    denom = smax(denom, 1e-20);
    return c[C_A1] * kSafe / denom;
}

double or_sst_pk(const double c[11], double nuT, double S, double k, double omega) {
    double Pk_raw = nuT * S * S;
    double Pk_lim = 10.0 * c[C_BS] * smax(k, 0.0) * smax(omega, 1e-20);
    return smin(Pk_raw, Pk_lim);
}

double or_sst_pomega(const double c[11], double nuT, double S, double k, double omega) {
    double S2 = S * S;
    double lim = 10.0 * c[C_BS] * smax(k, 0.0) * smax(omega, 1e-20) / smax(nuT, 1e-30);
    return smin(S2, lim);
}

/* ------------------------------------------------------------------------ */
/* Solver state                                                             */
/* ------------------------------------------------------------------------ */
typedef struct {
    const or_mesh_t *m;
    const or_bc_t *bc;
    const or_settings_t *s;
    const or_obs_t *obs;
    double nu;
    double c[11];
    int nc, nif, nf, nb;
    /* fields: interior [nc] and boundary [nb] */
    double *Ux, *Uy, *p, *k, *w, *nuT, *F1, *F2, *Pk, *CDkw;
    double *Uxb, *Uyb, *pb, *kb, *wb;
    int turbEstablished;
    /* planar omega balance (SIMPLESolver.cpp:92-153) */
    int planar;
    double *omSing, *omGrad, *omLap;
    /* workspaces */
    double *gx, *gy, *gx2, *gy2, *gx3, *gy3, *tmpc, *tmpb, *tsrc;
    double *phi, *phiBarOld, *phiPred, *pCoeff;
    double *aP, *aPu, *aPstore, *solveValues, *rbuf, *UxPrev, *UyPrev, *kOld, *wOld;
    double *pPrime, *pPrimeB, *strain, *wallOmega, *balance;
    double massFluxScale;
    char *isWallFace;
    lsys_t sys;
    kwork_t kw;
    amg_t amg;
} ctx_t;

/* greenGaussGrad (core/include/Field.hpp:178-212), 2D components. */
static void gg_grad(const ctx_t *X, const double *phi, const double *phib, double *gx, double *gy) {
    const or_mesh_t *m = X->m;
    for (int i = 0; i < X->nc; ++i) gx[i] = gy[i] = 0.0;
    for (int fi = 0; fi < X->nif; ++fi) {
        int o = m->owner[fi], n = m->neighbor[fi];
        double w = m->fweight[fi];
        double phiF = phi[o] * w + phi[n] * (1.0 - w);
        double Sx = m->fnx[fi] * m->farea[fi], Sy = m->fny[fi] * m->farea[fi];
        gx[o] = gx[o] + Sx * phiF;
        gy[o] = gy[o] + Sy * phiF;
        gx[n] = gx[n] - Sx * phiF;
        gy[n] = gy[n] - Sy * phiF;
    }
    for (int fi = X->nif; fi < X->nf; ++fi) {
        int o = m->owner[fi];
        double Sx = m->fnx[fi] * m->farea[fi], Sy = m->fny[fi] * m->farea[fi];
        double pb = phib[fi - X->nif];
        gx[o] = gx[o] + Sx * pb;
        gy[o] = gy[o] + Sy * pb;
    }
    for (int i = 0; i < X->nc; ++i) {
        double V = m->vol[i];
        if (V > 1e-30) {
            gx[i] = gx[i] / V;
            gy[i] = gy[i] / V;
        }
    }
}

/* applyVelocityBC (BoundaryCondition.hpp:279-310) */
static void apply_velocity_bc(ctx_t *X) {
    const or_mesh_t *m = X->m;
    for (int pi = 0; pi < m->n_patches; ++pi) {
        int t = X->bc->u_type[pi];
        for (int q = m->patch_face_ptr[pi]; q < m->patch_face_ptr[pi + 1]; ++q) {
            int fi = m->patch_face_idx[q], b = fi - X->nif, o = m->owner[fi];
            switch (t) {
            case OR_BC_WALL_NOSLIP:
                X->Uxb[b] = 0.0;
                X->Uyb[b] = 0.0;
                break;
            case OR_BC_WALL_MOVING:
            case OR_BC_INLET:
            case OR_BC_DIRICHLET:
                X->Uxb[b] = X->bc->u_val[2 * pi];
                X->Uyb[b] = X->bc->u_val[2 * pi + 1];
                break;
            case OR_BC_SYMMETRY: {
                double nx = m->fnx[fi], ny = m->fny[fi];
                double d = X->Ux[o] * nx + X->Uy[o] * ny;
                X->Uxb[b] = X->Ux[o] - nx * d;
                X->Uyb[b] = X->Uy[o] - ny * d;
                break;
            }
            default:
                X->Uxb[b] = X->Ux[o];
                X->Uyb[b] = X->Uy[o];
            }
        }
    }
}

/* applyPressureBC (BoundaryCondition.hpp:314-331) */
static void apply_pressure_bc(ctx_t *X) {
    const or_mesh_t *m = X->m;
    for (int pi = 0; pi < m->n_patches; ++pi) {
        int t = X->bc->p_type[pi];
        for (int q = m->patch_face_ptr[pi]; q < m->patch_face_ptr[pi + 1]; ++q) {
            int fi = m->patch_face_idx[q];
            X->pb[fi - X->nif] = (t == OR_BC_OUTLET || t == OR_BC_DIRICHLET) ? X->bc->p_val[pi]
                                                                             : X->p[m->owner[fi]];
        }
    }
}

/* applyKBC (BoundaryCondition.hpp:335-355) */
static void apply_k_bc(ctx_t *X) {
    const or_mesh_t *m = X->m;
    for (int pi = 0; pi < m->n_patches; ++pi) {
        int t = X->bc->k_type[pi];
        for (int q = m->patch_face_ptr[pi]; q < m->patch_face_ptr[pi + 1]; ++q) {
            int fi = m->patch_face_idx[q];
            double v;
            if (t == OR_BC_WALL_KOMEGA) v = 0.0;
            else if (t == OR_BC_DIRICHLET) v = X->bc->k_val[pi];
            else v = X->k[m->owner[fi]];
            X->kb[fi - X->nif] = v;
        }
    }
}

/* applyOmegaBC (BoundaryCondition.hpp:361-384) */
static void apply_omega_bc(ctx_t *X, double beta1) {
    const or_mesh_t *m = X->m;
    for (int pi = 0; pi < m->n_patches; ++pi) {
        int t = X->bc->w_type[pi];
        for (int q = m->patch_face_ptr[pi]; q < m->patch_face_ptr[pi + 1]; ++q) {
            int fi = m->patch_face_idx[q];
            double v;
            if (t == OR_BC_WALL_KOMEGA) {
                double y1 = smax(m->fdelta[fi], 1e-20);
                v = 60.0 * X->nu / (beta1 * y1 * y1);
            } else if (t == OR_BC_DIRICHLET) {
                v = X->bc->w_val[pi];
            } else {
                v = X->w[m->owner[fi]];
            }
            X->wb[fi - X->nif] = v;
        }
    }
}

/* SIMPLESolver::preparePlanarOmegaBalance (SIMPLESolver.cpp:92-153) */
static void prepare_planar(ctx_t *X) {
    const or_mesh_t *m = X->m;
    X->planar = 0;
    if (X->nf == X->nif) return;
    const int f0 = X->nif;
    const double inx = m->fnx[f0] * -1.0, iny = m->fny[f0] * -1.0;
    const double origin = m->fcx[f0] * inx + m->fcy[f0] * iny;
    double height = 0;
    int periodic = 0;
    for (int fi = X->nif; fi < X->nf; ++fi) {
        int pi = -1;
        for (int q = 0; q < m->n_patches && pi < 0; ++q)
            for (int r = m->patch_face_ptr[q]; r < m->patch_face_ptr[q + 1]; ++r)
                if (m->patch_face_idx[r] == fi) {
                    pi = q;
                    break;
                }
        if (pi < 0 || m->patch_type[pi] != OR_PATCH_WALL || X->bc->w_type[pi] != OR_BC_WALL_KOMEGA ||
            fabs(fabs(m->fnx[fi] * inx + m->fny[fi] * iny) - 1.0) > 1e-12)
            return;
        height = smax(height, (m->fcx[fi] * inx + m->fcy[fi] * iny) - origin);
    }
    if (!(height > 0)) return;
    const double tolerance = 1e-12 * height;
    for (int fi = 0; fi < X->nf; ++fi) {
        double ex = m->fdx[fi] - m->fnx[fi] * m->fdelta[fi];
        double ey = m->fdy[fi] - m->fny[fi] * m->fdelta[fi];
        if (sqrt(ex * ex + ey * ey) > 1e-9 * m->fdelta[fi] ||
            fabs(m->fshx[fi] * inx + m->fshy[fi] * iny) > tolerance)
            return;
        periodic = periodic || (m->fshx[fi] * m->fshx[fi] + m->fshy[fi] * m->fshy[fi]) > 0;
        if (fi >= X->nif) {
            const double y = (m->fcx[fi] * inx + m->fcy[fi] * iny) - origin;
            if (smin(fabs(y), fabs(y - height)) > tolerance) return;
            if ((y < height / 2 ? -1.0 : 1.0) * (m->fnx[fi] * inx + m->fny[fi] * iny) < 1 - 1e-12)
                return;
        }
    }
    if (!periodic) return;
    for (int c = 0; c < X->nc; ++c) {
        const double y = (m->ccx[c] * inx + m->ccy[c] * iny) - origin;
        if (!(y > 0 && y < height)) return;
        X->omSing[c] = 1 / (y * y) + 1 / ((height - y) * (height - y));
        X->omLap[c] = 6 * (pow(y, -4) + pow(height - y, -4));
    }
    for (int fi = 0; fi < X->nif; ++fi) {
        const double y = (m->fcx[fi] * inx + m->fcy[fi] * iny) - origin;
        if (!(y > 0 && y < height)) return;
        X->omGrad[fi] =
            2 * (pow(height - y, -3) - pow(y, -3)) * (m->fnx[fi] * inx + m->fny[fi] * iny);
    }
    X->planar = 1;
}

/* SIMPLESolver::resolvedWallOmega (SIMPLESolver.cpp:155-161) */
static double resolved_wall_omega(const ctx_t *X, double y) {
    const double coefficient = X->planar ? 6.0 : 60.0;
    return coefficient * X->nu / (X->c[C_B1] * y * y);
}

/* SSTModel::computeFields (SSTModel.cpp:157-241) */
static void sst_compute_fields(ctx_t *X) {
    const or_mesh_t *m = X->m;
    double *dux = X->gx, *duy = X->gy, *dvx = X->gx2, *dvy = X->gy2;
    gg_grad(X, X->Ux, X->Uxb, dux, duy);
    gg_grad(X, X->Uy, X->Uyb, dvx, dvy);
    for (int ci = 0; ci < X->nc; ++ci) {
        double S11 = dux[ci], S22 = dvy[ci];
        double S12 = 0.5 * (duy[ci] + dvx[ci]);
        X->strain[ci] = sqrt(2.0 * (S11 * S11 + S22 * S22 + 2.0 * (S12 * S12)));
    }
    double *dkx = X->gx, *dky = X->gy, *dwx = X->gx2, *dwy = X->gy2;
    gg_grad(X, X->k, X->kb, dkx, dky);
    gg_grad(X, X->w, X->wb, dwx, dwy);
    for (int ci = 0; ci < X->nc; ++ci) {
        double kc = smax(X->k[ci], 0.0);
        double wc = smax(X->w[ci], 1e-20);
        double y = smax(m->wall_dist[ci], 1e-20);
        double Sc = X->strain[ci];
        double omSafe = smax(wc, 1e-20);
        double CDkw = 2.0 * X->c[C_SW2] / omSafe * (dkx[ci] * dwx[ci] + dky[ci] * dwy[ci]);
        X->CDkw[ci] = CDkw;
        X->F1[ci] = (X->s->sst_variant == 2) ? 1.0 : or_sst_f1(X->c, kc, wc, y, X->nu, CDkw);
        X->F2[ci] = or_sst_f2(X->c, kc, wc, y, X->nu);
        X->nuT[ci] = or_sst_nut(X->c, X->s->sst_variant, kc, wc, Sc, X->F2[ci]);
        X->Pk[ci] = or_sst_pk(X->c, X->nuT[ci], Sc, kc, wc);
    }
}

/* SIMPLESolver::addConvectionCorrection (SIMPLESolver.cpp:495-532) */
static void convection_correction(ctx_t *X, const double *phi, const double *phib, int scheme) {
    if (scheme == OR_CONV_UPWIND) return;
    const or_mesh_t *m = X->m;
    const int limited = scheme == OR_CONV_LIMITED;
    double *gx = X->gx3, *gy = X->gy3;
    gg_grad(X, phi, phib, gx, gy);
    double *src = X->sys.source;
    for (int fi = 0; fi < X->nif; ++fi) {
        const double F = X->phi[fi];
        if (F == 0.0) continue;
        const int o = m->owner[fi], n = m->neighbor[fi];
        const int ownerUp = F > 0.0;
        const int up = ownerUp ? o : n, dn = ownerUp ? n : o;
        const double xUx = m->ccx[up], xUy = m->ccy[up];
        double phiF;
        if (!limited) {
            double xfx = ownerUp ? m->fcx[fi] : m->fcx[fi] - m->fshx[fi];
            double xfy = ownerUp ? m->fcy[fi] : m->fcy[fi] - m->fshy[fi];
            phiF = phi[up] + (gx[up] * (xfx - xUx) + gy[up] * (xfy - xUy));
        } else {
            double xDx = ownerUp ? m->ccx[n] + m->fshx[fi] : m->ccx[o] - m->fshx[fi];
            double xDy = ownerUp ? m->ccy[n] + m->fshy[fi] : m->ccy[o] - m->fshy[fi];
            const double dphi = phi[dn] - phi[up];
            double psi = 0.0;
            if (fabs(dphi) > 1e-300) {
                const double r =
                    2.0 * (gx[up] * (xDx - xUx) + gy[up] * (xDy - xUy)) / dphi - 1.0;
                psi = r > 0.0 ? (r * r + r) / (r * r + 1.0) : 0.0;
            }
            phiF = phi[up] + 0.5 * psi * dphi;
        }
        const double corr = F * (phiF - phi[up]);
        src[o] -= corr;
        src[n] += corr;
    }
}

/* SIMPLESolver::assembleMomentum (SIMPLESolver.cpp:303-482), components 0/1 */
/* The front and back faces of a one-layer extruded mesh (an OpenFOAM 2D
 * case), folded out of the face arrays: zero-gradient boundary faces with
 * no flux, each adding D = nuEff S / delta to the diagonal and D phi to the
 * source of its owner, phi the owner's value (the boundary-face loops of
 * SIMPLESolver.cpp:360-382, 983-1001, 1092-1113; such faces come after the
 * in-plane boundary faces in a case's face order). nuEff is nu + nuT for
 * momentum (F1 NULL) and nu + sigma(F1) nuT for k and omega. */
static void folded_faces(ctx_t *X, lsys_t *A, const double *phi, const double *F1, double s1,
                         double s2) {
    const double *fold = X->m->cell_folded;
    if (!fold) return;
    for (int ci = 0; ci < X->nc; ++ci) {
        const double nuEff = F1 ? X->nu + blend(s1, s2, F1[ci]) * X->nuT[ci] : X->nu + X->nuT[ci];
        for (int e = 0; e < 2; ++e) {
            const double Sf = fold[4 * ci + 2 * e], delta = fold[4 * ci + 2 * e + 1];
            if (!(Sf > 0.0)) continue;
            const double Db = nuEff * Sf / delta;
            A->diag[ci] += Db;
            A->source[ci] += Db * phi[ci];
        }
    }
}

static void assemble_momentum(ctx_t *X, int comp) {
    const or_mesh_t *m = X->m;
    lsys_t *A = &X->sys;
    ls_zero(A);
    const double nu = X->nu;
    for (int fi = 0; fi < X->nif; ++fi) {
        int o = m->owner[fi], n = m->neighbor[fi];
        double Sf = m->farea[fi];
        double delta = smax(m->fdelta[fi], 1e-20);
        double nuEff_o = nu + X->nuT[o];
        double nuEff_n = nu + X->nuT[n];
        double nuEff_f = m->fweight[fi] * nuEff_o + (1.0 - m->fweight[fi]) * nuEff_n;
        double Df = nuEff_f * Sf / delta;
        double mFlux = X->phi[fi];
        double cPos = smax(mFlux, 0.0), cNeg = smax(-mFlux, 0.0);
        A->diag[o] += Df + cPos;
        A->diag[n] += Df + cNeg;
        A->upper[fi] = -(Df + cNeg);
        A->lower[fi] = -(Df + cPos);
    }
    for (int fi = X->nif; fi < X->nf; ++fi) {
        int o = m->owner[fi];
        double Sf = m->farea[fi];
        double delta = smax(m->fdelta[fi], 1e-20);
        double nuEff = X->isWallFace[fi] ? nu : nu + X->nuT[o];
        double Db = nuEff * Sf / delta;
        double Ub = comp == 0 ? X->Uxb[fi - X->nif] : X->Uyb[fi - X->nif];
        double mFlux = X->phi[fi];
        if (mFlux >= 0) {
            A->diag[o] += mFlux + Db;
            A->source[o] += Db * Ub;
        } else {
            A->diag[o] += Db;
            A->source[o] += (Db - mFlux) * Ub;
        }
    }
    folded_faces(X, A, comp == 0 ? X->Ux : X->Uy, NULL, 0.0, 0.0);
    /* pressure gradient source */
    gg_grad(X, X->p, X->pb, X->gx, X->gy);
    for (int ci = 0; ci < X->nc; ++ci) {
        double dpdx = comp == 0 ? X->gx[ci] : X->gy[ci];
        A->source[ci] -= dpdx * m->vol[ci];
    }
    const double fb = comp == 0 ? X->s->body_force_x : X->s->body_force_y;
    if (fb != 0.0)
        for (int ci = 0; ci < X->nc; ++ci) A->source[ci] += fb * m->vol[ci];
    /* transpose stress with zero wall coefficient (StressOperators.hpp:31-81) */
    {
        double *dux = X->gx, *duy = X->gy, *dvx = X->gx2, *dvy = X->gy2;
        gg_grad(X, X->Ux, X->Uxb, dux, duy);
        gg_grad(X, X->Uy, X->Uyb, dvx, dvy);
        /* column `comp` of grad U: (dUx/dx_c, dUy/dx_c) */
        const double *ga = comp == 0 ? dux : duy, *gb = comp == 0 ? dvx : dvy;
        double *tsrc = X->tsrc;
        for (int ci = 0; ci < X->nc; ++ci) tsrc[ci] = 0.0;
        for (int fi = 0; fi < X->nif; ++fi) {
            int o = m->owner[fi], n = m->neighbor[fi];
            double w = m->fweight[fi];
            double cf = w * X->nuT[o] + (1.0 - w) * X->nuT[n];
            double gxf = ga[o] * w + ga[n] * (1.0 - w);
            double gyf = gb[o] * w + gb[n] * (1.0 - w);
            double flux = cf * (gxf * m->fnx[fi] + gyf * m->fny[fi]) * m->farea[fi];
            tsrc[o] += flux;
            tsrc[n] -= flux;
        }
        for (int pi = 0; pi < m->n_patches; ++pi) {
            const int isWall = m->patch_type[pi] == OR_PATCH_WALL;
            for (int q = m->patch_face_ptr[pi]; q < m->patch_face_ptr[pi + 1]; ++q) {
                int fi = m->patch_face_idx[q], o = m->owner[fi];
                double cf = isWall ? 0.0 : X->nuT[o];
                double flux = cf * (ga[o] * m->fnx[fi] + gb[o] * m->fny[fi]) * m->farea[fi];
                tsrc[o] += flux;
            }
        }
        for (int ci = 0; ci < X->nc; ++ci) A->source[ci] += tsrc[ci];
    }
    /* deferred higher-order convection correction */
    if (X->s->momentum_convection != OR_CONV_UPWIND) {
        const double *uc = comp == 0 ? X->Ux : X->Uy;
        const double *ub = comp == 0 ? X->Uxb : X->Uyb;
        convection_correction(X, uc, ub, X->s->momentum_convection);
    }
    /* under-relaxation */
    const double alphaU = X->s->alpha_u;
    for (int ci = 0; ci < X->nc; ++ci) {
        double phi_old = comp == 0 ? X->Ux[ci] : X->Uy[ci];
        X->aPu[ci] = A->diag[ci];
        A->source[ci] += (1.0 - alphaU) / alphaU * A->diag[ci] * phi_old;
        A->diag[ci] /= alphaU;
        X->aP[ci] = A->diag[ci];
    }
}

/* An experiment on the precision pattern, not part of the reference: with
 * rounding on, every assembled system (momentum, p', k and omega rows after
 * relaxation and wall constraints) is rounded to FP32 before it is used,
 * which is what storing an FP32 assembly costs at the least. Residuals,
 * norms, Krylov solves and observations stay FP64. Off by default; the flag
 * is process-wide, for single-threaded experiments. */
static int g_round_assembly = 0;
void or_set_assembly_rounding(int on) { g_round_assembly = on != 0; }
static void round_assembly(lsys_t *A, double *aP) {
    if (!g_round_assembly) return;
    for (int i = 0; i < A->n; ++i) {
        A->diag[i] = (double)(float)A->diag[i];
        A->source[i] = (double)(float)A->source[i];
        if (aP) aP[i] = A->diag[i];
    }
    for (int f = 0; f < A->nif; ++f) {
        A->upper[f] = (double)(float)A->upper[f];
        A->lower[f] = (double)(float)A->lower[f];
    }
}

/* interpolatedFluxes (SIMPLESolver.cpp:842-852) */
static void interpolated_fluxes(const ctx_t *X, double *out) {
    const or_mesh_t *m = X->m;
    for (int fi = 0; fi < X->nif; ++fi) {
        int o = m->owner[fi], n = m->neighbor[fi];
        double w = m->fweight[fi];
        double ufx = X->Ux[o] * w + X->Ux[n] * (1.0 - w);
        double ufy = X->Uy[o] * w + X->Uy[n] * (1.0 - w);
        out[fi] = (ufx * m->fnx[fi] + ufy * m->fny[fi]) * m->farea[fi];
    }
    for (int fi = X->nif; fi < X->nf; ++fi) {
        int b = fi - X->nif;
        out[fi] = (X->Uxb[b] * m->fnx[fi] + X->Uyb[b] * m->fny[fi]) * m->farea[fi];
    }
}

/* addRhieChow (SIMPLESolver.cpp:870-884) */
static void add_rhie_chow(ctx_t *X, const double *aP, double *phi) {
    const or_mesh_t *m = X->m;
    gg_grad(X, X->p, X->pb, X->gx, X->gy);
    for (int fi = 0; fi < X->nif; ++fi) {
        int o = m->owner[fi], n = m->neighbor[fi];
        double delta = smax(m->fdelta[fi], 1e-20);
        double dP_o = (fabs(aP[o]) > 1e-30) ? m->vol[o] / aP[o] : 0.0;
        double dP_n = (fabs(aP[n]) > 1e-30) ? m->vol[n] / aP[n] : 0.0;
        double w = m->fweight[fi];
        double dP_f = w * dP_o + (1.0 - w) * dP_n;
        double dd = smax(m->fdelta[fi], 1e-30);
        double ex = m->fdx[fi] / dd, ey = m->fdy[fi] / dd;
        double gbx = X->gx[o] * w + X->gx[n] * (1.0 - w);
        double gby = X->gy[o] * w + X->gy[n] * (1.0 - w);
        phi[fi] += -dP_f * m->farea[fi] * ((X->p[n] - X->p[o]) / delta - (gbx * ex + gby * ey));
    }
}

/* assemblePressureCorrection with stored predicted fluxes (SIMPLESolver.cpp:717-830) */
static void assemble_pressure(ctx_t *X, const double *aP, const double *predicted) {
    const or_mesh_t *m = X->m;
    lsys_t *A = &X->sys;
    ls_zero(A);
    for (int i = 0; i < X->nc; ++i) X->pPrime[i] = 0.0;
    for (int b = 0; b < X->nb; ++b) X->pPrimeB[b] = 0.0;
    X->massFluxScale = 0.0;
    for (int fi = 0; fi < X->nf; ++fi) X->pCoeff[fi] = 0.0;
    int hasOutlet = 0;
    for (int pi = 0; pi < m->n_patches; ++pi)
        if (m->patch_type[pi] == OR_PATCH_OUTLET && m->patch_face_ptr[pi + 1] > m->patch_face_ptr[pi])
            hasOutlet = 1;
    for (int fi = 0; fi < X->nif; ++fi) {
        int o = m->owner[fi], n = m->neighbor[fi];
        double Sf = m->farea[fi];
        double delta = smax(m->fdelta[fi], 1e-20);
        double dP_o = (fabs(aP[o]) > 1e-30) ? m->vol[o] / aP[o] : 0.0;
        double dP_n = (fabs(aP[n]) > 1e-30) ? m->vol[n] / aP[n] : 0.0;
        double dP_f = m->fweight[fi] * dP_o + (1.0 - m->fweight[fi]) * dP_n;
        double coeff = dP_f * Sf / delta;
        X->pCoeff[fi] = coeff;
        A->diag[o] += coeff;
        A->diag[n] += coeff;
        A->upper[fi] = -coeff;
        A->lower[fi] = -coeff;
        double massFlux = predicted[fi];
        A->source[o] -= massFlux;
        A->source[n] += massFlux;
        X->massFluxScale += fabs(massFlux);
    }
    for (int pi = 0; pi < m->n_patches; ++pi) {
        const int isOutlet = m->patch_type[pi] == OR_PATCH_OUTLET;
        for (int q = m->patch_face_ptr[pi]; q < m->patch_face_ptr[pi + 1]; ++q) {
            int fi = m->patch_face_idx[q], o = m->owner[fi];
            double Sf = m->farea[fi];
            double delta = smax(m->fdelta[fi], 1e-20);
            double massFlux = predicted[fi];
            A->source[o] -= massFlux;
            X->massFluxScale += fabs(massFlux);
            if (isOutlet) {
                double dP_o = (fabs(aP[o]) > 1e-30) ? m->vol[o] / aP[o] : 0.0;
                double coeff = dP_o * Sf / delta;
                X->pCoeff[fi] = coeff;
                A->diag[o] += coeff;
            }
        }
    }
    if (!hasOutlet && X->nc > 0) A->diag[0] += fabs(A->diag[0]) > 1e-30 ? A->diag[0] : 1.0;
}

/* assembleKEquation (SIMPLESolver.cpp:947-1023) */
static void assemble_k(ctx_t *X) {
    const or_mesh_t *m = X->m;
    lsys_t *A = &X->sys;
    ls_zero(A);
    const double nu = X->nu;
    for (int fi = 0; fi < X->nif; ++fi) {
        int o = m->owner[fi], n = m->neighbor[fi];
        double Sf = m->farea[fi];
        double delta = smax(m->fdelta[fi], 1e-20);
        double w = m->fweight[fi];
        double F1_f = w * X->F1[o] + (1.0 - w) * X->F1[n];
        double sk = blend(X->c[C_SK1], X->c[C_SK2], F1_f);
        double nuEff_o = nu + sk * X->nuT[o];
        double nuEff_n = nu + sk * X->nuT[n];
        double nuEff_f = w * nuEff_o + (1.0 - w) * nuEff_n;
        double Df = nuEff_f * Sf / delta;
        double mFlux = X->phi[fi];
        double cPos = smax(mFlux, 0.0), cNeg = smax(-mFlux, 0.0);
        A->diag[o] += Df + cPos;
        A->diag[n] += Df + cNeg;
        A->upper[fi] = -(Df + cNeg);
        A->lower[fi] = -(Df + cPos);
    }
    for (int fi = X->nif; fi < X->nf; ++fi) {
        int o = m->owner[fi];
        double Sf = m->farea[fi];
        double delta = smax(m->fdelta[fi], 1e-20);
        double sk = blend(X->c[C_SK1], X->c[C_SK2], X->F1[o]);
        double nuEff = nu + sk * X->nuT[o];
        double Db = nuEff * Sf / delta;
        double kb = X->kb[fi - X->nif];
        double mFlux = X->phi[fi];
        if (mFlux >= 0) {
            A->diag[o] += mFlux + Db;
            A->source[o] += Db * kb;
        } else {
            A->diag[o] += Db;
            A->source[o] += (Db - mFlux) * kb;
        }
    }
    folded_faces(X, A, X->k, X->F1, X->c[C_SK1], X->c[C_SK2]);
    for (int ci = 0; ci < X->nc; ++ci) {
        double vol = m->vol[ci];
        A->source[ci] += X->Pk[ci] * vol;
        double destruction = X->c[C_BS] * smax(X->w[ci], 1e-20);
        A->diag[ci] += destruction * vol;
    }
    convection_correction(X, X->k, X->kb, X->s->turbulence_convection);
    const double aK = X->s->alpha_k;
    for (int ci = 0; ci < X->nc; ++ci) {
        A->source[ci] += (1.0 - aK) / aK * A->diag[ci] * X->k[ci];
        A->diag[ci] /= aK;
    }
}

/* assembleOmegaEquation (SIMPLESolver.cpp:1030-1169) */
static void assemble_omega(ctx_t *X) {
    const or_mesh_t *m = X->m;
    lsys_t *A = &X->sys;
    ls_zero(A);
    const double nu = X->nu;
    if (X->planar)
        for (int ci = 0; ci < X->nc; ++ci) X->balance[ci] = 0;
    for (int fi = 0; fi < X->nif; ++fi) {
        int o = m->owner[fi], n = m->neighbor[fi];
        double Sf = m->farea[fi];
        double delta = smax(m->fdelta[fi], 1e-20);
        double w = m->fweight[fi];
        double F1_f = w * X->F1[o] + (1.0 - w) * X->F1[n];
        double sw = blend(X->c[C_SW1], X->c[C_SW2], F1_f);
        double nuEff_o = nu + sw * X->nuT[o];
        double nuEff_n = nu + sw * X->nuT[n];
        double nuEff_f = w * nuEff_o + (1.0 - w) * nuEff_n;
        double Df = nuEff_f * Sf / delta;
        double mFlux = X->phi[fi];
        double cPos = smax(mFlux, 0.0), cNeg = smax(-mFlux, 0.0);
        A->diag[o] += Df + cPos;
        A->diag[n] += Df + cNeg;
        A->upper[fi] = -(Df + cNeg);
        A->lower[fi] = -(Df + cPos);
        if (X->planar) {
            const double amplitude = 6 * nu / X->c[C_B1];
            const double difference = (X->omSing[n] - X->omSing[o]) / delta;
            const double correction =
                amplitude * Sf * ((nuEff_f - nu) * X->omGrad[fi] - nuEff_f * difference);
            X->balance[o] += correction;
            X->balance[n] -= correction;
        }
    }
    for (int fi = X->nif; fi < X->nf; ++fi) {
        int o = m->owner[fi];
        double Sf = m->farea[fi];
        double delta = smax(m->fdelta[fi], 1e-20);
        double sw = blend(X->c[C_SW1], X->c[C_SW2], X->F1[o]);
        double nuEff = nu + sw * X->nuT[o];
        double Db = nuEff * Sf / delta;
        double wb = X->wb[fi - X->nif];
        double mFlux = X->phi[fi];
        if (mFlux >= 0) {
            A->diag[o] += mFlux + Db;
            A->source[o] += Db * wb;
        } else {
            A->diag[o] += Db;
            A->source[o] += (Db - mFlux) * wb;
        }
    }
    folded_faces(X, A, X->w, X->F1, X->c[C_SW1], X->c[C_SW2]);
    for (int ci = 0; ci < X->nc; ++ci) {
        double vol = m->vol[ci];
        double F1 = X->F1[ci];
        double omC = smax(X->w[ci], 1e-20);
        double alphaB = blend(X->c[C_AL1], X->c[C_AL2], F1);
        double S = X->strain[ci];
        double production = or_sst_pomega(X->c, X->nuT[ci], S, X->k[ci], X->w[ci]);
        A->source[ci] += alphaB * production * vol;
        double betaB = blend(X->c[C_B1], X->c[C_B2], F1);
        A->diag[ci] += betaB * omC * vol;
        A->source[ci] += (1.0 - F1) * X->CDkw[ci] * vol;
        if (X->planar) {
            const double correction =
                X->balance[ci] + nu * (6 * nu / X->c[C_B1]) * X->omLap[ci] * vol;
            if (correction < 0) A->diag[ci] -= correction / omC;
            else A->source[ci] += correction;
        }
    }
    convection_correction(X, X->w, X->wb, X->s->turbulence_convection);
    const double aW = X->s->alpha_omega;
    for (int ci = 0; ci < X->nc; ++ci) {
        A->source[ci] += (1.0 - aW) / aW * A->diag[ci] * X->w[ci];
        A->diag[ci] /= aW;
    }
}

/* omegaWallValues (SIMPLESolver.cpp:665-684), resolved walls only */
static void omega_wall_values(ctx_t *X) {
    const or_mesh_t *m = X->m;
    for (int ci = 0; ci < X->nc; ++ci) X->wallOmega[ci] = NAN;
    for (int pi = 0; pi < m->n_patches; ++pi) {
        if (m->patch_type[pi] != OR_PATCH_WALL) continue;
        for (int q = m->patch_face_ptr[pi]; q < m->patch_face_ptr[pi + 1]; ++q) {
            int fi = m->patch_face_idx[q];
            const double y = smax(m->fdelta[fi], 1e-20);
            X->wallOmega[m->owner[fi]] = resolved_wall_omega(X, y);
        }
    }
}

/* constrainOmega (SIMPLESolver.cpp:686-704) */
static void constrain_omega(ctx_t *X) {
    const or_mesh_t *m = X->m;
    lsys_t *A = &X->sys;
    const double *v = X->wallOmega;
    for (int fi = 0; fi < X->nif; ++fi) {
        const int o = m->owner[fi], n = m->neighbor[fi];
        if (isfinite(v[o])) {
            A->source[n] -= A->lower[fi] * v[o];
            A->upper[fi] = A->lower[fi] = 0;
        } else if (isfinite(v[n])) {
            A->source[o] -= A->upper[fi] * v[n];
            A->upper[fi] = A->lower[fi] = 0;
        }
    }
    for (int ci = 0; ci < X->nc; ++ci)
        if (isfinite(v[ci])) A->source[ci] = A->diag[ci] * v[ci];
}

/* PreparedObservation::evaluate (ObservationOperator.hpp:678-772) for
 * velocity probes, pressure taps, skin friction, drag, separation points and
 * reattachment lengths. */
static void observe(const ctx_t *X, double *out) {
    const or_obs_t *ob = X->obs;
    const or_mesh_t *m = X->m;
    for (int i = 0; i < ob->n_obs; ++i) {
        const int type = ob->type[i];
        if (type == OR_OBS_VELOCITY || type == OR_OBS_PRESSURE) {
            const int pressure = type == OR_OBS_PRESSURE;
            const int comp = ob->component[i];
            const double *cell = pressure ? X->p : (comp == 0 ? X->Ux : X->Uy);
            const double *bnd = pressure ? X->pb : (comp == 0 ? X->Uxb : X->Uyb);
            const int zero = !pressure && comp == 2;
            const double center = zero ? 0.0 : cell[ob->cell[i]];
            double bx = 0, by = 0;
            for (int q = ob->nb_ptr[i]; q < ob->nb_ptr[i + 1]; ++q) {
                const double nbv = zero ? 0.0
                                        : (ob->nb_cell[q] < 0 ? bnd[ob->nb_face[q] - X->nif]
                                                              : cell[ob->nb_cell[q]]);
                const double difference = nbv - center;
                bx += ob->nb_w[q] * ob->nb_dx[q] * difference;
                by += ob->nb_w[q] * ob->nb_dy[q] * difference;
            }
            double gxv = 0, gyv = 0;
            if (fabs(ob->det[i]) > 1e-14 * (ob->axx[i] * ob->ayy[i] + 1e-300)) {
                gxv = (ob->ayy[i] * bx - ob->axy[i] * by) / ob->det[i];
                gyv = (ob->axx[i] * by - ob->axy[i] * bx) / ob->det[i];
            }
            const double value = center + gxv * ob->offx[i] + gyv * ob->offy[i];
            const double dynP = .5 * ob->ref_density[i] * ob->ref_velocity[i] * ob->ref_velocity[i];
            out[i] = pressure ? (value - ob->ref_pressure[i]) / smax(dynP, 1e-20) : value;
        } else if (type == OR_OBS_DRAG) {
            /* computeDrag's prepared form (ObservationOperator.hpp:728-740) */
            double force = 0;
            for (int q = ob->nb_ptr[i]; q < ob->nb_ptr[i + 1]; ++q) {
                const int fi = ob->nb_face[q], ow = m->owner[fi];
                force += X->p[ow] * m->fnx[fi] * m->farea[fi];
                const double delta = smax(m->fdelta[fi], 1e-20);
                const double tau = X->nu * (X->Ux[ow] - X->Uxb[fi - X->nif]) / delta;
                force += tau * m->farea[fi];
            }
            const double dynP = .5 * ob->ref_velocity[i] * ob->ref_velocity[i];
            out[i] = force / (dynP * ob->offx[i]);
        } else if (type == OR_OBS_SEPARATION || type == OR_OBS_REATTACHMENT) {
            /* computeSeparationPoint / computeReattachmentLength
             * (ObservationOperator.hpp:485-552, prepared form :741-767) */
            const int sep = type == OR_OBS_SEPARATION;
            const int q0 = ob->nb_ptr[i], q1 = ob->nb_ptr[i + 1];
            double value = (sep || q1 == q0) ? 0.0 : ob->nb_dx[q1 - 1];
            int have = 0;
            double prevTau = sep ? 1.0 : -1.0, prevX = 0.0;
            for (int q = q0; q < q1; ++q) {
                const int fi = ob->nb_face[q], ow = m->owner[fi];
                const double nx = m->fnx[fi], ny = m->fny[fi];
                const double un = X->Ux[ow] * nx + X->Uy[ow] * ny;
                const double tau = X->nu * (X->Ux[ow] - nx * un) / smax(m->fdelta[fi], 1e-20);
                const double x = ob->nb_dx[q];
                if (have && (sep ? (prevTau > 0 && tau <= 0) : (prevTau < 0 && tau >= 0))) {
                    const double den = sep ? prevTau - tau : tau - prevTau;
                    const double frac = fabs(den) > 1e-30 ? (sep ? prevTau : -prevTau) / den : 0.0;
                    value = prevX + frac * (x - prevX);
                    break;
                }
                prevTau = tau;
                prevX = x;
                have = 1;
            }
            out[i] = value;
        } else {
            double value = 0;
            const int fa = ob->face_a[i], fb = ob->face_b[i];
            for (int pass = 0; pass < 2; ++pass) {
                const int fi = pass == 0 ? fa : fb;
                if (fi < 0) break;
                const int ow = m->owner[fi];
                const double delta = smax(m->fdelta[fi], 1e-20);
                const double nx = m->fnx[fi], ny = m->fny[fi];
                const double un = X->Ux[ow] * nx + X->Uy[ow] * ny;
                const double utx = X->Ux[ow] - nx * un, uty = X->Uy[ow] - ny * un;
                const double tau = X->nu * sqrt(utx * utx + uty * uty) / delta;
                const double dynP = 0.5 * ob->ref_velocity[i] * ob->ref_velocity[i];
                const double cf = tau / smax(dynP, 1e-20);
                if (pass == 0) value = cf;
                else value = (1.0 - ob->t[i]) * value + ob->t[i] * cf;
            }
            out[i] = value;
        }
    }
}

/* ContractionEstimate (SIMPLESolver.cpp:32-74) over a sliding window of the
 * last 2m samples (older samples never enter rate() or remaining()). */
typedef struct {
    int m;
    long long n;
    double *ring;
} contraction_t;

static void ce_init(contraction_t *e, int window) {
    e->m = window < 2 ? 2 : window;
    e->n = 0;
    e->ring = (double *)calloc(2 * e->m, sizeof(double));
}
static void ce_push(contraction_t *e, double d) { e->ring[e->n++ % (2 * e->m)] = d; }
static double ce_at(const contraction_t *e, long long i) { return e->ring[i % (2 * e->m)]; }
static double ce_last(const contraction_t *e) { return e->n == 0 ? INFINITY : ce_at(e, e->n - 1); }
static double ce_rate(const contraction_t *e) {
    const long long n = e->n, m = e->m;
    if (n < 2 * m) return NAN;
    double dNew = 0.0, dOld = 0.0;
    for (long long i = n - m; i < n; ++i) dNew = smax(dNew, ce_at(e, i));
    for (long long i = n - 2 * m; i < n - m; ++i) dOld = smax(dOld, ce_at(e, i));
    if (dOld <= 0.0) return dNew > 0.0 ? INFINITY : 0.0;
    return pow(dNew / dOld, 1.0 / e->m);
}
static double ce_remaining(const contraction_t *e, double negligible) {
    const long long n = e->n, m = e->m;
    if (n < 2 * m) return INFINITY;
    double dNew = 0.0;
    for (long long i = n - m; i < n; ++i) dNew = smax(dNew, ce_at(e, i));
    if (dNew <= negligible) return dNew;
    const double r = ce_rate(e);
    if (!(r < 1.0)) return INFINITY;
    return dNew * r / (1.0 - r);
}

static double l1(int n, const double *v) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += fabs(v[i]);
    return s;
}

#define ALLOC(n) ((double *)calloc((n) > 0 ? (size_t)(n) : 1, sizeof(double)))

static void ctx_free(ctx_t *X) {
    double *arrs[] = {X->Ux, X->Uy, X->p, X->k, X->w, X->nuT, X->F1, X->F2, X->Pk, X->CDkw,
                      X->Uxb, X->Uyb, X->pb, X->kb, X->wb, X->omSing, X->omGrad, X->omLap,
                      X->gx, X->gy, X->gx2, X->gy2, X->gx3, X->gy3, X->tmpc, X->tmpb, X->tsrc,
                      X->phi, X->phiBarOld, X->phiPred, X->pCoeff, X->aP, X->aPu, X->aPstore,
                      X->solveValues, X->rbuf, X->UxPrev, X->UyPrev, X->kOld, X->wOld, X->pPrime,
                      X->pPrimeB, X->strain, X->wallOmega, X->balance, X->sys.diag, X->sys.upper,
                      X->sys.lower, X->sys.source};
    for (size_t i = 0; i < sizeof arrs / sizeof arrs[0]; ++i) free(arrs[i]);
    free(X->isWallFace);
    kwork_free(&X->kw);
    amg_free(&X->amg);
}

static void set_error(or_result_t *r, const char *msg) {
    snprintf(r->error, sizeof r->error, "%s", msg);
}

/* SIMPLESolver::solve (SIMPLESolver.cpp:1193-1716) driven as
 * ForwardModel::evaluateWithDiagnostics (ForwardModel.cpp:35-137). */
int or_evaluate(const or_mesh_t *mesh, const or_bc_t *bc, const or_settings_t *st, double nu,
                const double coeffs[11], const double init[5], const or_obs_t *obs,
                or_result_t *R) {
    return or_evaluate_ex(mesh, bc, st, nu, coeffs, init, obs, NULL, NULL, 0, 0.0, 0, R);
}

int or_evaluate_ex(const or_mesh_t *mesh, const or_bc_t *bc, const or_settings_t *st_in,
                   double nu, const double coeffs[11], const double init[5], const or_obs_t *obs,
                   const double *warm, double *save, int flags, double tol_override,
                   int max_iter_override, or_result_t *R) {
    or_settings_t st_local = *st_in;
    if (tol_override > 0.0) st_local.convergence_tol = tol_override;
    if (max_iter_override > 0) st_local.max_iterations = max_iter_override;
    const or_settings_t *st = &st_local;
    ctx_t Xs;
    ctx_t *X = &Xs;
    memset(X, 0, sizeof *X);
    X->m = mesh;
    X->bc = bc;
    X->s = st;
    X->obs = obs;
    X->nu = nu;
    memcpy(X->c, coeffs, sizeof X->c);
    const int nc = X->nc = mesh->n_cells, nif = X->nif = mesh->n_internal, nf = X->nf = mesh->n_faces;
    const int nb = X->nb = nf - nif;
    R->error[0] = 0;
    memset(R->krylov_iters, 0, sizeof R->krylov_iters);
    R->status = OR_UNKNOWN;
    R->loglik = -1e6;
    R->iters = 0;
    R->stalled = R->converged = R->diverged = 0;
    R->worst = NAN;
    R->rate = NAN;
    for (int i = 0; i < obs->n_obs; ++i) R->predictions[i] = NAN;
    if (st->momentum_solver == OR_SOLVER_AMG || st->turbulence_solver == OR_SOLVER_AMG) {
        set_error(R, "AMG is supported for the pressure correction only");
        return 1;
    }

    X->Ux = ALLOC(nc); X->Uy = ALLOC(nc); X->p = ALLOC(nc); X->k = ALLOC(nc); X->w = ALLOC(nc);
    X->nuT = ALLOC(nc); X->F1 = ALLOC(nc); X->F2 = ALLOC(nc); X->Pk = ALLOC(nc); X->CDkw = ALLOC(nc);
    X->Uxb = ALLOC(nb); X->Uyb = ALLOC(nb); X->pb = ALLOC(nb); X->kb = ALLOC(nb); X->wb = ALLOC(nb);
    X->omSing = ALLOC(nc); X->omGrad = ALLOC(nif); X->omLap = ALLOC(nc);
    X->gx = ALLOC(nc); X->gy = ALLOC(nc); X->gx2 = ALLOC(nc); X->gy2 = ALLOC(nc);
    X->gx3 = ALLOC(nc); X->gy3 = ALLOC(nc); X->tmpc = ALLOC(nc); X->tmpb = ALLOC(nb);
    X->tsrc = ALLOC(nc); X->phi = ALLOC(nf); X->phiBarOld = ALLOC(nf); X->phiPred = ALLOC(nf);
    X->pCoeff = ALLOC(nf); X->aP = ALLOC(nc); X->aPu = ALLOC(nc); X->aPstore = ALLOC(nc);
    X->solveValues = ALLOC(nc); X->rbuf = ALLOC(nc); X->UxPrev = ALLOC(nc); X->UyPrev = ALLOC(nc);
    X->kOld = ALLOC(nc); X->wOld = ALLOC(nc); X->pPrime = ALLOC(nc); X->pPrimeB = ALLOC(nb);
    X->strain = ALLOC(nc); X->wallOmega = ALLOC(nc); X->balance = ALLOC(nc);
    X->sys.n = nc;
    X->sys.nif = nif;
    X->sys.own = mesh->owner;
    X->sys.nbr = mesh->neighbor;
    X->sys.diag = ALLOC(nc); X->sys.upper = ALLOC(nif); X->sys.lower = ALLOC(nif);
    X->sys.source = ALLOC(nc);
    X->isWallFace = (char *)calloc(nf > 0 ? nf : 1, 1);
    for (int pi = 0; pi < mesh->n_patches; ++pi)
        if (mesh->patch_type[pi] == OR_PATCH_WALL)
            for (int q = mesh->patch_face_ptr[pi]; q < mesh->patch_face_ptr[pi + 1]; ++q)
                X->isWallFace[mesh->patch_face_idx[q]] = 1;
    kwork_alloc(&X->kw, nc);
    X->amg.needFullSetup = 1;
    X->amg.itersAfterSetup = -1;

    prepare_planar(X);

    const int nbs = nb > 1 ? nb : 1;
    if (warm) {
        /* warm start: the state as stored (ForwardModel.cpp:85-88); the
         * sensitivity path re-applies the boundary conditions with this
         * solve's beta1 (ParameterSensitivity.cpp:790-794) */
        double *dst[10] = {X->Ux, X->Uy, X->p, X->k, X->w, X->nuT, X->F1, X->F2, X->Pk, X->CDkw};
        double *bdst[5] = {X->Uxb, X->Uyb, X->pb, X->kb, X->wb};
        for (int a = 0; a < 10; ++a) memcpy(dst[a], warm + (size_t)a * nc, sizeof(double) * nc);
        for (int a = 0; a < 5; ++a)
            memcpy(bdst[a], warm + (size_t)10 * nc + (size_t)a * nbs, sizeof(double) * nb);
        X->turbEstablished = warm[(size_t)10 * nc + (size_t)5 * nbs] != 0.0;
        if (flags & OR_FLAG_WARM_BCS) {
            apply_velocity_bc(X);
            apply_pressure_bc(X);
            apply_k_bc(X);
            apply_omega_bc(X, X->c[C_B1]);
        }
    } else {
    /* initUniform (SIMPLESolver.cpp:267-298) */
    for (int i = 0; i < nc; ++i) {
        X->Ux[i] = init[0];
        X->Uy[i] = init[1];
        X->p[i] = init[2];
        X->k[i] = init[3];
        X->w[i] = init[4];
        X->F1[i] = 1.0;
        X->F2[i] = 1.0;
        X->Pk[i] = 0.0;
        X->CDkw[i] = 0.0;
    }
    for (int b = 0; b < nb; ++b) {
        X->Uxb[b] = init[0];
        X->Uyb[b] = init[1];
        X->pb[b] = init[2];
        X->kb[b] = init[3];
        X->wb[b] = init[4];
    }
    X->turbEstablished = 0;
    for (int ci = 0; ci < nc; ++ci) {
        double y = smax(mesh->wall_dist[ci], 1e-20);
        double omWall = resolved_wall_omega(X, y);
        X->w[ci] = smax(omWall, init[4]);
    }
    for (int ci = 0; ci < nc; ++ci)
        X->nuT[ci] = X->c[C_A1] * init[3] / smax(X->c[C_A1] * X->w[ci], 1e-20);
    apply_velocity_bc(X);
    apply_pressure_bc(X);
    apply_k_bc(X);
    apply_omega_bc(X, 0.075);
    }

    /* solve() */
    double lastKChange = 1.0, lastOmChange = 1.0;
    const double tol = st->convergence_tol;
    const double negligible = 0.01 * tol;
    contraction_t estU, estK, estOm, estQ;
    ce_init(&estU, st->convergence_window);
    ce_init(&estK, st->convergence_window);
    ce_init(&estOm, st->convergence_window);
    ce_init(&estQ, st->convergence_window);
    const int nObs = obs->n_obs;
    /* the observation convergence monitor (ForwardModel.cpp:60-72); a
     * ParameterSensitivity solver has none */
    const int nMon = (flags & OR_FLAG_NO_MONITOR) ? 0 : nObs;
    double *q = ALLOC(nObs), *qPrev = ALLOC(nObs);
    int haveQPrev = 0;
    double worst = INFINITY;
    double pressureFluxReference = 0.0;
    double normUx0 = 1, normUy0 = 1, normP0 = 1;
    int converged = 0, diverged = 0, finalIter = 0;
    double finalWorst = 0, rateU = 0;
    int failed = 0;

    interpolated_fluxes(X, X->phi);
    for (int iter = 0; iter < st->max_iterations; ++iter) {
        for (int ci = 0; ci < nc; ++ci) {
            X->UxPrev[ci] = X->Ux[ci];
            X->UyPrev[ci] = X->Uy[ci];
        }
        double momResid = 0.0, momScale = 0.0;
        interpolated_fluxes(X, X->phiBarOld);
        const int turbActive = iter >= st->turb_start_iter;
        const int turbUpdate =
            turbActive && ((iter - st->turb_start_iter) % st->turb_update_interval == 0);
        if (turbUpdate) {
            sst_compute_fields(X);
            if (!X->turbEstablished && iter >= st->turb_start_iter + st->nut_floor_iters)
                X->turbEstablished = 1;
            const double nuTMin = X->turbEstablished ? 0.0 : 0.1 * nu;
            for (int ci = 0; ci < nc; ++ci) X->nuT[ci] = smax(X->nuT[ci], nuTMin);
        }
        sres_t resUx = {0, 0, 0, 0}, resUy = {0, 0, 0, 0}, resUz = {0, 0, 0, 0};
        /* x momentum */
        assemble_momentum(X, 0);
        round_assembly(&X->sys, X->aP);
        memcpy(X->solveValues, X->Ux, sizeof(double) * nc);
        ls_residual(&X->sys, X->solveValues, X->rbuf);
        momResid = l1(nc, X->rbuf);
        for (int ci = 0; ci < nc; ++ci) {
            const double contribution =
                X->aPu[ci] * sqrt(X->Ux[ci] * X->Ux[ci] + X->Uy[ci] * X->Uy[ci]);
            momScale += contribution;
        }
        if (linear_solve(st->momentum_solver, &X->amg, &X->sys, X->solveValues, st->inner_iterations,
                         st->inner_tolerance, &X->kw, &resUx)) {
            set_error(R, "linear system coefficients and source must be finite");
            failed = 1;
            break;
        }
        R->krylov_iters[0] += resUx.iterations;
        memcpy(X->Ux, X->solveValues, sizeof(double) * nc);
        memcpy(X->aPstore, X->aP, sizeof(double) * nc);
        /* y momentum */
        assemble_momentum(X, 1);
        round_assembly(&X->sys, X->aP);
        memcpy(X->solveValues, X->Uy, sizeof(double) * nc);
        ls_residual(&X->sys, X->solveValues, X->rbuf);
        momResid = smax(momResid, l1(nc, X->rbuf));
        if (linear_solve(st->momentum_solver, &X->amg, &X->sys, X->solveValues, st->inner_iterations,
                         st->inner_tolerance, &X->kw, &resUy)) {
            set_error(R, "linear system coefficients and source must be finite");
            failed = 1;
            break;
        }
        R->krylov_iters[1] += resUy.iterations;
        memcpy(X->Uy, X->solveValues, sizeof(double) * nc);
        /* z momentum is identically zero on a 2D mesh: zero source, zero
         * residual, the solver returns at r0 = 0 */
        apply_velocity_bc(X);

        /* predicted flux and pressure correction */
        interpolated_fluxes(X, X->phiPred);
        add_rhie_chow(X, X->aPstore, X->phiPred);
        const double relaxCorr = 1.0 - st->alpha_u;
        for (int fi = 0; fi < nif; ++fi) X->phiPred[fi] += relaxCorr * (X->phi[fi] - X->phiBarOld[fi]);
        assemble_pressure(X, X->aPstore, X->phiPred);
        const double contScaled = l1(nc, X->sys.source) / smax(X->massFluxScale, 1e-300);
        round_assembly(&X->sys, NULL);
        for (int i = 0; i < nc; ++i) X->solveValues[i] = 0.0;
        sres_t resP;
        if (linear_solve(st->pressure_solver, &X->amg, &X->sys, X->solveValues, st->inner_iterations,
                         st->inner_tolerance, &X->kw, &resP)) {
            set_error(R, "linear system coefficients and source must be finite");
            failed = 1;
            break;
        }
        R->krylov_iters[2] += resP.iterations;
        memcpy(X->pPrime, X->solveValues, sizeof(double) * nc);
        for (int pi = 0; pi < mesh->n_patches; ++pi) {
            if (mesh->patch_type[pi] == OR_PATCH_OUTLET) continue;
            for (int qq = mesh->patch_face_ptr[pi]; qq < mesh->patch_face_ptr[pi + 1]; ++qq) {
                int fi = mesh->patch_face_idx[qq];
                X->pPrimeB[fi - nif] = X->pPrime[mesh->owner[fi]];
            }
        }
        /* correctVelocity / correctPressure (SIMPLESolver.cpp:916-940) */
        gg_grad(X, X->pPrime, X->pPrimeB, X->gx, X->gy);
        for (int ci = 0; ci < nc; ++ci) {
            double rAP = (fabs(X->aPstore[ci]) > 1e-30) ? 1.0 / X->aPstore[ci] : 0.0;
            double vol = mesh->vol[ci];
            X->Ux[ci] -= rAP * X->gx[ci] * vol;
            X->Uy[ci] -= rAP * X->gy[ci] * vol;
        }
        apply_velocity_bc(X);
        for (int ci = 0; ci < nc; ++ci) X->p[ci] += st->alpha_p * X->pPrime[ci];
        apply_pressure_bc(X);
        for (int fi = 0; fi < nf; ++fi) {
            if (fi >= nif) X->phi[fi] = X->phiPred[fi] + X->pCoeff[fi] * X->pPrime[mesh->owner[fi]];
            else
                X->phi[fi] = X->phiPred[fi] - X->pCoeff[fi] * (X->pPrime[mesh->neighbor[fi]] -
                                                               X->pPrime[mesh->owner[fi]]);
        }

        /* turbulence */
        if (turbUpdate) {
            memcpy(X->kOld, X->k, sizeof(double) * nc);
            memcpy(X->wOld, X->w, sizeof(double) * nc);
            assemble_k(X);
            round_assembly(&X->sys, NULL);
            memcpy(X->solveValues, X->k, sizeof(double) * nc);
            sres_t resK;
            if (linear_solve(st->turbulence_solver, &X->amg, &X->sys, X->solveValues,
                             st->inner_iterations, st->inner_tolerance, &X->kw, &resK)) {
                set_error(R, "linear system coefficients and source must be finite");
                failed = 1;
                break;
            }
            R->krylov_iters[3] += resK.iterations;
            memcpy(X->k, X->solveValues, sizeof(double) * nc);
            for (int ci = 0; ci < nc; ++ci) X->k[ci] = smax(st->k_min, smin(st->k_max, X->k[ci]));
            for (int b = 0; b < nb; ++b) X->kb[b] = smax(st->k_min, smin(st->k_max, X->kb[b]));
            apply_k_bc(X);
            assemble_omega(X);
            omega_wall_values(X);
            constrain_omega(X);
            round_assembly(&X->sys, NULL);
            memcpy(X->solveValues, X->w, sizeof(double) * nc);
            for (int ci = 0; ci < nc; ++ci)
                if (isfinite(X->wallOmega[ci])) X->solveValues[ci] = X->wallOmega[ci];
            sres_t resOm;
            if (linear_solve(st->turbulence_solver, &X->amg, &X->sys, X->solveValues,
                             st->inner_iterations, st->inner_tolerance, &X->kw, &resOm)) {
                set_error(R, "linear system coefficients and source must be finite");
                failed = 1;
                break;
            }
            R->krylov_iters[4] += resOm.iterations;
            memcpy(X->w, X->solveValues, sizeof(double) * nc);
            for (int ci = 0; ci < nc; ++ci) X->w[ci] = smax(st->omega_min, smin(1e15, X->w[ci]));
            for (int b = 0; b < nb; ++b) X->wb[b] = smax(st->omega_min, smin(1e15, X->wb[b]));
            for (int ci = 0; ci < nc; ++ci)
                if (isfinite(X->wallOmega[ci])) X->w[ci] = X->wallOmega[ci];
            apply_omega_bc(X, X->c[C_B1]);
            double kMaxVal = 1e-30, omMaxVal = 1e-30, kMaxDiff = 0.0, omMaxDiff = 0.0;
            for (int ci = 0; ci < nc; ++ci) {
                kMaxVal = smax(kMaxVal, fabs(X->k[ci]));
                omMaxVal = smax(omMaxVal, fabs(X->w[ci]));
                kMaxDiff = smax(kMaxDiff, fabs(X->k[ci] - X->kOld[ci]));
                omMaxDiff = smax(omMaxDiff, fabs(X->w[ci] - X->wOld[ci]));
            }
            const double kChangeNorm = kMaxDiff / kMaxVal;
            const double omChangeNorm = omMaxDiff / omMaxVal;
            lastKChange = kChangeNorm;
            lastOmChange = omChangeNorm;
            ce_push(&estK, kChangeNorm);
            ce_push(&estOm, omChangeNorm);
        }

        /* convergence bookkeeping (SIMPLESolver.cpp:1545-1705) */
        const double eUx = resUx.initial_res, eUy = resUy.initial_res, eUz = resUz.initial_res;
        const double eP = resP.initial_res;
        const double eK = turbActive ? lastKChange : 0.0;
        const double eOm = turbActive ? lastOmChange : 0.0;
        double uMax = 1e-300, duMax = 0.0;
        for (int ci = 0; ci < nc; ++ci) {
            uMax = smax(uMax, sqrt(X->Ux[ci] * X->Ux[ci] + X->Uy[ci] * X->Uy[ci]));
            double dx = X->Ux[ci] - X->UxPrev[ci], dy = X->Uy[ci] - X->UyPrev[ci];
            duMax = smax(duMax, sqrt(dx * dx + dy * dy));
        }
        const double momentumScaled = momResid / smax(momScale, 1e-300);
        const double continuityScaled = contScaled;
        const double changeU = duMax / uMax;
        ce_push(&estU, changeU);
        const double remainingU = ce_remaining(&estU, negligible);
        if (nMon > 0 && iter % (st->monitor_interval > 1 ? st->monitor_interval : 1) == 0) {
            observe(X, q);
            int finite = 1;
            for (int i = 0; i < nObs; ++i) finite = finite && isfinite(q[i]);
            if (!finite) {
                set_error(R, "convergence monitor returned a nonfinite observation");
                failed = 1;
                break;
            }
            if (haveQPrev) {
                double qMax = 0.0, dq = 0.0;
                for (int i = 0; i < nObs; ++i) qMax = smax(qMax, fabs(q[i]));
                for (int i = 0; i < nObs; ++i) {
                    const double floor_ = sqrt(obs->variance[i]);
                    double den = smax(smax(fabs(q[i]), floor_), 1e-300);
                    dq = smax(dq, fabs(q[i] - qPrev[i]) / den);
                }
                ce_push(&estQ, dq);
            }
            memcpy(qPrev, q, sizeof(double) * nObs);
            haveQPrev = 1;
        }
        worst = smax(smax(smax(momentumScaled, continuityScaled), changeU), remainingU);
        if (turbActive) {
            worst = smax(worst, lastKChange);
            worst = smax(worst, ce_remaining(&estK, negligible));
            worst = smax(worst, lastOmChange);
            worst = smax(worst, ce_remaining(&estOm, negligible));
        }
        if (nMon > 0) {
            worst = smax(worst, ce_last(&estQ));
            worst = smax(worst, ce_remaining(&estQ, negligible));
        }
        if (!isfinite(worst)) worst = INFINITY;
        if (iter == 0) {
            normUx0 = smax(eUx, 1e-30);
            normUy0 = smax(eUy, 1e-30);
            normP0 = smax(eP, 1e-30);
            pressureFluxReference = X->massFluxScale;
        } else if (iter < 50) {
            normP0 = smax(normP0, eP);
            pressureFluxReference = smax(pressureFluxReference, X->massFluxScale);
        }
        const double normMom0 = smax(normUx0, normUy0);
        const double nUx = normMom0 < 1e-20 ? 0.0 : eUx / normMom0;
        const double nUy = normMom0 < 1e-20 ? 0.0 : eUy / normMom0;
        const double pressureReference = smax(normP0, pressureFluxReference);
        const double nP = pressureReference < 1e-20 ? 0.0 : eP / pressureReference;
        double maxRes = smax(smax(nUx, nUy), nP);
        if (turbActive) maxRes = smax(smax(maxRes, lastKChange), lastOmChange);
        if (!isfinite(eUx) || !isfinite(eUy) || !isfinite(eUz) || !isfinite(eP) || !isfinite(eK) ||
            !isfinite(eOm) || !isfinite(lastKChange) || !isfinite(lastOmChange) ||
            !isfinite(momentumScaled) || !isfinite(continuityScaled) || !isfinite(changeU)) {
            maxRes = INFINITY;
            worst = INFINITY;
        }
        const int turbScheduled = st->turb_start_iter < st->max_iterations;
        finalWorst = worst;
        rateU = ce_rate(&estU);
        if (worst < tol && iter > 0 && (!turbScheduled || (turbActive && X->turbEstablished))) {
            converged = 1;
            finalIter = iter;
            break;
        }
        if (isnan(maxRes) || isinf(maxRes) || maxRes > st->divergence_limit) {
            diverged = 1;
            finalIter = iter;
            break;
        }
        finalIter = iter + 1;
    }
    if (!failed && !converged && !diverged) finalIter = st->max_iterations;
    const int stalled = !failed && !converged && !diverged && isfinite(rateU) && rateU >= 1.0;

    /* a solve that threw leaves the default ConvergenceHistory behind
     * (FlowFields.hpp:53-66): zero iterations, measure and rate */
    R->iters = failed ? 0 : finalIter;
    R->worst = failed ? 0.0 : finalWorst;
    R->rate = failed ? 0.0 : rateU;
    R->converged = converged;
    R->diverged = diverged;
    R->stalled = stalled;
    if (failed) {
        R->status = OR_UNKNOWN;
        R->loglik = -1e6;
    } else if (diverged) {
        R->status = OR_DIVERGED;
        R->loglik = -1e6;
    } else {
        R->status = converged ? OR_CONVERGED : OR_UNCONVERGED;
        observe(X, R->predictions);
        /* ObservationOperator::logLikelihood (ObservationOperator.hpp:592-607) */
        double ll = 0.0;
        for (int i = 0; i < nObs; ++i) {
            if (!isfinite(R->predictions[i])) {
                set_error(R, "Likelihood predictions must be finite");
                failed = 1;
                break;
            }
            const double var = obs->variance[i];
            double diff = R->predictions[i] - obs->observed[i];
            ll += -0.5 * diff * diff / var;
            if (!isfinite(ll)) {
                set_error(R, "Gaussian log-likelihood exceeds finite range");
                failed = 1;
                break;
            }
        }
        if (failed) {
            R->status = OR_UNKNOWN;
            R->loglik = -1e6;
            for (int i = 0; i < nObs; ++i) R->predictions[i] = NAN;
        } else {
            R->loglik = ll;
        }
    }
    double *outs[] = {R->Ux, R->Uy, R->p, R->k, R->omega, R->nuT, R->F1, R->F2, R->Pk, R->CDkw};
    double *srcs[] = {X->Ux, X->Uy, X->p, X->k, X->w, X->nuT, X->F1, X->F2, X->Pk, X->CDkw};
    for (int i = 0; i < 10; ++i)
        if (outs[i]) memcpy(outs[i], srcs[i], sizeof(double) * nc);
    if (R->flux) memcpy(R->flux, X->phi, sizeof(double) * nf);
    if (save) {
        for (int a = 0; a < 10; ++a) memcpy(save + (size_t)a * nc, srcs[a], sizeof(double) * nc);
        double *bsrc[5] = {X->Uxb, X->Uyb, X->pb, X->kb, X->wb};
        for (int a = 0; a < 5; ++a) {
            memcpy(save + (size_t)10 * nc + (size_t)a * nbs, bsrc[a], sizeof(double) * nb);
            for (int b = nb; b < nbs; ++b) save[(size_t)10 * nc + (size_t)a * nbs + b] = 0.0;
        }
        save[(size_t)10 * nc + (size_t)5 * nbs] = X->turbEstablished ? 1.0 : 0.0;
        save[(size_t)10 * nc + (size_t)5 * nbs + 1] = 0.0;
    }
    free(q);
    free(qPrev);
    free(estU.ring);
    free(estK.ring);
    free(estOm.ring);
    free(estQ.ring);
    ctx_free(X);
    return 0;
}

/*
 * CPU oracle for the ml_prec iterative-solve path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's serial algorithms, against which
 * the HIP product path is tested: only tests, the smoke check and the
 * benchmark's CPU-baseline leg load it.  Citations are reference file:line.
 *
 * Compiled with -ffp-contract=off and no -march, so every a*b+c is a rounded
 * multiply followed by a rounded add, exactly as the reference's default
 * x86-64 -O2 build evaluates it.
 *
 * How each part is pinned (the reference itself cannot be built here: every
 * source includes the autoconf-generated config/config.h, which only exists
 * as config.h.in):
 *   multiply / sym_multiply / residual  -- dense A*x on tiny grids; x=1 gives
 *                                          row sums; sym == full for symmetric A
 *   ddot / dnrm2 / daxpy / dscal        -- closed forms (sum of squares of
 *                                          0..n-1), exact small-integer cases
 *   PCG / GMRES                         -- scipy / numpy.linalg.solve on small
 *                                          SPD systems; identity-preconditioned
 *                                          PCG iterates vs textbook CG
 *   uh_cheb construction                -- invariants: map is a permutation,
 *                                          M+n+Md=N, per-row drop budget and
 *                                          maximality, tails have one link,
 *                                          L*U == modified matrix, dense Schur
 *                                          complement, closed-form bounds
 *   uh_cheb apply                       -- dense recursion: B_l = U^-1 diag(I,
 *                                          (I - p_k(B_{l+1} S)) S^-1) L^-1 with
 *                                          the scaled Chebyshev polynomial p_k
 *   multi_split construction / apply    -- the same invariants; apply vs dense
 *                                          nested Richardson
 *   GMRES update_x                      -- hand-checked against the reference's
 *                                          loop bound (gmres.cc:156-159)
 */
#include "oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <float.h>

#define NO_INDEX 0xffffffffu

/* ------------------------------------------------------------------ */
/* small helpers                                                       */
/* ------------------------------------------------------------------ */
static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void set_err(char* err, size_t errlen, const char* fmt, ...) {
    if (!err || !errlen) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, errlen, fmt, ap);
    va_end(ap);
}

static void* xmalloc(size_t n) {
    void* p = malloc(n ? n : 1);
    if (!p) { fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n); abort(); }
    return p;
}
static void* xcalloc(size_t n, size_t s) {
    void* p = calloc(n ? n : 1, s ? s : 1);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

typedef struct { uint32_t* p; size_t n, cap; } vu32;
typedef struct { double* p; size_t n, cap; } vf64;

static void vu32_push(vu32* v, uint32_t x) {
    if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 16; v->p = realloc(v->p, v->cap * sizeof(uint32_t)); }
    v->p[v->n++] = x;
}
static void vf64_push(vf64* v, double x) {
    if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 16; v->p = realloc(v->p, v->cap * sizeof(double)); }
    v->p[v->n++] = x;
}

/* is_equal: |x-y| <= DBL_EPSILON (include/tools.h:28-30) */
static int is_equal(double x, double y) { return fabs(x - y) <= DBL_EPSILON; }

static void mat_free(or_mat* A) {
    free(A->ia); free(A->ja); free(A->a);
    free(A->iasym); free(A->jasym); free(A->asym);
    memset(A, 0, sizeof(*A));
}

static void mat_from_arrays(or_mat* A, uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a) {
    memset(A, 0, sizeof(*A));
    A->nrow = A->ncol = n;
    uint32_t nnz = ia[n];
    A->ia = xmalloc((size_t)(n + 1) * sizeof(uint32_t));
    A->ja = xmalloc((size_t)nnz * sizeof(uint32_t));
    A->a  = xmalloc((size_t)nnz * sizeof(double));
    memcpy(A->ia, ia, (size_t)(n + 1) * sizeof(uint32_t));
    memcpy(A->ja, ja, (size_t)nnz * sizeof(uint32_t));
    memcpy(A->a, a, (size_t)nnz * sizeof(double));
}

static uint32_t mat_nnz(const or_mat* A) { return (A->nrow && A->ncol) ? A->ia[A->nrow] : 0; }

/* ------------------------------------------------------------------ */
/* BLAS-1: modules/vector/blas.cc:43-82 (the no-BLAS build)            */
/* ------------------------------------------------------------------ */
void or_daxpy(double alpha, const double* x, double* y, uint32_t n) {
    if (!is_equal(alpha, 1.))                         /* blas.cc:44-46 */
        for (uint32_t i = 0; i < n; i++) y[i] += alpha * x[i];
    else                                              /* blas.cc:47-49 */
        for (uint32_t i = 0; i < n; i++) y[i] += x[i];
}
/* dscal(Vector&) skips alpha == 1 (blas.cc:77-82) */
void or_dscal(double alpha, double* x, uint32_t n) {
    if (n && !is_equal(alpha, 1.))
        for (uint32_t i = 0; i < n; i++) x[i] *= alpha;
}
double or_ddot(const double* x, const double* y, uint32_t n) {     /* blas.cc:57-62 */
    double s = 0.;
    for (uint32_t i = 0; i < n; i++) s += x[i] * y[i];
    return s;
}
double or_dnrm2(const double* x, uint32_t n) { return sqrt(or_ddot(x, x, n)); }  /* blas.cc:64-66 */

/* ------------------------------------------------------------------ */
/* SpMV: modules/matrix/csrmatrix.cc:260-300, skyline.cc:86-155        */
/* ------------------------------------------------------------------ */
void or_multiply(const or_mat* A, const double* v, double* res) {
    if (A->nrow == 0) return;
    memset(res, 0, (size_t)A->nrow * sizeof(double));          /* csrmatrix.cc:265 */
    for (uint32_t i = 0; i < A->nrow; i++)                     /* csrmatrix.cc:271-273 */
        for (uint32_t j = A->ia[i]; j < A->ia[i + 1]; j++)
            res[i] += A->a[j] * v[A->ja[j]];
}

void or_residual(const or_mat* A, const double* b, const double* x, double* r) {
    or_multiply(A, x, r);                                      /* csrmatrix.cc:297-299 */
    for (uint32_t i = 0; i < A->nrow; i++) r[i] = b[i] - r[i];
}

static uint32_t upper_bound_u32(const uint32_t* p, uint32_t lo, uint32_t hi, uint32_t key) {
    while (lo < hi) { uint32_t mid = lo + (hi - lo) / 2; if (!(key < p[mid])) lo = mid + 1; else hi = mid; }
    return lo;
}
static uint32_t lower_bound_u32(const uint32_t* p, uint32_t lo, uint32_t hi, uint32_t key) {
    while (lo < hi) { uint32_t mid = lo + (hi - lo) / 2; if (p[mid] < key) lo = mid + 1; else hi = mid; }
    return lo;
}

int or_optimize_storage(or_mat* A) {                           /* skyline.cc:86-113 */
    uint32_t n = A->nrow, nnz = mat_nnz(A);
    free(A->iasym); free(A->jasym); free(A->asym);
    A->iasym = xmalloc((size_t)(n + 1) * sizeof(uint32_t));
    A->jasym = xmalloc((size_t)((nnz + n) >> 1) * sizeof(uint32_t) + 4);
    A->asym  = xmalloc((size_t)((nnz + n) >> 1) * sizeof(double) + 8);
    A->iasym[0] = 0;
    uint32_t ind = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t start = A->ia[i], end = A->ia[i + 1];
        uint32_t middle = upper_bound_u32(A->ja, start + 1, end, i);
        for (uint32_t j = start; j < middle; j++, ind++) { A->jasym[ind] = A->ja[j]; A->asym[ind] = A->a[j]; }
        A->iasym[i + 1] = ind;
    }
    return 0;
}

void or_sym_multiply(const or_mat* A, const double* v, double* res) {   /* skyline.cc:115-155 */
    memset(res, 0, (size_t)A->nrow * sizeof(double));
    if (A->iasym) {
        for (uint32_t i = 0; i < A->nrow; i++) {
            uint32_t start = A->iasym[i], end = A->iasym[i + 1];
            res[i] += A->asym[start] * v[i];
            for (uint32_t j_ = start + 1; j_ < end; j_++) {
                uint32_t j = A->jasym[j_];
                res[i] += A->asym[j_] * v[j];
                res[j] += A->asym[j_] * v[i];
            }
        }
    } else {
        for (uint32_t i = 0; i < A->nrow; i++) {
            uint32_t start = A->ia[i], end = A->ia[i + 1];
            uint32_t middle = upper_bound_u32(A->ja, start + 1, end, i);
            res[i] += A->a[start] * v[i];
            for (uint32_t j_ = start + 1; j_ < middle; j_++) {
                uint32_t j = A->ja[j_];
                double val = A->a[j_];
                res[i] += val * v[j];
                res[j] += val * v[i];
            }
        }
    }
}

/* SkylineMatrix::index (skyline.cc:15-28) */
static uint32_t sky_index(const or_mat* A, uint32_t i, uint32_t j) {
    if (i == j) return A->ia[i];
    uint32_t s = A->ia[i] + 1, e = A->ia[i + 1];
    uint32_t it = lower_bound_u32(A->ja, s, e, j);
    if (it != e && !(j < A->ja[it])) return it;
    return NO_INDEX;
}

/* generate_x0 (modules/solvers/solvers.cc:9-21): glibc srandom(3)/random() */
void or_generate_x0(double* x, uint32_t n) {
    srandom(3);
    for (uint32_t i = 0; i < n; i++)
        x[i] = 20. * (random() - 0.5 * RAND_MAX) / RAND_MAX + 100;
}

/* cheb(x,k) (modules/prec/misc/misc.cc:232-242), host libm */
double or_cheb(double x, uint32_t k) {
    switch (k) {
        case 0: return 1;
        case 1: return x;
        case 2: return 2 * x * x - 1;
        case 3: return x * (4 * x * x - 3);
        default: return cosh(k * acosh(x));
    }
}

/* check_and_replace_eps (solvers.cc:49-58) */
static void check_and_replace_eps(double init_norm, double* eps) {
    const double MIN_EPS = 1e-14;
    if (init_norm * (*eps) < MIN_EPS) *eps = MIN_EPS / init_norm;
}

/* ------------------------------------------------------------------ */
/* sorted set of uint32 with duplicate-free use (include/svector.h)    */
/* ------------------------------------------------------------------ */
typedef struct { uint32_t* p; uint32_t n, cap; } sset;
static void sset_insert(sset* s, uint32_t x) {                  /* svector.h:190-208, 292-295 */
    uint32_t pos = lower_bound_u32(s->p, 0, s->n, x);
    if (s->n == s->cap) { s->cap = s->cap ? 2 * s->cap : 8; s->p = realloc(s->p, s->cap * sizeof(uint32_t)); }
    memmove(s->p + pos + 1, s->p + pos, (size_t)(s->n - pos) * sizeof(uint32_t));
    s->p[pos] = x;
    s->n++;
}

/* ------------------------------------------------------------------ */
/* link status objects                                                 */
/* ------------------------------------------------------------------ */
enum { PRESENT = 0, REMOVED = 1, ABSENT = 2 };

typedef struct {
    int kind;                /* 1 = cheb (two-sided counters), 2 = multi_split (one-sided flags) */
    const or_mat* A;
    uint32_t* a;             /* cheb: counters / references (cheb_misc.h:33-48) */
    unsigned char* c;        /* multi_split: flags (multi_split_misc.h:11-17) */
    uint32_t row;
} ltype_t;

static uint32_t ltc_index(const ltype_t* t, uint32_t i, uint32_t j) {   /* cheb_misc.h:26-34 */
    if (i > j) { uint32_t s = i; i = j; j = s; }
    return sky_index(t->A, i, j);
}

static void ltc_init(ltype_t* t, const or_mat* A) {                     /* cheb_misc.h:45-58 */
    t->kind = 1; t->A = A; t->c = NULL; t->row = NO_INDEX;
    uint32_t nnz = mat_nnz(A);
    t->a = xmalloc((size_t)nnz * sizeof(uint32_t) + 4);
    for (uint32_t i = 0; i < A->nrow; i++)
        for (uint32_t j_ = A->ia[i] + 1; j_ < A->ia[i + 1]; j_++) {
            uint32_t j = A->ja[j_];
            t->a[j_] = (j < i) ? ltc_index(t, j, i) : 2;
        }
}

static void lt_copy(ltype_t* dst, const ltype_t* src) {
    *dst = *src;
    uint32_t nnz = mat_nnz(src->A);
    if (src->a) { dst->a = xmalloc((size_t)nnz * sizeof(uint32_t) + 4); memcpy(dst->a, src->a, (size_t)nnz * sizeof(uint32_t)); }
    if (src->c) { dst->c = xmalloc((size_t)nnz + 1); memcpy(dst->c, src->c, nnz); }
}
static void lt_free(ltype_t* t) { free(t->a); free(t->c); t->a = NULL; t->c = NULL; }

static uint32_t ltc_real(const ltype_t* t, uint32_t j_) {              /* cheb_misc.h:37-41 */
    return (t->A->ja[j_] < t->row) ? t->a[j_] : j_;
}
static int lt_stat(const ltype_t* t, uint32_t j_) {
    if (t->kind == 1) return t->a[ltc_real(t, j_)] ? PRESENT : REMOVED;   /* cheb_misc.h:81 */
    return t->c[j_] ? PRESENT : REMOVED;                                 /* multi_split_misc.h:20 */
}
static int lt_stat_ij(const ltype_t* t, uint32_t i, uint32_t j) {      /* multi_split_misc.h:21-26 */
    uint32_t ind = sky_index(t->A, i, j);
    if (ind == NO_INDEX) return ABSENT;
    return t->c[ind] ? PRESENT : REMOVED;
}

/* ------------------------------------------------------------------ */
/* hierarchy                                                           */
/* ------------------------------------------------------------------ */
typedef struct {
    uint32_t N, nnz, M, Md;
    uint32_t ncheb, niter;
    double alpha, beta, lmin, lmax;      /* uh_cheb (cheb_prec.h:36-38) */
    double q, eps;                       /* multi_split (multi_split_prec.h:21-23) */
    or_mat A;                            /* level matrix for l >= 1 */
    or_mat L, U;
    uint32_t *map, *rmap;
    double *dval, *aux;
    double *w, *tmp, *x2, *u1, *u0, *r, *F;
} lev_t;

struct or_prec {
    int kind;
    int use_tails;
    uint32_t nlevels;
    uint32_t nalloc;
    lev_t* lv;
    or_mat A0;
    /* multi_split options */
    int inner_dynamic, order;
    uint32_t degree_max, coarse_n;
    /* diag */
    double* d;
    uint32_t n;
    /* GS: the lower triangle with its diagonal first (gs_prec.cc:26-65) */
    or_mat Lgs;
    /* composite: the two stages; sym: sub[0] is the uh_cheb hierarchy of Asym */
    or_prec* sub[2];
    or_mat Asym;
    double sym_eps;
};

static void lev_free(lev_t* L) {
    mat_free(&L->A); mat_free(&L->L); mat_free(&L->U);
    free(L->map); free(L->rmap); free(L->dval); free(L->aux);
    free(L->w); free(L->tmp); free(L->x2); free(L->u1); free(L->u0); free(L->r); free(L->F);
}

void or_prec_free(or_prec* P) {
    if (!P) return;
    for (uint32_t l = 0; l < P->nalloc; l++) lev_free(&P->lv[l]);
    free(P->lv);
    mat_free(&P->A0);
    mat_free(&P->Lgs);
    mat_free(&P->Asym);
    or_prec_free(P->sub[0]);
    or_prec_free(P->sub[1]);
    free(P->d);
    free(P);
}

/* psort: stable insertion sort by |a| (cheb_constr.cc:26-36, multi_split_constr.cc:26-36) */
static void psort_abs(const double* a, uint32_t n, uint32_t* sorted) {
    sorted[0] = 0;
    for (uint32_t i = 1; i < n; i++) {
        double v = fabs(a[i]);
        int j;
        for (j = (int)i - 1; j >= 0 && fabs(a[sorted[j]]) > v; j--) sorted[j + 1] = sorted[j];
        sorted[j + 1] = i;
    }
}

/* to_remove, dynamic variant (cheb_prec.h:57-85) */
static int to_remove(double c, double v, double beta, double* s) {
    double t = -2 * v / (c * (beta - 1));
    if (*s + t > 1) return 0;
    *s += t;
    return 1;
}

/* construct_sparse_lu (modules/prec/misc/sparse_lu.cc:6-133) */
static void construct_sparse_lu(const or_mat* A, const uint32_t* map, const uint32_t* rmap,
                                uint32_t Md, uint32_t M, ltype_t* lt, double beta, const double* aux,
                                or_mat* nA, or_mat* U, or_mat* L) {
    const uint32_t N = A->nrow;
    uint32_t n = N - M - Md;
    int* jr = xmalloc((size_t)N * sizeof(int));
    for (uint32_t i = 0; i < N; i++) jr[i] = -1;
    sset jw = {0};
    vf64 w = {0};
    vu32 Lia = {0}, Lja = {0}, Uia = {0}, Uja = {0}, Aia = {0}, Aja = {0};
    vf64 La = {0}, Ua = {0}, Aa = {0};

    vu32_push(&Lia, 0); vu32_push(&Uia, 0); vu32_push(&Aia, 0);
    for (uint32_t i = 0; i < N - Md; i++) {
        for (uint32_t k = 0; k < jw.n; k++) jr[jw.p[k]] = -1;       /* :41-45 */
        jw.n = 0;
        w.n = 0;
        uint32_t arow = map[i];
        jr[i] = 0;                                                  /* :50-53 */
        sset_insert(&jw, i);
        vf64_push(&w, aux[map[i]]);
        uint32_t max_num = 1;
        lt->row = arow;
        for (uint32_t j_ = A->ia[arow] + 1; j_ < A->ia[arow + 1]; j_++) {   /* :56-74 */
            uint32_t j = A->ja[j_];
            if (lt_stat(lt, j_) == PRESENT) {
                uint32_t new_j = rmap[j];
                jr[new_j] = (int)max_num++;
                sset_insert(&jw, new_j);
                double z = A->a[j_] / beta;
                vf64_push(&w, z);
                w.p[0] -= z;
            }
        }
        uint32_t m = i < M ? i : M;                                 /* :76-105 */
        uint32_t k = jw.p[0];
        while (k < m) {
            double lik = w.p[jr[k]] / Ua.p[Uia.p[k]];
            vu32_push(&Lja, k);
            vf64_push(&La, lik);
            for (uint32_t j_ = Uia.p[k] + 1; j_ < Uia.p[k + 1]; j_++) {
                uint32_t j = Uja.p[j_];
                if (jr[j] != -1) {
                    w.p[jr[j]] -= lik * Ua.p[j_];
                } else {
                    jr[j] = (int)max_num++;
                    sset_insert(&jw, j);
                    vf64_push(&w, -lik * Ua.p[j_]);
                }
            }
            k = jw.p[upper_bound_u32(jw.p, 0, jw.n, k)];
        }
        vu32_push(&Lia, (uint32_t)Lja.n);
        if (i < M) {                                                /* :108-115 */
            for (uint32_t it = lower_bound_u32(jw.p, 0, jw.n, i); it < jw.n; it++) {
                vu32_push(&Uja, jw.p[it]);
                vf64_push(&Ua, w.p[jr[jw.p[it]]]);
            }
            vu32_push(&Uia, (uint32_t)Uja.n);
        } else {                                                    /* :116-131 */
            vu32_push(&Aja, i - M);
            vf64_push(&Aa, w.p[0]);
            for (uint32_t it = lower_bound_u32(jw.p, 0, jw.n, M); it < jw.n; it++)
                if (jw.p[it] != i) {
                    vu32_push(&Aja, jw.p[it] - M);
                    vf64_push(&Aa, w.p[jr[jw.p[it]]]);
                }
            vu32_push(&Aia, (uint32_t)Aja.n);
        }
    }
    mat_free(L); mat_free(U); mat_free(nA);
    L->nrow = L->ncol = N - Md; L->ia = Lia.p; L->ja = Lja.p; L->a = La.p;
    U->nrow = M; U->ncol = N - Md; U->ia = Uia.p; U->ja = Uja.p; U->a = Ua.p;
    nA->nrow = nA->ncol = n; nA->ia = Aia.p; nA->ja = Aja.p; nA->a = Aa.p;
    if (!L->ja) L->ja = xmalloc(4);
    if (!L->a) L->a = xmalloc(8);
    if (!U->ja) U->ja = xmalloc(4);
    if (!U->a) U->a = xmalloc(8);
    if (!nA->ja) nA->ja = xmalloc(4);
    if (!nA->a) nA->a = xmalloc(8);
    free(jr); free(jw.p); free(w.p);
}

static void lev_alloc_scratch_cheb(lev_t* li, uint32_t N, uint32_t n) {
    li->w = xcalloc(N, sizeof(double));
    if (n) {
        li->tmp = xcalloc(n, sizeof(double));
        li->x2 = xcalloc(n, sizeof(double));
        li->u0 = xcalloc(n, sizeof(double));
        li->u1 = xcalloc(n, sizeof(double));
    }
}

/* ---------------- uh_cheb ---------------- */

/* Prec::construct_permutation (modules/prec/cheb/cheb_mark.cc:8-92) */
static void cheb_construct_permutation(const or_prec* P, const or_mat* A, const ltype_t* lt_in,
                                       int* nlinks, uint32_t* Md, uint32_t* M, uint32_t* map) {
    uint32_t N = A->nrow;
    ltype_t lt;
    lt_copy(&lt, lt_in);
    uint32_t pind = 0;
    unsigned char* marked = xcalloc(N, 1);
    if (P->use_tails) {
        uint32_t i0 = NO_INDEX, i1 = NO_INDEX;
        for (uint32_t i = 0; i < N; i++)
            if (nlinks[i] == 1) {
                i0 = i;
                do {
                    map[pind++] = i0;
                    marked[i0] = 1;
                    lt.row = i0;
                    for (uint32_t j_ = A->ia[i0] + 1; j_ < A->ia[i0 + 1]; j_++)
                        if (lt_stat(&lt, j_) == PRESENT) { i1 = A->ja[j_]; break; }
                    lt.a[ltc_index(&lt, i0, i1)] = 0;                 /* remove(i0,i1) */
                    nlinks[i0] = -1;
                    nlinks[i1]--;
                    i0 = i1;
                } while (nlinks[i0] == 1);
                if (nlinks[i0] == 0) {
                    map[pind++] = i0;
                    marked[i0] = 1;
                    nlinks[i0] = -1;
                }
            }
    }
    *M = pind;
    uint32_t eind = N - 1;
    for (uint32_t i = 0; i < N; i++)
        if (!marked[i]) {
            if (nlinks[i] == 0) map[eind--] = i;
            else map[pind++] = i;
        }
    *Md = N - pind;
    free(marked);
    lt_free(&lt);
}

/* Prec::construct_level (modules/prec/cheb/cheb_constr.cc:38-161) */
static int cheb_construct_level(or_prec* P, uint32_t level, const or_mat* A, char* err, size_t errlen) {
    lev_t* li = &P->lv[level];
    li->N = A->nrow;
    li->nnz = mat_nnz(A);
    uint32_t N = li->N;
    int* nlinks = xcalloc(N, sizeof(int));
    ltype_t lt;
    ltc_init(&lt, A);
    li->aux = xmalloc((size_t)N * sizeof(double));
    double* aux = li->aux;
    uint32_t maxrow = 1;
    for (uint32_t i = 0; i < N; i++) if (A->ia[i + 1] - A->ia[i] > maxrow) maxrow = A->ia[i + 1] - A->ia[i];
    uint32_t* sorted = xmalloc((size_t)(maxrow + 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < N; i++) {                               /* :62-103 */
        uint32_t rstart = A->ia[i], rend = A->ia[i + 1];
        uint32_t nrz = rend - rstart - 1;
        nlinks[i] += (int)nrz;
        aux[i] = 0.0;
        for (uint32_t j_ = rstart; j_ < rend; j_++) aux[i] += A->a[j_];
        psort_abs(A->a + rstart + 1, nrz, sorted);
        double s = 0.;
        lt.row = i;
        for (uint32_t _j = 0; _j < nrz; _j++) {
            double v = A->a[rstart + 1 + sorted[_j]];
            if (to_remove(aux[i], v, li->beta, &s)) {
                uint32_t j_ = rstart + 1 + sorted[_j];
                uint32_t j = A->ja[j_];
                if (--lt.a[ltc_real(&lt, j_)] == 0) {             /* mark(j_) == REMOVED */
                    nlinks[i]--;
                    nlinks[j]--;
                }
            } else {
                break;
            }
        }
    }
    free(sorted);

    li->map = xmalloc((size_t)N * sizeof(uint32_t) + 4);
    cheb_construct_permutation(P, A, &lt, nlinks, &li->Md, &li->M, li->map);
    li->rmap = xmalloc((size_t)N * sizeof(uint32_t) + 4);
    for (uint32_t i = 0; i < N; i++) li->rmap[li->map[i]] = i;      /* :121-124 */

    lev_t* ln = &P->lv[level + 1];
    construct_sparse_lu(A, li->map, li->rmap, li->Md, li->M, &lt, li->beta, aux, &ln->A, &li->U, &li->L);

    li->dval = xmalloc((size_t)li->Md * sizeof(double) + 8);          /* :136-141 */
    for (uint32_t i = 0; i < li->Md; i++) li->dval[i] = 1. / aux[li->map[i + (N - li->Md)]];

    free(nlinks);
    lt_free(&lt);

    uint32_t n = ln->A.nrow;
    lev_alloc_scratch_cheb(li, N, n);
    if (n) {
        if (level + 1 >= P->nlevels) { set_err(err, errlen, "Too many levels: %u", level + 1); return 1; }
        return cheb_construct_level(P, level + 1, &ln->A, err, errlen);
    }
    P->nlevels = level + 1;
    return 0;
}

or_prec* or_cheb_build(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                       const double* sigmas, uint32_t nsig, const uint32_t* niters, uint32_t nnit,
                       int use_tails, int optimize_storage, char* err, size_t errlen) {
    if (!nsig || !nnit) { set_err(err, errlen, "empty sigmas/niters"); return NULL; }
    or_prec* P = xcalloc(1, sizeof(or_prec));
    P->kind = OR_PREC_UH_CHEB;
    P->use_tails = use_tails;
    P->nlevels = 30;                                                 /* cheb_constr.cc:175-187 */
    P->nalloc = 31;
    P->lv = xcalloc(P->nalloc, sizeof(lev_t));
    mat_from_arrays(&P->A0, n, ia, ja, a);
    for (uint32_t l = 0; l < P->nlevels; l++) {
        P->lv[l].alpha = 1.0;
        P->lv[l].beta = (l < nsig) ? sigmas[l] : sigmas[nsig - 1];
    }
    if (cheb_construct_level(P, 0, &P->A0, err, errlen)) { or_prec_free(P); return NULL; }
    for (uint32_t l = 0; l < P->nlevels; l++)                         /* :223-224 */
        P->lv[l].ncheb = (l >= nnit) ? niters[nnit - 1] : niters[l];
    lev_t* last = &P->lv[P->nlevels - 1];                            /* :227-236 */
    last->lmin = last->alpha;
    last->lmax = last->beta;
    for (int l = (int)P->nlevels - 2; l >= 0; l--) {
        lev_t* li = &P->lv[l];
        lev_t* ln = &P->lv[l + 1];
        double cs = or_cheb((ln->lmax + ln->lmin) / (ln->lmax - ln->lmin), li->ncheb);
        li->lmin = li->alpha * (1 - 1 / cs);
        li->lmax = li->beta * (1 + 1 / cs);
    }
    if (optimize_storage)                                            /* :239-244, 260-263 */
        for (uint32_t l = 1; l < P->nlevels; l++) or_optimize_storage(&P->lv[l].A);
    return P;
}

/* Prec::solve (modules/prec/cheb/cheb_solve.cc:13-123) */
static void cheb_solve(const or_prec* P, uint32_t level, const double* f, double* x) {
    const lev_t* li = &P->lv[level];
    uint32_t N = li->N, M = li->M, Md = li->Md;
    const uint32_t* map = li->map;
    const or_mat* U = &li->U;
    const or_mat* L = &li->L;
    double* w = li->w;
    for (uint32_t i = 0; i < N - Md; i++) {                          /* :33-38 */
        w[i] = f[map[i]];
        for (uint32_t j_ = L->ia[i]; j_ < L->ia[i + 1]; j_++) w[i] -= L->a[j_] * w[L->ja[j_]];
    }
    double* x2 = li->x2;
    if (level < P->nlevels - 1) {
        const lev_t* ln = &P->lv[level + 1];
        uint32_t n = ln->N;
        uint32_t ncheb = li->ncheb;
        double* F = xmalloc((size_t)n * sizeof(double));             /* :45-46 */
        memcpy(F, &w[M], (size_t)n * sizeof(double));
        const or_mat* A = &ln->A;
        double lmin = ln->lmin, lmax = ln->lmax;
        double eta = (lmax + lmin) / (lmax - lmin);
        double* tmp = li->tmp;
        memcpy(tmp, F, (size_t)n * sizeof(double));
        double alpha, beta;
        double* u1 = li->u1;
        alpha = 2 / (lmax + lmin);                                   /* :61-69 */
        if (ncheb > 1) { cheb_solve(P, level + 1, tmp, u1); or_dscal(alpha, u1, n); }
        else { cheb_solve(P, level + 1, tmp, x2); or_dscal(alpha, x2, n); }
        if (ncheb > 1) {                                             /* :72-82 */
            alpha = 4 / (lmax - lmin) * or_cheb(eta, 1) / or_cheb(eta, 2);
            beta = or_cheb(eta, 0) / or_cheb(eta, 2);
            or_sym_multiply(A, u1, tmp);
            or_daxpy(-1, F, tmp, n);
            cheb_solve(P, level + 1, tmp, x2);
            or_dscal(-alpha, x2, n);
            or_daxpy(1 + beta, u1, x2, n);
        }
        double* u0 = li->u0;
        /* the reference swaps buffers (u0<->x2, u1<->u0); we rotate pointers locally */
        double *pu0 = u0, *pu1 = u1, *px2 = x2;
        for (uint32_t i = 3; i <= ncheb; i++) {                      /* :86-101 */
            double* t0 = pu0; pu0 = px2; px2 = t0;                 /* u0.swap(x2) */
            double* t1 = pu1; pu1 = pu0; pu0 = t1;                 /* u1.swap(u0) */
            alpha = 4 / (lmax - lmin) * or_cheb(eta, i - 1) / or_cheb(eta, i);
            beta = or_cheb(eta, i - 2) / or_cheb(eta, i);
            or_sym_multiply(A, pu1, tmp);
            or_daxpy(-1, F, tmp, n);
            cheb_solve(P, level + 1, tmp, px2);
            for (uint32_t k = 0; k < n; k++) px2[k] = pu1[k] - alpha * px2[k] + beta * (pu1[k] - pu0[k]);
        }
        for (uint32_t i = 0; i < n; i++) x[map[i + M]] = px2[i];     /* :103-104 */
        free(F);
        /* keep the level's buffers consistent with the reference's swapped vectors */
        if (px2 != x2) {
            lev_t* mli = (lev_t*)li;
            mli->x2 = px2; mli->u1 = pu1; mli->u0 = pu0;
        }
    }
    for (int i = (int)M - 1; i >= 0; i--) {                          /* :108-115 */
        double z = w[i];
        for (uint32_t j_ = U->ia[i] + 1; j_ < U->ia[i + 1]; j_++) z -= U->a[j_] * x[map[U->ja[j_]]];
        x[map[i]] = z / U->a[U->ia[i]];
    }
    const double* dval = li->dval;                                   /* :118-122 */
    for (uint32_t i = 0; i < Md; i++) {
        uint32_t j = map[(N - Md) + i];
        x[j] = dval[i] * f[j];
    }
}

/* ---------------- multi_split ---------------- */

/* order_none (multi_split_mark.cc:8-25) */
static void ms_order_none(const or_mat* A, const int* nin, const int* nout, uint32_t* Md, uint32_t* M, uint32_t* map) {
    uint32_t N = A->nrow, pind = 0;
    *M = 0;
    uint32_t eind = N - 1;
    for (uint32_t i = 0; i < N; i++)
        if (nin[i] == 0 && nout[i] == 0) map[eind--] = i;
        else map[pind++] = i;
    *Md = N - pind;
}

/* order_original (multi_split_mark.cc:27-154) */
static int ms_order_original(const or_prec* P, const or_mat* A, const ltype_t* lt_in, int* nin, int* nout,
                             uint32_t* Md, uint32_t* M, uint32_t* map, char* err, size_t errlen) {
    uint32_t N = A->nrow;
    ltype_t lt;
    lt_copy(&lt, lt_in);
    uint32_t pind = 0;
    unsigned char* marked = xcalloc(N, 1);
    if (P->use_tails) {
        uint32_t i0 = NO_INDEX, i1 = NO_INDEX;
        for (uint32_t i = 0; i < N; i++)
            if (nout[i] == 1 && nin[i] <= 1) {
                i0 = i;
                do {
                    if (nout[i0] != 1) break;
                    uint32_t j_;
                    for (j_ = A->ia[i0] + 1; j_ < A->ia[i0 + 1]; j_++)
                        if (lt_stat(&lt, j_) == PRESENT) break;
                    if (j_ == A->ia[i0 + 1]) {
                        set_err(err, errlen, "Row %u: remaining link was not found", i0);
                        free(marked); lt_free(&lt); return 1;
                    }
                    i1 = A->ja[j_];
                    int aji_present = (lt_stat_ij(&lt, i1, i0) == PRESENT);
                    if (!(nin[i0] == 0 || (nin[i0] == 1 && aji_present))) break;
                    map[pind++] = i0;
                    marked[i0] = 1;
                    lt.c[j_] = 0;
                    if (aji_present) {
                        lt.c[sky_index(A, i1, i0)] = 0;
                        nin[i0]--;
                        nout[i1]--;
                    }
                    nout[i0] = -1;
                    nin[i1]--;
                    i0 = i1;
                } while (1);
                if (nout[i0] == 0 && nin[i0] == 0) {
                    map[pind++] = i0;
                    marked[i0] = 1;
                    nout[i0] = -1;
                }
            }
    }
    const int max_links = 2;                                         /* :127-139 */
    for (uint32_t i = 0; i < N; i++)
        if (marked[i] == 0 && (nin[i] <= max_links || nout[i] <= max_links)) {
            map[pind++] = i;
            nin[i] = 0;
            nout[i] = -1;
            marked[i] = 1;
        }
    *M = pind;
    uint32_t eind = N - 1;
    for (uint32_t i = 0; i < N; i++)
        if (marked[i] == 0) {
            if (nin[i] == 0 && nout[i] == 0) map[eind--] = i;
            else map[pind++] = i;
        }
    *Md = N - pind;
    free(marked);
    lt_free(&lt);
    return 0;
}

/* order_simple_1 (multi_split_mark.cc:156-191) */
static void ms_order_simple_1(const or_prec* P, const or_mat* A, const int* nin, const int* nout,
                              uint32_t* Md, uint32_t* M, uint32_t* map) {
    uint32_t N = A->nrow;
    const int MAX_USEFUL_DEGREE = 2;
    vu32 deg[4] = {{0}, {0}, {0}, {0}};
    for (uint32_t i = 0; i < N; i++)
        if (nout[i] <= MAX_USEFUL_DEGREE) vu32_push(&deg[nout[i]], i);
        else vu32_push(&deg[3], i);
    uint32_t pind = 0;
    for (size_t k = 0; k < deg[0].n; k++)
        if (nin[deg[0].p[k]] != 0) map[pind++] = deg[0].p[k];
    for (uint32_t i = 1; i < 4; i++) {
        for (size_t k = 0; k < deg[i].n; k++) map[pind++] = deg[i].p[k];
        if (i == P->degree_max) *M = pind;
    }
    uint32_t md = pind;
    for (size_t k = 0; k < deg[0].n; k++)
        if (nin[deg[0].p[k]] == 0) map[pind++] = deg[0].p[k];
    *Md = N - md;
    for (int i = 0; i < 4; i++) free(deg[i].p);
}

/* order_block (multi_split_mark.cc:193-252): hard-coded 60x220x85 mesh */
static int ms_order_block(const or_mat* A, const int* nin, const int* nout, uint32_t* Md, uint32_t* M,
                          uint32_t* map, char* err, size_t errlen) {
    const uint32_t nx = 60, ny = 220, nz = 85, BS = 5;
    uint32_t N = A->nrow;
    if (N != nx * ny * nz) { set_err(err, errlen, "Wrong mesh size: %u (expected 1122000)", N); return 1; }
    unsigned char* marked = xcalloc(N, 1);
    uint32_t pind = 0;
    for (uint32_t K = 0; K < nz; K += BS)
        for (uint32_t J = 0; J < ny; J += BS)
            for (uint32_t I = 0; I < nx; I += BS) {
                uint32_t kend = (K + BS < nz ? K + BS : nz) - 1;
                uint32_t jend = (J + BS < ny ? J + BS : ny) - 1;
                uint32_t iend = (I + BS < nx ? I + BS : nx) - 1;
                for (uint32_t k = K + 1; k < kend; k++)
                    for (uint32_t j = J + 1; j < jend; j++)
                        for (uint32_t i = I + 1; i < iend; i++) {
                            uint32_t id = k * ny * nx + j * nx + i;
                            if (nin[id] || nout[id]) { map[pind++] = id; marked[id] = 1; }
                        }
            }
    *M = pind;
    uint32_t eind = N - 1;
    for (uint32_t i = 0; i < N; i++)
        if (marked[i] == 0) {
            if (nin[i] == 0 && nout[i] == 0) map[eind--] = i;
            else map[pind++] = i;
        }
    *Md = N - pind;
    free(marked);
    return 0;
}

/* MultiSplitPrec::construct_level (multi_split_constr.cc:39-168), built without UMFPACK */
static int ms_construct_level(or_prec* P, uint32_t level, const or_mat* A, char* err, size_t errlen) {
    lev_t* li = &P->lv[level];
    li->N = A->nrow;
    li->nnz = mat_nnz(A);
    if (level == P->nlevels - 1) {                                   /* :46-49 */
        set_err(err, errlen, "Too many levels: %u", P->nlevels);
        return 1;
    }
    if (li->N <= P->coarse_n) {                                      /* :51-55 */
        P->nlevels = level + 1;
        P->coarse_n = li->N;
        return 0;
    }
    uint32_t N = li->N;
    int* nout = xcalloc(N, sizeof(int));
    int* nin = xcalloc(N, sizeof(int));
    ltype_t lt = {0};
    lt.kind = 2; lt.A = A; lt.row = NO_INDEX;
    lt.c = xmalloc((size_t)li->nnz + 1);
    memset(lt.c, 1, li->nnz);
    uint32_t maxrow = 1;
    for (uint32_t i = 0; i < N; i++) if (A->ia[i + 1] - A->ia[i] > maxrow) maxrow = A->ia[i + 1] - A->ia[i];
    uint32_t* sorted = xmalloc((size_t)(maxrow + 1) * sizeof(uint32_t));
    li->aux = xmalloc((size_t)N * sizeof(double) + 8);
    double* aux = li->aux;
    for (uint32_t i = 0; i < N; i++) {                               /* :70-113 */
        uint32_t rstart = A->ia[i], rend = A->ia[i + 1];
        uint32_t nrz = rend - rstart - 1;
        nout[i] = (int)nrz;
        aux[i] = A->a[rstart];
        for (uint32_t j_ = rstart + 1; j_ < rend; j_++) {
            nin[A->ja[j_]]++;
            aux[i] += A->a[j_];
        }
        psort_abs(A->a + rstart + 1, nrz, sorted);
        double s = li->q / (1 - li->q) * aux[i];
        for (uint32_t k = 0; k < nrz; k++) {
            uint32_t j_ = rstart + 1 + sorted[k];
            double aij = -A->a[j_];
            if (aij <= s) {
                uint32_t j = A->ja[j_];
                lt.c[j_] = 0;
                nout[i]--;
                nin[j]--;
                s -= aij;
            } else {
                break;
            }
        }
        aux[i] *= 1. / (1 - li->q);
    }
    free(sorted);
    li->map = xmalloc((size_t)N * sizeof(uint32_t) + 4);
    int rc = 0;
    switch (P->order) {
        case OR_ORDER_NONE: ms_order_none(A, nin, nout, &li->Md, &li->M, li->map); break;
        case OR_ORDER_ORIGINAL: rc = ms_order_original(P, A, &lt, nin, nout, &li->Md, &li->M, li->map, err, errlen); break;
        case OR_ORDER_BLOCK: rc = ms_order_block(A, nin, nout, &li->Md, &li->M, li->map, err, errlen); break;
        case OR_ORDER_SIMPLE_1: ms_order_simple_1(P, A, nin, nout, &li->Md, &li->M, li->map); break;
        default: set_err(err, errlen, "unknown order"); rc = 1;
    }
    free(nin); free(nout);
    if (rc) { lt_free(&lt); return rc; }
    li->rmap = xmalloc((size_t)N * sizeof(uint32_t) + 4);
    for (uint32_t i = 0; i < N; i++) li->rmap[li->map[i]] = i;
    lev_t* ln = &P->lv[level + 1];
    construct_sparse_lu(A, li->map, li->rmap, li->Md, li->M, &lt, 1.0, aux, &ln->A, &li->U, &li->L);
    lt_free(&lt);
    li->dval = xmalloc((size_t)li->Md * sizeof(double) + 8);         /* :147-151 */
    for (uint32_t i = 0; i < li->Md; i++) li->dval[i] = 1. / aux[li->map[i + (N - li->Md)]];
    li->u0 = xcalloc(N, sizeof(double));
    li->r = xcalloc(N, sizeof(double));
    li->w = xcalloc(N, sizeof(double));
    uint32_t n = ln->A.nrow;
    if (n) {
        li->x2 = xcalloc(n, sizeof(double));
        li->F = xcalloc(n, sizeof(double));
        return ms_construct_level(P, level + 1, &ln->A, err, errlen);
    }
    P->nlevels = level + 1;
    return 0;
}

or_prec* or_msplit_build(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                         const double* qs, uint32_t nq, const uint32_t* niters, uint32_t nnit,
                         const double* epss, uint32_t neps, int inner_dynamic,
                         int order, int use_tails, uint32_t degree_max, uint32_t max_levels,
                         uint32_t coarse_n, char* err, size_t errlen) {
    if (!nq || (!inner_dynamic && !nnit) || (inner_dynamic && !neps) || max_levels == 0) {
        set_err(err, errlen, "bad multi_split parameters");
        return NULL;
    }
    or_prec* P = xcalloc(1, sizeof(or_prec));
    P->kind = OR_PREC_MULTI_SPLIT;
    P->use_tails = use_tails;
    P->inner_dynamic = inner_dynamic;
    P->order = order;
    P->degree_max = degree_max;
    P->coarse_n = coarse_n;
    P->nlevels = max_levels;
    P->nalloc = max_levels + 1;
    P->lv = xcalloc(P->nalloc, sizeof(lev_t));
    mat_from_arrays(&P->A0, n, ia, ja, a);
    for (uint32_t l = 0; l < P->nlevels; l++)                        /* :255-266 */
        P->lv[l].q = (l >= nq) ? qs[nq - 1] : qs[l];
    P->lv[0].niter = 1;
    for (uint32_t l = 1; l < P->nlevels; l++) {
        if (!inner_dynamic) P->lv[l].niter = (l >= nnit) ? niters[nnit - 1] : niters[l];
        else P->lv[l].eps = (l >= neps) ? epss[neps - 1] : epss[l];
    }
    if (ms_construct_level(P, 0, &P->A0, err, errlen)) { or_prec_free(P); return NULL; }
    lev_t* last = &P->lv[P->nlevels - 1];                            /* :270-275 */
    if (last->M == 0) { last->niter = 0; last->eps = 0.0; last->q = 0.0; }
    return P;
}

static void ms_solve_diagonal(const lev_t* li, const double* f, double* x) {   /* :12-22 */
    for (uint32_t i = 0; i < li->Md; i++) {
        uint32_t j = li->map[(li->N - li->Md) + i];
        x[j] = li->dval[i] * f[j];
    }
}
static void ms_solve_L(const lev_t* li, const double* f, double* w) {          /* :25-36 */
    const or_mat* L = &li->L;
    for (uint32_t i = 0; i < li->N - li->Md; i++) {
        w[i] = f[li->map[i]];
        for (uint32_t j_ = L->ia[i]; j_ < L->ia[i + 1]; j_++) w[i] -= L->a[j_] * w[L->ja[j_]];
    }
}
static void ms_solve_U(const lev_t* li, const double* w, double* x) {          /* :39-53 */
    const or_mat* U = &li->U;
    for (int i = (int)li->M - 1; i >= 0; i--) {
        double z = w[i];
        for (uint32_t j_ = U->ia[i] + 1; j_ < U->ia[i + 1]; j_++) z -= U->a[j_] * x[li->map[U->ja[j_]]];
        x[li->map[i]] = z / U->a[U->ia[i]];
    }
}

/* MultiSplitPrec::solve (multi_split_solve.cc:59-157) */
static void ms_solve(const or_prec* P, uint32_t level, const double* f, double* x) {
    const lev_t* li = &P->lv[level];
    uint32_t N = li->N, M = li->M;
    uint32_t n = (level < P->nlevels - 1) ? P->lv[level + 1].N : 0;
    uint32_t niter = li->niter;
    const or_mat* A = level ? &li->A : &P->A0;
    double *r = li->r, *w = li->w, *u0 = li->u0, *x2 = li->x2, *F = li->F;
    memcpy(r, f, (size_t)N * sizeof(double));                        /* :101 */
    double init_norm = 0;
    if (li->eps) init_norm = or_dnrm2(r, N);
    ms_solve_L(li, r, w);
    if (level < P->nlevels - 1) {
        memcpy(F, &w[M], (size_t)n * sizeof(double));
        ms_solve(P, level + 1, F, x2);
        for (uint32_t i = 0; i < n; i++) x[li->map[i + M]] = x2[i];
    }
    ms_solve_U(li, w, x);
    ms_solve_diagonal(li, r, x);
    if (level == P->nlevels - 1) or_residual(A, f, x, r);           /* :112-115 */
    for (uint32_t i = 2; li->eps || i <= niter; i++) {               /* :119-149 */
        or_residual(A, f, x, r);
        if (li->eps) {
            double norm = or_dnrm2(r, N);
            if (norm / init_norm < li->eps) break;
        }
        ms_solve_L(li, r, w);
        if (level < P->nlevels - 1) {
            memcpy(F, &w[M], (size_t)n * sizeof(double));
            ms_solve(P, level + 1, F, x2);
            for (uint32_t k = 0; k < n; k++) u0[li->map[k + M]] = x2[k];
        }
        ms_solve_U(li, w, u0);
        ms_solve_diagonal(li, r, u0);
        or_daxpy(1., u0, x, N);
    }
}

/* ---------------- diag / id (modules/prec/diag/diag_prec.cc, id/id_prec.h) ---------------- */
or_prec* or_diag_build(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a) {
    or_prec* P = xcalloc(1, sizeof(or_prec));
    P->kind = OR_PREC_DIAG;
    P->n = n;
    P->d = xmalloc((size_t)n * sizeof(double) + 8);
    for (uint32_t i = 0; i < n; i++) P->d[i] = 1. / a[ia[i]];       /* diag_prec.cc:13-21 (Skyline: diagonal first) */
    (void)ja;
    return P;
}
or_prec* or_id_build(uint32_t n) {
    or_prec* P = xcalloc(1, sizeof(or_prec));
    P->kind = OR_PREC_ID;
    P->n = n;
    return P;
}

/* GSPrec(A): L keeps the diagonal, then the entries with column < i (gs_prec.cc:26-65) */
or_prec* or_gs_build(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a, char* err, size_t errlen) {
    or_prec* P = xcalloc(1, sizeof(or_prec));
    P->kind = OR_PREC_GS;
    P->n = n;
    or_mat* L = &P->Lgs;
    L->nrow = L->ncol = n;
    L->ia = xmalloc(((size_t)n + 1) * sizeof(uint32_t));
    L->ja = xmalloc(((size_t)ia[n] + 1) * sizeof(uint32_t));
    L->a = xmalloc(((size_t)ia[n] + 1) * sizeof(double));
    uint32_t ind = 0;
    L->ia[0] = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (!(a[ia[i]] > 0)) {
            set_err(err, errlen, "Something is very strange: (%u) has %g on diagonal, aborting...", i, a[ia[i]]);
            or_prec_free(P);
            return NULL;
        }
        L->ja[ind] = i;
        L->a[ind++] = a[ia[i]];
        for (uint32_t j = ia[i] + 1; j < ia[i + 1] && ja[j] <= i; j++) {   /* upper_bound(i) over the sorted rest */
            L->ja[ind] = ja[j];
            L->a[ind++] = a[j];
        }
        L->ia[i + 1] = ind;
    }
    return P;
}

/* GSPrec::solve (gs_prec.cc:9-23) */
static void gs_solve(const or_prec* P, const double* f, double* x) {
    const or_mat* L = &P->Lgs;
    for (uint32_t i = 0; i < P->n; i++) {
        double d = f[i];
        for (uint32_t j = L->ia[i] + 1; j < L->ia[i + 1]; j++) d -= L->a[j] * x[L->ja[j]];
        x[i] = d / L->a[L->ia[i]];
    }
}

or_prec* or_composite_make(or_prec* first, or_prec* second) {
    or_prec* P = xcalloc(1, sizeof(or_prec));
    P->kind = OR_PREC_COMPOSITE;
    P->n = first->kind == OR_PREC_GS || first->kind == OR_PREC_DIAG || first->kind == OR_PREC_ID ? first->n
                                                                                                  : first->A0.nrow;
    P->sub[0] = first;
    P->sub[1] = second;
    return P;
}

/* element (r, c) of a Skyline matrix (diagonal first, then ascending columns); 0 when absent */
static double sky_at(const or_mat* A, uint32_t r, uint32_t c) {
    if (A->ja[A->ia[r]] == c) return A->a[A->ia[r]];
    uint32_t lo = A->ia[r] + 1, hi = A->ia[r + 1];
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (A->ja[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    return (lo < A->ia[r + 1] && A->ja[lo] == c) ? A->a[lo] : 0.0;
}

/* SymPrec(A, cfg) (sym_prec.cc:17-38): Asym = A, then row by row each off-diagonal a_ij is raised to
 * Asym(j, i) when that is larger (rows before i are already symmetrised), so Asym = max(A, A^T) off
 * the diagonal; the uh_cheb hierarchy is built on Asym. */
or_prec* or_sym_build(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                      const double* sigmas, uint32_t nsig, const uint32_t* niters, uint32_t nnit,
                      int use_tails, int optimize_storage, char* err, size_t errlen) {
    or_prec* P = xcalloc(1, sizeof(or_prec));
    P->kind = OR_PREC_SYM;
    P->n = n;
    or_mat* S = &P->Asym;
    S->nrow = S->ncol = n;
    const size_t nnz = ia[n];
    S->ia = xmalloc(((size_t)n + 1) * sizeof(uint32_t));
    S->ja = xmalloc((nnz + 1) * sizeof(uint32_t));
    S->a = xmalloc((nnz + 1) * sizeof(double));
    memcpy(S->ia, ia, ((size_t)n + 1) * sizeof(uint32_t));
    memcpy(S->ja, ja, nnz * sizeof(uint32_t));
    memcpy(S->a, a, nnz * sizeof(double));
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t j = S->ia[i] + 1; j < S->ia[i + 1]; j++) {
            double d = sky_at(S, S->ja[j], i);
            if (d > S->a[j]) S->a[j] = d;
        }
    P->sub[0] = or_cheb_build(n, S->ia, S->ja, S->a, sigmas, nsig, niters, nnit, use_tails, optimize_storage, err,
                              errlen);
    if (!P->sub[0]) {
        or_prec_free(P);
        return NULL;
    }
    P->sym_eps = 1e-14;
    return P;
}

int or_sym_matrix(const or_prec* P, double* out) {
    if (P->kind != OR_PREC_SYM) return 1;
    memcpy(out, P->Asym.a, (size_t)P->Asym.ia[P->n] * sizeof(double));
    return 0;
}

int or_prec_apply(const or_prec* P, const double* f, double* x) {
    switch (P->kind) {
        case OR_PREC_UH_CHEB: cheb_solve(P, 0, f, x); return 0;
        case OR_PREC_MULTI_SPLIT: ms_solve(P, 0, f, x); return 0;
        case OR_PREC_DIAG: for (uint32_t i = 0; i < P->n; i++) x[i] = f[i] * P->d[i]; return 0;
        case OR_PREC_ID: memcpy(x, f, (size_t)P->n * sizeof(double)); return 0;
        case OR_PREC_GS: gs_solve(P, f, x); return 0;
        case OR_PREC_COMPOSITE: {                      /* composite_prec.cc:9-16 */
            double* x1 = xcalloc((size_t)P->n + 1, sizeof(double));
            int rc = or_prec_apply(P->sub[0], f, x);
            if (!rc) {
                memcpy(x1, x, (size_t)P->n * sizeof(double));
                rc = or_prec_apply(P->sub[1], x1, x);
            }
            free(x1);
            return rc;
        }
        case OR_PREC_SYM: {                            /* sym_prec.cc:8-15: PCGSolver(Asym, f, *B, x, 1e-14) */
            char e[256];
            or_stats st;
            return or_pcg(&P->Asym, f, P->sub[0], x, &st, P->sym_eps, 0, 0, e, sizeof(e));
        }
    }
    return 1;
}

uint32_t or_prec_nlevels(const or_prec* P) { return P->lv ? P->nlevels : 0; }

int or_prec_level_info(const or_prec* P, uint32_t l, uint32_t* ints9, double* dbls5) {
    if (!P->lv || l >= P->nlevels) return 1;
    const lev_t* li = &P->lv[l];
    uint32_t nnext = (l + 1 < P->nalloc) ? P->lv[l + 1].A.nrow : 0;
    uint32_t nnzn = (l + 1 < P->nalloc && P->lv[l + 1].A.ia) ? mat_nnz(&P->lv[l + 1].A) : 0;
    if (l + 1 >= P->nlevels) { nnext = 0; nnzn = 0; }
    ints9[0] = li->N; ints9[1] = li->nnz; ints9[2] = li->M; ints9[3] = li->Md;
    ints9[4] = (P->kind == OR_PREC_UH_CHEB) ? li->ncheb : li->niter;
    ints9[5] = li->L.ia ? mat_nnz(&li->L) : 0;
    ints9[6] = (li->U.ia && li->U.nrow) ? li->U.ia[li->U.nrow] : 0;
    ints9[7] = nnext; ints9[8] = nnzn;
    dbls5[0] = li->alpha;
    dbls5[1] = (P->kind == OR_PREC_UH_CHEB) ? li->beta : li->q;
    dbls5[2] = li->lmin; dbls5[3] = li->lmax; dbls5[4] = li->eps;
    return 0;
}

int or_prec_level_array(const or_prec* P, uint32_t l, int which, void* out) {
    if (!P->lv || l >= P->nlevels) return 1;
    const lev_t* li = &P->lv[l];
    const or_mat* An = &P->lv[l + 1].A;
    int has_next = (l + 1 < P->nlevels);
    uint32_t N = li->N, nL = li->N - li->Md;
    switch (which) {
        case 0: memcpy(out, li->map, (size_t)N * 4); return 0;
        case 1: if (li->L.ia) memcpy(out, li->L.ia, (size_t)(nL + 1) * 4); return 0;
        case 2: if (li->L.ia) memcpy(out, li->L.ja, (size_t)li->L.ia[nL] * 4); return 0;
        case 3: if (li->L.ia) memcpy(out, li->L.a, (size_t)li->L.ia[nL] * 8); return 0;
        case 4: if (li->U.ia) memcpy(out, li->U.ia, (size_t)(li->M + 1) * 4); return 0;
        case 5: if (li->U.ia) memcpy(out, li->U.ja, (size_t)li->U.ia[li->M] * 4); return 0;
        case 6: if (li->U.ia) memcpy(out, li->U.a, (size_t)li->U.ia[li->M] * 8); return 0;
        case 7: memcpy(out, li->dval, (size_t)li->Md * 8); return 0;
        case 8: if (has_next) memcpy(out, An->ia, (size_t)(An->nrow + 1) * 4); return 0;
        case 9: if (has_next) memcpy(out, An->ja, (size_t)mat_nnz(An) * 4); return 0;
        case 10: if (has_next) memcpy(out, An->a, (size_t)mat_nnz(An) * 8); return 0;
        case 11: memcpy(out, li->rmap, (size_t)N * 4); return 0;
        case 12: memcpy(out, li->aux, (size_t)N * 8); return 0;
    }
    return 1;
}

/* ------------------------------------------------------------------ */
/* solvers                                                             */
/* ------------------------------------------------------------------ */

/* PCGSolver (modules/solvers/pcg.cc:8-102), NORM_L2 */
int or_pcg(const or_mat* A, const double* b, const or_prec* B, double* x, or_stats* st,
           double eps, int x0_mode, uint32_t maxit, char* err, size_t errlen) {
    double gtime = now_s();
    uint32_t n = A->nrow;
    double* r = xmalloc((size_t)n * 8 + 8);
    double* z = xmalloc((size_t)n * 8 + 8);
    double* p = xmalloc((size_t)n * 8 + 8);
    double* Ap = xmalloc((size_t)n * 8 + 8);
    double alpha, beta, app, rz0, rz1, norm, init_norm;
    double mult = 0, inv = 0, delta;
    int rc = 0;
    if (x0_mode == 0) or_generate_x0(x, n);                          /* :24 */
    or_residual(A, b, x, r);
    norm = init_norm = or_dnrm2(r, n);
    or_prec_apply(B, r, z);
    memcpy(p, z, (size_t)n * 8);
    rz0 = or_ddot(r, z, n);
    uint32_t niter = 0;
    check_and_replace_eps(init_norm, &eps);                          /* :39 */
    while (norm / init_norm > eps) {                                 /* :42-76 */
        if (maxit && niter >= maxit) { set_err(err, errlen, "maxit %u reached", maxit); rc = 3; break; }
        delta = now_s();
        or_multiply(A, p, Ap);
        mult += now_s() - delta;
        app = or_ddot(Ap, p, n);
        if (app <= 0) {
            set_err(err, errlen, "<Ap, p> = %g: |Ap| = %g, |p| = %g", app, or_dnrm2(Ap, n), or_dnrm2(p, n));
            rc = 1;
            break;
        }
        alpha = rz0 / app;
        or_daxpy(alpha, p, x, n);
        or_daxpy(-alpha, Ap, r, n);
        norm = or_dnrm2(r, n);
        delta = now_s();
        or_prec_apply(B, r, z);
        inv += now_s() - delta;
        rz1 = or_ddot(r, z, n);
        beta = rz1 / rz0;
        or_dscal(beta, p, n);
        or_daxpy(1, z, p, n);
        rz0 = rz1;
        niter++;
    }
    if (rc == 0 && niter == 0) { set_err(err, errlen, "No need to iterate"); rc = 2; }   /* :94-96 */
    if (st) {
        st->niter = niter;
        st->t_sol = now_s() - gtime;
        st->t_mult = mult;
        st->t_prec = inv;
        st->init_norm = init_norm;
        st->final_norm = norm;
        st->eps_used = eps;
    }
    free(r); free(z); free(p); free(Ap);
    return rc;
}

/* update_x (modules/solvers/gmres.cc:144-168).  The back substitution loop stops at
 * j < n after n is decremented, exactly as the reference writes it. */
static void gmres_update_x(uint32_t nn, const or_prec* B, double** v, double** hh, double* rs,
                           double* x, double* z, uint32_t dim) {
    if (nn == 0) return;
    nn--;
    rs[nn] /= hh[nn][nn];
    for (uint32_t i = 1; i <= nn; i++) {
        int k = (int)(nn - i);
        for (uint32_t j = (uint32_t)k + 1; j < nn; j++) rs[k] -= hh[k][j] * rs[j];
        rs[k] /= hh[k][k];
    }
    double* tmp = xcalloc(dim, sizeof(double));
    for (uint32_t j = 0; j <= nn; j++) or_daxpy(rs[j], v[j], tmp, dim);
    or_prec_apply(B, tmp, z);
    or_daxpy(1.0, z, x, dim);
    free(tmp);
}

/* GMRESSolver (modules/solvers/gmres.cc:13-141), restart 30, classical Gram-Schmidt */
int or_gmres(const or_mat* A, const double* b, const or_prec* B, double* x, or_stats* st,
             double eps, int x0_mode, uint32_t maxit, char* err, size_t errlen) {
    double gtime = now_s();
    uint32_t n = A->nrow;
    const uint32_t m = 30;
    double s[30], c[30], rs[30];
    double* H = xcalloc((size_t)(m + 1) * m, sizeof(double));
    double* hh[30];
    for (uint32_t k = 0; k < m; k++) hh[k] = &H[k * m];
    double* v[30];
    for (uint32_t k = 0; k < m; k++) v[k] = xcalloc(n, sizeof(double));
    double* z = xcalloc(n, sizeof(double));
    double norm, init_norm, dtmp;
    int rc = 0;
    if (x0_mode == 0) or_generate_x0(x, n);
    or_residual(A, b, x, v[0]);
    norm = init_norm = or_dnrm2(v[0], n);
    check_and_replace_eps(init_norm, &eps);
    uint32_t niter = 0;
    while (1) {
        if (norm / init_norm <= eps) break;
        or_dscal(1. / norm, v[0], n);
        rs[0] = norm;
        uint32_t i = 0;
        for (; i < m - 1;) {
            or_prec_apply(B, v[i], z);
            or_multiply(A, z, v[i + 1]);
            for (uint32_t k = 0; k <= i; k++) hh[k][i] = or_ddot(v[i + 1], v[k], n);
            for (uint32_t k = 0; k <= i; k++) or_daxpy(-hh[k][i], v[k], v[i + 1], n);
            hh[i + 1][i] = or_dnrm2(v[i + 1], n);
            or_dscal(1. / hh[i + 1][i], v[i + 1], n);
            for (uint32_t k = 1; k <= i; k++) {
                uint32_t k1 = k - 1;
                dtmp = hh[k1][i];
                hh[k1][i] = c[k1] * dtmp + s[k1] * hh[k][i];
                hh[k][i] = -s[k1] * dtmp + c[k1] * hh[k][i];
            }
            dtmp = sqrt(hh[i][i] * hh[i][i] + hh[i + 1][i] * hh[i + 1][i]);
            dtmp = 1.0 / dtmp;
            c[i] = hh[i][i] * dtmp;
            s[i] = hh[i + 1][i] * dtmp;
            hh[i][i] = c[i] * hh[i][i] + s[i] * hh[i + 1][i];
            rs[i + 1] = -s[i] * rs[i];
            rs[i] *= c[i];
            norm = fabs(rs[i + 1]);
            niter++;
            i++;
            if (norm / init_norm <= eps) break;
        }
        gmres_update_x(i, B, v, hh, rs, x, z, n);
        if (norm / init_norm <= eps) break;
        if (maxit && niter >= maxit) { set_err(err, errlen, "maxit %u reached", maxit); rc = 3; break; }
        or_residual(A, b, x, v[0]);
    }
    if (st) {
        st->niter = niter;
        st->t_sol = now_s() - gtime;
        st->init_norm = init_norm;
        st->final_norm = norm;
        st->eps_used = eps;
    }
    for (uint32_t k = 0; k < m; k++) free(v[k]);
    free(z); free(H);
    return rc;
}

/* ChebSolver (modules/solvers/cheb.cc:9-100), NORM_L2; lmin/lmax are Prec::lmin()/lmax() */
int or_cheb_solver(const or_mat* A, double lmin, double lmax, const double* b, const or_prec* B, double* x,
                   or_stats* st, double eps, int x0_mode, uint32_t maxit, char* err, size_t errlen) {
    double gtime = now_s();
    uint32_t n = A->nrow;
    double* r = xmalloc((size_t)n * 8 + 8);
    double* u0 = xcalloc(n, 8);
    double* u1 = xcalloc(n, 8);
    double* xc = x;                                   /* the buffers rotate like the reference's swaps */
    double alpha, beta, norm, init_norm;
    double eta = (lmax + lmin) / (lmax - lmin);
    int rc = 0;
    if (x0_mode == 0) or_generate_x0(x, n);
    or_residual(A, b, x, r);
    norm = init_norm = or_dnrm2(r, n);
    uint32_t niter = 1;
    check_and_replace_eps(init_norm, &eps);
    alpha = 2 / (lmax + lmin);                        /* :37-47 */
    memcpy(u1, x, (size_t)n * 8);
    or_prec_apply(B, r, x);
    or_dscal(alpha, x, n);
    or_daxpy(1, u1, x, n);
    or_residual(A, b, x, r);
    norm = or_dnrm2(r, n);
    niter++;
    double *pu0 = u0, *pu1 = u1, *px = x;
    while (norm / init_norm > eps) {                  /* :50-78 */
        if (maxit && niter - 1 >= maxit) { set_err(err, errlen, "maxit %u reached", maxit); rc = 3; break; }
        double* t0 = pu0; pu0 = px; px = t0;          /* u0.swap(x) */
        double* t1 = pu1; pu1 = pu0; pu0 = t1;        /* u1.swap(u0) */
        alpha = 4 / (lmax - lmin) * or_cheb(eta, niter - 1) / or_cheb(eta, niter);
        beta = or_cheb(eta, niter - 2) / or_cheb(eta, niter);
        or_prec_apply(B, r, px);
        for (uint32_t k = 0; k < n; k++) px[k] = pu1[k] + alpha * px[k] + beta * (pu1[k] - pu0[k]);
        or_residual(A, b, px, r);
        norm = or_dnrm2(r, n);
        niter++;
    }
    niter--;
    if (px != xc) memcpy(xc, px, (size_t)n * 8);
    if (st) {
        st->niter = niter;
        st->t_sol = now_s() - gtime;
        st->init_norm = init_norm;
        st->final_norm = norm;
        st->eps_used = eps;
    }
    /* free the two buffers that are not the caller's x */
    double* bufs[3] = {u0, u1, xc};
    for (int i = 0; i < 3; i++) if (bufs[i] != xc) free(bufs[i]);
    free(r);
    return rc;
}

/* SimpleSolver (modules/solvers/simple_solver.cc:8-88), NORM_L2 */
int or_simple_solver(const or_mat* A, const double* b, const or_prec* B, double* x, or_stats* st,
                     double eps, int x0_mode, uint32_t maxit, char* err, size_t errlen) {
    double gtime = now_s();
    uint32_t n = A->nrow;
    double* r = xmalloc((size_t)n * 8 + 8);
    double* z = xmalloc((size_t)n * 8 + 8);
    double norm, init_norm;
    int rc = 0;
    if (x0_mode == 0) or_generate_x0(x, n);
    or_residual(A, b, x, r);
    norm = init_norm = or_dnrm2(r, n);
    uint32_t niter = 0, ninv = 0;
    check_and_replace_eps(init_norm, &eps);
    while (norm / init_norm > eps) {
        if (maxit && niter >= maxit) { set_err(err, errlen, "maxit %u reached", maxit); rc = 3; break; }
        or_prec_apply(B, r, z);
        if (niter) ninv++;                            /* the first apply counts as construction (:45-52) */
        or_daxpy(1., z, x, n);
        or_residual(A, b, x, r);
        norm = or_dnrm2(r, n);
        niter++;
    }
    if (rc == 0 && !ninv) { set_err(err, errlen, "No need to iterate"); rc = 2; }   /* :65-82 */
    if (st) {
        st->niter = niter;
        st->t_sol = now_s() - gtime;
        st->init_norm = init_norm;
        st->final_norm = norm;
        st->eps_used = eps;
    }
    free(r); free(z);
    return rc;
}

/* ------------------------------------------------------------------ */
/* raw-array wrappers                                                  */
/* ------------------------------------------------------------------ */
static void mat_view(or_mat* A, uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a) {
    memset(A, 0, sizeof(*A));
    A->nrow = A->ncol = n;
    A->ia = (uint32_t*)ia; A->ja = (uint32_t*)ja; A->a = (double*)a;
}

static void stats_out(const or_stats* st, double* out6) {
    if (!out6) return;
    out6[0] = st->niter; out6[1] = st->t_sol; out6[2] = st->init_norm;
    out6[3] = st->final_norm; out6[4] = st->eps_used; out6[5] = st->t_prec;
}

int or_pcg_csr(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
               const double* b, const or_prec* B, double* x, double* stats6,
               double eps, int x0_mode, uint32_t maxit, char* err, size_t errlen) {
    or_mat A; or_stats st = {0};
    mat_view(&A, n, ia, ja, a);
    int rc = or_pcg(&A, b, B, x, &st, eps, x0_mode, maxit, err, errlen);
    stats_out(&st, stats6);
    return rc;
}

int or_gmres_csr(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                 const double* b, const or_prec* B, double* x, double* stats6,
                 double eps, int x0_mode, uint32_t maxit, char* err, size_t errlen) {
    or_mat A; or_stats st = {0};
    mat_view(&A, n, ia, ja, a);
    int rc = or_gmres(&A, b, B, x, &st, eps, x0_mode, maxit, err, errlen);
    stats_out(&st, stats6);
    return rc;
}

int or_cheb_solver_csr(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a, double lmin,
                       double lmax, const double* b, const or_prec* B, double* x, double* stats6, double eps,
                       int x0_mode, uint32_t maxit, char* err, size_t errlen) {
    or_mat A; or_stats st = {0};
    mat_view(&A, n, ia, ja, a);
    int rc = or_cheb_solver(&A, lmin, lmax, b, B, x, &st, eps, x0_mode, maxit, err, errlen);
    stats_out(&st, stats6);
    return rc;
}

int or_simple_solver_csr(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a, const double* b,
                         const or_prec* B, double* x, double* stats6, double eps, int x0_mode, uint32_t maxit,
                         char* err, size_t errlen) {
    or_mat A; or_stats st = {0};
    mat_view(&A, n, ia, ja, a);
    int rc = or_simple_solver(&A, b, B, x, &st, eps, x0_mode, maxit, err, errlen);
    stats_out(&st, stats6);
    return rc;
}

void or_spmv_csr(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                 const double* x, double* y) {
    or_mat A;
    mat_view(&A, n, ia, ja, a);
    or_multiply(&A, x, y);
}

void or_sym_spmv_csr(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                     const double* x, double* y) {
    or_mat A;
    mat_view(&A, n, ia, ja, a);
    or_sym_multiply(&A, x, y);
}

/*
 * CPU oracle for the ml_prec iterative-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * serial algorithms, against which the HIP product path is tested.  Only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark load
 * it; the product library never links or calls it.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the reference repository root).  Floating-point operations are kept in the
 * reference's order and the file is compiled without FMA contraction, so
 * results are the ones the reference's -O2 x86-64 build computes.
 *
 * Parity status of each part (how it is pinned) is listed in the header
 * comment of oracle.c.
 */
#ifndef ML_PREC_ORACLE_H
#define ML_PREC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Skyline CSR: diagonal first, then ascending columns (modules/matrix/matrix.h:269-291). */
typedef struct {
    uint32_t nrow, ncol;
    uint32_t *ia, *ja;
    double *a;
    size_t cap;                 /* capacity of ja/a while building */
    /* lower-half storage created by optimize_storage (modules/matrix/skyline.cc:86-113) */
    uint32_t *iasym, *jasym;
    double *asym;
} or_mat;

typedef struct {
    double t_const, t_sol, t_mult, t_prec;
    uint32_t niter;
    double init_norm, final_norm;     /* recurrence norms the stop test used */
    double eps_used;
} or_stats;

/* ---- linear algebra (modules/matrix/csrmatrix.cc, skyline.cc, modules/vector/blas.cc) ---- */
void   or_multiply(const or_mat* A, const double* v, double* res);              /* 'o' */
void   or_sym_multiply(const or_mat* A, const double* v, double* res);          /* 's' */
void   or_residual(const or_mat* A, const double* b, const double* x, double* r);
double or_ddot(const double* x, const double* y, uint32_t n);
double or_dnrm2(const double* x, uint32_t n);
void   or_daxpy(double alpha, const double* x, double* y, uint32_t n);
void   or_dscal(double alpha, double* x, uint32_t n);
void   or_generate_x0(double* x, uint32_t n);
int    or_optimize_storage(or_mat* A);
double or_cheb(double x, uint32_t k);

/* ---- preconditioners ---- */
typedef struct or_prec or_prec;   /* opaque: uh_cheb or multi_split hierarchy */

enum { OR_PREC_UH_CHEB = 1, OR_PREC_MULTI_SPLIT = 2, OR_PREC_DIAG = 3, OR_PREC_ID = 4, OR_PREC_GS = 5,
       OR_PREC_COMPOSITE = 6, OR_PREC_SYM = 7 };
enum { OR_ORDER_NONE = 0, OR_ORDER_ORIGINAL = 1, OR_ORDER_BLOCK = 2, OR_ORDER_SIMPLE_1 = 3 };

/* uh_cheb Prec(A, cfg) (modules/prec/cheb/cheb_constr.cc:163-248).  Returns NULL and fills
 * err on failure (the reference throws). */
or_prec* or_cheb_build(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                       const double* sigmas, uint32_t nsig, const uint32_t* niters, uint32_t nnit,
                       int use_tails, int optimize_storage, char* err, size_t errlen);

/* MultiSplitPrec(A, cfg) with the config-file values already parsed
 * (modules/prec/multi_split/multi_split_constr.cc:170-299). inner_dynamic selects level_eps. */
or_prec* or_msplit_build(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                         const double* qs, uint32_t nq, const uint32_t* niters, uint32_t nnit,
                         const double* epss, uint32_t neps, int inner_dynamic,
                         int order, int use_tails, uint32_t degree_max, uint32_t max_levels,
                         uint32_t coarse_n, char* err, size_t errlen);

or_prec* or_diag_build(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a);
or_prec* or_id_build(uint32_t n);
/* GSPrec(A) (modules/prec/gs/gs_prec.cc:26-65): forward Gauss-Seidel on the lower triangle. */
or_prec* or_gs_build(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a, char* err, size_t errlen);
/* CompositePrec (modules/prec/composite/composite_prec.cc:18-36): x = B2(B1(f)); takes ownership of
 * both (the reference builds B1 = GSPrec(A), B2 = MultiSplitPrec(A, cfg)). */
or_prec* or_composite_make(or_prec* first, or_prec* second);
/* SymPrec(A, cfg) (modules/prec/sym/sym_prec.cc:17-38): uh_cheb built on max(A, A^T) off the diagonal;
 * solve runs PCG on that matrix to 1e-14 from generate_x0. */
or_prec* or_sym_build(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                      const double* sigmas, uint32_t nsig, const uint32_t* niters, uint32_t nnit,
                      int use_tails, int optimize_storage, char* err, size_t errlen);
/* the symmetrised matrix of a SymPrec (same pattern as A): copies nnz values into out */
int or_sym_matrix(const or_prec* P, double* out);

void     or_prec_free(or_prec* P);
int      or_prec_apply(const or_prec* P, const double* f, double* x);   /* PrecBase::solve */
uint32_t or_prec_nlevels(const or_prec* P);
/* ints: N, nnz, M, Md, ncheb|niter, nnz(L), nnz(U), n_next, nnz(A_next)
 * dbls: alpha, beta|q, lmin, lmax, eps */
int      or_prec_level_info(const or_prec* P, uint32_t l, uint32_t* ints9, double* dbls5);
/* which: 0 map, 1 L.ia, 2 L.ja, 3 L.a, 4 U.ia, 5 U.ja, 6 U.a, 7 dval,
 *        8 A_next.ia, 9 A_next.ja, 10 A_next.a, 11 rmap, 12 aux.  Copies into out. */
int      or_prec_level_array(const or_prec* P, uint32_t l, int which, void* out);

/* ---- solvers (modules/solvers/pcg.cc, gmres.cc) ----
 * x0_mode 0: x0 from generate_x0 (the reference's behaviour); 1: x0 taken from x on entry.
 * Return 0 on success, nonzero on the reference's exceptions (message in err). */
int or_pcg(const or_mat* A, const double* b, const or_prec* B, double* x, or_stats* st,
           double eps, int x0_mode, uint32_t maxit, char* err, size_t errlen);
int or_gmres(const or_mat* A, const double* b, const or_prec* B, double* x, or_stats* st,
             double eps, int x0_mode, uint32_t maxit, char* err, size_t errlen);

int or_cheb_solver(const or_mat* A, double lmin, double lmax, const double* b, const or_prec* B, double* x,
                   or_stats* st, double eps, int x0_mode, uint32_t maxit, char* err, size_t errlen);
int or_simple_solver(const or_mat* A, const double* b, const or_prec* B, double* x, or_stats* st,
                     double eps, int x0_mode, uint32_t maxit, char* err, size_t errlen);

/* convenience wrappers taking raw CSR arrays (used through ctypes) */
int or_cheb_solver_csr(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a, double lmin,
                       double lmax, const double* b, const or_prec* B, double* x, double* stats_out6, double eps,
                       int x0_mode, uint32_t maxit, char* err, size_t errlen);
int or_simple_solver_csr(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a, const double* b,
                         const or_prec* B, double* x, double* stats_out6, double eps, int x0_mode,
                         uint32_t maxit, char* err, size_t errlen);
int or_pcg_csr(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
               const double* b, const or_prec* B, double* x, double* stats_out6,
               double eps, int x0_mode, uint32_t maxit, char* err, size_t errlen);
int or_gmres_csr(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                 const double* b, const or_prec* B, double* x, double* stats_out6,
                 double eps, int x0_mode, uint32_t maxit, char* err, size_t errlen);
void or_spmv_csr(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                 const double* x, double* y);
void or_sym_spmv_csr(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                     const double* x, double* y);

#ifdef __cplusplus
}
#endif
#endif

/*
 * mlp.h -- C ABI of the MI355X (gfx950) implementation of ml_prec's
 * iterative-solve path: CSR SpMV, the multilevel uh_cheb / multi_split
 * preconditioner apply, PCG and GMRES with their BLAS-1 reductions.
 *
 * Conventions
 *   - every entry point returns int status (MLP_OK = 0) and, on failure,
 *     leaves a message for mlp_last_error() (thread-local); no C++ exception
 *     crosses this boundary;
 *   - pointers named d_* are device pointers (HBM); pointers without the
 *     prefix are host pointers;
 *   - all work of a context is ordered on the context's HIP stream; a context
 *     is not thread-safe, like the reference's solvers (whose scratch vectors
 *     are `mutable` members, modules/prec/cheb/cheb_prec.h:35-36);
 *   - matrices use the reference's Skyline CSR layout: uint32 ia[n+1],
 *     uint32 ja[nnz], double a[nnz], diagonal first in every row, then
 *     ascending columns (modules/matrix/matrix.h:187-291).
 *
 * Each declaration cites the reference interface it replaces.
 */
#ifndef ML_PREC_MLP_H
#define ML_PREC_MLP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference throws Exception for the same cases) ---- */
#define MLP_OK              0
#define MLP_ERR_INVALID     1   /* bad argument / size mismatch (ASSERT_SIZE, include/exception.h) */
#define MLP_ERR_HIP         2   /* HIP runtime failure */
#define MLP_ERR_NOT_SPD     3   /* "<Ap, p> <= 0"            modules/solvers/pcg.cc:51-56 */
#define MLP_ERR_NO_ITER     4   /* "No need to iterate"      modules/solvers/pcg.cc:94-96 */
#define MLP_ERR_MAXIT       5   /* iteration cap reached (no counterpart: the reference loops on) */
#define MLP_ERR_CONSTRUCT   6   /* hierarchy construction ("Too many levels", ...) */
#define MLP_ERR_CONFIG      7   /* prec config file errors   multi_split_constr.cc:204-228 */
#define MLP_ERR_NO_CONV     8   /* flexible PCG: the FP64 true residual missed eps after the restarts
                                   (no counterpart: the reference has no flexible PCG) */

/* ---- precision of a device preconditioner hierarchy ---- */
#define MLP_FP64 0
#define MLP_FP32 1
/* FP32 hierarchy whose level 0 keeps the vectors of its Chebyshev iteration on level 1 (F, u1, a1 u1,
 * x2; cheb_solve.cc:45-101) in BF16; the kernels compute in FP32 and round to BF16 only where they store
 * those vectors.  uh_cheb with two or more levels and ncheb = 2 on level 0. */
#define MLP_FP32_BF16 2

/* ---- preconditioner kinds ---- */
#define MLP_PREC_UH_CHEB     1  /* Prec, "uh_cheb"           modules/prec/cheb/cheb_prec.h:13 */
#define MLP_PREC_MULTI_SPLIT 2  /* MultiSplitPrec            modules/prec/multi_split/multi_split_prec.h:15 */
#define MLP_PREC_DIAG        3  /* DiagPrec                  modules/prec/diag/diag_prec.cc:7-23 */
#define MLP_PREC_ID          4  /* IdPrec                    modules/prec/id/id_prec.h:6-19 */
#define MLP_PREC_GS          5  /* GSPrec                    modules/prec/gs/gs_prec.cc:9-65 */
#define MLP_PREC_COMPOSITE   6  /* CompositePrec             modules/prec/composite/composite_prec.cc:9-36 */
#define MLP_PREC_SYM         7  /* SymPrec ("sym_split")     modules/prec/sym/sym_prec.cc:8-38 */

/* ---- multi_split elimination orders (ConfigFile key elim_order) ---- */
#define MLP_ORDER_NONE     0
#define MLP_ORDER_ORIGINAL 1
#define MLP_ORDER_BLOCK    2
#define MLP_ORDER_SIMPLE_1 3

typedef struct mlp_ctx_s*  mlp_ctx;    /* device, stream, scratch */
typedef struct mlp_mat_s*  mlp_mat;    /* a matrix resident in HBM */
typedef struct mlp_hier_s* mlp_hier;   /* a hierarchy built on the host */
typedef struct mlp_prec_s* mlp_prec;   /* a hierarchy resident in HBM */

const char* mlp_last_error(void);

/* ---- context ---- */
int mlp_ctx_create(int device, mlp_ctx* out);
int mlp_ctx_destroy(mlp_ctx ctx);
int mlp_ctx_sync(mlp_ctx ctx);
int mlp_ctx_stream(mlp_ctx ctx, void** hip_stream_out);

/* ---- device memory helpers (for callers without their own allocator) ---- */
/* Device memory on ctx's device.  Blocks the library releases (these, matrices', preconditioners' and
 * contexts' own) are kept per device and size class, up to 8 GiB per device, and handed out again; a release
 * first waits for the device, as hipFree does.  mlp_dev_free takes blocks of mlp_dev_alloc only. */
int mlp_dev_alloc(mlp_ctx ctx, size_t bytes, void** d_out);
int mlp_dev_free(mlp_ctx ctx, void* d_ptr);
int mlp_copy_h2d(mlp_ctx ctx, void* d_dst, const void* src, size_t bytes);
int mlp_copy_d2h(mlp_ctx ctx, void* dst, const void* d_src, size_t bytes);

/* ---- matrices ----
 * mlp_csr_upload replaces holding a CSRMatrix/SkylineMatrix for the solve
 * (modules/matrix/matrix.h:187-291); the device copy is a sliced-ELL layout
 * (64-row slices, column-major within a slice) with the same per-row order. */
int mlp_csr_upload(mlp_ctx ctx, uint32_t n, const uint32_t* ia, const uint32_t* ja,
                   const double* a, mlp_mat* out);
int mlp_mat_destroy(mlp_mat m);
int mlp_mat_info(mlp_mat m, uint32_t* n, uint32_t* nnz, uint64_t* device_bytes);
/* y = A x: multiply(A, v, res, 'o')   modules/matrix/csrmatrix.cc:260-291 */
int mlp_spmv(mlp_ctx ctx, mlp_mat A, const double* d_x, double* d_y);
/* r = b - A x: residual(A, b, x, r)   modules/matrix/csrmatrix.cc:293-300 */
int mlp_residual(mlp_ctx ctx, mlp_mat A, const double* d_b, const double* d_x, double* d_r);

/* ---- BLAS-1: modules/vector/blas.cc:57-82 (ddot, dnrm2, daxpy, dscal) ----
 * daxpy/dscal keep the reference's alpha == 1 shortcuts (blas.cc:43-49, 77-82). */
int mlp_ddot(mlp_ctx ctx, uint32_t n, const double* d_x, const double* d_y, double* out);
int mlp_dnrm2(mlp_ctx ctx, uint32_t n, const double* d_x, double* out);
int mlp_daxpy(mlp_ctx ctx, uint32_t n, double alpha, const double* d_x, double* d_y);
int mlp_dscal(mlp_ctx ctx, uint32_t n, double alpha, double* d_x);

/* ---- host construction of a hierarchy ----
 * Prec(const SkylineMatrix&, const Config&)       modules/prec/cheb/cheb_constr.cc:163-248 */
int mlp_hier_build_uh_cheb(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                           const double* sigmas, uint32_t nsigmas,
                           const uint32_t* niters, uint32_t nniters,
                           int use_tails, mlp_hier* out);
/* The same construction on ctx's device, every level as kernels with the level matrices kept in device
 * memory: row sums, the magnitude order of the links (psort, cheb_constr.cc:26-36), the budget marks
 * (cheb_constr.cc:62-103, to_remove cheb_prec.h:71-99), the two-sided counters (cheb_misc.h:45-96), the
 * tail order (construct_permutation, cheb_mark.cc:8-92), the sparse LU and the next level
 * (construct_sparse_lu, sparse_lu.cc:6-133).  The level arrays are copied to the host hierarchy, which is
 * identical to mlp_hier_build_uh_cheb's. */
int mlp_hier_build_uh_cheb_dev(mlp_ctx ctx, uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                               const double* sigmas, uint32_t nsigmas,
                               const uint32_t* niters, uint32_t nniters,
                               int use_tails, mlp_hier* out);
/* The device construction reading level 0 from the matrix handle: from device memory when A was assembled
 * there (mlp_spe_matrix_dev), else from A's host copy.  Prec(A, cfg), cheb_constr.cc:163-248. */
int mlp_hier_build_uh_cheb_mat(mlp_ctx ctx, mlp_mat A, const double* sigmas, uint32_t nsigmas,
                               const uint32_t* niters, uint32_t nniters, int use_tails, mlp_hier* out);

typedef struct {
    const double* q;        uint32_t nq;          /* key q            */
    const uint32_t* niters; uint32_t nniters;     /* key level_niter  */
    const double* eps;      uint32_t neps;        /* key level_eps    */
    int inner_dynamic;                            /* key inner_iter   */
    int order;                                    /* key elim_order   */
    int use_tails;                                /* Config.use_tails */
    uint32_t degree_max;                          /* key degree_max   */
    uint32_t max_levels;                          /* key max_levels   */
    uint32_t coarse_n;                            /* key coarse_n     */
} mlp_msplit_params;

/* MultiSplitPrec(const SkylineMatrix&, const Config&)  multi_split_constr.cc:170-299.
 * config_path != NULL: read the prec config file with the reference's keys and
 * defaults (modules/config_file/config_file.h:54-116); params then only supply
 * use_tails.  A missing file keeps params, as the reference keeps its CLI values. */
int mlp_hier_build_multi_split(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                               const char* config_path, const mlp_msplit_params* params,
                               mlp_hier* out);
int mlp_hier_destroy(mlp_hier h);
int mlp_hier_nlevels(mlp_hier h, uint32_t* out);

typedef struct {
    uint32_t N, nnz, M, Md;        /* level sizes (Level::N, nnz, M, Md) */
    uint32_t nit;                  /* ncheb (uh_cheb) or niter (multi_split) */
    uint32_t nnz_L, nnz_U;
    uint32_t n_next, nnz_next;     /* the Schur complement A_{l+1} */
    double alpha, beta_or_q, lmin, lmax, eps;
} mlp_level_info;
int mlp_hier_level_info(mlp_hier h, uint32_t level, mlp_level_info* out);
/* which: 0 map, 1 L.ia, 2 L.ja, 3 L.a, 4 U.ia, 5 U.ja, 6 U.a, 7 dval,
 *        8 A_next.ia, 9 A_next.ja, 10 A_next.a, 11 rmap, 12 aux */
int mlp_hier_level_array(mlp_hier h, uint32_t level, int which, void* out);

/* ---- device preconditioners ----
 * A hierarchy as plain per-level arrays: what Prec::Level / MultiSplitPrec::Level
 * hold (cheb_prec.h:17-40, multi_split_prec.h:19-44).  A_next is levels[l+1].A. */
typedef struct {
    uint32_t N, M, Md;
    uint32_t nit;                            /* ncheb or niter */
    double lmin, lmax;                       /* this level's spectral bounds (uh_cheb) */
    double eps;                              /* multi_split level_eps (0 = fixed) */
    const uint32_t* map;                     /* N */
    const uint32_t *L_ia, *L_ja; const double* L_a;   /* N-Md rows */
    const uint32_t *U_ia, *U_ja; const double* U_a;   /* M rows, pivot first */
    const double* dval;                      /* Md */
    const uint32_t *A_ia, *A_ja; const double* A_a;   /* this level's matrix (unused for level 0) */
} mlp_level_desc;

int mlp_prec_upload(mlp_ctx ctx, int kind, uint32_t nlevels, const mlp_level_desc* levels,
                    mlp_mat A0, int precision, mlp_prec* out);
int mlp_prec_from_hier(mlp_ctx ctx, mlp_hier h, mlp_mat A0, int precision, mlp_prec* out);
/* Autotuned launch schedule (no counterpart in the reference, which has one CPU schedule): the hierarchy's
 * device programs built with each candidate schedule -- the default (uh_cheb: the longest level chain whose
 * merged components fit one-wavefront blocks of 64 nodes), the chain in two-wavefront blocks of up to 128
 * nodes (also what MLP_CHAIN_BS=128 asks of every build), level kernels with and without the fused last
 * pair, every level program in 64- or 256-thread blocks -- each timed over `reps` applies on ctx's stream
 * (HIP events, after one untimed apply), the fastest kept.  FP64 applies are bitwise the same under every
 * schedule.  report (optional) receives each distinct candidate's schedule and time (at most 8). */
#define MLP_SCHED_CHAIN 1u        /* the level chain runs the last levels */
#define MLP_SCHED_PAIR  2u        /* the last two levels run fused */
#define MLP_SCHED_WIDE_CHAIN 4u   /* the chain runs in two-wavefront blocks (components of up to 128 nodes) */
typedef struct {
    uint32_t ncand;           /* distinct candidates timed (at most 8 reported) */
    uint32_t chosen;          /* index of the one kept */
    uint32_t sched[8];        /* MLP_SCHED_* bits | forced level block size << 8 (0: chosen per level) */
    double us[8];             /* microseconds per apply */
} mlp_tune_report;
int mlp_prec_from_hier_tuned(mlp_ctx ctx, mlp_hier h, mlp_mat A0, int precision, uint32_t reps,
                             mlp_tune_report* report, mlp_prec* out);
int mlp_prec_simple(mlp_ctx ctx, int kind, mlp_mat A0, mlp_prec* out);   /* DIAG or ID */
/* GSPrec(const SkylineMatrix& A)  gs_prec.cc:26-65: one forward Gauss-Seidel sweep (lower triangle
 * with the diagonal; a non-positive diagonal is an error as in the reference). */
int mlp_prec_gs(mlp_ctx ctx, mlp_mat A, mlp_prec* out);
/* CompositePrec(A, cfg)  composite_prec.cc:18-36: GSPrec(A) followed by MultiSplitPrec(A, cfg); ms is the
 * multi_split hierarchy built on A (mlp_hier_build_multi_split with the same matrix). */
int mlp_prec_composite(mlp_ctx ctx, mlp_mat A, mlp_hier ms, mlp_prec* out);
/* SymPrec(A, cfg)  sym_prec.cc:17-38: uh_cheb (sigmas, niters, use_tails as for mlp_hier_build_uh_cheb)
 * built on max(A, A^T) off the diagonal; each apply is a PCG solve of that matrix to 1e-14 from
 * generate_x0 (sym_prec.cc:8-15).  Its apply synchronises with the host (no graph capture). */
int mlp_prec_sym(mlp_ctx ctx, mlp_mat A, const double* sigmas, uint32_t nsigmas, const uint32_t* niters,
                 uint32_t nniters, int use_tails, mlp_prec* out);
/* A copy of a uh_cheb or multi_split preconditioner on context ctx (same device) that shares src's device data
 * (level programs and matrices, which stay alive while any copy does) and has level vectors of its own, so that
 * solves on different contexts can apply one hierarchy at the same time (no counterpart in the reference, whose
 * solve() mutates per-level scratch and is not reentrant, cheb_prec.h:35-36). */
int mlp_prec_clone(mlp_ctx ctx, mlp_prec src, mlp_prec* out);
/* SymPrec::set_eps  sym_prec.h:26-28: tolerance of the inner PCG (default 1e-14) */
int mlp_prec_sym_set_eps(mlp_prec p, double eps);
int mlp_prec_destroy(mlp_prec p);
/* x = B f: PrecBase::solve(f, x)   modules/prec/prec_base.h:8-12,
 * Prec::solve cheb_solve.cc:9-123, MultiSplitPrec::solve multi_split_solve.cc:59-157 */
int mlp_prec_apply(mlp_ctx ctx, mlp_prec p, const double* d_f, double* d_x);

/* ---- solvers: PCGSolver / GMRESSolver / ChebSolver / SimpleSolver   modules/solvers/solvers.h:31-45 ---- */
#define MLP_X0_ZERO      0   /* x0 = 0 */
#define MLP_X0_GIVEN     1   /* x0 = d_x on entry */
#define MLP_X0_REFERENCE 2   /* generate_x0: srandom(3), x = 20(random()-RAND_MAX/2)/RAND_MAX + 100
                                modules/solvers/solvers.cc:9-21 */
typedef struct {
    double eps;               /* relative tolerance on the recurrence residual */
    uint32_t maxit;           /* 0 = no cap */
    int x0_mode;
    int flexible;             /* PCG: Polak-Ribiere beta (for an inexact preconditioner); once the
                                 recurrence residual meets eps the FP64 true residual b - A x is
                                 computed, and while it is above eps PCG restarts from it, at most 8
                                 times; still above eps after that: MLP_ERR_NO_CONV (x is the last
                                 iterate, stats filled) */
    uint32_t true_res_every;  /* PCG: recompute r = b - A x every k iterations (0 = never) */
    uint32_t check_every;     /* iterations enqueued between host convergence checks (0 = 16) */
    int use_graph;            /* capture one iteration in a hipGraph and replay it */
} mlp_solve_opts;

typedef struct {
    uint32_t niter;
    double init_norm, final_norm, eps_used;
    double true_rel_res;      /* ||b - A x|| / ||b - A x0|| in FP64 after the solve */
    double t_solve_ms;        /* wall time of the solve on the host clock */
    int status;
} mlp_solve_stats;

int mlp_pcg(mlp_ctx ctx, mlp_mat A, mlp_prec B, const double* d_b, double* d_x,
            const mlp_solve_opts* opts, mlp_solve_stats* stats);
int mlp_gmres(mlp_ctx ctx, mlp_mat A, mlp_prec B, const double* d_b, double* d_x,
              const mlp_solve_opts* opts, mlp_solve_stats* stats);
/* mlp_pcg for callers that cannot allocate device memory (hot serving loops, code that must not call
 * hipMalloc): no counterpart in the reference.  It runs mlp_pcg's kernels in mlp_pcg's order, so x, the
 * iteration count and the stats equal mlp_pcg's, but it allocates nothing: the solve state, the reduction
 * partial sums and the five PCG vectors live in one static device array per GPU (47.4 MB, held from the
 * library's load until the process ends) and the iterations are plain launches (no graph is instantiated;
 * opts->use_graph is ignored).  Limits: n <= 1,179,648 rows (MLP_ERR_INVALID beyond); x0_mode ZERO or
 * GIVEN; preconditioners whose apply synchronises with the host (SymPrec, dynamic multi_split) are refused.
 * Every call on one GPU shares that array: calls that overlap on one GPU (two host threads, or two
 * streams) race and must be serialised by the caller. */
int mlp_pcg_noalloc(mlp_ctx ctx, mlp_mat A, mlp_prec B, const double* d_b, double* d_x,
                    const mlp_solve_opts* opts, mlp_solve_stats* stats);
/* Ensembles (no counterpart in the reference, which solves one system per run): `count` independent
 * PCGSolver calls (pcg.cc:8-102) run concurrently.  Every job needs its own context (stream, workspace,
 * device state) holding its matrix and preconditioner.  Each job's x and stats are bitwise those of a
 * single mlp_pcg call with the same opts.  Returns MLP_OK, or the first failing job's code (every job's
 * own code is in job.status; the jobs after a failing one still run). */
typedef struct {
    mlp_ctx ctx;
    mlp_mat A;
    mlp_prec B;
    const double* d_b;
    double* d_x;
    mlp_solve_stats stats;    /* out */
    int status;               /* out */
} mlp_pcg_job;
int mlp_pcg_batch(uint32_t count, mlp_pcg_job* jobs, const mlp_solve_opts* opts);

/* Many right-hand sides of one system (no counterpart in the reference, which
 * solves one at a time; for e.g. one matrix with 32 right-hand sides): nrhs
 * (1..64) PCG solves with matrix A and preconditioner B (uh_cheb, on ctx),
 * d_b and d_x holding nrhs device vectors of length n one after another.
 * Every kernel launch of an iteration covers all right-hand sides; each solve
 * runs PCGSolver's steps (pcg.cc:8-102) on its own vectors and stops on its
 * own test, giving bitwise the result of mlp_pcg on that right-hand side
 * (flexible solves take mlp_pcg's true-residual restarts one by one after the
 * batch).  stats has nrhs entries; the return code is the first failing
 * right-hand side's. */
int mlp_pcg_multi(mlp_ctx ctx, mlp_mat A, mlp_prec B, uint32_t nrhs, const double* d_b, double* d_x,
                  const mlp_solve_opts* opts, mlp_solve_stats* stats);
/* Ensembles in one launch per step (no counterpart in the reference): nseg (1..256) independent systems held
 * as one block-diagonal matrix A (segment y = rows seg_off[y] .. seg_off[y+1]-1, host array; no entry
 * couples two segments, e.g. mlp_spe_ensemble_matrix_dev) with one preconditioner B built on A (a uh_cheb
 * hierarchy of such a matrix is the union of the systems' own hierarchies; when the systems' hierarchies
 * have equal depth it applies each system's own preconditioner exactly).  Each segment runs PCGSolver's
 * steps (pcg.cc:8-102) with its own state, its own reductions (the grid and order of a solve of that system
 * alone) and its own stop test, so its x and stats are bitwise those of mlp_pcg on that system with that
 * preconditioner; every kernel launch of an iteration, the preconditioner's included, covers all segments.
 * Flexible solves then take mlp_pcg's true-residual restarts (every segment that needs one in one batched
 * phase per round).  stats has nseg entries; the return code is the first failing segment's. */
int mlp_pcg_ensemble(mlp_ctx ctx, mlp_mat A, mlp_prec B, uint32_t nseg, const uint32_t* seg_off, const double* d_b,
                     double* d_x, const mlp_solve_opts* opts, mlp_solve_stats* stats);
/* Several such ensembles at once, each on its own context (e.g. the groups of an ensemble whose systems'
 * hierarchies have different depths), driven from the calling thread: their iteration chunks are enqueued
 * round-robin and polled through events.  Every job's x, stats and status are those of mlp_pcg_ensemble on
 * it alone; returns MLP_OK or the first failing job's code. */
typedef struct {
    mlp_ctx ctx;
    mlp_mat A;
    mlp_prec B;
    uint32_t nseg;
    const uint32_t* seg_off;  /* host, nseg + 1 */
    const double* d_b;
    double* d_x;
    mlp_solve_stats* stats;   /* nseg entries (out) */
    int status;               /* out */
} mlp_ens_job;
int mlp_pcg_ensemble_batch(uint32_t count, mlp_ens_job* jobs, const mlp_solve_opts* opts);
/* Adjoint gradient for the two-point-flux assembly of modules/mesh/spe.cc:69-131 (no counterpart in the
 * reference; for differentiating a misfit through the solve).  Parameters theta_i = ln kx_i per cell
 * (ky = kx, kz = kz_ratio kx); fx, fy, fz are the per-direction face factors (1/h^2 in spe.cc's scaling);
 * cells are numbered x fastest.  Writes grad_i = -lam^T (dA/dtheta_i) p, which is dL/dtheta_i when
 * A p = b and A lam = dL/dp (A symmetric).  All vectors are device pointers of nx*ny*nz doubles. */
int mlp_spe_adjoint_grad(mlp_ctx ctx, uint32_t nx, uint32_t ny, uint32_t nz, double fx, double fy, double fz,
                         double kz_ratio, const double* d_kx, const double* d_p, const double* d_lam, double* d_grad);
/* ChebSolver(A, lmin, lmax, b, B, x, eps, norm, silent)  modules/solvers/cheb.cc:9-100 (NORM_L2).
 * [lmin, lmax] bound the spectrum of B A, e.g. Prec::lmin()/lmax() (cheb_prec.h:122-123) = level 0 of
 * the hierarchy.  stats->niter follows the reference's count (cheb.cc:81).  opts->flexible and
 * opts->true_res_every are ignored; the residual is recomputed every iteration as in the reference. */
int mlp_cheb_solver(mlp_ctx ctx, mlp_mat A, mlp_prec B, double lmin, double lmax, const double* d_b, double* d_x,
                    const mlp_solve_opts* opts, mlp_solve_stats* stats);
/* SimpleSolver(A, b, B, x, stats, eps, norm, silent)  modules/solvers/simple_solver.cc:8-88 (NORM_L2):
 * x += B (b - A x).  Returns MLP_ERR_NO_ITER ("No need to iterate", simple_solver.cc:80-82) when fewer
 * than two iterations ran; x holds the result all the same. */
int mlp_simple_solver(mlp_ctx ctx, mlp_mat A, mlp_prec B, const double* d_b, double* d_x,
                      const mlp_solve_opts* opts, mlp_solve_stats* stats);

/* ---- matrix assembly ---- */
/* Seven-point matrix of an nx*ny*nz grid (cell i + nx*(j + ny*k)) in Skyline CSR,
 * SPEMesh::construct_matrix (modules/mesh/spe.cc:69-131, XYZ order): row = diagonal
 * c + the face values of the present neighbours, then -v for the neighbours z-, y-,
 * x-, x+, y+, z+, with v = 2/(1/k_a + 1/k_b) / h^2 on the face's axis (kind
 * MLP_ASM_SPE).  kx = ky per cell, kz = kz_ratio * kx (the reference reads kx, ky,
 * kz from spe_perm.dat).  MLP_ASM_TPFA scales the harmonic mean by area/spacing and
 * adds Dirichlet half-cells on x = 0 and x = max to the diagonal.  ia holds n+1
 * entries, ja and a mlp_spe_nnz(nx, ny, nz); host arrays. */
#define MLP_ASM_SPE 0
#define MLP_ASM_TPFA 1
uint64_t mlp_spe_nnz(uint32_t nx, uint32_t ny, uint32_t nz);
int mlp_spe_assemble(int kind, uint32_t nx, uint32_t ny, uint32_t nz, const double* kx, double kz_ratio,
                     double c, double hx, double hy, double hz, uint32_t* ia, uint32_t* ja, double* a);
/* The same matrix assembled on ctx's device from device-resident kx, as a matrix handle: its Skyline CSR
 * stays in device memory (mlp_hier_build_uh_cheb_mat reads it), the solve layout (seven-point when every
 * dimension is at least 2) is written by the assembly, and the host copy is fetched only when something
 * needs it (DiagPrec, GSPrec, CompositePrec, SymPrec, mlp_mat_get_csr).  Bit for bit mlp_spe_assemble's
 * matrix. */
int mlp_spe_matrix_dev(mlp_ctx ctx, int kind, uint32_t nx, uint32_t ny, uint32_t nz, const double* d_kx,
                       double kz_ratio, double c, double hx, double hy, double hz, mlp_mat* out);
/* An ensemble of nsys such systems with one nx*ny footprint and depths nz[0..nsys-1] (host array), stacked
 * along z into one block-diagonal matrix (no entry couples two systems; block s is bit for bit the matrix
 * mlp_spe_matrix_dev assembles for system s alone).  d_kx holds the systems' fields one after another. */
int mlp_spe_ensemble_matrix_dev(mlp_ctx ctx, int kind, uint32_t nx, uint32_t ny, uint32_t nsys, const uint32_t* nz,
                                const double* d_kx, double kz_ratio, double c, double hx, double hy, double hz,
                                mlp_mat* out);
/* A's Skyline CSR into host arrays (n+1, nnz, nnz): CSRMatrix::get_ia/get_ja/get_a, matrix.h:204-210 */
int mlp_mat_get_csr(mlp_mat A, uint32_t* ia, uint32_t* ja, double* a);

/* ---- diagnostics (no counterpart in the reference; used for roofline reporting) ---- */
/* algorithmic HBM bytes of one apply (sum over level visits) and the kernels the last apply launched
 * (0 before the first apply) */
int mlp_prec_bytes(mlp_prec p, double* bytes_per_apply, uint32_t* launches_per_apply);
/* The launch schedule a preconditioner was built with (no counterpart in the reference): MLP_SCHED_* bits |
 * forced level block size << 8 (0 for preconditioners other than uh_cheb / multi_split hierarchies). */
int mlp_prec_schedule(mlp_prec p, uint32_t* sched);
/* Average device time of `reps` back-to-back SpMV launches y = A x (two HIP
 * events on the context stream around the whole run).  flush_bytes > 0 makes
 * the launches rotate over device copies of A totalling at least flush_bytes,
 * so that each launch reads its matrix from HBM rather than from L2 or the
 * 256 MB Infinity Cache; 0 reuses A (cache-warm). */
int mlp_time_spmv(mlp_ctx ctx, mlp_mat A, const double* d_x, double* d_y, uint32_t reps,
                  size_t flush_bytes, double* avg_ms);
/* Average device time of `reps` preconditioner applies x = B f. */
int mlp_time_prec_apply(mlp_ctx ctx, mlp_prec p, const double* d_f, double* d_x, uint32_t reps,
                        double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif

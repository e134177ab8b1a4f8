// Device kernels for the iterative-solve path (gfx950, wave64).
//
// Floating-point kernels are compiled with -ffp-contract=off: each a*b+c is
// a rounded multiply followed by a rounded add, in the reference's operand
// order, so FP64 results match the reference's CPU loops bit for bit wherever
// the operation order is per-row (SpMV, triangular sweeps, vector updates).
// Reductions use a fixed tree and are deterministic run to run.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bfloat16.h>
#include <cstdint>

namespace mlp {

constexpr uint32_t kPad = 0xffffffffu;
constexpr uint32_t kGlobalBit = 0x80000000u;
constexpr int kBS = 256;                 // threads per block for flat kernels
constexpr int kRedBlocks = 16384;        // max blocks of a one- or two-value reduction grid
constexpr int kMdotBlocks = 1024;        // max blocks of the Gram-Schmidt multi-dot (32 partials each)
constexpr size_t kPartials = 32768;      // doubles of partial-sum scratch (2 * kRedBlocks = 32 * kMdotBlocks)
constexpr uint32_t kRedGrid = 1024;      // blocks of a streaming reduction over n rows: clamp(ceil(n / kBS), 1, kRedGrid)

// Sliced ELL (64-row slices, column-major inside a slice).  Column indices
// are either uint32 (col) or, when every |column - row| fits, int16 deltas
// from the row (dcol, kPad16 marks padding), which cuts the index stream in
// half; exactly one of col / dcol is set.
constexpr int16_t kPad16 = -32768;
template <class VT> struct SellView {
    uint32_t n = 0;
    const uint32_t* off64 = nullptr;
    const uint32_t* col = nullptr;
    const int16_t* dcol = nullptr;
    const VT* val = nullptr;
    // seven-point layout (plan.h:StencilLayout), used instead of the slices when sd is set
    const uint8_t* spres = nullptr;
    const VT *sd = nullptr, *slx = nullptr, *sly = nullptr, *slz = nullptr;
    uint32_t ssy = 0, ssz = 0;
};


// Device-resident state of a PCG/GMRES solve.  Host code never needs these
// values between iterations; it polls `done` every few iterations.
struct SolveState {
    double rzs[2];                       // (r, z) of the iteration with niter of that parity
    double app, alpha, beta, rr, norm, init_norm, eps;
    uint32_t niter, done, status, maxit;
    uint32_t alpha_one, malpha_one, beta_one;
    uint32_t pad;                        // Chebyshev/simple: niter offset of the maxit test
    unsigned counter[8];                 // last-block counters of the fused reductions
};

// Batched right-hand sides: a launch with gridDim.y = R runs R independent
// problems on one matrix and preconditioner.  Problem y = blockIdx.y finds
// each of its vectors y * stride elements after problem 0's (the strides by
// vector kind below), its solve state at st + y and its partial sums
// y * part doubles on.  With R = 1 every offset is 0.
struct Batch {
    uint32_t n = 0;        // the length-n solve vectors
    uint32_t part = 0;     // partial-sum buffers
    // ragged batches (an ensemble of systems in one block-diagonal matrix): solve y owns rows
    // [soff[y], soff[y+1]) of one vector; null for batches of equal-length vectors
    const uint32_t* soff = nullptr;
};
struct LevelBatch {
    uint32_t N = 0;        // level vectors: the input f, the output x, the parent's u, F and us
    uint32_t nodes = 0;    // wt and the global-block scratch gx
    uint32_t n = 0;        // kept-row vectors: F, x2, u1, u0
    uint32_t st = 0;       // solve states, in 32-bit words (the done-flag guard)
};
// p + off, leaving null pointers null
template <class T>
__device__ __forceinline__ T* shift(T* p, size_t off) { return p ? p + off : p; }

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool guarded(const uint32_t* g) { return g && *((volatile const uint32_t*)g); }

__device__ __forceinline__ bool dev_is_one(double a) {   // include/tools.h:28-30
    return fabs(a - 1.0) <= 2.220446049250313e-16;
}

// acc - v x in a row update.  FP64 rounds the product and the difference
// separately (the reference's arithmetic, no contraction); FP32 fuses them,
// one instruction per entry on the latency-bound chains.
template <class VT> __device__ __forceinline__ VT msub(VT acc, VT v, VT x);
template <> __device__ __forceinline__ double msub<double>(double acc, double v, double x) { return acc - v * x; }
template <> __device__ __forceinline__ float msub<float>(float acc, float v, float x) { return fmaf(-v, x, acc); }

// x / u_ii at the end of a U row.  FP64 divides (IEEE, the reference's
// arithmetic); an FP32 hierarchy stores 1 / u_ii (rounded from the FP64
// reciprocal) and multiplies, one instruction instead of a division sequence
// on every step of the latency-bound U chains.
template <class VT> __device__ __forceinline__ VT udiv(VT z, VT d);
template <> __device__ __forceinline__ double udiv<double>(double z, double d) { return z / d; }
template <> __device__ __forceinline__ float udiv<float>(float z, float rcp) { return z * rcp; }

template <int BS>
__device__ __forceinline__ double block_sum(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < BS / 64; i++) s += sh[i];
    return s;   // valid in thread 0
}

// Every block stores its partial sums; a one-block finalize kernel adds
// them in block order (deterministic for a fixed grid).  No device-scope
// fence is needed: the kernel boundary publishes the partials.
template <int BS, int NV>
__device__ __forceinline__ void block_partials(const double (&part)[NV], double* partials) {
    __shared__ double sh[BS / 64];
    for (int k = 0; k < NV; k++) {
        const double b = block_sum<BS>(part[k], sh);
        if (threadIdx.x == 0) partials[NV * blockIdx.x + k] = b;
    }
}

// sum of partials[NV*b + k] over b < nb, same order for every call
template <int BS>
__device__ __forceinline__ double sum_partials(const double* partials, uint32_t nb, int NV, int k, double* sh) {
    double s = 0.0;
    for (uint32_t b = threadIdx.x; b < nb; b += BS) s += partials[NV * b + k];
    return block_sum<BS>(s, sh);
}

// st->done read once per block, so that every thread of a block takes the
// same branch even while another block of the same launch sets it
__device__ __forceinline__ bool block_done(const SolveState* st) {
    __shared__ uint32_t d;
    if (threadIdx.x == 0) d = *((volatile const uint32_t*)&st->done);
    __syncthreads();
    const bool r = d != 0;
    __syncthreads();
    return r;
}

// Solve y of a batch: its first row, its length and the reduction grid that a
// solve of that length alone launches.  Batches of equal-length vectors: solve
// y at y * n with the launch's own grid.  Ragged batches (SEG): segment
// [soff[y], soff[y+1]) with grid clamp(ceil(len / kBS), 1, kRedGrid), so that
// a segment's rows go to the same threads and blocks, and its partial sums are
// added in the same order, as in a solve of that system alone.
struct Span {
    size_t base;
    uint32_t len, rg;
};
template <bool SEG> __device__ __forceinline__ Span span_of(const Batch& bt, uint32_t y) {
    if constexpr (SEG) {
        const uint32_t b = bt.soff[y], len = bt.soff[y + 1] - b;
        const uint32_t t = (len + kBS - 1) / kBS;
        return {b, len, t < 1 ? 1u : (t > kRedGrid ? kRedGrid : t)};
    } else {
        return {size_t(y) * bt.n, bt.n, gridDim.x};
    }
}

// sum_partials, returned to every thread of the block
template <int BS>
__device__ __forceinline__ double total_of_partials(const double* partials, uint32_t nb, int NV, int k, double* sh) {
    const double t = sum_partials<BS>(partials, nb, NV, k, sh);
    __shared__ double out;
    if (threadIdx.x == 0) out = t;
    __syncthreads();
    const double r = out;
    __syncthreads();
    return r;
}

// Workgroups are dispatched to the 8 XCDs round-robin (block b runs on XCD
// b % 8), and each XCD has its own L2.  xcd_block maps b to a logical block so
// that every XCD gets one contiguous run of logical blocks: the neighbouring
// rows an SpMV gathers from x are then mostly in that XCD's L2.  A bijection
// on [0, G) for any G.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t G) {
    constexpr uint32_t X = 8;
    const uint32_t q = G / X, rem = G % X, x = b % X;
    return x * q + (x < rem ? x : rem) + b / X;
}

// y_r = sum_k a_rk * (S ? s*x_k : x_k), in storage order, from 0.
// Entries are fetched in groups of 8: column indices and values together
// (padding slots hold a valid value 0), then the x gathers, so that a row's
// loads are in flight before the first product; the sum itself still runs
// entry by entry and skips padding.
// The same row sum on the seven-point layout: the diagonal, then z-, y-, x-,
// x+, y+, z+ (the row's storage order), the upper values read from the
// neighbouring rows' lower values.  Every load in bounds is issued at once,
// without waiting for the presence bits, which only select the terms.
template <class VT, class XT, bool SCALE>
__device__ __forceinline__ VT stencil_row(const SellView<VT>& A, const XT* __restrict__ x, uint32_t r, VT s) {
    const uint32_t m = A.spres[r];
    const uint32_t c[6] = {r - A.ssz, r - A.ssy, r - 1, r + 1, r + A.ssy, r + A.ssz};
    const bool inb[6] = {r >= A.ssz, r >= A.ssy, r >= 1, r + 1 < A.n, r + A.ssy < A.n, r + A.ssz < A.n};
    VT v[6];
    XT xv[6];
    v[0] = A.slz[r];
    v[1] = A.sly[r];
    v[2] = A.slx[r];
    v[3] = inb[3] ? A.slx[r + 1] : VT(0);
    v[4] = inb[4] ? A.sly[r + A.ssy] : VT(0);
    v[5] = inb[5] ? A.slz[r + A.ssz] : VT(0);
#pragma unroll
    for (int t = 0; t < 6; t++) xv[t] = inb[t] ? x[c[t]] : XT(0);
    VT t0 = VT(x[r]);
    if (SCALE) t0 = t0 * s;
    VT acc = VT(0);
    acc += A.sd[r] * t0;
#pragma unroll
    for (int t = 0; t < 6; t++) {
        if ((m >> t) & 1u) {
            VT u = VT(xv[t]);
            if (SCALE) u = u * s;
            acc += v[t] * u;
        }
    }
    return acc;
}

// Rows r and r+1 of the FP64 seven-point layout, each summed with stencil_row's
// operations in its order.  VEC: each pair of neighbouring values is one
// 16-byte load; that needs r, n, ssy and ssz even and 16-byte aligned arrays
// (pair_vec), and then every load is in bounds for both rows.  Otherwise the
// rows are summed one by one, row r+1 only when two.  x0, x1: x[r], x[r+1].
struct RowPair {
    double y0, y1, x0, x1;
};
typedef double dbl2 __attribute__((ext_vector_type(2)));

__host__ __device__ __forceinline__ bool pair_vec(const SellView<double>& A, uintptr_t vecs) {
    const uintptr_t a = uintptr_t(A.sd) | uintptr_t(A.slx) | uintptr_t(A.sly) | uintptr_t(A.slz) | vecs;
    return ((A.n | A.ssy | A.ssz) & 1u) == 0 && (a & 15u) == 0 && (uintptr_t(A.spres) & 1u) == 0;
}

template <bool VEC>
__device__ __forceinline__ RowPair stencil_pair(const SellView<double>& A, const double* __restrict__ x, uint32_t r,
                                                bool two) {
    RowPair o{0.0, 0.0, 0.0, 0.0};
    if constexpr (!VEC) {
        o.y0 = stencil_row<double, double, false>(A, x, r, 1.0);
        o.x0 = x[r];
        if (two) {
            o.y1 = stencil_row<double, double, false>(A, x, r + 1, 1.0);
            o.x1 = x[r + 1];
        }
    } else {
        const uint32_t n = A.n, sy = A.ssy, sz = A.ssz;
        const dbl2 z2{0.0, 0.0};
        auto l2 = [](const double* p) { return *(const dbl2*)p; };
        const uint32_t m2 = *(const uint16_t*)(A.spres + r);
        const dbl2 dd = l2(A.sd + r), vz = l2(A.slz + r), vy = l2(A.sly + r), vx = l2(A.slx + r);
        const double vx2 = r + 2 < n ? A.slx[r + 2] : 0.0;
        const dbl2 vyp = r + sy < n ? l2(A.sly + r + sy) : z2;
        const dbl2 vzp = r + sz < n ? l2(A.slz + r + sz) : z2;
        const dbl2 xc = l2(x + r);
        const double xm1 = r >= 1 ? x[r - 1] : 0.0;
        const double xp2 = r + 2 < n ? x[r + 2] : 0.0;
        const dbl2 xmy = r >= sy ? l2(x + r - sy) : z2;
        const dbl2 xpy = r + sy < n ? l2(x + r + sy) : z2;
        const dbl2 xmz = r >= sz ? l2(x + r - sz) : z2;
        const dbl2 xpz = r + sz < n ? l2(x + r + sz) : z2;
        // row r: diagonal, then z-, y-, x-, x+, y+, z+
        {
            const uint32_t m = m2 & 0xffu;
            const double v[6] = {vz.x, vy.x, vx.x, vx.y, vyp.x, vzp.x};
            const double xv[6] = {xmz.x, xmy.x, xm1, xc.y, xpy.x, xpz.x};
            double acc = 0.0;
            acc += dd.x * xc.x;
#pragma unroll
            for (int t = 0; t < 6; t++)
                if ((m >> t) & 1u) acc += v[t] * xv[t];
            o.y0 = acc;
        }
        {
            const uint32_t m = m2 >> 8;
            const double v[6] = {vz.y, vy.y, vx.y, vx2, vyp.y, vzp.y};
            const double xv[6] = {xmz.y, xmy.y, xc.x, xp2, xpy.y, xpz.y};
            double acc = 0.0;
            acc += dd.y * xc.y;
#pragma unroll
            for (int t = 0; t < 6; t++)
                if ((m >> t) & 1u) acc += v[t] * xv[t];
            o.y1 = acc;
        }
        o.x0 = xc.x;
        o.x1 = xc.y;
    }
    return o;
}

// SCALE: the row of A (s x).  An FP64 row scales every entry first (the
// reference's dscal then multiply, bit for bit); an FP32 row scales its sum.
template <class VT, class XT, bool SCALE>
__device__ __forceinline__ VT sell_row(const SellView<VT>& A, const XT* __restrict__ x, uint32_t r, double s) {
    constexpr bool kScaleEntries = sizeof(VT) == sizeof(double);
    constexpr uint32_t G = 8;
    const uint32_t sl = r >> 6, lane = r & 63;
    const uint32_t b = A.off64[sl], e = A.off64[sl + 1];
    VT acc = VT(0);
    for (uint32_t k = b; k < e; k += G) {
        uint32_t c[G];
        bool ok[G];
        VT v[G];
        XT xv[G];
        if (A.dcol) {
#pragma unroll
            for (uint32_t j = 0; j < G; j++) {
                const int16_t d = k + j < e ? A.dcol[size_t(k + j) * 64 + lane] : kPad16;
                ok[j] = d != kPad16;
                c[j] = uint32_t(int32_t(r) + int32_t(d));
            }
        } else {
#pragma unroll
            for (uint32_t j = 0; j < G; j++) {
                c[j] = k + j < e ? A.col[size_t(k + j) * 64 + lane] : kPad;
                ok[j] = c[j] != kPad;
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < G; j++) v[j] = k + j < e ? A.val[size_t(k + j) * 64 + lane] : VT(0);   // not gated on c
#pragma unroll
        for (uint32_t j = 0; j < G; j++) xv[j] = ok[j] ? x[c[j]] : XT(0);
#pragma unroll
        for (uint32_t j = 0; j < G; j++) {
            if (ok[j]) {
                VT t = VT(xv[j]);
                if (SCALE && kScaleEntries) t = VT(double(t) * s);   // the reference's dscal, then the row
                acc += v[j] * t;
            }
        }
    }
    // below FP64 the Chebyshev coefficient is applied once to the row, in FP64: A (s u) = s (A u)
    if (SCALE && !kScaleEntries) acc = VT(double(acc) * s);
    return acc;
}

// Row r of a matrix in either layout (the kernels that read a top-level
// matrix; the level kernels' fused residuals only ever see slices).
template <class VT, class XT>
__device__ __forceinline__ VT mat_row(const SellView<VT>& A, const XT* __restrict__ x, uint32_t r) {
    return A.sd ? stencil_row<VT, XT, false>(A, x, r, VT(1)) : sell_row<VT, XT, false>(A, x, r, 1.0);
}

// ---------------------------------------------------------------------------
// SpMV family
// ---------------------------------------------------------------------------
template <class VT, class XT, class YT>
__global__ void __launch_bounds__(kBS) k_spmv(SellView<VT> A, const XT* __restrict__ x, YT* __restrict__ y,
                                              const uint32_t* guard) {
    if (guarded(guard)) return;
    const uint32_t r = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (r >= A.n) return;
    y[r] = YT(mat_row<VT, XT>(A, x, r));
}

// y = A x and r = b - A x on the FP64 seven-point layout, two rows per thread
// (grid: ceil(n / (2 kBS))); every row as k_spmv / k_residual compute it.
// VEC: pair_vec holds for A, x, y and b (the host checks it).
template <bool RES, bool VEC>
__global__ void __launch_bounds__(kBS) k_spmv_pairs(SellView<double> A, const double* __restrict__ b,
                                                    const double* __restrict__ x, double* __restrict__ y) {
    const uint32_t r = 2 * (xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
    if (r >= A.n) return;
    if constexpr (VEC) {
        const RowPair o = stencil_pair<true>(A, x, r, true);       // n even: both rows exist
        dbl2 v{o.y0, o.y1};
        if (RES) v = *(const dbl2*)(b + r) - v;
        *(dbl2*)(y + r) = v;
    } else {
        const bool two = r + 1 < A.n;
        const RowPair o = stencil_pair<false>(A, x, r, two);
        y[r] = RES ? b[r] - o.y0 : o.y0;
        if (two) y[r + 1] = RES ? b[r + 1] - o.y1 : o.y1;
    }
}

// r = b - A x   (residual, csrmatrix.cc:293-300)
template <class VT, class XT>
__global__ void __launch_bounds__(kBS) k_residual(SellView<VT> A, const XT* __restrict__ b, const XT* __restrict__ x,
                                                  XT* __restrict__ rr, const uint32_t* guard) {
    if (guarded(guard)) return;
    const uint32_t r = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (r >= A.n) return;
    VT acc = mat_row<VT, XT>(A, x, r);
    rr[r] = XT(VT(b[r]) - acc);
}

// ---------------------------------------------------------------------------
// level programs: triangular sweeps over the eliminated rows
// (cheb_solve.cc:33-38, 108-122; multi_split_solve.cc:12-53)
// ---------------------------------------------------------------------------
constexpr int kLBS = 256;              // largest block of the level kernels
// The level kernels are instantiated for 64-thread blocks (one wavefront:
// the compiler drops the workgroup barriers of the depth-level loops, so a
// level costs an LDS round trip, not a hardware barrier) and for kLBS.

template <class VT> struct ProgView {
    uint32_t nblk = 0, nflat = 0, nodes = 0;
    uint32_t max_nodes = 0, max_Lent = 0, max_Uent = 0, max_Unodes = 0, max_levels = 0;
    const uint32_t* blk = nullptr;     // 12 per block (see plan.h)
    const uint32_t* lvlL = nullptr;
    const uint32_t* lvlU = nullptr;
    const uint32_t* qrow = nullptr;    // level row read at consumption position q
    const uint32_t* out = nullptr;
    const uint32_t* Lptr = nullptr;
    const uint32_t* Lloc = nullptr;
    const VT* Lval = nullptr;
    const uint32_t* Ldep = nullptr;    // depth level of each node within its block
    const uint32_t* Udep = nullptr;    // U depth level of each U-list row within its block
    const uint32_t* uloc = nullptr;
    const uint32_t* uxidx = nullptr;
    const uint32_t* Uptr = nullptr;
    const uint32_t* Ucol = nullptr;
    const VT* Uval = nullptr;
    const VT* Udiag = nullptr;
    const uint32_t* flatq = nullptr;   // F slot written by flat row i
    VT* gx = nullptr;                  // x of the global-memory blocks (nodes long; scratch of the instance)
};

// header word 1 flag of a block whose values stay in global memory (a
// component too large for LDS)
constexpr uint32_t kGlobalBlk = 0x80000000u;

// Value of kept row r of the level output, produced by the iteration on the
// next level (cheb_solve.cc:61-104, multi_split_solve.cc:106-108, 131-133).
enum { OP_NONE = 0, OP_COPY = 1, OP_SCALE = 2, OP_CHEB2 = 3, OP_CHEBK = 4 };
// CV: the type the Chebyshev vectors x2, u1, u0 are stored in (the level's
// own type, or BF16 for the finest level's vectors of a BF16 variant).
template <class VT, class CV = VT> struct Combine {
    const CV* x2 = nullptr;
    const CV* u1 = nullptr;
    const CV* u0 = nullptr;
    double a = 0.0, b = 0.0;             // Chebyshev coefficients, applied in FP64 (one rounding to VT)
    int skip = 0, one = 0;
};

template <class VT, int OP, class CV>
__device__ __forceinline__ VT combine(const Combine<VT, CV>& C, uint32_t r) {
    if (OP == OP_COPY) return C.x2[r];
    if (OP == OP_SCALE) return C.skip ? VT(C.x2[r]) : VT(double(C.x2[r]) * C.a); // dscal(alpha, x2)
    if (OP == OP_CHEB2) {                                                          // dscal(-alpha,x2); daxpy(1+beta,u1,x2)
        const double v = C.skip ? double(C.x2[r]) : double(C.x2[r]) * C.a;
        const double u = double(C.u1[r]);
        return VT(C.one ? v + u : v + C.b * u);
    }
    if (OP == OP_CHEBK) {
        const double u1 = double(C.u1[r]), x2 = double(C.x2[r]), u0 = double(C.u0[r]);
        return VT(u1 - C.a * x2 + C.b * (u1 - u0));
    }
    return VT(0);
}

// Where a level reads its right-hand side; q is the consumption position and
// j = qrow[q] the level row.
//  IN_PLAIN: f[j] (a vector in level order).
//  IN_PERM:  f[q] (a vector already stored in consumption order: the kept-row
//            right-hand side F written by the parent level).
//  IN_RES, IN_RES_SCALE: the Chebyshev residual of the parent level fused into
//            the read, f = (A (s u))_j - F[q], the row taken from a copy of A
//            stored in consumption order; IN_RES_SCALE also stores
//            us[j] = s u[j] when us is given (cheb_solve.cc:61-78, 96-97; the
//            chain recomputes s u[j] where its combine reads it instead).
// The kernels read every position exactly once, so each residual row is
// computed once, with the same operations as a separate SpMV.
enum { IN_PLAIN = 0, IN_PERM = 1, IN_RES = 2, IN_RES_SCALE = 3 };
// FT: the type the parent's vectors (f, u, F, us) are stored in.
template <class VT, class FT, int IN> struct InSrc {
    static constexpr bool kNeedRow = IN == IN_PLAIN || IN == IN_RES_SCALE;
    const FT* f = nullptr;
    SellView<VT> A{};
    const FT* u = nullptr;
    const FT* F = nullptr;
    FT* us = nullptr;
    double s = 1.0;                      // Chebyshev coefficient (FP64)
};

template <class VT, class FT, int IN>
__device__ __forceinline__ VT fetch(const InSrc<VT, FT, IN>& S, uint32_t q, uint32_t j) {
    if (IN == IN_PLAIN) return VT(S.f[j]);
    if (IN == IN_PERM) return VT(S.f[q]);
    const VT acc = sell_row<VT, FT, IN == IN_RES_SCALE>(S.A, S.u, q, S.s);
    if (IN == IN_RES_SCALE && S.us) S.us[j] = FT(double(S.u[j]) * S.s);
    return acc - VT(S.F[q]);
}

template <class VT, class FT, int IN>
__device__ __forceinline__ VT fetch_at(const ProgView<VT>& P, const InSrc<VT, FT, IN>& S, uint32_t q) {
    const uint32_t j = InSrc<VT, FT, IN>::kNeedRow ? P.qrow[q] : 0u;
    return fetch(S, q, j);
}

// LDS carve-up shared by the host (sizes) and the kernels (offsets)
struct LdsPlan {
    uint32_t sw, sx, lval, uval, uref, sz, sdiag, lptr, lloc, uptr, ucol, uloc, lvl, total;
};
template <class VT>
__host__ __device__ inline LdsPlan lds_plan(uint32_t nodes, uint32_t lent, uint32_t uent, uint32_t unodes,
                                            uint32_t levels, bool down, bool up) {
    LdsPlan p{};
    uint32_t o = 0;
    auto take = [&](uint32_t bytes) { uint32_t r = o; o += (bytes + 7u) & ~7u; return r; };
    p.sw = take(nodes * sizeof(VT));
    p.sx = up ? take(nodes * sizeof(VT)) : 0;
    p.lval = down ? take(lent * sizeof(VT)) : 0;
    p.uval = up ? take(uent * sizeof(VT)) : 0;
    p.uref = up ? take(uent * sizeof(VT)) : 0;
    p.sz = up ? take(unodes * sizeof(VT)) : 0;
    p.sdiag = up ? take(unodes * sizeof(VT)) : 0;
    p.lptr = down ? take((nodes + 1) * 4) : 0;
    p.lloc = down ? take(lent * 4) : 0;
    p.uptr = up ? take((unodes + 1) * 4) : 0;
    p.ucol = up ? take(uent * 4) : 0;
    p.uloc = up ? take(unodes * 4) : 0;
    p.lvl = take(levels * 4);            // depth-level end offsets of the sweep being run
    p.total = o;
    return p;
}

// Forward sweep of one block: stage the right-hand side and the L rows in
// LDS, then one depth level per barrier.  Leaves w in sw.
template <class VT, class FT, int IN>
__device__ __forceinline__ void block_down(const ProgView<VT>& P, const uint32_t* B, const InSrc<VT, FT, IN>& f,
                                           unsigned char* lds, const LdsPlan& lp) {
    VT* sw = reinterpret_cast<VT*>(lds + lp.sw);
    VT* sval = reinterpret_cast<VT*>(lds + lp.lval);
    uint32_t* sptr = reinterpret_cast<uint32_t*>(lds + lp.lptr);
    uint32_t* sloc = reinterpret_cast<uint32_t*>(lds + lp.lloc);
    uint32_t* slv = reinterpret_cast<uint32_t*>(lds + lp.lvl);
    const uint32_t base = B[0], S = B[1], e0 = B[6], ne = B[7] - B[6], nl = B[3] - B[2];
    for (uint32_t p = threadIdx.x; p < S; p += blockDim.x) {
        sw[p] = fetch_at(P, f, base + p);
        sptr[p] = P.Lptr[base + p] - e0;
    }
    if (threadIdx.x == 0) sptr[S] = ne;
    for (uint32_t q = threadIdx.x; q < ne; q += blockDim.x) {
        sloc[q] = P.Lloc[e0 + q];
        sval[q] = P.Lval[e0 + q];
    }
    for (uint32_t li = threadIdx.x; li < nl; li += blockDim.x) slv[li] = P.lvlL[B[2] + li];
    __syncthreads();
    uint32_t s = 0;
    for (uint32_t li = 0; li < nl; li++) {
        const uint32_t e = slv[li];
        for (uint32_t p = s + threadIdx.x; p < e; p += blockDim.x) {
            VT acc = sw[p];
            for (uint32_t q = sptr[p]; q < sptr[p + 1]; q++) acc = msub(acc, sval[q], sw[sloc[q]]);
            sw[p] = acc;
        }
        s = e;
        __syncthreads();
    }
}

// Back sweep of one block over its eliminated rows in U order.  Each row's w
// comes from LDS (after block_down) or from wt; kept columns take their value
// from the combine op.  Results land in sz.
template <class VT, int OP, bool W_IN_LDS, class CV>
__device__ __forceinline__ void block_up(const ProgView<VT>& P, const uint32_t* B, const VT* __restrict__ wt,
                                         const Combine<VT, CV>& C, unsigned char* lds, const LdsPlan& lp) {
    VT* sw = reinterpret_cast<VT*>(lds + lp.sw);
    VT* sx = reinterpret_cast<VT*>(lds + lp.sx);
    VT* sval = reinterpret_cast<VT*>(lds + lp.uval);
    VT* sref = reinterpret_cast<VT*>(lds + lp.uref);
    VT* sz = reinterpret_cast<VT*>(lds + lp.sz);
    VT* sdiag = reinterpret_cast<VT*>(lds + lp.sdiag);
    uint32_t* sptr = reinterpret_cast<uint32_t*>(lds + lp.uptr);
    uint32_t* scol = reinterpret_cast<uint32_t*>(lds + lp.ucol);
    uint32_t* sloc = reinterpret_cast<uint32_t*>(lds + lp.uloc);
    uint32_t* slv = reinterpret_cast<uint32_t*>(lds + lp.lvl);
    const uint32_t base = B[0], e0 = B[8], ne = B[9] - B[8], ub = B[10], nu = B[11], nl = B[5] - B[4];
    for (uint32_t u = threadIdx.x; u < nu; u += blockDim.x) {
        const uint32_t loc = P.uloc[ub + u];
        sloc[u] = loc;
        sz[u] = W_IN_LDS ? sw[loc] : wt[base + loc];
        sdiag[u] = P.Udiag[ub + u];
        sptr[u] = P.Uptr[ub + u] - e0;
    }
    if (threadIdx.x == 0) sptr[nu] = ne;
    for (uint32_t q = threadIdx.x; q < ne; q += blockDim.x) {
        const uint32_t c = P.Ucol[e0 + q];
        scol[q] = c;
        sval[q] = P.Uval[e0 + q];
        if (OP != OP_NONE && (c & kGlobalBit)) sref[q] = combine<VT, OP>(C, c & ~kGlobalBit);
    }
    for (uint32_t li = threadIdx.x; li < nl; li += blockDim.x) slv[li] = P.lvlU[B[4] + li];
    __syncthreads();
    uint32_t s = 0;
    for (uint32_t li = 0; li < nl; li++) {
        const uint32_t e = slv[li];
        for (uint32_t u = s + threadIdx.x; u < e; u += blockDim.x) {
            VT z = sz[u];
            for (uint32_t q = sptr[u]; q < sptr[u + 1]; q++) {
                const uint32_t c = scol[q];
                const VT xv = (c & kGlobalBit) ? sref[q] : sx[c];
                z = msub(z, sval[q], xv);
            }
            const VT v = udiv(z, sdiag[u]);
            sx[sloc[u]] = v;
            sz[u] = v;
        }
        s = e;
        __syncthreads();
    }
}

// End of one depth level of an owned-node sweep.  A one-wavefront block only
// needs the compiler not to move LDS accesses across it: the LDS executes a
// wavefront's accesses in program order, so a value written at one level is
// seen by every lane that reads it at a later level.
template <int BS>
__device__ __forceinline__ void level_sync() {
    if constexpr (BS <= 64) __builtin_amdgcn_wave_barrier();
    else __syncthreads();
}

// Owned-node sweeps, used when a block has no more nodes (U-list rows) than
// threads: thread p owns node p, reads its right-hand side, depth level, entry
// range and first two entries straight from global memory once, and at its
// depth level gathers the solved values it needs from LDS.  A node's value is
// read only by nodes of later levels, after the barrier that ends its own
// level, so nothing is staged before the sweep.  Returns the thread's solved
// value (0 for threads without a node); sw holds every node's w afterwards.
template <int BS, class VT, class FT, int IN>
__device__ __forceinline__ VT owned_down(const ProgView<VT>& P, const uint32_t* B, const InSrc<VT, FT, IN>& f,
                                         VT* sw) {
    const uint32_t base = B[0], S = B[1], nl = B[3] - B[2];
    const uint32_t p = threadIdx.x;
    uint32_t my = nl, q0 = 0, q1 = 0, l0 = p, l1 = p;
    VT acc = VT(0), v0 = VT(0), v1 = VT(0);
    if (p < S) {
        my = P.Ldep[base + p];
        q0 = P.Lptr[base + p];
        q1 = P.Lptr[base + p + 1];
        acc = fetch_at(P, f, base + p);
        if (q1 > q0) { l0 = P.Lloc[q0]; v0 = P.Lval[q0]; }
        if (q1 > q0 + 1) { l1 = P.Lloc[q0 + 1]; v1 = P.Lval[q0 + 1]; }
    }
    // absent entries read the thread's own slot and are selected away, so the
    // level body has no branches on the entry count
    const bool h0 = q1 > q0, h1 = q1 > q0 + 1, more = q1 > q0 + 2;
    for (uint32_t li = 0; li < nl; li++) {
        if (li == my) {
            const VT x0 = sw[l0], x1 = sw[l1];
            acc = h0 ? msub(acc, v0, x0) : acc;
            acc = h1 ? msub(acc, v1, x1) : acc;
            if (more)
                for (uint32_t q = q0 + 2; q < q1; q++) acc = msub(acc, P.Lval[q], sw[P.Lloc[q]]);
            sw[p] = acc;
        }
        level_sync<BS>();
    }
    return acc;
}

// U sweep of the owned-node path.  w of the row comes from sw (after
// owned_down in the same launch) or from wt; kept columns from the combine op.
// Solved values land in sx (local L positions); returns the thread's value.
template <int BS, class VT, int OP, bool W_IN_LDS, class CV>
__device__ __forceinline__ VT owned_up(const ProgView<VT>& P, const uint32_t* B, const VT* __restrict__ wt,
                                       const Combine<VT, CV>& C, const VT* sw, VT* sx) {
    const uint32_t base = B[0], ub = B[10], nu = B[11], nl = B[5] - B[4];
    const uint32_t u = threadIdx.x;
    uint32_t my = nl, q0 = 0, q1 = 0, c0 = kGlobalBit, c1 = kGlobalBit, loc = 0;
    VT z = VT(0), d = VT(1), v0 = VT(0), v1 = VT(0), r0 = VT(0), r1 = VT(0);
    if (u < nu) {
        my = P.Udep[ub + u];
        loc = P.uloc[ub + u];
        d = P.Udiag[ub + u];
        q0 = P.Uptr[ub + u];
        q1 = P.Uptr[ub + u + 1];
        z = W_IN_LDS ? sw[loc] : wt[base + loc];
        if (q1 > q0) {
            c0 = P.Ucol[q0];
            v0 = P.Uval[q0];
            if (OP != OP_NONE && (c0 & kGlobalBit)) r0 = combine<VT, OP>(C, c0 & ~kGlobalBit);
        }
        if (q1 > q0 + 1) {
            c1 = P.Ucol[q0 + 1];
            v1 = P.Uval[q0 + 1];
            if (OP != OP_NONE && (c1 & kGlobalBit)) r1 = combine<VT, OP>(C, c1 & ~kGlobalBit);
        }
    }
    // kept columns and absent entries read the thread's own slot and are
    // selected away, so the level body has no branches on the entry kinds
    const bool h0 = q1 > q0, h1 = q1 > q0 + 1, more = q1 > q0 + 2;
    const bool g0 = (c0 & kGlobalBit) != 0, g1 = (c1 & kGlobalBit) != 0;
    const uint32_t a0 = g0 ? loc : c0, a1 = g1 ? loc : c1;
    for (uint32_t li = 0; li < nl; li++) {
        if (li == my) {
            const VT s0 = sx[a0], s1 = sx[a1];
            const VT x0 = g0 ? r0 : s0, x1 = g1 ? r1 : s1;
            z = h0 ? msub(z, v0, x0) : z;
            z = h1 ? msub(z, v1, x1) : z;
            if (more)
                for (uint32_t q = q0 + 2; q < q1; q++) {
                    const uint32_t c = P.Ucol[q];
                    const VT xv = (c & kGlobalBit) ? combine<VT, OP>(C, c & ~kGlobalBit) : sx[c];
                    z = msub(z, P.Uval[q], xv);
                }
            z = udiv(z, d);
            sx[loc] = z;
        }
        level_sync<BS>();
    }
    return z;
}

// Sweeps of a block too large for LDS: the same depth levels and per-row
// arithmetic as block_down / block_up, with every value in global memory
// (writes of one depth level are read by the next after the workgroup
// barrier).  global_down leaves w of every node in gw[base + p].
template <class VT, class FT, int IN>
__device__ __forceinline__ void global_down(const ProgView<VT>& P, const uint32_t* B, const InSrc<VT, FT, IN>& f,
                                            VT* gw) {
    const uint32_t base = B[0], S = B[1] & ~kGlobalBlk, nl = B[3] - B[2];
    for (uint32_t p = threadIdx.x; p < S; p += blockDim.x) gw[base + p] = fetch_at(P, f, base + p);
    __syncthreads();
    uint32_t s = 0;
    for (uint32_t li = 0; li < nl; li++) {
        const uint32_t e = P.lvlL[B[2] + li];
        for (uint32_t p = s + threadIdx.x; p < e; p += blockDim.x) {
            VT acc = gw[base + p];
            for (uint32_t q = P.Lptr[base + p]; q < P.Lptr[base + p + 1]; q++) acc = msub(acc, P.Lval[q], gw[base + P.Lloc[q]]);
            gw[base + p] = acc;
        }
        s = e;
        __syncthreads();
    }
}

// U sweep of a global-memory block: w of a row from ws[base + loc], solved
// values into P.gx[base + loc] (ws may be P.gx itself: a row reads its own w
// before it writes its x there, and other rows read only solved columns).
template <class VT, int OP, class CV>
__device__ __forceinline__ void global_up(const ProgView<VT>& P, const uint32_t* B, const VT* ws,
                                          const Combine<VT, CV>& C) {
    VT* gx = P.gx;
    const uint32_t base = B[0], ub = B[10], nl = B[5] - B[4];
    uint32_t s = 0;
    for (uint32_t li = 0; li < nl; li++) {
        const uint32_t e = P.lvlU[B[4] + li];
        for (uint32_t u = s + threadIdx.x; u < e; u += blockDim.x) {
            const uint32_t loc = P.uloc[ub + u];
            VT z = ws[base + loc];
            for (uint32_t q = P.Uptr[ub + u]; q < P.Uptr[ub + u + 1]; q++) {
                const uint32_t c = P.Ucol[q];
                const VT xv = (c & kGlobalBit) ? combine<VT, OP>(C, c & ~kGlobalBit) : gx[base + c];
                z = msub(z, P.Uval[q], xv);
            }
            gx[base + loc] = udiv(z, P.Udiag[ub + u]);
        }
        s = e;
        __syncthreads();
    }
}

template <class XT, int MODE>
__device__ __forceinline__ void put(XT* x, uint32_t j, XT v) {
    if (MODE == 0) x[j] = v;
    else x[j] = x[j] + v;
}

// Loads the block header (kBlkFields words) into registers.
struct BlkHdr {
    uint32_t w[12];
};
__device__ __forceinline__ BlkHdr load_hdr(const uint32_t* blk) {
    BlkHdr h;
    const uint32_t* B = blk + 12 * blockIdx.x;
#pragma unroll
    for (int k = 0; k < 12; k++) h.w[k] = B[k];
    return h;
}

// Forward part of a level: component blocks run the L sweep and write w of
// the eliminated rows (wt) and the right-hand side F of the kept rows they
// feed; the remaining blocks move the right-hand side of the other kept rows
// into F and apply the isolated-node diagonal x[j] = dval * f[j] (DMODE 1
// adds it to x).
template <class VT, class FT, int IN, class XT, int DMODE, int BS, bool BAT, class FV = VT>
__global__ void __launch_bounds__(BS) k_level_down(ProgView<VT> P, InSrc<VT, FT, IN> f, VT* __restrict__ wt,
                                                     FV* __restrict__ F, uint32_t Md, const VT* __restrict__ dval,
                                                     XT* __restrict__ x, const uint32_t* guard, LevelBatch lb) {
    if constexpr (BAT) {                 // batched right-hand sides (compiled out for single solves)
        const size_t y = blockIdx.y;
        f.f = shift(f.f, y * lb.N); f.u = shift(f.u, y * lb.N); f.F = shift(f.F, y * lb.N); f.us = shift(f.us, y * lb.N);
        wt = shift(wt, y * lb.nodes); P.gx = shift(P.gx, y * lb.nodes); F = shift(F, y * lb.n);
        x = shift(x, y * lb.N); guard = shift(guard, y * lb.st);
    }
    // the solve's done flag: loaded now, tested before the first store, so that
    // its latency overlaps the block's own loads (a converged iteration only
    // writes the apply's scratch, never an output)
    const uint32_t gflag = guard ? *guard : 0u;
    if (blockIdx.x < P.nblk) {
        extern __shared__ __align__(16) unsigned char lds_raw[];
        const LdsPlan lp = lds_plan<VT>(P.max_nodes, P.max_Lent, P.max_Uent, P.max_Unodes, P.max_levels, true, false);
        const BlkHdr h = load_hdr(P.blk);
        const uint32_t base = h.w[0], S = h.w[1] & ~kGlobalBlk;
        if (h.w[1] & kGlobalBlk) {
            global_down(P, h.w, f, wt);
            if (gflag) return;
            for (uint32_t p = threadIdx.x; p < S; p += blockDim.x) {
                const uint32_t o = P.out[base + p];
                if (o & kGlobalBit) F[o & ~kGlobalBit] = FV(wt[base + p]);
            }
            return;
        }
        // the store targets of this thread's first node, fetched ahead of the sweep
        const uint32_t o0 = threadIdx.x < S ? P.out[base + threadIdx.x] : 0u;
        if (S <= blockDim.x) {
            const VT v = owned_down<BS>(P, h.w, f, reinterpret_cast<VT*>(lds_raw + lp.sw));
            if (gflag) return;
            if (threadIdx.x < S) {
                if (o0 & kGlobalBit) F[o0 & ~kGlobalBit] = FV(v);
                else wt[base + threadIdx.x] = v;
            }
            return;
        }
        block_down(P, h.w, f, lds_raw, lp);
        if (gflag) return;
        const VT* sw = reinterpret_cast<const VT*>(lds_raw + lp.sw);
        for (uint32_t p = threadIdx.x; p < S; p += blockDim.x) {
            const uint32_t o = p == threadIdx.x ? o0 : P.out[base + p];
            if (o & kGlobalBit) F[o & ~kGlobalBit] = FV(sw[p]);
            else wt[base + p] = sw[p];
        }
    } else {
        const uint32_t i = (blockIdx.x - P.nblk) * blockDim.x + threadIdx.x;
        const uint32_t q = P.nodes + i;
        if (i < P.nflat) {
            const VT v = fetch_at(P, f, q);
            if (!gflag) F[P.flatq[i]] = FV(v);
        } else if (i < P.nflat + Md) {
            const uint32_t j = P.qrow[q];
            const XT v = XT(dval[i - P.nflat] * fetch(f, q, j));
            if (!gflag) put<XT, DMODE>(x, j, v);
        }
    }
}

// Backward part of a level: component blocks run the U sweep (kept columns
// from the combine op) and write their rows of x; the remaining blocks write
// the kept rows x[map[M+r]] = combine(r) (UMODE 1 adds into x).
template <class VT, class XT, int OP, int UMODE, int BS, bool BAT, class CV = VT>
__global__ void __launch_bounds__(BS) k_level_up(ProgView<VT> P, const VT* __restrict__ wt, Combine<VT, CV> C,
                                                   const uint32_t* __restrict__ mapM, uint32_t n, XT* __restrict__ x,
                                                   const uint32_t* guard, LevelBatch lb) {
    if constexpr (BAT) {                 // batched right-hand sides (compiled out for single solves)
        const size_t y = blockIdx.y;
        wt = shift(wt, y * lb.nodes); P.gx = shift(P.gx, y * lb.nodes);
        C.x2 = shift(C.x2, y * lb.n); C.u1 = shift(C.u1, y * lb.n); C.u0 = shift(C.u0, y * lb.n);
        x = shift(x, y * lb.N); guard = shift(guard, y * lb.st);
    }
    // the solve's done flag: loaded now, tested before the first store, so that
    // its latency overlaps the block's own loads (a converged iteration only
    // writes the apply's scratch, never an output)
    const uint32_t gflag = guard ? *guard : 0u;
    if (blockIdx.x < P.nblk) {
        extern __shared__ __align__(16) unsigned char lds_raw[];
        const LdsPlan lp = lds_plan<VT>(P.max_nodes, P.max_Lent, P.max_Uent, P.max_Unodes, P.max_levels, false, true);
        const BlkHdr h = load_hdr(P.blk);
        const uint32_t ub = h.w[10], nu = h.w[11];
        if (h.w[1] & kGlobalBlk) {
            global_up<VT, OP>(P, h.w, wt, C);
            if (gflag) return;
            for (uint32_t u = threadIdx.x; u < nu; u += blockDim.x)
                put<XT, UMODE>(x, P.uxidx[ub + u], XT(P.gx[h.w[0] + P.uloc[ub + u]]));
            return;
        }
        const uint32_t x0 = threadIdx.x < nu ? P.uxidx[ub + threadIdx.x] : 0u;
        if (nu <= blockDim.x) {
            const VT v = owned_up<BS, VT, OP, false>(P, h.w, wt, C, nullptr, reinterpret_cast<VT*>(lds_raw + lp.sx));
            if (gflag) return;
            if (threadIdx.x < nu) put<XT, UMODE>(x, x0, XT(v));
            return;
        }
        block_up<VT, OP, false>(P, h.w, wt, C, lds_raw, lp);
        if (gflag) return;
        const VT* sz = reinterpret_cast<const VT*>(lds_raw + lp.sz);
        for (uint32_t u = threadIdx.x; u < nu; u += blockDim.x)
            put<XT, UMODE>(x, u == threadIdx.x ? x0 : P.uxidx[ub + u], XT(sz[u]));
    } else {
        const uint32_t r = (blockIdx.x - P.nblk) * blockDim.x + threadIdx.x;
        if (r < n) {
            const XT v = XT(combine<VT, OP>(C, r));
            if (!gflag) put<XT, UMODE>(x, mapM[r], v);
        }
    }
}

// A whole level without kept rows (the coarsest level): L sweep, U sweep and
// diagonal in one launch.  MODE 1 adds the result into x.
template <class VT, class FT, int IN, class XT, int MODE, int BS, bool BAT>
__global__ void __launch_bounds__(BS) k_level_solve(ProgView<VT> P, InSrc<VT, FT, IN> f, uint32_t Md,
                                                      const VT* __restrict__ dval, XT* __restrict__ x,
                                                      const uint32_t* guard, LevelBatch lb) {
    if constexpr (BAT) {                 // batched right-hand sides (compiled out for single solves)
        const size_t y = blockIdx.y;
        f.f = shift(f.f, y * lb.N); f.u = shift(f.u, y * lb.N); f.F = shift(f.F, y * lb.N); f.us = shift(f.us, y * lb.N);
        P.gx = shift(P.gx, y * lb.nodes); x = shift(x, y * lb.N); guard = shift(guard, y * lb.st);
    }
    // the solve's done flag: loaded now, tested before the first store, so that
    // its latency overlaps the block's own loads (a converged iteration only
    // writes the apply's scratch, never an output)
    const uint32_t gflag = guard ? *guard : 0u;
    if (blockIdx.x < P.nblk) {
        extern __shared__ __align__(16) unsigned char lds_raw[];
        const LdsPlan lp = lds_plan<VT>(P.max_nodes, P.max_Lent, P.max_Uent, P.max_Unodes, P.max_levels, true, true);
        const BlkHdr h = load_hdr(P.blk);
        const uint32_t ub = h.w[10], nu = h.w[11];
        if (h.w[1] & kGlobalBlk) {
            Combine<VT> none;
            global_down(P, h.w, f, P.gx);
            global_up<VT, OP_NONE>(P, h.w, P.gx, none);
            if (gflag) return;
            for (uint32_t u = threadIdx.x; u < nu; u += blockDim.x)
                put<XT, MODE>(x, P.uxidx[ub + u], XT(P.gx[h.w[0] + P.uloc[ub + u]]));
            return;
        }
        const uint32_t x0 = threadIdx.x < nu ? P.uxidx[ub + threadIdx.x] : 0u;
        if (h.w[1] <= blockDim.x && nu <= blockDim.x) {
            VT* sw = reinterpret_cast<VT*>(lds_raw + lp.sw);
            owned_down<BS>(P, h.w, f, sw);
            Combine<VT> none;
            const VT v = owned_up<BS, VT, OP_NONE, true>(P, h.w, nullptr, none, sw,
                                                         reinterpret_cast<VT*>(lds_raw + lp.sx));
            if (gflag) return;
            if (threadIdx.x < nu) put<XT, MODE>(x, x0, XT(v));
            return;
        }
        block_down(P, h.w, f, lds_raw, lp);
        Combine<VT> none;
        block_up<VT, OP_NONE, true>(P, h.w, nullptr, none, lds_raw, lp);
        if (gflag) return;
        const VT* sz = reinterpret_cast<const VT*>(lds_raw + lp.sz);
        for (uint32_t u = threadIdx.x; u < nu; u += blockDim.x)
            put<XT, MODE>(x, u == threadIdx.x ? x0 : P.uxidx[ub + u], XT(sz[u]));
    } else {
        const uint32_t k = (blockIdx.x - P.nblk) * blockDim.x + threadIdx.x;
        if (k < Md) {
            const uint32_t q = P.nodes + P.nflat + k;
            const uint32_t j = P.qrow[q];
            const XT v = XT(dval[k] * fetch(f, q, j));
            if (!gflag) put<XT, MODE>(x, j, v);
        }
    }
}

// ---------------------------------------------------------------------------
// fused last pair: a uh_cheb level l and the coarsest level l+1 (plan.h:
// PairProgram).  Two launches per visit of level l with two Chebyshev steps
// (cheb_solve.cc:33-104 on level l, :108-122 on level l+1); every row keeps
// the arithmetic of k_level_down / k_level_solve / k_level_up.
// ---------------------------------------------------------------------------
// VT: level l's values; CT: the coarsest level's (FP64 in every hierarchy, so
// that an FP32 hierarchy still solves its coarsest level in FP64).
template <class VT, class CT> struct PairView {
    uint32_t nblk = 0, nodes = 0, nodes2 = 0, nflat2 = 0, Md = 0;
    const uint32_t* blk = nullptr;
    const uint32_t* qrow = nullptr;
    const uint32_t *Ldep = nullptr, *Lptr = nullptr, *Lloc = nullptr, *kept2 = nullptr;
    const VT* Lval = nullptr;
    const uint32_t *Udep = nullptr, *uloc = nullptr, *uxidx = nullptr, *Uptr = nullptr, *Ucol = nullptr;
    const VT *Uval = nullptr, *Udiag = nullptr;
    const uint32_t *src2 = nullptr, *row2 = nullptr, *xl2 = nullptr, *Ldep2 = nullptr, *Lptr2 = nullptr,
                   *Lloc2 = nullptr;
    const CT *Lval2 = nullptr, *d2 = nullptr;
    const uint32_t *Udep2 = nullptr, *uloc2 = nullptr, *Uptr2 = nullptr, *Ucol2 = nullptr;
    const CT *Uval2 = nullptr, *Udiag2 = nullptr;
    const uint32_t *frow2 = nullptr, *fxl = nullptr;
    const CT* fd2 = nullptr;
    const VT* dval = nullptr;          // level l's isolated rows
};
constexpr uint32_t kPairHdr = 12;          // header words per block of a pair program
constexpr uint32_t kPairIso2 = 0x80000000u;
constexpr uint32_t kPairNoRef = 0xffffffffu;

// Per-thread data of an owned-node L sweep (as owned_down): thread p owns
// node p of the n nodes at base.  Loaded at kernel start, so that the loads
// overlap whatever runs before the sweep.
template <class VT> struct OwnL {
    uint32_t my, q0, q1, l0, l1;
    VT v0, v1;
};
template <class VT>
__device__ __forceinline__ OwnL<VT> own_L(uint32_t n, uint32_t nl, uint32_t base, const uint32_t* __restrict__ Ldep,
                                          const uint32_t* __restrict__ Lptr, const uint32_t* __restrict__ Lloc,
                                          const VT* __restrict__ Lval) {
    const uint32_t p = threadIdx.x;
    OwnL<VT> o{nl, 0u, 0u, p, p, VT(0), VT(0)};
    if (p < n) {
        o.my = Ldep[base + p];
        o.q0 = Lptr[base + p];
        o.q1 = Lptr[base + p + 1];
        if (o.q1 > o.q0) { o.l0 = Lloc[o.q0]; o.v0 = Lval[o.q0]; }
        if (o.q1 > o.q0 + 1) { o.l1 = Lloc[o.q0 + 1]; o.v1 = Lval[o.q0 + 1]; }
    }
    return o;
}
// The sweep: at its depth level a node subtracts its L entries from acc;
// sw holds every node's w afterwards.
template <int BS, class VT>
__device__ __forceinline__ VT sweep_L(const OwnL<VT>& o, uint32_t nl, const uint32_t* __restrict__ Lloc,
                                      const VT* __restrict__ Lval, VT acc, VT* sw) {
    const bool h0 = o.q1 > o.q0, h1 = o.q1 > o.q0 + 1, more = o.q1 > o.q0 + 2;
    for (uint32_t li = 0; li < nl; li++) {
        if (li == o.my) {
            const VT x0 = sw[o.l0], x1 = sw[o.l1];
            acc = h0 ? msub(acc, o.v0, x0) : acc;
            acc = h1 ? msub(acc, o.v1, x1) : acc;
            if (more)
                for (uint32_t q = o.q0 + 2; q < o.q1; q++) acc = msub(acc, Lval[q], sw[Lloc[q]]);
            sw[threadIdx.x] = acc;
        }
        level_sync<BS>();
    }
    return acc;
}

// Per-thread data of an owned-node U sweep (as owned_up) over the U-list
// entries [ub, ub + nu): columns are local, or (KEPT, flagged kGlobalBit)
// kept rows whose values come from an LDS array at the sweep.
template <class VT> struct OwnU {
    uint32_t my, q0, q1, loc, c0, c1;
    VT d, v0, v1;
};
template <class VT>
__device__ __forceinline__ OwnU<VT> own_U(uint32_t nu, uint32_t nl, uint32_t ub, const uint32_t* __restrict__ Udep,
                                          const uint32_t* __restrict__ uloc, const VT* __restrict__ Udiag,
                                          const uint32_t* __restrict__ Uptr, const uint32_t* __restrict__ Ucol,
                                          const VT* __restrict__ Uval) {
    const uint32_t u = threadIdx.x;
    OwnU<VT> o{nl, 0u, 0u, 0u, 0u, 0u, VT(1), VT(0), VT(0)};
    if (u < nu) {
        o.my = Udep[ub + u];
        o.loc = uloc[ub + u];
        o.d = Udiag[ub + u];
        o.q0 = Uptr[ub + u];
        o.q1 = Uptr[ub + u + 1];
        if (o.q1 > o.q0) { o.c0 = Ucol[o.q0]; o.v0 = Uval[o.q0]; }
        if (o.q1 > o.q0 + 1) { o.c1 = Ucol[o.q0 + 1]; o.v1 = Uval[o.q0 + 1]; }
    }
    return o;
}
// The sweep from the row's w z; solved values land in sx[loc].
template <int BS, class VT, bool KEPT>
__device__ __forceinline__ VT sweep_U(const OwnU<VT>& o, uint32_t nl, const uint32_t* __restrict__ Ucol,
                                      const VT* __restrict__ Uval, VT z, const VT* kept, VT* sx) {
    const bool h0 = o.q1 > o.q0, h1 = o.q1 > o.q0 + 1, more = o.q1 > o.q0 + 2;
    const bool g0 = KEPT && h0 && (o.c0 & kGlobalBit) != 0, g1 = KEPT && h1 && (o.c1 & kGlobalBit) != 0;
    VT r0 = VT(0), r1 = VT(0);
    if (KEPT) {
        if (g0) r0 = kept[o.c0 & ~kGlobalBit];
        if (g1) r1 = kept[o.c1 & ~kGlobalBit];
    }
    const uint32_t a0 = (g0 || !h0) ? o.loc : o.c0, a1 = (g1 || !h1) ? o.loc : o.c1;
    for (uint32_t li = 0; li < nl; li++) {
        if (li == o.my) {
            const VT s0 = sx[a0], s1 = sx[a1];
            const VT x0 = g0 ? r0 : s0, x1 = g1 ? r1 : s1;
            z = h0 ? msub(z, o.v0, x0) : z;
            z = h1 ? msub(z, o.v1, x1) : z;
            if (more)
                for (uint32_t q = o.q0 + 2; q < o.q1; q++) {
                    const uint32_t c = Ucol[q];
                    VT xv;
                    if (KEPT && (c & kGlobalBit)) xv = kept[c & ~kGlobalBit];
                    else xv = sx[c];
                    z = msub(z, Uval[q], xv);
                }
            z = udiv(z, o.d);
            sx[o.loc] = z;
        }
        level_sync<BS>();
    }
    return z;
}

// K1: level l's L sweep, then step 1 on level l+1 for the kept rows of the
// block: writes w of level l's eliminated rows (wt), F of every kept row (at
// its level-(l+1) consumption position), u1 = B_{l+1} F (level-(l+1) rows)
// and x = dval f of level l's isolated rows.
template <class VT, class CT, class FT, int IN, class XT, int BS, bool BAT>
__global__ void __launch_bounds__(BS) k_pair_down(PairView<VT, CT> P, InSrc<VT, FT, IN> f, VT* __restrict__ wt,
                                                    VT* __restrict__ F, VT* __restrict__ u1, XT* __restrict__ x,
                                                    const uint32_t* guard, LevelBatch lb) {
    if constexpr (BAT) {
        const size_t y = blockIdx.y;
        f.f = shift(f.f, y * lb.N); f.u = shift(f.u, y * lb.N); f.F = shift(f.F, y * lb.N); f.us = shift(f.us, y * lb.N);
        wt = shift(wt, y * lb.nodes); F = shift(F, y * lb.n); u1 = shift(u1, y * lb.n);
        x = shift(x, y * lb.N); guard = shift(guard, y * lb.st);
    }
    const uint32_t gflag = guard ? *guard : 0u;
    if (blockIdx.x < P.nblk) {
        __shared__ VT sw[BS];
        __shared__ CT s2w[BS], s2x[BS];
        const uint32_t* B = P.blk + kPairHdr * blockIdx.x;
        uint32_t h[kPairHdr];
#pragma unroll
        for (int k = 0; k < int(kPairHdr); k++) h[k] = B[k];
        const uint32_t base = h[0], S = h[1], base2 = h[6], nk = h[7];
        const uint32_t t = threadIdx.x;
        // every per-node load of both levels first, so that they overlap
        const OwnL<VT> oa = own_L(S, h[2], base, P.Ldep, P.Lptr, P.Lloc, P.Lval);
        const OwnL<CT> ob = own_L(nk, h[8], base2, P.Ldep2, P.Lptr2, P.Lloc2, P.Lval2);
        const OwnU<CT> ou = own_U(h[10], h[11], h[9], P.Udep2, P.uloc2, P.Udiag2, P.Uptr2, P.Ucol2, P.Uval2);
        uint32_t src = 0, row = 0;
        CT dd = CT(0);
        if (t < nk) {
            src = P.src2[base2 + t];
            row = P.row2[base2 + t];
            dd = P.d2[base2 + t];
        }
        const uint32_t vrow = t < h[10] ? P.row2[base2 + ou.loc] & ~kPairIso2 : 0u;
        const uint32_t k2 = t < S ? P.kept2[base + t] : kPairNoRef;
        VT acc = VT(0);
        if (t < S) acc = fetch(f, base + t, P.qrow[base + t]);
        const VT w = sweep_L<BS>(oa, h[2], P.Lloc, P.Lval, acc, sw);
        if (!gflag && t < S && k2 == kPairNoRef) wt[base + t] = w;
        // step 1 on level l+1: its right-hand side is the w of the block's kept rows
        const VT fv = t < nk ? sw[src] : VT(0);
        if (!gflag && t < nk) F[base2 + t] = fv;
        const CT f2 = CT(fv);
        sweep_L<BS>(ob, h[8], P.Lloc2, P.Lval2, f2, s2w);
        const CT z = t < h[10] ? s2w[ou.loc] : CT(0);
        const CT v = sweep_U<BS, CT, false>(ou, h[11], P.Ucol2, P.Uval2, z, (const CT*)nullptr, s2x);
        if (gflag) return;
        if (t < h[10]) u1[vrow] = VT(v);
        if (t < nk && (row & kPairIso2)) u1[row & ~kPairIso2] = VT(dd * f2);
    } else {
        const uint32_t i = (blockIdx.x - P.nblk) * BS + threadIdx.x;
        if (i < P.nflat2) {
            const uint32_t q = P.nodes + i;
            const VT v = fetch(f, q, P.qrow[q]);
            if (!gflag) {
                F[P.nodes2 + i] = v;
                u1[P.frow2[i]] = VT(P.fd2[i] * CT(v));
            }
        } else if (i < P.nflat2 + P.Md) {
            const uint32_t k = i - P.nflat2, q = P.nodes + i, j = P.qrow[q];
            const XT v = XT(P.dval[k] * fetch(f, q, j));
            if (!gflag) x[j] = v;
        }
    }
}

// step 2's combine of a kept row: x2 = (-alpha) x2 + (1+beta) u1s (cheb_solve.cc:79-81), as combine<OP_CHEB2>
template <class VT>
__device__ __forceinline__ VT pair_cheb2(VT x2, VT u, double ma, int skip, double opb, int opb_one) {
    const double v = skip ? double(x2) : double(x2) * ma;
    const double w = double(u);
    return VT(opb_one ? v + w : v + opb * w);
}

// K2: step 2 on level l+1 from the residual A_{l+1} (s u1) - F (A2 stored in
// level (l+1)'s consumption order), the Chebyshev combine of step 2
// (cheb_solve.cc:79-81) for every kept row, and level l's U sweep.
template <class VT, class CT, class XT, bool SCALE, int BS, bool BAT>
__global__ void __launch_bounds__(BS) k_pair_up(PairView<VT, CT> P, SellView<CT> A2, const VT* __restrict__ u1, double s,
                                                  const VT* __restrict__ F, const VT* __restrict__ wt, double ma,
                                                  int skip, double opb, int opb_one, XT* __restrict__ x,
                                                  const uint32_t* guard, LevelBatch lb) {
    if constexpr (BAT) {
        const size_t y = blockIdx.y;
        u1 = shift(u1, y * lb.n); F = shift(F, y * lb.n); wt = shift(wt, y * lb.nodes);
        x = shift(x, y * lb.N); guard = shift(guard, y * lb.st);
    }
    const uint32_t gflag = guard ? *guard : 0u;
    if (blockIdx.x < P.nblk) {
        __shared__ CT s2w[BS], s2x[BS];
        __shared__ VT s2c[BS], sx[BS];
        const uint32_t* B = P.blk + kPairHdr * blockIdx.x;
        uint32_t h[kPairHdr];
#pragma unroll
        for (int k = 0; k < int(kPairHdr); k++) h[k] = B[k];
        const uint32_t base = h[0], base2 = h[6], nk = h[7];
        const uint32_t t = threadIdx.x;
        // every per-node load first: level l+1's sweeps and level l's U sweep with its w
        const OwnL<CT> ob = own_L(nk, h[8], base2, P.Ldep2, P.Lptr2, P.Lloc2, P.Lval2);
        const OwnU<CT> o2 = own_U(h[10], h[11], h[9], P.Udep2, P.uloc2, P.Udiag2, P.Uptr2, P.Ucol2, P.Uval2);
        const OwnU<VT> oa = own_U(h[4], h[5], h[3], P.Udep, P.uloc, P.Udiag, P.Uptr, P.Ucol, P.Uval);
        const VT za = t < h[4] ? wt[base + oa.loc] : VT(0);
        const uint32_t xa = t < h[4] ? P.uxidx[h[3] + t] : 0u;
        uint32_t row = 0, xl = 0;
        CT dd = CT(0), f2 = CT(0), us = CT(0);
        if (t < nk) {
            row = P.row2[base2 + t];
            xl = P.xl2[base2 + t];
            dd = P.d2[base2 + t];
            const uint32_t r = row & ~kPairIso2;
            f2 = sell_row<CT, VT, SCALE>(A2, u1, base2 + t, s) - CT(F[base2 + t]);
            us = SCALE ? CT(double(u1[r]) * s) : CT(u1[r]);
        }
        sweep_L<BS>(ob, h[8], P.Lloc2, P.Lval2, f2, s2w);
        if (t < nk && (row & kPairIso2)) s2x[t] = dd * f2;
        const CT z2 = t < h[10] ? s2w[o2.loc] : CT(0);
        sweep_U<BS, CT, false>(o2, h[11], P.Ucol2, P.Uval2, z2, (const CT*)nullptr, s2x);
        // the value of every kept row: the combine of step 2
        VT c = VT(0);
        if (t < nk) {
            c = VT(pair_cheb2(s2x[t], us, ma, skip, opb, opb_one));
            s2c[t] = c;
        }
        __syncthreads();
        const VT v = sweep_U<BS, VT, true>(oa, h[5], P.Ucol, P.Uval, za, (const VT*)s2c, sx);
        if (gflag) return;
        if (t < nk) x[xl] = XT(c);
        if (t < h[4]) x[xa] = XT(v);
    } else {
        const uint32_t i = (blockIdx.x - P.nblk) * BS + threadIdx.x;
        if (i < P.nflat2) {
            const uint32_t q = P.nodes2 + i, r = P.frow2[i];
            const CT f2 = sell_row<CT, VT, SCALE>(A2, u1, q, s) - CT(F[q]);
            const CT x2 = P.fd2[i] * f2;
            const CT us = SCALE ? CT(double(u1[r]) * s) : CT(u1[r]);
            const XT v = XT(VT(pair_cheb2(x2, us, ma, skip, opb, opb_one)));
            if (!gflag) x[P.fxl[i]] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// chain of uh_cheb levels a..c (plan.h: ChainProgram): descent and ascent
// launches of Prec::solve (cheb_solve.cc:13-123).  BS threads per block: one
// wavefront (64; depth levels end with a wave barrier) or two (128, for
// hierarchies with components of 65..128 nodes; depth levels end with a
// workgroup barrier).  Thread t owns slot t of every level (node slots are the
// first ones, so the owned-node sweeps above apply); slots t + BS (rows without
// L entries of large components, at most 2 BS slots per level) are handled in a
// second pass.  Every row keeps the arithmetic of k_level_down / k_level_up /
// k_pair_down / k_pair_up.
// ---------------------------------------------------------------------------
constexpr int kChainMax = 4;               // levels a..c-1 of one chain
constexpr uint32_t kChainBS = 64;          // block size of the one-wavefront chain; the wide chain has 2 kChainBS
constexpr uint32_t kChainHdrW = 8;         // header words per block and level

// A block-uniform word of program data (never written by the kernel), read
// through the scalar cache so that it occupies a scalar register.
__device__ __forceinline__ uint32_t uload(const uint32_t* p) {
    return *(const __attribute__((address_space(4))) uint32_t*)p;
}

template <class T> struct ChainLv {
    const uint32_t *hdr = nullptr, *qrow = nullptr, *src = nullptr, *xl = nullptr;
    const uint32_t *Lmeta = nullptr, *Lptr = nullptr, *Lloc = nullptr;
    const uint32_t *Umeta = nullptr, *uxidx = nullptr, *Uptr = nullptr, *Ucol = nullptr;
    const T *Lv2 = nullptr, *Lval = nullptr, *Uv1 = nullptr, *Uval = nullptr, *Udiag = nullptr, *dval = nullptr;
    uint32_t nslots = 0, npt = 0;
};

// A chain node's sweep data from its packed word (plan.h: ChainLevelProgram::Lmeta / Umeta) and inline
// values: one round of loads with no dependence between them.  Entries past the first two (L) or the first
// (U) are read through Lptr / Uptr, which only such rows load.
// both inline values of a node in one load (the arrays are 256-byte aligned, a node's pair at 2 x its slot)
__device__ __forceinline__ void load_pair(const double* p, double& a, double& b) {
    const dbl2 v = *reinterpret_cast<const dbl2*>(p);
    a = v.x;
    b = v.y;
}
__device__ __forceinline__ void load_pair(const float* p, float& a, float& b) {
    typedef float flt2 __attribute__((ext_vector_type(2)));
    const flt2 v = *reinterpret_cast<const flt2*>(p);
    a = v.x;
    b = v.y;
}
// A chain node's L row: entries 0 and 1 from the packed word, entry 2 (rows of three entries occur on every
// level of the SPE10 hierarchies) loaded with them, so that the sweep reads no global memory for rows of up to
// three entries; entries past the third through Lloc / Lval at the sweep.
template <class VT> struct ChainOwnL {
    OwnL<VT> o;
    uint32_t l2;
    VT v2;
};
// (the values load with the word, unconditionally: an absent entry's value is stored as 0 and selected away by
// the sweep's entry count, as in own_L / own_U)
template <class VT>
__device__ __forceinline__ ChainOwnL<VT> chain_own_L(uint32_t n, uint32_t nl, uint32_t base, const ChainLv<VT>& L) {
    const uint32_t p = threadIdx.x;
    ChainOwnL<VT> e{OwnL<VT>{nl, 0u, 0u, p, p, VT(0), VT(0)}, p, VT(0)};
    OwnL<VT>& o = e.o;
    if (p < n) {
        const uint32_t m = L.Lmeta[base + p], q0 = L.Lptr[base + p];
        load_pair(L.Lv2 + 2 * size_t(base + p), o.v0, o.v1);
        const uint32_t c = (m >> 8) & 0xffu;
        o.my = m & 0xffu;
        o.q0 = q0;
        o.q1 = q0 + c;
        if (c > 0) o.l0 = (m >> 16) & 0xffu;
        if (c > 1) o.l1 = m >> 24;
        if (c > 2) {
            e.l2 = L.Lloc[q0 + 2];
            e.v2 = L.Lval[q0 + 2];
        }
    }
    return e;
}
// sweep_L for a chain node: its entries in storage order, the first three from registers
template <int BS, class VT>
__device__ __forceinline__ VT chain_sweep_L(const ChainOwnL<VT>& e, uint32_t nl, const uint32_t* __restrict__ Lloc,
                                            const VT* __restrict__ Lval, VT acc, VT* sw) {
    const OwnL<VT>& o = e.o;
    const bool h0 = o.q1 > o.q0, h1 = o.q1 > o.q0 + 1, h2 = o.q1 > o.q0 + 2, more = o.q1 > o.q0 + 3;
    for (uint32_t li = 0; li < nl; li++) {
        if (li == o.my) {
            const VT x0 = sw[o.l0], x1 = sw[o.l1];
            acc = h0 ? msub(acc, o.v0, x0) : acc;
            acc = h1 ? msub(acc, o.v1, x1) : acc;
            if (h2) {
                acc = msub(acc, e.v2, sw[e.l2]);
                if (more)
                    for (uint32_t q = o.q0 + 3; q < o.q1; q++) acc = msub(acc, Lval[q], sw[Lloc[q]]);
            }
            sw[threadIdx.x] = acc;
        }
        level_sync<BS>();
    }
    return acc;
}
template <class VT>
__device__ __forceinline__ OwnU<VT> chain_own_U(uint32_t nu, uint32_t nl, uint32_t ub, const ChainLv<VT>& L) {
    const uint32_t u = threadIdx.x;
    OwnU<VT> o{nl, 0u, 0u, 0u, 0u, 0u, VT(1), VT(0), VT(0)};
    if (u < nu) {
        const uint32_t m = L.Umeta[ub + u];
        o.d = L.Udiag[ub + u];
        o.v0 = L.Uv1[ub + u];
        const uint32_t c = (m >> 8) & 0xffu;
        o.my = m & 0xffu;
        o.loc = (m >> 16) & 0xffu;
        o.q0 = c > 1 ? L.Uptr[ub + u] : 0u;
        o.q1 = o.q0 + c;
        if (c > 0) o.c0 = m >> 24;
        if (c > 1) { o.c1 = L.Ucol[o.q0 + 1]; o.v1 = L.Uval[o.q0 + 1]; }
    }
    return o;
}
// a chain level's vectors as the parent of the next level (cheb_prec.h:36), and
// the combine of its second Chebyshev step (cheb_solve.cc:79-81)
template <class VT> struct ChainVec {
    VT *F = nullptr, *u1 = nullptr, *x2 = nullptr, *wt = nullptr;
    double a1 = 1.0;                       // the first step's scaling of u1 (cheb_solve.cc:67-68)
    int s1_skip = 1;                       // a1 == 1: u1 unscaled
    double ma2 = 0.0, opb2 = 0.0;
    int s2_skip = 0, opb2_one = 0;
    uint32_t n = 0, N = 0;                 // batched strides: the next level's size, this level's
};
// u1 as step 2's combine reads it: a1 u1 when the first step scales, formed here with the same single product
// (rounded to VT) that IN_RES_SCALE stores, so that no pass writes it out
template <class VT>
__device__ __forceinline__ VT chain_u1s(const ChainVec<VT>& V, size_t k) {
    const VT u = V.u1[k];
    return V.s1_skip ? u : VT(double(u) * V.a1);
}

template <class VT, class CT> struct ChainView {
    ChainLv<VT> L[kChainMax];
    ChainVec<VT> V[kChainMax];
    ChainLv<CT> C;                         // the coarsest level
    SellView<CT> A2;                       // its matrix, rows in its chain order (step 2's residual)
    double a1c = 1.0;                      // level c-1's first-step scaling
    uint32_t nv = 0;                       // levels a..c-1
    uint32_t st = 0;                       // solve-state stride, in words (batched guards)
};

// The j-th level one chain launch runs (j < NL): with NL = kChainMax the loops
// run over every chain level with compile-time indices and skip those outside
// i0..nv-1; with NL = 1 the launch runs level i0 = nv-1 alone (run-time index).
// chain_top: level i0.
template <int NL>
__device__ __forceinline__ uint32_t chain_level(int j, uint32_t i0) {
    static_assert(NL == 1 || NL == kChainMax, "a chain launch runs the last level alone or any levels");
    return NL == kChainMax ? uint32_t(j) : i0 + uint32_t(j);
}
template <int NL>
__device__ __forceinline__ bool chain_skip(uint32_t i, uint32_t i0, uint32_t nv) {
    return (NL == kChainMax && i < i0) || i >= nv;
}
template <int NL>
__device__ __forceinline__ bool chain_top(uint32_t i, uint32_t i0, int j) {
    return NL == kChainMax ? i == i0 : j == 0;
}
__device__ __forceinline__ uint32_t chain_next(uint32_t i) { return i + 1 < uint32_t(kChainMax) ? i + 1 : i; }

// Descent from chain level i0: level i0 reads its input from f, levels
// i0+1..nv-1 (first visits) the F of the level above from LDS, then step 1 on
// the coarsest level.  Writes F of every level below i0, wt of every level,
// the isolated rows' x (level i0's into out0, the others' into the parent's
// u1) and u1 of the coarsest level.  NL >= nv - i0 bounds the chain levels
// one launch runs: its per-level registers are sized for NL levels (a descent
// from the last chain level, NL = 1, keeps one level's node data live beside
// the residual row instead of kChainMax levels').
template <class VT, class CT, class FT, int IN, class XT, bool BAT, int BS, int NL>
__global__ void __launch_bounds__(BS) k_chain_down(ChainView<VT, CT> V, uint32_t i0, InSrc<VT, FT, IN> f,
                                                           XT* __restrict__ out0, uint32_t Ntop, uint32_t nblk,
                                                           const uint32_t* guard) {
    const size_t y = BAT ? blockIdx.y : 0;
    if constexpr (BAT) {
        f.f = shift(f.f, y * Ntop); f.u = shift(f.u, y * Ntop); f.F = shift(f.F, y * Ntop); f.us = shift(f.us, y * Ntop);
        out0 = shift(out0, y * Ntop); guard = shift(guard, y * V.st);
    }
    const uint32_t gflag = guard ? *guard : 0u;
    const uint32_t nv = V.nv, t = threadIdx.x;
    // the last level's vectors (the coarsest level's parent)
    VT *pF = nullptr, *pu1 = nullptr;
#pragma unroll
    for (int i = 0; i < kChainMax; i++)
        if (uint32_t(i) + 1 == nv) { pF = V.V[i].F + y * V.V[i].n; pu1 = V.V[i].u1 + y * V.V[i].n; }
    if (blockIdx.x < nblk) {
        __shared__ VT sv[2][2 * BS];
        __shared__ CT cw[2 * BS], cx[2 * BS];
        const uint32_t b = blockIdx.x;
        // every level's per-slot loads first, so that they overlap
        uint32_t base[NL], S[NL], nn[NL], nkeep[NL], nlL[NL], qr[NL], sr[NL];
        VT dv[NL];
        ChainOwnL<VT> ol[NL];
        VT acc = VT(0);
#pragma unroll
        for (int j = 0; j < NL; j++) {
            const uint32_t i = chain_level<NL>(j, i0);
            base[j] = S[j] = nn[j] = nkeep[j] = nlL[j] = qr[j] = sr[j] = 0;
            dv[j] = VT(0);
            ol[j] = ChainOwnL<VT>{OwnL<VT>{0u, 0u, 0u, t, t, VT(0), VT(0)}, t, VT(0)};
            if (chain_skip<NL>(i, i0, nv)) continue;
            const ChainLv<VT>& L = V.L[i];
            const uint32_t* H = L.hdr + kChainHdrW * b;
            base[j] = uload(H); S[j] = uload(H + 1); nn[j] = uload(H + 2); nkeep[j] = nn[j] + uload(H + 3);
            nlL[j] = uload(H + 4);
            ol[j] = chain_own_L(nn[j], nlL[j], base[j], L);
            if (t < S[j]) {
                qr[j] = L.qrow[base[j] + t];
                if (!chain_top<NL>(i, i0, j)) sr[j] = L.src[base[j] + t];
                if (t >= nkeep[j]) dv[j] = L.dval[base[j] + t];
                if (chain_top<NL>(i, i0, j)) acc = fetch(f, base[j] + t, qr[j]);
            }
        }
        const ChainLv<CT>& C = V.C;
        const uint32_t* Hc = C.hdr + kChainHdrW * b;
        const uint32_t baseC = uload(Hc), SC = uload(Hc + 1), nnC = uload(Hc + 2), nlLC = uload(Hc + 4),
                       ubC = uload(Hc + 5), nuC = uload(Hc + 6), nlUC = uload(Hc + 7);
        const ChainOwnL<CT> olC = chain_own_L(nnC, nlLC, baseC, C);
        const OwnU<CT> ouC = chain_own_U(nuC, nlUC, ubC, C);
        uint32_t qrC = 0, srC = 0, uxC = 0;
        CT dvC = CT(0);
        if (t < SC) {
            qrC = C.qrow[baseC + t];
            srC = C.src[baseC + t];
            if (t >= nnC) dvC = C.dval[baseC + t];
        }
        if (t < nuC) uxC = C.uxidx[ubC + t];
        if (gflag) return;
        int cur = 0;
#pragma unroll
        for (int j = 0; j < NL; j++) {
            const uint32_t i = chain_level<NL>(j, i0);
            if (chain_skip<NL>(i, i0, nv)) continue;
            const ChainLv<VT>& L = V.L[i];
            const ChainVec<VT>& P = V.V[i > 0 ? i - 1 : 0];     // the parent's vectors (levels below i0)
            VT* pFi = P.F + y * P.n;
            VT* pu1i = P.u1 + y * P.n;
            VT* s = sv[cur];
            const VT* p = sv[cur ^ 1];
            VT v = VT(0);
            if (t < S[j]) {
                v = chain_top<NL>(i, i0, j) ? acc : p[sr[j]];
                if (!chain_top<NL>(i, i0, j)) pFi[base[j] + t] = v;
                if (t >= nkeep[j]) {
                    if (chain_top<NL>(i, i0, j)) out0[qr[j]] = XT(dv[j] * v);
                    else pu1i[qr[j]] = VT(dv[j] * v);
                }
                if (t >= nn[j]) s[t] = v;
            }
            for (uint32_t q = t + BS; q < S[j]; q += BS) {   // slots beyond the wavefront: no nodes
                const uint32_t r = L.qrow[base[j] + q];
                const VT w = chain_top<NL>(i, i0, j) ? fetch(f, base[j] + q, r) : p[L.src[base[j] + q]];
                if (!chain_top<NL>(i, i0, j)) pFi[base[j] + q] = w;
                if (q >= nkeep[j]) {
                    if (chain_top<NL>(i, i0, j)) out0[r] = XT(L.dval[base[j] + q] * w);
                    else pu1i[r] = VT(L.dval[base[j] + q] * w);
                }
                s[q] = w;
            }
            const VT w = chain_sweep_L<BS>(ol[j], nlL[j], L.Lloc, L.Lval, v, s);
            if (t < nn[j]) V.V[i].wt[y * V.V[i].N + base[j] + t] = w;
            level_sync<BS>();
            cur ^= 1;
        }
        // step 1 on the coarsest level from the F of the last chain level
        const VT* p = sv[cur ^ 1];
        VT v = VT(0);
        if (t < SC) {
            v = p[srC];
            pF[baseC + t] = v;
            if (t >= nnC) pu1[qrC] = VT(dvC * CT(v));
        }
        for (uint32_t q = t + BS; q < SC; q += BS) {
            const VT w = p[C.src[baseC + q]];
            pF[baseC + q] = w;
            pu1[C.qrow[baseC + q]] = VT(C.dval[baseC + q] * CT(w));
        }
        const CT f2 = CT(v);
        chain_sweep_L<BS>(olC, nlLC, C.Lloc, C.Lval, f2, cw);
        level_sync<BS>();
        const CT z = t < nuC ? cw[ouC.loc] : CT(0);
        const CT x = sweep_U<BS, CT, false>(ouC, nlUC, C.Ucol, C.Uval, z, (const CT*)nullptr, cx);
        if (t < nuC) pu1[uxC] = VT(x);
    } else {
        // passthrough rows: kept rows without L entries down to the level where they are isolated
        const uint32_t tp = (blockIdx.x - nblk) * BS + t;
        bool live = true;
        VT v = VT(0);
#pragma unroll
        for (int j = 0; j < NL; j++) {
            const uint32_t i = chain_level<NL>(j, i0);
            if (chain_skip<NL>(i, i0, nv) || !live) continue;
            const ChainLv<VT>& L = V.L[i];
            if (tp >= L.npt) { live = false; continue; }
            const uint32_t q = L.nslots + tp;
            const ChainVec<VT>& P = V.V[i > 0 ? i - 1 : 0];
            if (chain_top<NL>(i, i0, j)) {
                v = fetch(f, q, L.qrow[q]);
                if (gflag) return;
            } else {
                P.F[y * P.n + q] = v;
            }
            const uint32_t nxt = NL != 1 && i + 1 < nv ? V.L[chain_next(i)].npt : V.C.npt;
            if (tp >= nxt) {                  // isolated here
                if (chain_top<NL>(i, i0, j)) out0[L.qrow[q]] = XT(L.dval[q] * v);
                else P.u1[y * P.n + L.qrow[q]] = VT(L.dval[q] * v);
                live = false;
            }
        }
        if (live && tp < V.C.npt) {           // isolated on the coarsest level
            const uint32_t q = V.C.nslots + tp;
            pF[q] = v;
            pu1[V.C.qrow[q]] = VT(V.C.dval[q] * CT(v));
        }
    }
}

// Ascent to chain level ie: step 2 on the coarsest level from the residual
// A_c (a1 u1) - F, then for levels nv-1 .. ie the combine of the kept rows and
// the U sweep.  Every level above the coarsest but ie completes its second
// visit here, so its isolated rows' values come from x2 of its parent, where
// the descent of that visit wrote them; level ie's rows are written to out.
// NL >= nv - ie, as for the descent.
template <class VT, class CT, class XT, bool SCALE, bool BAT, int BS, int NL>
__global__ void __launch_bounds__(BS) k_chain_up(ChainView<VT, CT> V, uint32_t ie, XT* __restrict__ out,
                                                         uint32_t Ntop, uint32_t nblk, const uint32_t* guard) {
    const size_t y = BAT ? blockIdx.y : 0;
    if constexpr (BAT) { out = shift(out, y * Ntop); guard = shift(guard, y * V.st); }
    const uint32_t gflag = guard ? *guard : 0u;
    const uint32_t nv = V.nv, t = threadIdx.x;
    const VT *pF = nullptr, *pu1 = nullptr;
#pragma unroll
    for (int i = 0; i < kChainMax; i++)
        if (uint32_t(i) + 1 == nv) { pF = V.V[i].F + y * V.V[i].n; pu1 = V.V[i].u1 + y * V.V[i].n; }
    const ChainLv<CT>& C = V.C;
    if (blockIdx.x < nblk) {
        __shared__ VT sx[2][2 * BS];
        __shared__ CT cw[2 * BS], cx[2 * BS];
        const uint32_t b = blockIdx.x;
        const uint32_t* Hc = C.hdr + kChainHdrW * b;
        const uint32_t baseC = uload(Hc), SC = uload(Hc + 1), nnC = uload(Hc + 2), nlLC = uload(Hc + 4),
                       ubC = uload(Hc + 5), nuC = uload(Hc + 6), nlUC = uload(Hc + 7);
        // every per-slot load first: the coarsest level's residual and sweeps, each level's U sweep with its
        // w, the next level's rows with their combine operand, the isolated rows' values
        const ChainOwnL<CT> olC = chain_own_L(nnC, nlLC, baseC, C);
        const OwnU<CT> ouC = chain_own_U(nuC, nlUC, ubC, C);
        CT f2 = CT(0), dvC = CT(0);
        if (t < SC) {
            f2 = sell_row<CT, VT, SCALE>(V.A2, pu1, baseC + t, V.a1c) - CT(pF[baseC + t]);
            if (t >= nnC) dvC = C.dval[baseC + t];
        }
        uint32_t base[NL], S[NL], nkeep[NL], nu[NL], nlU[NL], ux[NL], Sn[NL], srn[NL], xln[NL];
        VT z[NL], xiso[NL], usv[NL];
        OwnU<VT> ou[NL];
        CT usC = CT(0);
#pragma unroll
        for (int j = 0; j < NL; j++) {
            const uint32_t i = chain_level<NL>(j, ie);
            base[j] = S[j] = nkeep[j] = nu[j] = nlU[j] = ux[j] = Sn[j] = srn[j] = xln[j] = 0;
            z[j] = xiso[j] = usv[j] = VT(0);
            ou[j] = OwnU<VT>{0u, 0u, 0u, 0u, 0u, 0u, VT(1), VT(0), VT(0)};
            if (chain_skip<NL>(i, ie, nv)) continue;
            const ChainLv<VT>& L = V.L[i];
            const ChainVec<VT>& Vi = V.V[i];
            const uint32_t* H = L.hdr + kChainHdrW * b;
            const uint32_t ub = uload(H + 5);
            base[j] = uload(H); S[j] = uload(H + 1); nkeep[j] = uload(H + 2) + uload(H + 3); nu[j] = uload(H + 6);
            nlU[j] = uload(H + 7);
            ou[j] = chain_own_U(nu[j], nlU[j], ub, L);
            if (t < nu[j]) {
                z[j] = Vi.wt[y * Vi.N + base[j] + ou[j].loc];
                if (chain_top<NL>(i, ie, j)) ux[j] = L.uxidx[ub + t];
            }
            if (!chain_top<NL>(i, ie, j) && t < S[j] && t >= nkeep[j]) {
                const ChainVec<VT>& P = V.V[i > 0 ? i - 1 : 0];
                xiso[j] = P.x2[y * P.n + L.qrow[base[j] + t]];
            }
            // the next level's rows (the kept rows of this level)
            const bool last = NL == 1 || i + 1 == nv;   // NL = 1: level nv-1 only
            const ChainLv<VT>& Ln = V.L[chain_next(i)];
            const uint32_t* qn = last ? C.qrow : Ln.qrow;
            const uint32_t* sn = last ? C.src : Ln.src;
            const uint32_t* xn = last ? C.xl : Ln.xl;
            const uint32_t* Hn = (last ? C.hdr : Ln.hdr) + kChainHdrW * b;
            Sn[j] = uload(Hn + 1);
            if (t < Sn[j]) {
                const uint32_t q = uload(Hn) + t, r = qn[q];
                srn[j] = sn[q];
                if (chain_top<NL>(i, ie, j)) xln[j] = xn[q];
                if (last) usC = SCALE ? CT(double(pu1[r]) * V.a1c) : CT(pu1[r]);
                else usv[j] = chain_u1s(Vi, y * Vi.n + r);
            }
        }
        if (gflag) return;
        // step 2 on the coarsest level
        chain_sweep_L<BS>(olC, nlLC, C.Lloc, C.Lval, f2, cw);
        if (t < SC && t >= nnC) cx[t] = dvC * f2;
        for (uint32_t q = t + BS; q < SC; q += BS) {
            const CT g = sell_row<CT, VT, SCALE>(V.A2, pu1, baseC + q, V.a1c) - CT(pF[baseC + q]);
            cx[q] = C.dval[baseC + q] * g;
        }
        level_sync<BS>();
        const CT zc = t < nuC ? cw[ouC.loc] : CT(0);
        sweep_U<BS, CT, false>(ouC, nlUC, C.Ucol, C.Uval, zc, (const CT*)nullptr, cx);
        level_sync<BS>();
        int cur = 0;
#pragma unroll
        for (int j = NL - 1; j >= 0; j--) {
            const uint32_t i = chain_level<NL>(j, ie);
            if (chain_skip<NL>(i, ie, nv)) continue;
            const ChainLv<VT>& L = V.L[i];
            const ChainVec<VT>& Vi = V.V[i];
            const bool last = NL == 1 || i + 1 == nv;   // NL = 1: level nv-1 only
            VT* s = sx[cur];
            const VT* xp = sx[cur ^ 1];
            // the combine of this level's kept rows (the next level's rows)
            if (t < Sn[j]) {
                const VT c = last ? VT(pair_cheb2(cx[t], usC, Vi.ma2, Vi.s2_skip, Vi.opb2, Vi.opb2_one))
                                  : VT(pair_cheb2(xp[t], usv[j], Vi.ma2, Vi.s2_skip, Vi.opb2, Vi.opb2_one));
                s[srn[j]] = c;
                if (chain_top<NL>(i, ie, j)) out[xln[j]] = XT(c);
            }
            if (Sn[j] > BS) {
                const ChainLv<VT>& Ln = V.L[chain_next(i)];
                const uint32_t* qn = last ? C.qrow : Ln.qrow;
                const uint32_t* sn = last ? C.src : Ln.src;
                const uint32_t* xn = last ? C.xl : Ln.xl;
                const uint32_t b0 = uload((last ? C.hdr : Ln.hdr) + kChainHdrW * b);
                for (uint32_t q = t + BS; q < Sn[j]; q += BS) {
                    const uint32_t r = qn[b0 + q];
                    VT c;
                    if (last) {
                        const CT us = SCALE ? CT(double(pu1[r]) * V.a1c) : CT(pu1[r]);
                        c = VT(pair_cheb2(cx[q], us, Vi.ma2, Vi.s2_skip, Vi.opb2, Vi.opb2_one));
                    } else {
                        c = VT(pair_cheb2(xp[q], chain_u1s(Vi, y * Vi.n + r), Vi.ma2, Vi.s2_skip, Vi.opb2, Vi.opb2_one));
                    }
                    s[sn[b0 + q]] = c;
                    if (chain_top<NL>(i, ie, j)) out[xn[b0 + q]] = XT(c);
                }
            }
            // the isolated rows of a level completing its second visit
            if (!chain_top<NL>(i, ie, j)) {
                if (t < S[j] && t >= nkeep[j]) s[t] = xiso[j];
                const ChainVec<VT>& P = V.V[i > 0 ? i - 1 : 0];
                for (uint32_t q = t + BS; q < S[j]; q += BS)
                    if (q >= nkeep[j]) s[q] = P.x2[y * P.n + L.qrow[base[j] + q]];
            }
            level_sync<BS>();
            const VT x = sweep_U<BS, VT, false>(ou[j], nlU[j], L.Ucol, L.Uval, z[j], (const VT*)nullptr, s);
            if (chain_top<NL>(i, ie, j) && t < nu[j]) out[ux[j]] = XT(x);
            level_sync<BS>();
            cur ^= 1;
        }
    } else {
        // passthrough rows that reach level ie + 1: their combines from the deepest level up
        const uint32_t tp = (blockIdx.x - nblk) * BS + t;
        bool have = false;
        VT v = VT(0);
        if (gflag) return;
#pragma unroll
        for (int j = NL - 1; j >= 0; j--) {
            const uint32_t i = chain_level<NL>(j, ie);
            if (chain_skip<NL>(i, ie, nv)) continue;
            const ChainVec<VT>& Vi = V.V[i];
            const bool last = NL == 1 || i + 1 == nv;   // NL = 1: level nv-1 only
            const ChainLv<VT>& Ln = V.L[chain_next(i)];
            const uint32_t npn = last ? C.npt : Ln.npt;
            if (tp >= npn) continue;                  // the row does not reach level i+1
            const uint32_t qn = (last ? C.nslots : Ln.nslots) + tp;
            const uint32_t r = last ? C.qrow[qn] : Ln.qrow[qn];
            if (last) {                               // isolated on the coarsest level
                const CT g = sell_row<CT, VT, SCALE>(V.A2, pu1, qn, V.a1c) - CT(pF[qn]);
                const CT xc = C.dval[qn] * g;
                const CT us = SCALE ? CT(double(pu1[r]) * V.a1c) : CT(pu1[r]);
                v = VT(pair_cheb2(xc, us, Vi.ma2, Vi.s2_skip, Vi.opb2, Vi.opb2_one));
            } else {
                if (!have) v = Vi.x2[y * Vi.n + r];   // isolated on level i+1
                v = VT(pair_cheb2(v, chain_u1s(Vi, y * Vi.n + r), Vi.ma2, Vi.s2_skip, Vi.opb2, Vi.opb2_one));
            }
            have = true;
        }
        if (have) out[V.L[ie].qrow[V.L[ie].nslots + tp]] = XT(v);
    }
}

// step 2: x2 = (-alpha) x2 ; x2 += (1+beta) u1   (cheb_solve.cc:79-81), then store or scatter
template <class VT, class XT, bool STORE, bool BAT>
__global__ void __launch_bounds__(kBS) k_cheb2(uint32_t n, const VT* __restrict__ x2, const VT* __restrict__ u1s,
                                               double ma, int skip, double opb, int opb_one,
                                               const uint32_t* __restrict__ mapM, XT* __restrict__ x,
                                               VT* __restrict__ x2out, const uint32_t* guard, LevelBatch lb) {
    if constexpr (BAT) {                 // batched right-hand sides (compiled out for single solves)
        const size_t y = blockIdx.y;
        x2 = shift(x2, y * lb.n); u1s = shift(u1s, y * lb.n); x2out = shift(x2out, y * lb.n);
        x = shift(x, y * lb.N); guard = shift(guard, y * lb.st);
    }
    if (guarded(guard)) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = double(x2[i]);
    if (!skip) v = v * ma;
    const double u = double(u1s[i]);
    const VT w = VT(opb_one ? v + u : v + opb * u);
    if (STORE) x2out[i] = w;
    else x[mapM[i]] = XT(w);
}

// steps >= 3: x2 = u1 - alpha x2 + beta (u1 - u0)   (cheb_solve.cc:99-100)
template <class VT, class XT, bool STORE, bool BAT>
__global__ void __launch_bounds__(kBS) k_cheb_step(uint32_t n, const VT* __restrict__ u1, const VT* __restrict__ u0,
                                                   const VT* __restrict__ x2, double a, double b,
                                                   const uint32_t* __restrict__ mapM, XT* __restrict__ x,
                                                   VT* __restrict__ x2out, const uint32_t* guard, LevelBatch lb) {
    if constexpr (BAT) {                 // batched right-hand sides (compiled out for single solves)
        const size_t y = blockIdx.y;
        u1 = shift(u1, y * lb.n); u0 = shift(u0, y * lb.n); x2 = shift(x2, y * lb.n); x2out = shift(x2out, y * lb.n);
        x = shift(x, y * lb.N); guard = shift(guard, y * lb.st);
    }
    if (guarded(guard)) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double p1 = double(u1[i]), p0 = double(u0[i]);
    const VT v = VT(p1 - a * double(x2[i]) + b * (p1 - p0));
    if (STORE) x2out[i] = v;
    else x[mapM[i]] = XT(v);
}

// DiagPrec / IdPrec (diag_prec.cc:7-23, id_prec.h:6-19)
__global__ void __launch_bounds__(kBS) k_diag_apply(uint32_t n, const double* __restrict__ d, const double* __restrict__ f,
                                                    double* __restrict__ x, const uint32_t* guard) {
    if (guarded(guard)) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = d ? f[i] * d[i] : f[i];
}

// ---------------------------------------------------------------------------
// BLAS-1 (blas.cc:43-82) and the PCG iteration (pcg.cc:42-76)
//
// The streaming kernels store per-block partial sums; k_finalize adds them
// and updates the scalar recurrences of the device-resident SolveState.
// ---------------------------------------------------------------------------
enum { FIN_DOT = 0, FIN_XR = 2, FIN_XR_NOTEST = 3, FIN_INIT = 6, FIN_TRUE_RES = 7, FIN_STEP = 8, FIN_START = 9,
       FIN_RESTART = 10 };

__global__ void __launch_bounds__(kBS) k_finalize(int op, uint32_t nb, const double* __restrict__ partials,
                                                  SolveState* st, double* out, Batch bt) {
    partials += size_t(blockIdx.y) * bt.part;
    if (st) st += blockIdx.y;
    if (bt.soff) {                                            // a ragged batch: this segment's own grid
        const uint32_t len = bt.soff[blockIdx.y + 1] - bt.soff[blockIdx.y], t = (len + kBS - 1) / kBS;
        nb = t < 1 ? 1u : (t > kRedGrid ? kRedGrid : t);
    }
    if (op != FIN_DOT && op != FIN_INIT && op != FIN_START && st->done) return;
    __shared__ double sh[kBS / 64];
    const int NV = 1;
    const double t0 = sum_partials<kBS>(partials, nb, NV, 0, sh);
    const double t1 = NV == 2 ? sum_partials<kBS>(partials, nb, NV, 1, sh) : 0.0;
    if (threadIdx.x != 0) return;
    switch (op) {
        case FIN_DOT: out[blockIdx.y] = t0; break;
        case FIN_RESTART:                                    // a flexible restart's first norm (pcg_phase)
            st->rr = t0;
            st->norm = sqrt(t0);
            if (!(st->norm / st->init_norm > st->eps)) st->done = 1;
            break;
        case FIN_INIT: st->rzs[st->niter & 1] = t0; break;
        case FIN_XR:                                         // norm, niter, stop test   pcg.cc:42, 61, 75
        case FIN_XR_NOTEST:
            st->rr = t0;
            st->norm = sqrt(t0);
            st->niter += 1;
            if (op == FIN_XR && !(st->norm / st->init_norm > st->eps)) st->done = 1;
            else if (st->maxit && st->niter >= st->maxit) { st->status = 5; st->done = 1; }
            break;
        case FIN_TRUE_RES:
            st->rr = t0;
            st->norm = sqrt(t0);
            if (!(st->norm / st->init_norm > st->eps)) st->done = 1;
            break;
        case FIN_START: {                                    // init norm, eps floor, first test   pcg.cc:25-31
            const double nrm = sqrt(t0);
            st->init_norm = nrm;
            st->norm = nrm;
            double e = st->eps;
            if (nrm * e < 1e-14) e = 1e-14 / nrm;            // check_and_replace_eps, solvers.cc:49-58
            st->eps = e;
            if (!(nrm / nrm > e)) { st->status = 4; st->done = 1; }
            break;
        }
        case FIN_STEP:                                       // Chebyshev / simple iteration: norm, count, test
            st->rr = t0;
            st->norm = sqrt(t0);
            st->niter += 1;
            if (!(st->norm / st->init_norm > st->eps)) st->done = 1;
            else if (st->maxit && st->niter - st->pad >= st->maxit) { st->status = 5; st->done = 1; }
            break;
    }
}

__global__ void __launch_bounds__(kBS) k_dot(uint32_t n, const double* __restrict__ x, const double* __restrict__ y,
                                             double* partials) {
    double s = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += x[i] * y[i];
    double part[1] = {s};
    block_partials<kBS, 1>(part, partials);
}

// The start of a batch of solves (grid.y = right-hand side): r = b - A x and
// the partials of (r, r), with the per-row and per-block operations of
// k_residual and k_dot, so each solve starts from the same values as alone.
__global__ void __launch_bounds__(kBS) k_residual_batch(SellView<double> A, const double* __restrict__ b,
                                                        const double* __restrict__ x, double* __restrict__ rr,
                                                        Batch bt) {
    const size_t y = blockIdx.y;
    b += y * bt.n; x += y * bt.n; rr += y * bt.n;
    const uint32_t r = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (r >= A.n) return;
    rr[r] = b[r] - mat_row<double, double>(A, x, r);
}
template <bool SEG = false>
__global__ void __launch_bounds__(kBS) k_dot_batch(uint32_t n, const double* __restrict__ x, double* partials,
                                                   Batch bt) {
    const size_t y = blockIdx.y;
    const Span sp = span_of<SEG>(bt, uint32_t(y));
    if (SEG && blockIdx.x >= sp.rg) return;
    x += sp.base; partials += y * bt.part;
    if (SEG) n = sp.len;
    double s = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += sp.rg * blockDim.x) s += x[i] * x[i];
    double part[1] = {s};
    block_partials<kBS, 1>(part, partials);
}

__global__ void __launch_bounds__(kBS) k_axpy(uint32_t n, double a, int one, const double* __restrict__ x,
                                              double* __restrict__ y) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        y[i] = one ? y[i] + x[i] : y[i] + a * x[i];
}

__global__ void __launch_bounds__(kBS) k_scal(uint32_t n, double a, double* __restrict__ x) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = x[i] * a;
}

__global__ void __launch_bounds__(kBS) k_copy(uint32_t n, const double* __restrict__ x, double* __restrict__ y,
                                              const uint32_t* guard) {
    if (guarded(guard)) return;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = x[i];
}

// Ap = A p ; partials of (Ap, p).  A ragged batch multiplies its segment's
// rows of the block-diagonal matrix (global rows and columns).
template <bool SEG = false>
__global__ void __launch_bounds__(kBS) k_pcg_spmv_dot(SellView<double> A, const double* __restrict__ p,
                                                      double* __restrict__ Ap, const SolveState* st, double* partials,
                                                      Batch bt) {
    const uint32_t y = blockIdx.y;
    const Span sp = span_of<SEG>(bt, y);
    if (SEG && blockIdx.x >= sp.rg) return;
    if (!SEG) { p += sp.base; Ap += sp.base; }
    st += y; partials += size_t(y) * bt.part;
    if (st->done) return;
    double s = 0.0;
    // grid-stride over 512-row tiles, XCD-grouped within each sweep of the grid;
    // two rows per thread, each thread adding its rows' products in row order
    // (the same order in both matrix layouts)
    const uint32_t len = SEG ? sp.len : A.n, r0 = SEG ? uint32_t(sp.base) : 0u;
    const uint32_t tiles = (len + 2 * blockDim.x - 1) / (2 * blockDim.x);
    const bool vec = A.sd && (r0 & 1u) == 0 && pair_vec(A, uintptr_t(p) | uintptr_t(Ap));
    for (uint32_t t = xcd_block(blockIdx.x, sp.rg); t < tiles; t += sp.rg) {
        const uint32_t q = 2 * (t * blockDim.x + threadIdx.x);
        if (q >= len) continue;
        const bool two = q + 1 < len;
        const uint32_t r = r0 + q;
        RowPair o;
        if (vec) {
            o = stencil_pair<true>(A, p, r, two);
        } else if (A.sd) {
            o = stencil_pair<false>(A, p, r, two);
        } else {
            o = RowPair{sell_row<double, double, false>(A, p, r, 1.0), 0.0, p[r], 0.0};
            if (two) {
                o.y1 = sell_row<double, double, false>(A, p, r + 1, 1.0);
                o.x1 = p[r + 1];
            }
        }
        if (vec && two) {
            *(dbl2*)(Ap + r) = dbl2{o.y0, o.y1};
        } else {
            Ap[r] = o.y0;
            if (two) Ap[r + 1] = o.y1;
        }
        s += o.y0 * o.x0;
        if (two) s += o.y1 * o.x1;
    }
    double part[1] = {s};
    block_partials<kBS, 1>(part, partials);
}

// alpha = rz / (Ap, p) (pcg.cc:51-57), summed by every block from the
// partials of k_pcg_spmv_dot in the same fixed order; then x += alpha p,
// r -= alpha Ap and partials of (r, r) (pcg.cc:58-61).  Block 0 records alpha
// (and "<Ap, p> <= 0", which stops the solve) in the state.
template <bool SEG = false>
__global__ void __launch_bounds__(kBS) k_pcg_xr(uint32_t n, const double* __restrict__ p, const double* __restrict__ Ap,
                                                double* __restrict__ x, double* __restrict__ r, SolveState* st,
                                                const double* __restrict__ app_partials, uint32_t nb,
                                                double* partials, Batch bt) {
    const uint32_t y = blockIdx.y;
    const Span sp = span_of<SEG>(bt, y);
    if (SEG && blockIdx.x >= sp.rg) return;
    if (SEG) { n = sp.len; nb = sp.rg; }
    p += sp.base; Ap += sp.base; x += sp.base; r += sp.base; st += y;
    app_partials += size_t(y) * bt.part; partials += size_t(y) * bt.part;
    if (block_done(st)) return;
    __shared__ double sh[kBS / 64];
    const double app = total_of_partials<kBS>(app_partials, nb, 1, 0, sh);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (!(app > 0)) {
        if (lead) {
            st->app = app;
            st->status = 3;
            st->done = 1;
        }
        return;
    }
    const double a = st->rzs[st->niter & 1] / app, ma = -a;
    const int one = dev_is_one(a), mone = dev_is_one(-a);
    if (lead) {
        st->app = app;
        st->alpha = a;
        st->alpha_one = one;
        st->malpha_one = mone;
    }
    double s = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += sp.rg * blockDim.x) {
        x[i] = one ? x[i] + p[i] : x[i] + a * p[i];
        const double ri = mone ? r[i] + Ap[i] : r[i] + ma * Ap[i];
        r[i] = ri;
        s += ri * ri;
    }
    double part[1] = {s};
    block_partials<kBS, 1>(part, partials);
}

// partials of (r, z) and, flexible, (r, z - z_prev)   (pcg.cc:68)
template <bool FLEX, bool SEG = false>
__global__ void __launch_bounds__(kBS) k_pcg_rz(uint32_t n, const double* __restrict__ r, const double* __restrict__ z,
                                                const double* __restrict__ zp, const SolveState* st, double* partials,
                                                Batch bt) {
    const uint32_t y = blockIdx.y;
    const Span sp = span_of<SEG>(bt, y);
    if (SEG && blockIdx.x >= sp.rg) return;
    if (SEG) n = sp.len;
    r += sp.base; z += sp.base; zp += sp.base; st += y; partials += size_t(y) * bt.part;
    if (st->done) return;
    double s0 = 0.0, s1 = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += sp.rg * blockDim.x) {
        s0 += r[i] * z[i];
        if (FLEX) s1 += r[i] * (z[i] - zp[i]);
    }
    if (FLEX) {
        double part[2] = {s0, s1};
        block_partials<kBS, 2>(part, partials);
    } else {
        double part[1] = {s0};
        block_partials<kBS, 1>(part, partials);
    }
}

// beta = rz1 / rz0 (pcg.cc:68-69; flexible: (r, z - z_prev) / rz0), summed by
// every block from the partials of k_pcg_rz; then p = beta p + z (dscal +
// daxpy(1, z, p), pcg.cc:71-72); flexible keeps z for the next beta.  The new
// rz goes to the state slot of this iteration's parity, so no block reads a
// slot another block writes.
template <bool FLEX, bool SEG = false>
__global__ void __launch_bounds__(kBS) k_pcg_p(uint32_t n, const double* __restrict__ z, double* __restrict__ p,
                                               double* __restrict__ zp, SolveState* st,
                                               const double* __restrict__ rz_partials, uint32_t nb, Batch bt) {
    const uint32_t y = blockIdx.y;
    const Span sp = span_of<SEG>(bt, y);
    if (SEG && blockIdx.x >= sp.rg) return;
    if (SEG) { n = sp.len; nb = sp.rg; }
    z += sp.base; p += sp.base; zp += sp.base; st += y;
    rz_partials += size_t(y) * bt.part;
    if (block_done(st)) return;
    __shared__ double sh[kBS / 64];
    const double t0 = total_of_partials<kBS>(rz_partials, nb, FLEX ? 2 : 1, 0, sh);
    const double t1 = FLEX ? total_of_partials<kBS>(rz_partials, nb, 2, 1, sh) : 0.0;
    const uint32_t k = st->niter;
    const double b = (FLEX ? t1 : t0) / st->rzs[(k + 1) & 1];
    const int one = dev_is_one(b);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->rzs[k & 1] = t0;
        st->beta = b;
        st->beta_one = one;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += sp.rg * blockDim.x) {
        const double pi = one ? p[i] : p[i] * b;
        p[i] = pi + z[i];
        if (FLEX) zp[i] = z[i];
    }
}

// p = z ; partials of (r, z)   (pcg.cc:32-33)
template <bool FLEX, bool SEG = false>
__global__ void __launch_bounds__(kBS) k_pcg_init(uint32_t n, const double* __restrict__ r, const double* __restrict__ z,
                                                  double* __restrict__ p, double* __restrict__ zp, const SolveState* st,
                                                  double* partials, Batch bt) {
    const uint32_t y = blockIdx.y;
    const Span sp = span_of<SEG>(bt, y);
    if (SEG && blockIdx.x >= sp.rg) return;
    if (SEG) n = sp.len;
    r += sp.base; z += sp.base; p += sp.base; zp += sp.base; st += y;
    partials += size_t(y) * bt.part;
    if (st->done) return;
    double s = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += sp.rg * blockDim.x) {
        p[i] = z[i];
        if (FLEX) zp[i] = z[i];
        s += r[i] * z[i];
    }
    double part[1] = {s};
    block_partials<kBS, 1>(part, partials);
}

// true residual replacement: r = b - A x ; partials of (r, r)
template <bool SEG = false>
__global__ void __launch_bounds__(kBS) k_pcg_true_res(SellView<double> A, const double* __restrict__ b,
                                                      const double* __restrict__ x, double* __restrict__ r,
                                                      const SolveState* st, double* partials, Batch bt) {
    const uint32_t y = blockIdx.y;
    const Span sp = span_of<SEG>(bt, y);
    if (SEG && blockIdx.x >= sp.rg) return;
    if (!SEG) { b += sp.base; x += sp.base; r += sp.base; }
    st += y; partials += size_t(y) * bt.part;
    if (st->done) return;
    double s = 0.0;
    const uint32_t len = SEG ? sp.len : A.n, r0 = SEG ? uint32_t(sp.base) : 0u;
    const uint32_t tiles = (len + blockDim.x - 1) / blockDim.x;
    for (uint32_t t = xcd_block(blockIdx.x, sp.rg); t < tiles; t += sp.rg) {
        const uint32_t i = t * blockDim.x + threadIdx.x;
        if (i < len) {
            const double v = b[r0 + i] - mat_row<double, double>(A, x, r0 + i);
            r[r0 + i] = v;
            s += v * v;
        }
    }
    double part[1] = {s};
    block_partials<kBS, 1>(part, partials);
}

// ChebSolver step 1 (cheb.cc:37-44): x = B r scaled by alpha (dscal, skipped
// when alpha == 1), then x += u1 (daxpy with alpha 1)
__global__ void __launch_bounds__(kBS) k_cheb_first(uint32_t n, double* __restrict__ x, const double* __restrict__ u1,
                                                    double alpha, int alpha_one) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double v = alpha_one ? x[i] : x[i] * alpha;
        x[i] = v + u1[i];
    }
}

// ChebSolver steps 2+ (cheb.cc:59-69): x = u1 + alpha x + beta (u1 - u0) with
// alpha, beta of the current iteration count taken from a table
// coef[2k], coef[2k+1] computed on the host
__global__ void __launch_bounds__(kBS) k_cheb_outer(uint32_t n, double* __restrict__ x, const double* __restrict__ u1,
                                                    const double* __restrict__ u0, const double* __restrict__ coef,
                                                    const SolveState* st) {
    if (st->done) return;
    const double a = coef[2 * st->niter], b = coef[2 * st->niter + 1];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        x[i] = u1[i] + a * x[i] + b * (u1[i] - u0[i]);
}

// SimpleSolver update x += z (daxpy with alpha 1, simple_solver.cc:52)
__global__ void __launch_bounds__(kBS) k_add(uint32_t n, const double* __restrict__ z, double* __restrict__ x,
                                             const SolveState* st) {
    if (st->done) return;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = x[i] + z[i];
}

// Adjoint gradient of the 7-point two-point-flux matrix (the reference's
// assembly, modules/mesh/spe.cc:69-131: a face between cells a and b
// contributes T = 2/(1/k_a + 1/k_b) * fac to both diagonals and -T to both
// off-diagonals) with respect to theta = ln kx per cell, ky = kx and
// kz = kzr * kx:  g_i = -lam^T (dA/dtheta_i) p
//                    = -sum_{faces f of i} dT_f/dtheta_i (lam_a - lam_b)(p_a - p_b),
// dT_f/dtheta_i = T_f^2 / (2 fac k_i) with k_i the cell's permeability along
// the face.  Cells are numbered x fastest, then y, then z.
__global__ void __launch_bounds__(kBS) k_spe_grad(uint32_t nx, uint32_t ny, uint32_t nz, double fx, double fy,
                                                  double fz, double kzr, const double* __restrict__ kx,
                                                  const double* __restrict__ p, const double* __restrict__ lam,
                                                  double* __restrict__ grad) {
    const uint32_t n = nx * ny * nz;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
    const double ki = kx[i], pi = p[i], li = lam[i];
    double g = 0.0;
    auto face = [&](uint32_t j, double fac, double kr) {
        const double ka = kr * ki, kb = kr * kx[j];
        const double T = 2.0 / (1.0 / ka + 1.0 / kb) * fac;
        g -= T * T / (2.0 * fac * ka) * (li - lam[j]) * (pi - p[j]);
    };
    if (z > 0) face(i - nx * ny, fz, kzr);
    if (y > 0) face(i - nx, fy, 1.0);
    if (x > 0) face(i - 1, fx, 1.0);
    if (x + 1 < nx) face(i + 1, fx, 1.0);
    if (y + 1 < ny) face(i + nx, fy, 1.0);
    if (z + 1 < nz) face(i + nx * ny, fz, kzr);
    grad[i] = g;
}

// One level of the forward Gauss-Seidel sweep (GSPrec::solve, gs_prec.cc:9-23):
// the rows of a level read only x of rows in earlier levels.  Row k of the
// level order is row rows[k]; its lower entries in ascending column order.
__global__ void __launch_bounds__(kBS) k_gs_level(uint32_t r0, uint32_t r1, const uint32_t* __restrict__ rows,
                                                  const uint32_t* __restrict__ ptr, const uint32_t* __restrict__ col,
                                                  const double* __restrict__ val, const double* __restrict__ diag,
                                                  const double* __restrict__ f, double* __restrict__ x,
                                                  const uint32_t* guard) {
    if (guarded(guard)) return;
    const uint32_t k = r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= r1) return;
    const uint32_t i = rows[k];
    double d = f[i];
    for (uint32_t q = ptr[k]; q < ptr[k + 1]; q++) d -= val[q] * x[col[q]];
    x[i] = d / diag[k];
}

// The same level with the lower entries in W fixed slots per level-order
// position (slot t of position k at t * nrow + k, cnt[k] of them used), so the
// row, its entries and its diagonal load together and only x[col] and f[row]
// wait for them: two dependent load rounds instead of four.
template <int W>
__global__ void __launch_bounds__(kBS) k_gs_level_ell(uint32_t r0, uint32_t r1, uint32_t nrow,
                                                      const uint32_t* __restrict__ rows,
                                                      const uint8_t* __restrict__ cnt,
                                                      const uint32_t* __restrict__ col,
                                                      const double* __restrict__ val,
                                                      const double* __restrict__ diag,
                                                      const double* __restrict__ f, double* __restrict__ x,
                                                      const uint32_t* guard) {
    if (guarded(guard)) return;
    const uint32_t k = r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= r1) return;
    const uint32_t i = rows[k], c = cnt[k];
    const double dg = diag[k];
    uint32_t cl[W];
    double vl[W];
#pragma unroll
    for (int t = 0; t < W; t++) {
        if (uint32_t(t) < c) {
            cl[t] = col[size_t(t) * nrow + k];
            vl[t] = val[size_t(t) * nrow + k];
        }
    }
    double d = f[i];
#pragma unroll
    for (int t = 0; t < W; t++)
        if (uint32_t(t) < c) d -= vl[t] * x[cl[t]];
    x[i] = d / dg;
}

// The forward Gauss-Seidel sweep of a matrix whose lower triangle is a
// seven-point grid (row i = x + nx (y + ny z) has lower entries at most at
// i - nx ny, i - nx, i - 1), by 8x8x8 bricks: a launch runs every brick of one
// brick wavefront (bx + by + bz = const), whose -x, -y, -z neighbour bricks
// ran in earlier launches; inside a brick the cells go by diagonals
// (lx + ly + lz = s) through LDS.  Each row is the row-pointer sweep's
// operation sequence: d = f_i, minus a_ij x_j over its lower entries in
// ascending column order (z, y, x), then d / a_ii.  pres bits: 1 z, 2 y, 4 x.
constexpr int kBrick = 8;
__global__ void __launch_bounds__(kBrick * kBrick * kBrick)
k_gs_brick(uint32_t nx, uint32_t ny, uint32_t nz, const uint32_t* __restrict__ bricks, uint32_t b0,
           const double* __restrict__ diag, const double* __restrict__ vz, const double* __restrict__ vy,
           const double* __restrict__ vx, const uint8_t* __restrict__ pres, const double* __restrict__ f,
           double* __restrict__ x, const uint32_t* guard) {
    if (guarded(guard)) return;
    __shared__ double sx[kBrick][kBrick][kBrick];
    const uint32_t code = bricks[b0 + blockIdx.x];
    const uint32_t lx = threadIdx.x & 7, ly = (threadIdx.x >> 3) & 7, lz = threadIdx.x >> 6;
    const uint32_t cx = (code & 0x3ffu) * kBrick + lx, cy = ((code >> 10) & 0x3ffu) * kBrick + ly,
                   cz = (code >> 20) * kBrick + lz;
    const bool live = cx < nx && cy < ny && cz < nz;
    const size_t nxy = size_t(nx) * ny;
    const size_t i = live ? cx + size_t(nx) * cy + nxy * cz : 0;
    uint32_t pr = 0;
    double fi = 0.0, dg = 1.0, az = 0.0, ay = 0.0, ax = 0.0, ez = 0.0, ey = 0.0, ex = 0.0;
    if (live) {
        pr = pres[i];
        fi = f[i];
        dg = diag[i];
        if (pr & 1u) { az = vz[i]; if (lz == 0) ez = x[i - nxy]; }   // neighbours outside the brick are final
        if (pr & 2u) { ay = vy[i]; if (ly == 0) ey = x[i - nx]; }
        if (pr & 4u) { ax = vx[i]; if (lx == 0) ex = x[i - 1]; }
    }
    const uint32_t mine = lx + ly + lz;
    for (uint32_t st = 0; st <= 3 * (kBrick - 1); st++) {
        if (live && mine == st) {
            double d = fi;
            if (pr & 1u) d -= az * (lz ? sx[lz - 1][ly][lx] : ez);
            if (pr & 2u) d -= ay * (ly ? sx[lz][ly - 1][lx] : ey);
            if (pr & 4u) d -= ax * (lx ? sx[lz][ly][lx - 1] : ex);
            const double xi = d / dg;
            sx[lz][ly][lx] = xi;
            x[i] = xi;
        }
        __syncthreads();
    }
}

// GMRES: the i+1 classical Gram-Schmidt dots (v_{i+1}, v_k), k = 0..i (gmres.cc:80-82).
// The k loops are unrolled to the 32-slot maximum under a k < nv test, so the
// sums stay in registers and the nv loads of a row are in flight together.
__global__ void __launch_bounds__(kBS) k_mdot(uint32_t n, uint32_t nv, const double* __restrict__ V, size_t ld,
                                              const double* __restrict__ w, double* partials) {
    double s[32];
#pragma unroll
    for (uint32_t k = 0; k < 32; k++) s[k] = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double wi = w[i];
#pragma unroll
        for (uint32_t k = 0; k < 32; k++)
            if (k < nv) s[k] += wi * V[k * ld + i];
    }
    __shared__ double sh[kBS / 64];
#pragma unroll
    for (uint32_t k = 0; k < 32; k++) {
        if (k < nv) {
            const double b = block_sum<kBS>(s[k], sh);
            if (threadIdx.x == 0) partials[32 * blockIdx.x + k] = b;
        }
    }
}

// out[k] = the sum of the nb partials of dot k; block k adds dot k
__global__ void __launch_bounds__(kBS) k_mdot_finalize(uint32_t nb, uint32_t nv, const double* __restrict__ partials,
                                                       double* out) {
    __shared__ double sh[kBS / 64];
    const uint32_t k = blockIdx.x;
    if (k >= nv) return;
    const double a = sum_partials<kBS>(partials, nb, 32, int(k), sh);
    if (threadIdx.x == 0) out[k] = a;
}

// The Gram-Schmidt update y -= sum_k h_k v_k with h from device memory, one
// daxpy(-h_k) at a time in k order (gmres.cc:83-84), and the partials of
// (y, y) with k_dot's grid-stride order and block sums, so the norm that
// follows is the one dnrm2(y) gives (gmres.cc:85).
__global__ void __launch_bounds__(kBS) k_cgs_update(uint32_t n, const double* __restrict__ V, size_t ld,
                                                   const double* __restrict__ h, uint32_t nk, double* __restrict__ y,
                                                   double* partials) {
    double s = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double acc = y[i];
        for (uint32_t k = 0; k < nk; k++) {
            const double c = -h[k], v = V[k * ld + i];
            acc = dev_is_one(c) ? acc + v : acc + c * v;            // daxpy's alpha == 1 path, blas.cc:31-38
        }
        y[i] = acc;
        s += acc * acc;
    }
    double part[1] = {s};
    block_partials<kBS, 1>(part, partials);
}

struct MaxpyCoef {
    double c[32];
    int one[32];
    uint32_t nk;
};

// y = y0 + sum_k c_k v_k, added one term at a time in k order like a sequence
// of daxpy calls (gmres.cc:83-84, 165-166); y0 = 0 when `zero`.
__global__ void __launch_bounds__(kBS) k_maxpy(uint32_t n, const double* __restrict__ V, size_t ld, MaxpyCoef cf,
                                               double* __restrict__ y, int zero) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double acc = zero ? 0.0 : y[i];
        for (uint32_t k = 0; k < cf.nk; k++) {
            const double v = V[k * ld + i];
            acc = cf.one[k] ? acc + v : acc + cf.c[k] * v;
        }
        y[i] = acc;
    }
}

}  // namespace mlp

// Device kernels for the iterative-solve path (gfx950, wave64).
//
// Floating-point kernels are compiled with -ffp-contract=off: each a*b+c is
// a rounded multiply followed by a rounded add, in the reference's operand
// order, so FP64 results match the reference's CPU loops bit for bit wherever
// the operation order is per-row (SpMV, triangular sweeps, vector updates).
// Reductions use a fixed tree and are deterministic run to run.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace mlp {

constexpr uint32_t kPad = 0xffffffffu;
constexpr uint32_t kGlobalBit = 0x80000000u;
constexpr int kBS = 256;                 // threads per block for flat kernels
constexpr int kRedBlocks = 1024;         // max blocks of a reduction grid
constexpr int kPartialsPerBlock = 32;    // partial sums a block may store (k_mdot: up to 32)

template <class VT> struct SellView {
    uint32_t n = 0;
    const uint32_t* off64 = nullptr;
    const uint32_t* col = nullptr;
    const VT* val = nullptr;
};

template <class VT> struct TailView {
    uint32_t nblk = 0;
    const uint32_t* blk = nullptr;
    const uint32_t* lvlL = nullptr;
    const uint32_t* lvlU = nullptr;
    const uint32_t* fidx = nullptr;
    const uint32_t* Lptr = nullptr;
    const uint32_t* Lloc = nullptr;
    const VT* Lval = nullptr;
    const uint32_t* uloc = nullptr;
    const uint32_t* uxidx = nullptr;
    const uint32_t* Uptr = nullptr;
    const uint32_t* Ucol = nullptr;
    const VT* Uval = nullptr;
    const VT* Udiag = nullptr;
};

// Device-resident state of a PCG/GMRES solve.  Host code never needs these
// values between iterations; it polls `done` every few iterations.
struct SolveState {
    double rz, app, alpha, beta, rr, norm, init_norm, eps;
    double rz_prev;
    uint32_t niter, done, status, maxit;
    uint32_t alpha_one, malpha_one, beta_one, pad;
    unsigned counter[8];                 // last-block counters of the fused reductions
};

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool guarded(const uint32_t* g) { return g && *((volatile const uint32_t*)g); }

__device__ __forceinline__ bool dev_is_one(double a) {   // include/tools.h:28-30
    return fabs(a - 1.0) <= 2.220446049250313e-16;
}

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

// Every block stores its partial sums; the last block to finish adds the
// partials in block order and runs `fin` on thread 0.  Deterministic for a
// fixed grid size.
template <int BS, int NV, class Fin>
__device__ __forceinline__ void grid_reduce(const double (&part)[NV], double* partials, unsigned* counter, Fin fin) {
    __shared__ double sh[BS / 64];
    __shared__ int last;
    double bs[NV];
    for (int k = 0; k < NV; k++) bs[k] = block_sum<BS>(part[k], sh);
    if (threadIdx.x == 0) {
        for (int k = 0; k < NV; k++) partials[NV * blockIdx.x + k] = bs[k];
        __threadfence();
        unsigned t = atomicAdd(counter, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double tot[NV];
    volatile const double* vp = partials;
    for (int k = 0; k < NV; k++) {
        double s = 0.0;
        for (unsigned b = threadIdx.x; b < gridDim.x; b += BS) s += vp[NV * b + k];
        tot[k] = block_sum<BS>(s, sh);
    }
    if (threadIdx.x == 0) {
        fin(tot);
        *counter = 0;
        __threadfence();
    }
}

// y_r = sum_k a_rk * (S ? s*x_k : x_k), in storage order, from 0
template <class VT, class XT, bool SCALE>
__device__ __forceinline__ VT sell_row(const SellView<VT>& A, const XT* __restrict__ x, uint32_t r, VT s) {
    const uint32_t sl = r >> 6, lane = r & 63;
    const uint32_t b = A.off64[sl], e = A.off64[sl + 1];
    VT acc = VT(0);
    for (uint32_t k = b; k < e; k++) {
        const size_t idx = size_t(k) * 64 + lane;
        const uint32_t c = A.col[idx];
        if (c != kPad) {
            VT xv = VT(x[c]);
            if (SCALE) xv = xv * s;
            acc += A.val[idx] * xv;
        }
    }
    return acc;
}

// ---------------------------------------------------------------------------
// SpMV family
// ---------------------------------------------------------------------------
template <class VT, class XT, class YT>
__global__ void __launch_bounds__(kBS) k_spmv(SellView<VT> A, const XT* __restrict__ x, YT* __restrict__ y,
                                              const uint32_t* guard) {
    if (guarded(guard)) return;
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= A.n) return;
    y[r] = YT(sell_row<VT, XT, false>(A, x, r, VT(1)));
}

// r = b - A x   (residual, csrmatrix.cc:293-300)
template <class VT, class XT>
__global__ void __launch_bounds__(kBS) k_residual(SellView<VT> A, const XT* __restrict__ b, const XT* __restrict__ x,
                                                  XT* __restrict__ rr, const uint32_t* guard) {
    if (guarded(guard)) return;
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= A.n) return;
    VT acc = sell_row<VT, XT, false>(A, x, r, VT(1));
    rr[r] = XT(VT(b[r]) - acc);
}

// tmp = A (s u) - F ; us = s u     (Chebyshev step 2, cheb_solve.cc:61-78)
// tmp = A u - F                    (steps >= 3, cheb_solve.cc:96-97)
template <class VT, bool SCALE>
__global__ void __launch_bounds__(kBS) k_cheb_res(SellView<VT> A, const VT* __restrict__ u, const VT* __restrict__ F,
                                                  VT* __restrict__ tmp, VT* __restrict__ us, VT s,
                                                  const uint32_t* guard) {
    if (guarded(guard)) return;
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= A.n) return;
    VT acc = sell_row<VT, VT, SCALE>(A, u, r, s);
    tmp[r] = acc - F[r];
    if (SCALE) us[r] = u[r] * s;
}

// ---------------------------------------------------------------------------
// triangular sweeps over the eliminated rows (cheb_solve.cc:33-38, 108-115)
// ---------------------------------------------------------------------------
template <class VT, class FT>
__global__ void k_tail_L(TailView<VT> T, const FT* __restrict__ f, VT* __restrict__ wt, const uint32_t* guard) {
    if (guarded(guard)) return;
    extern __shared__ unsigned char smem_raw[];
    VT* sw = reinterpret_cast<VT*>(smem_raw);
    const uint32_t* bk = T.blk + 6 * blockIdx.x;
    const uint32_t base = bk[0];
    uint32_t s = base;
    for (uint32_t li = bk[2]; li < bk[3]; li++) {
        const uint32_t e = T.lvlL[li];
        for (uint32_t t = s + threadIdx.x; t < e; t += blockDim.x) {
            VT acc = VT(f[T.fidx[t]]);
            for (uint32_t q = T.Lptr[t]; q < T.Lptr[t + 1]; q++) acc -= T.Lval[q] * sw[T.Lloc[q]];
            sw[t - base] = acc;
            wt[t] = acc;
        }
        s = e;
        __syncthreads();
    }
}

// MODE 0: x[map[i]] = v; global references read `xin`.
// MODE 1: x[map[i]] += v (the multi_split correction step); global references
//         read the correction vector `xin`.
template <class VT, class XT, int MODE>
__global__ void k_tail_U(TailView<VT> T, const VT* __restrict__ wt, const XT* xin, XT* xout, const uint32_t* guard) {
    if (guarded(guard)) return;
    extern __shared__ unsigned char smem_raw[];
    VT* sx = reinterpret_cast<VT*>(smem_raw);
    const uint32_t* bk = T.blk + 6 * blockIdx.x;
    const uint32_t base = bk[0];
    uint32_t s = base;
    for (uint32_t li = bk[4]; li < bk[5]; li++) {
        const uint32_t e = T.lvlU[li];
        for (uint32_t u = s + threadIdx.x; u < e; u += blockDim.x) {
            const uint32_t loc = T.uloc[u];
            VT z = wt[base + loc];
            for (uint32_t q = T.Uptr[u]; q < T.Uptr[u + 1]; q++) {
                const uint32_t c = T.Ucol[q];
                const VT xv = (c & kGlobalBit) ? VT(xin[c & ~kGlobalBit]) : sx[c];
                z -= T.Uval[q] * xv;
            }
            const VT v = z / T.Udiag[u];
            sx[loc] = v;
            const uint32_t xi = T.uxidx[u];
            if (MODE == 0) xout[xi] = XT(v);
            else xout[xi] = xout[xi] + XT(v);
        }
        s = e;
        __syncthreads();
    }
}

// rows M..M+n-1 of the L solve (the next level's right-hand side F) and the
// isolated-node diagonal (cheb_solve.cc:118-122)
template <class VT, class FT, class XT, int MODE>
__global__ void __launch_bounds__(kBS) k_schur_diag(uint32_t n, uint32_t Md, uint32_t M, uint32_t N,
                                                    const uint32_t* __restrict__ map,
                                                    const uint32_t* __restrict__ Sptr, const uint32_t* __restrict__ Scol,
                                                    const VT* __restrict__ Sval, const VT* __restrict__ wt,
                                                    const FT* __restrict__ f, VT* __restrict__ F,
                                                    const VT* __restrict__ dval, XT* __restrict__ x,
                                                    const uint32_t* guard) {
    if (guarded(guard)) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        VT acc = VT(f[map[M + i]]);
        for (uint32_t q = Sptr[i]; q < Sptr[i + 1]; q++) acc -= Sval[q] * wt[Scol[q]];
        F[i] = acc;
    } else if (i < n + Md) {
        const uint32_t k = i - n;
        const uint32_t j = map[N - Md + k];
        const VT v = dval[k] * VT(f[j]);
        if (MODE == 0) x[j] = XT(v);
        else x[j] = x[j] + XT(v);
    }
}

// multi_split prolongation: x[map[M+i]] = x2[i] (MODE 0) or uc[..] = x2, x[..] += x2 (MODE 1)
template <class VT, class XT, int MODE>
__global__ void __launch_bounds__(kBS) k_scatter(uint32_t n, const VT* __restrict__ x2, const uint32_t* __restrict__ mapM,
                                                 XT* __restrict__ uc, XT* __restrict__ x, const uint32_t* guard) {
    if (guarded(guard)) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = mapM[i];
    const VT v = x2[i];
    if (MODE == 0) {
        x[j] = XT(v);
    } else {
        uc[j] = XT(v);
        x[j] = x[j] + XT(v);
    }
}

// ncheb == 1: x[map[M+i]] = alpha * x2[i]   (cheb_solve.cc:66-68, 103-104)
template <class VT, class XT>
__global__ void __launch_bounds__(kBS) k_cheb1(uint32_t n, const VT* __restrict__ x2, VT a, int skip,
                                               const uint32_t* __restrict__ mapM, XT* __restrict__ x,
                                               const uint32_t* guard) {
    if (guarded(guard)) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    VT v = x2[i];
    if (!skip) v = v * a;
    x[mapM[i]] = XT(v);
}

// step 2: x2 = (-alpha) x2 ; x2 += (1+beta) u1   (cheb_solve.cc:79-81), then store or scatter
template <class VT, class XT, bool STORE>
__global__ void __launch_bounds__(kBS) k_cheb2(uint32_t n, const VT* __restrict__ x2, const VT* __restrict__ u1s,
                                               VT ma, int skip, VT opb, int opb_one,
                                               const uint32_t* __restrict__ mapM, XT* __restrict__ x,
                                               VT* __restrict__ x2out, const uint32_t* guard) {
    if (guarded(guard)) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    VT v = x2[i];
    if (!skip) v = v * ma;
    const VT u = u1s[i];
    v = opb_one ? v + u : v + opb * u;
    if (STORE) x2out[i] = v;
    else x[mapM[i]] = XT(v);
}

// steps >= 3: x2 = u1 - alpha x2 + beta (u1 - u0)   (cheb_solve.cc:99-100)
template <class VT, class XT, bool STORE>
__global__ void __launch_bounds__(kBS) k_cheb_step(uint32_t n, const VT* __restrict__ u1, const VT* __restrict__ u0,
                                                   const VT* __restrict__ x2, VT a, VT b,
                                                   const uint32_t* __restrict__ mapM, XT* __restrict__ x,
                                                   VT* __restrict__ x2out, const uint32_t* guard) {
    if (guarded(guard)) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const VT v = u1[i] - a * x2[i] + b * (u1[i] - u0[i]);
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
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBS) k_dot(uint32_t n, const double* __restrict__ x, const double* __restrict__ y,
                                             double* partials, unsigned* counter, double* out) {
    double s = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += x[i] * y[i];
    double part[1] = {s};
    grid_reduce<kBS, 1>(part, partials, counter, [=](const double* t) { *out = t[0]; });
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

// Ap = A p ; app = (Ap, p) ; alpha = rz / app
__global__ void __launch_bounds__(kBS) k_pcg_spmv_dot(SellView<double> A, const double* __restrict__ p,
                                                      double* __restrict__ Ap, SolveState* st, double* partials) {
    if (st->done) return;
    double s = 0.0;
    // grid-stride over rows: the stride is a multiple of 64, so lane == row % 64
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < A.n; r += gridDim.x * blockDim.x) {
        const double v = sell_row<double, double, false>(A, p, r, 1.0);
        Ap[r] = v;
        s += v * p[r];
    }
    double part[1] = {s};
    grid_reduce<kBS, 1>(part, partials, &st->counter[0], [=](const double* t) {
        st->app = t[0];
        if (t[0] <= 0) {
            st->status = 3;
            st->done = 1;
        } else {
            const double a = st->rz / t[0];
            st->alpha = a;
            st->alpha_one = dev_is_one(a);
            st->malpha_one = dev_is_one(-a);
        }
    });
}

// x += alpha p ; r -= alpha Ap ; norm = ||r|| ; niter++ ; stop test (pcg.cc:58-61, 42).
// TEST = false leaves the stop test to a following true-residual replacement.
template <bool TEST>
__global__ void __launch_bounds__(kBS) k_pcg_xr(uint32_t n, const double* __restrict__ p, const double* __restrict__ Ap,
                                                double* __restrict__ x, double* __restrict__ r, SolveState* st,
                                                double* partials) {
    if (st->done) return;
    const double a = st->alpha, ma = -a;
    const int one = st->alpha_one, mone = st->malpha_one;
    double s = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] = one ? x[i] + p[i] : x[i] + a * p[i];
        const double ri = mone ? r[i] + Ap[i] : r[i] + ma * Ap[i];
        r[i] = ri;
        s += ri * ri;
    }
    double part[1] = {s};
    grid_reduce<kBS, 1>(part, partials, &st->counter[1], [=](const double* t) {
        st->rr = t[0];
        st->norm = sqrt(t[0]);
        st->niter += 1;
        if (TEST && !(st->norm / st->init_norm > st->eps)) st->done = 1;
        else if (st->maxit && st->niter >= st->maxit) { st->status = 5; st->done = 1; }
    });
}

// rz1 = (r, z) ; beta = rz1 / rz0  (pcg.cc:68-69).  Flexible: beta = (r, z - z_prev) / rz0.
template <bool FLEX>
__global__ void __launch_bounds__(kBS) k_pcg_rz(uint32_t n, const double* __restrict__ r, const double* __restrict__ z,
                                                const double* __restrict__ zp, SolveState* st, double* partials) {
    if (st->done) return;
    double s0 = 0.0, s1 = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        s0 += r[i] * z[i];
        if (FLEX) s1 += r[i] * (z[i] - zp[i]);
    }
    double part[2] = {s0, s1};
    grid_reduce<kBS, 2>(part, partials, &st->counter[2], [=](const double* t) {
        const double b = (FLEX ? t[1] : t[0]) / st->rz;
        st->beta = b;
        st->beta_one = dev_is_one(b);
        st->rz = t[0];
    });
}

// p = beta p + z  (dscal + daxpy(1, z, p), pcg.cc:71-72); flexible keeps z for the next beta
template <bool FLEX>
__global__ void __launch_bounds__(kBS) k_pcg_p(uint32_t n, const double* __restrict__ z, double* __restrict__ p,
                                               double* __restrict__ zp, SolveState* st) {
    if (st->done) return;
    const double b = st->beta;
    const int one = st->beta_one;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double pi = one ? p[i] : p[i] * b;
        p[i] = pi + z[i];
        if (FLEX) zp[i] = z[i];
    }
}

// p = z ; rz = (r, z)   (pcg.cc:32-33)
template <bool FLEX>
__global__ void __launch_bounds__(kBS) k_pcg_init(uint32_t n, const double* __restrict__ r, const double* __restrict__ z,
                                                  double* __restrict__ p, double* __restrict__ zp, SolveState* st,
                                                  double* partials) {
    double s = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        p[i] = z[i];
        if (FLEX) zp[i] = z[i];
        s += r[i] * z[i];
    }
    double part[1] = {s};
    grid_reduce<kBS, 1>(part, partials, &st->counter[3], [=](const double* t) { st->rz = t[0]; });
}

// true residual replacement: r = b - A x ; norm = ||r|| ; stop test
__global__ void __launch_bounds__(kBS) k_pcg_true_res(SellView<double> A, const double* __restrict__ b,
                                                      const double* __restrict__ x, double* __restrict__ r,
                                                      SolveState* st, double* partials) {
    if (st->done) return;
    double s = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += gridDim.x * blockDim.x) {
        const double v = b[i] - sell_row<double, double, false>(A, x, i, 1.0);
        r[i] = v;
        s += v * v;
    }
    double part[1] = {s};
    grid_reduce<kBS, 1>(part, partials, &st->counter[4], [=](const double* t) {
        st->rr = t[0];
        st->norm = sqrt(t[0]);
        if (!(st->norm / st->init_norm > st->eps)) st->done = 1;
    });
}

// GMRES: the i+1 classical Gram-Schmidt dots (v_{i+1}, v_k), k = 0..i (gmres.cc:80-82)
__global__ void __launch_bounds__(kBS) k_mdot(uint32_t n, uint32_t nv, const double* __restrict__ V, size_t ld,
                                              const double* __restrict__ w, double* partials, unsigned* counter,
                                              double* out) {
    double s[32];
    for (uint32_t k = 0; k < 32; k++) s[k] = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double wi = w[i];
        for (uint32_t k = 0; k < nv; k++) s[k] += wi * V[k * ld + i];
    }
    __shared__ double sh[kBS / 64];
    __shared__ int last;
    for (uint32_t k = 0; k < nv; k++) {
        double b = block_sum<kBS>(s[k], sh);
        if (threadIdx.x == 0) partials[32 * blockIdx.x + k] = b;
    }
    if (threadIdx.x == 0) {
        __threadfence();
        unsigned t = atomicAdd(counter, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    volatile const double* vp = partials;
    for (uint32_t k = 0; k < nv; k++) {
        double a = 0.0;
        for (unsigned b = threadIdx.x; b < gridDim.x; b += kBS) a += vp[32 * b + k];
        a = block_sum<kBS>(a, sh);
        if (threadIdx.x == 0) out[k] = a;
    }
    if (threadIdx.x == 0) {
        *counter = 0;
        __threadfence();
    }
}

// streams through a buffer to evict the SpMV operands from L2 and the Infinity Cache
__global__ void __launch_bounds__(kBS) k_flush(size_t n, double* __restrict__ buf) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        buf[i] = buf[i] * 0.5 + 1.0;
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

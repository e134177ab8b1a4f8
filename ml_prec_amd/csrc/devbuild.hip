// uh_cheb hierarchy construction on the device: the steps of cheb_steps.h
// launched as kernels on one stream, every level's matrix kept in device
// memory for the next level, the level arrays copied to the host Hierarchy
// (which the planner and the tests read).
#define MLP_HD __device__
#include "cheb_steps.h"
#include "devbuild.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "par.h"

namespace mlp {
namespace {

constexpr uint32_t kStepBS = 256;
constexpr uint32_t kScanChunk = 2048;              // elements per block of the scan (8 per thread)
constexpr uint32_t kSegBlocks = 1024;              // grid of the round kernels (a stride loop over the round)

struct DevOps {
    __device__ static uint32_t add(uint32_t* p, uint32_t v) { return atomicAdd(p, v); }
    __device__ static int sub(int* p, int v) { return atomicSub(p, v); }
    __device__ static int max(int* p, int v) { return atomicMax(p, v); }
    __device__ static uint32_t maxu(uint32_t* p, uint32_t v) { return atomicMax(p, v); }
    __device__ static void fence() { __threadfence(); }
};

template <class F>
__global__ void __launch_bounds__(kStepBS) k_step(uint32_t n, F f) {
    const uint32_t i = blockIdx.x * kStepBS + threadIdx.x;
    if (i < n) f(i);
}

// f(q) for bnd[r] <= q < bnd[r + 1], the bounds written by the launch before
template <class F>
__global__ void __launch_bounds__(kStepBS) k_step_seg(const uint32_t* bnd, uint32_t r, F f) {
    const uint32_t b = bnd[r], e = bnd[r + 1];
    for (uint32_t q = b + blockIdx.x * kStepBS + threadIdx.x; q < e; q += gridDim.x * kStepBS) f(q);
}

// exclusive scan, pass 1: each block scans kScanChunk elements (8 consecutive per
// thread), writes the block-local prefix and its total
__global__ void __launch_bounds__(kStepBS) k_scan_part(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                       uint32_t n, uint32_t* __restrict__ bsum) {
    __shared__ uint32_t s[kStepBS];
    const uint32_t base = blockIdx.x * kScanChunk + threadIdx.x * 8;
    uint32_t v[8], t = 0;
    for (int k = 0; k < 8; k++) {
        v[k] = base + k < n ? in[base + k] : 0u;
        t += v[k];
    }
    s[threadIdx.x] = t;
    __syncthreads();
    for (uint32_t off = 1; off < kStepBS; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = s[threadIdx.x] - t;
    for (int k = 0; k < 8; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == kStepBS - 1) bsum[blockIdx.x] = s[kStepBS - 1];
}

// pass 2 (one block): the block totals' exclusive scan in place, the grand total at bsum[nb]
__global__ void __launch_bounds__(kStepBS) k_scan_top(uint32_t* bsum, uint32_t nb) {
    __shared__ uint32_t s[kStepBS];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += kStepBS) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t t = i < nb ? bsum[i] : 0u;
        s[threadIdx.x] = t;
        __syncthreads();
        for (uint32_t off = 1; off < kStepBS; off <<= 1) {
            const uint32_t add = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < nb) bsum[i] = carry + s[threadIdx.x] - t;
        carry += s[kStepBS - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

// pass 3: add each block's offset; out[n] = the total
__global__ void __launch_bounds__(kStepBS) k_scan_fix(uint32_t* __restrict__ out, uint32_t n,
                                                      const uint32_t* __restrict__ bsum, uint32_t nb) {
    const uint32_t off = bsum[blockIdx.x];
    const uint32_t base = blockIdx.x * kScanChunk;
    for (uint32_t k = threadIdx.x; k < kScanChunk; k += kStepBS)
        if (base + k < n) out[base + k] += off;
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = bsum[nb];
}

struct SetZero {
    uint32_t* p;
    __device__ void operator()(uint32_t i) const { p[i] = 0u; }
};

#define DEVBUILD_CHECK(expr)                                                                     \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            throw ::mlp::Error(std::string("HIP error: ") + hipGetErrorString(e_) + " (" #expr ")", 2); \
    } while (0)

// Memory a device keeps for its constructions (until the process ends): device
// blocks carved in order by each build, and the pinned staging buffers.  One
// build at a time uses them (the lock is held for the whole build).
struct BuildMemory {
    std::mutex m;
    std::vector<std::pair<void*, size_t>> blocks;
    char* pin = nullptr;                       // bounce buffer of large single copies
    size_t pin_n = 0;
    char* gpin = nullptr;                      // staging of a level's batched copies
    size_t gpin_n = 0;
};
BuildMemory& build_memory(int device) {
    static std::mutex g;
    static std::vector<std::unique_ptr<BuildMemory>> per;
    std::lock_guard<std::mutex> hold(g);
    if (device < 0) device = 0;
    if (size_t(device) >= per.size()) per.resize(size_t(device) + 1);
    if (!per[device]) per[device] = std::make_unique<BuildMemory>();
    return *per[device];
}
char* grow_pinned(char*& p, size_t& have, size_t want) {
    if (have < want) {
        if (p) DEVBUILD_CHECK(hipHostFree(p));
        p = nullptr;
        have = 0;
        DEVBUILD_CHECK(hipHostMalloc((void**)&p, want, hipHostMallocDefault));
        have = want;
    }
    return p;
}

// Runs the steps as kernels on one stream, with the device's build memory.
struct DevExec {
    hipStream_t s;
    BuildMemory& bm;
    std::unique_lock<std::mutex> hold;
    size_t blk = 0;                            // block being carved
    char* cur = nullptr;
    size_t left = 0;
    uint32_t* bsum = nullptr;                  // scan block totals
    uint32_t bsum_cap = 0;
    static int device_of() {
        int d = 0;
        DEVBUILD_CHECK(hipGetDevice(&d));
        return d;
    }
    explicit DevExec(hipStream_t st) : s(st), bm(build_memory(device_of())), hold(bm.m) {}
    ~DevExec() { (void)hipStreamSynchronize(s); }
    static constexpr size_t kPinChunk = size_t(32) << 20;
    char* bounce(size_t bytes) { return grow_pinned(bm.pin, bm.pin_n, std::min(bytes, kPinChunk)); }
    template <class T> T* alloc(size_t n) {
        const size_t b = (std::max<size_t>(n, 1) * sizeof(T) + 255) & ~size_t(255);
        while (b > left) {
            if (cur) blk++;                                   // the next kept block, or a new one
            if (blk == bm.blocks.size()) {
                const size_t sz = std::max<size_t>(b, size_t(64) << 20);
                void* p = nullptr;
                DEVBUILD_CHECK(hipMalloc(&p, sz));
                bm.blocks.push_back({p, sz});
            }
            cur = static_cast<char*>(bm.blocks[blk].first);
            left = bm.blocks[blk].second;
        }
        T* r = reinterpret_cast<T*>(cur);
        cur += b;
        left -= b;
        return r;
    }
    template <class T> void zero(T* p, size_t n) { DEVBUILD_CHECK(hipMemsetAsync(p, 0, n * sizeof(T), s)); }
    template <class F> void each(uint32_t n, const F& f) {
        if (!n) return;
        hipLaunchKernelGGL(k_step<F>, dim3((n + kStepBS - 1) / kStepBS), dim3(kStepBS), 0, s, n, f);
        DEVBUILD_CHECK(hipGetLastError());
    }
    template <class F> void each_seg(const uint32_t* bnd, uint32_t r, const F& f) {
        hipLaunchKernelGGL(k_step_seg<F>, dim3(kSegBlocks), dim3(kStepBS), 0, s, bnd, r, f);
        DEVBUILD_CHECK(hipGetLastError());
    }
    void scan(const uint32_t* in, uint32_t* out, uint32_t n) {
        if (n == 0) {
            each(1, SetZero{out});
            return;
        }
        const uint32_t nb = (n + kScanChunk - 1) / kScanChunk;
        if (nb + 1 > bsum_cap) {
            bsum_cap = std::max<uint32_t>(nb + 1, 1024);
            bsum = alloc<uint32_t>(bsum_cap);
        }
        hipLaunchKernelGGL(k_scan_part, dim3(nb), dim3(kStepBS), 0, s, in, out, n, bsum);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kStepBS), 0, s, bsum, nb);
        hipLaunchKernelGGL(k_scan_fix, dim3(nb), dim3(kStepBS), 0, s, out, n, (const uint32_t*)bsum, nb);
        DEVBUILD_CHECK(hipGetLastError());
    }
    // copies of 64 KiB and more go through the pinned buffer in chunks, each spread over the host threads
    template <class T> void get(T* dst, const T* src, size_t n) {
        const size_t bytes = n * sizeof(T);
        if (bytes < (size_t(1) << 16)) {
            DEVBUILD_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
            DEVBUILD_CHECK(hipStreamSynchronize(s));
            return;
        }
        char* b = bounce(bytes);
        const char* from = reinterpret_cast<const char*>(src);
        char* to = reinterpret_cast<char*>(dst);
        for (size_t off = 0; off < bytes; off += kPinChunk) {
            const size_t len = std::min(kPinChunk, bytes - off);
            DEVBUILD_CHECK(hipMemcpyAsync(b, from + off, len, hipMemcpyDeviceToHost, s));
            DEVBUILD_CHECK(hipStreamSynchronize(s));
            parallel_for(len, size_t(1) << 20, [&](uint64_t x0, uint64_t x1) { std::memcpy(to + off + x0, b + x0, x1 - x0); });
        }
    }
    // every copy into one pinned buffer behind one wait, then spread over the host threads into place
    void get_many(const std::vector<steps::Fetch>& items) {
        size_t total = 0;
        for (const auto& f : items) total += (f.bytes + 255) & ~size_t(255);
        if (!total) return;
        char* gpin = grow_pinned(bm.gpin, bm.gpin_n, total);
        std::vector<size_t> off(items.size());
        size_t o = 0;
        for (size_t k = 0; k < items.size(); k++) {
            off[k] = o;
            DEVBUILD_CHECK(hipMemcpyAsync(gpin + o, items[k].src, items[k].bytes, hipMemcpyDeviceToHost, s));
            o += (items[k].bytes + 255) & ~size_t(255);
        }
        DEVBUILD_CHECK(hipStreamSynchronize(s));
        for (size_t k = 0; k < items.size(); k++) {
            char* to = static_cast<char*>(items[k].dst);
            const char* from = gpin + off[k];
            parallel_for(items[k].bytes, size_t(1) << 20,
                         [&](uint64_t x0, uint64_t x1) { std::memcpy(to + x0, from + x0, x1 - x0); });
        }
    }
    template <class T> T get1(const T* src) {
        T v;
        get(&v, src, 1);
        return v;
    }
    template <class T> T* put(const T* src, size_t n) {
        T* d = alloc<T>(n);
        const size_t bytes = n * sizeof(T);
        char* b = bounce(bytes);
        const char* from = reinterpret_cast<const char*>(src);
        char* to = reinterpret_cast<char*>(d);
        for (size_t off = 0; off < bytes; off += kPinChunk) {
            const size_t len = std::min(kPinChunk, bytes - off);
            parallel_for(len, size_t(1) << 20, [&](uint64_t x0, uint64_t x1) { std::memcpy(b + x0, from + off + x0, x1 - x0); });
            DEVBUILD_CHECK(hipMemcpyAsync(to + off, b, len, hipMemcpyHostToDevice, s));
            DEVBUILD_CHECK(hipStreamSynchronize(s));
        }
        return d;
    }
};

struct DeviceChebBuilder : ChebLevelBuilder {
    hipStream_t s;
    DevCSRView dA;                              // level 0 on the device, or all null (then A is copied up)
    DeviceChebBuilder(hipStream_t st, const DevCSRView& d) : s(st), dA(d) {}
    void run(const HostCSR& A, bool use_tails, uint32_t max_levels, Hierarchy& H) override {
        DevExec x(s);
        steps::LevelMatrix M0;
        M0.n = A.nrow;
        if (dA.ia) {
            M0.nnz = dA.nnz;
            M0.ia = dA.ia;
            M0.ja = dA.ja;
            M0.a = dA.a;
        } else {
            M0.nnz = A.nnz();
            M0.ia = x.put(A.ia.data(), A.ia.size());
            M0.ja = x.put(A.ja.data(), A.ja.size());
            M0.a = x.put(A.a.data(), A.a.size());
        }
        steps::build_levels_steps<DevExec, DevOps>(x, M0, use_tails, max_levels, H);
    }
};

// ---------------------------------------------------------------------------
// seven-point assembly (modules/mesh/spe.cc:69-131), the host's spe_assemble
// row by row: same operations in the same order, so the same bits
// ---------------------------------------------------------------------------
struct SpeRows {
    int spe;
    uint32_t nx, ny, nz;
    const double* kx;
    double kz_ratio, c, hx, hy, hz;
    // an ensemble stacked along z: plane k belongs to the system of planes zfirst[k] .. zlast[k], and no
    // z-link crosses from one system into the next (null: one system of nz planes)
    const uint32_t* zfirst;
    const uint32_t* zlast;
    __device__ bool zlow(uint32_t k) const { return zfirst ? k > zfirst[k] : k > 0; }
    __device__ bool zhigh(uint32_t k) const { return zlast ? k < zlast[k] : k + 1 < nz; }
    __device__ double face(double ka, double kb, int axis) const {
        const double m = 2 / (1 / ka + 1 / kb);
        if (spe) return axis == 0 ? m / (hx * hx) : axis == 1 ? m / (hy * hy) : m / (hz * hz);
        return axis == 0 ? m * (hy * hz / hx) : axis == 1 ? m * (hx * hz / hy) : m * (hx * hy / hz);
    }
};

struct SpeCount {
    SpeRows R;
    uint32_t* cnt;
    __device__ void operator()(uint32_t r) const {
        const uint32_t i = r % R.nx, j = (r / R.nx) % R.ny, k = r / (R.nx * R.ny);
        cnt[r] = 1 + R.zlow(k) + (j > 0) + (i > 0) + (i + 1 < R.nx) + (j + 1 < R.ny) + R.zhigh(k);
    }
};

// row r: diagonal first, then z-, y-, x-, x+, y+, z+; optionally the seven-point
// layout's arrays (diagonal, the three lower values, presence bits 1 z- 2 y- 4 x-
// 8 x+ 16 y+ 32 z+)
struct SpeFill {
    SpeRows R;
    const uint32_t* ia;
    uint32_t* ja;
    double* a;
    uint8_t* pres;
    double *d, *lx, *ly, *lz;
    __device__ void operator()(uint32_t r) const {
        const uint32_t nx = R.nx, ny = R.ny, nz = R.nz, sy = nx, sz = nx * ny;
        const uint32_t i = r % nx, j = (r / nx) % ny, k = r / sz;
        const double* kx = R.kx;
        uint32_t p = ia[r];
        const uint32_t dpos = p;
        double diag = R.c;
        if (!R.spe) {
            double tb = 0.0;
            if (i == 0) tb += kx[r] * (R.hy * R.hz) / (R.hx / 2);
            if (i + 1 == nx) {
                if (nx > 1) tb += kx[r] * (R.hy * R.hz) / (R.hx / 2);
                else tb *= 2.0;
            }
            diag = diag + tb;
        }
        ja[p] = r;
        p++;
        double low[3] = {0.0, 0.0, 0.0};                      // z-, y-, x-
        uint32_t m = 0;
        auto add = [&](uint32_t q, double v) {
            ja[p] = q;
            a[p] = -v;
            p++;
            diag = diag + v;
        };
        const double kzr = R.kz_ratio * kx[r];
        if (R.zlow(k)) { const double v = R.face(kzr, R.kz_ratio * kx[r - sz], 2); add(r - sz, v); low[0] = -v; m |= 1; }
        if (j > 0) { const double v = R.face(kx[r], kx[r - sy], 1); add(r - sy, v); low[1] = -v; m |= 2; }
        if (i > 0) { const double v = R.face(kx[r], kx[r - 1], 0); add(r - 1, v); low[2] = -v; m |= 4; }
        if (i + 1 < nx) { add(r + 1, R.face(kx[r], kx[r + 1], 0)); m |= 8; }
        if (j + 1 < ny) { add(r + sy, R.face(kx[r], kx[r + sy], 1)); m |= 16; }
        if (R.zhigh(k)) { add(r + sz, R.face(kzr, R.kz_ratio * kx[r + sz], 2)); m |= 32; }
        a[dpos] = diag;
        if (pres) {
            pres[r] = uint8_t(m);
            d[r] = diag;
            lz[r] = low[0];
            ly[r] = low[1];
            lx[r] = low[2];
        }
    }
};

}  // namespace

std::unique_ptr<ChebLevelBuilder> device_cheb_builder(void* stream, const DevCSRView& level0) {
    return std::make_unique<DeviceChebBuilder>(static_cast<hipStream_t>(stream), level0);
}

void spe_assemble_device(void* stream, int spe, uint32_t nx, uint32_t ny, uint32_t nz, const double* kx,
                         double kz_ratio, double c, double hx, double hy, double hz, uint32_t* ia, uint32_t* ja,
                         double* a, const SpeStencilOut* st, const uint32_t* zfirst, const uint32_t* zlast) {
    const uint64_t n = uint64_t(nx) * ny * nz;
    if (!n) throw Error("grid dimensions must be positive", 1);
    DevExec x(static_cast<hipStream_t>(stream));
    const SpeRows R{spe, nx, ny, nz, kx, kz_ratio, c, hx, hy, hz, zfirst, zlast};
    uint32_t* cnt = x.alloc<uint32_t>(n);
    x.each(uint32_t(n), SpeCount{R, cnt});
    x.scan(cnt, ia, uint32_t(n));
    if (st) x.each(uint32_t(n), SpeFill{R, ia, ja, a, st->pres, st->d, st->lx, st->ly, st->lz});
    else x.each(uint32_t(n), SpeFill{R, ia, ja, a, nullptr, nullptr, nullptr, nullptr, nullptr});
    DEVBUILD_CHECK(hipStreamSynchronize(x.s));
}

}  // namespace mlp

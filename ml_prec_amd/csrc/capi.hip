// C ABI of the MI355X iterative-solve path: device matrices, device
// preconditioner hierarchies, PCG and GMRES.  See include/mlp.h.
#include "../../include/mlp.h"

#include "assembly.h"
#include "devbuild.h"
#include "hierarchy.h"
#include "kernels.h"
#include "mlp_error.h"
#include "par.h"
#include "plan.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <exception>
#include <chrono>
#include <thread>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <functional>
#include <list>
#include <mutex>
#include <type_traits>
#include <string>
#include <unordered_map>
#include <vector>

namespace mlp {

thread_local std::string g_last_error;
void set_last_error(const std::string& m) { g_last_error = m; }

namespace {

inline uint32_t cdiv(uint64_t a, uint64_t b) { return uint32_t((a + b - 1) / b); }
inline bool host_is_one(double a) { return std::fabs(a - 1.0) <= 2.220446049250313e-16; }   // include/tools.h:28-30

// ---------------------------------------------------------------------------
// device memory
// ---------------------------------------------------------------------------
// Device memory released by the library is kept per device and size class and handed out again, so that
// repeated set-ups (the solves of an adjoint loop, finite-difference checks, ensembles) do not pay hipMalloc
// and hipFree (each hipFree waits for the whole device).  A release waits for its devices once, which is what
// hipFree did, so no kernel still running can see its memory reused.  At most kCap bytes per device are kept.
struct MemCache {
    static constexpr size_t kCap = size_t(8) << 30;
    std::mutex m;
    std::unordered_map<uint64_t, std::vector<void*>> free;   // key: device << 56 | size class
    std::unordered_map<void*, std::pair<size_t, int>> live;  // blocks handed out by mlp_dev_alloc
    size_t cached[64] = {};
    static size_t size_class(size_t bytes) {
        if (bytes <= 4096) return 4096;
        size_t e = 63 - size_t(__builtin_clzll(bytes));
        const size_t step = size_t(1) << (e - 2);                 // four classes per octave (at most 25% waste)
        return (bytes + step - 1) & ~(step - 1);
    }
    static uint64_t key(int dev, size_t cls) { return (uint64_t(dev) << 56) | cls; }
};
inline MemCache& mem_cache() {
    static MemCache c;
    return c;
}
inline int current_device() {
    int d = 0;
    MLP_HIP_CHECK(hipGetDevice(&d));
    return d;
}
// a block of at least `bytes` on the current device
inline void* cache_alloc(size_t bytes, size_t* cls_out) {
    MemCache& c = mem_cache();
    const int dev = current_device();
    const size_t cls = MemCache::size_class(std::max<size_t>(bytes, 1));
    *cls_out = cls;
    {
        std::lock_guard<std::mutex> hold(c.m);
        auto it = c.free.find(MemCache::key(dev, cls));
        if (it != c.free.end() && !it->second.empty()) {
            void* p = it->second.back();
            it->second.pop_back();
            if (dev < 64) c.cached[dev] -= cls;
            return p;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, cls);
    if (e == hipErrorOutOfMemory) {                     // give the kept blocks of this device back, then retry
        (void)hipGetLastError();
        std::lock_guard<std::mutex> hold(c.m);
        for (auto& kv : c.free)
            if (int(kv.first >> 56) == dev) {
                for (void* q : kv.second) (void)hipFree(q);
                kv.second.clear();
            }
        if (dev < 64) c.cached[dev] = 0;
        e = hipMalloc(&p, cls);
    }
    MLP_HIP_CHECK(e);
    return p;
}
struct CachedBlock {
    void* p;
    size_t cls;
    int dev;
};
// blocks back into the cache, after the work on their devices has finished
inline void cache_release(std::vector<CachedBlock>& blocks) {
    if (blocks.empty()) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    uint64_t synced = 0;
    for (const CachedBlock& b : blocks)
        if (b.dev < 64 && !(synced >> b.dev & 1)) {
            if (b.dev != cur) (void)hipSetDevice(b.dev);
            (void)hipDeviceSynchronize();
            synced |= uint64_t(1) << b.dev;
        }
    (void)hipSetDevice(cur);
    MemCache& c = mem_cache();
    std::lock_guard<std::mutex> hold(c.m);
    for (const CachedBlock& b : blocks) {
        if (b.dev >= 64 || c.cached[b.dev] + b.cls > MemCache::kCap) {
            if (b.dev != cur) (void)hipSetDevice(b.dev);
            (void)hipFree(b.p);
            if (b.dev != cur) (void)hipSetDevice(cur);
            continue;
        }
        c.free[MemCache::key(b.dev, b.cls)].push_back(b.p);
        c.cached[b.dev] += b.cls;
    }
    blocks.clear();
}

struct DevMem {
    std::vector<CachedBlock> ptrs;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() { release(); }
    void release() { cache_release(ptrs); }
    template <class T> T* alloc(size_t n) {
        size_t cls = 0;
        void* p = cache_alloc(std::max<size_t>(n, 1) * sizeof(T), &cls);
        ptrs.push_back({p, cls, current_device()});
        return static_cast<T*>(p);
    }
    // synchronous: the host vectors are usually temporaries
    template <class T, class Al> T* upload(const std::vector<T, Al>& v, hipStream_t s) {
        T* d = alloc<T>(v.size());
        if (!v.empty()) {
            MLP_HIP_CHECK(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
            MLP_HIP_CHECK(hipStreamSynchronize(s));
        }
        return d;
    }
    template <class T, class S, class Al> T* upload_as(const std::vector<S, Al>& v, hipStream_t s) {
        if constexpr (std::is_same_v<T, S>) {
            return upload(v, s);
        } else {
            HostVec<T> c(v.size());
            parallel_for(v.size(), 1 << 16, [&](uint64_t b, uint64_t e) {
                for (uint64_t i = b; i < e; i++) c[i] = T(v[i]);
            });
            return upload(c, s);
        }
    }
};

// dst = src[0, n), copied by the host threads (the first touch of dst's pages is spread over them too)
template <class T> void par_copy(HostVec<T>& dst, const T* src, size_t n) {
    dst.resize(n);
    parallel_for(n, 1 << 18, [&](uint64_t b, uint64_t e) { std::memcpy(dst.data() + b, src + b, (e - b) * sizeof(T)); });
}

// handle ids key the graph cache; unique across contexts, never reused
inline uint64_t new_handle_id() {
    static std::atomic<uint64_t> next{1};
    return next.fetch_add(1);
}

template <class VT> struct DevSell {
    SellView<VT> v;
    uint64_t nent = 0;
    uint32_t nnz = 0;
    bool stencil = false;                // v uses the seven-point layout
    // bytes of the stored matrix (the layout's arrays)
    uint64_t stored_bytes() const {
        if (stencil) return uint64_t(v.n) * (4 * sizeof(VT) + 1);
        return nent * ((v.dcol ? sizeof(int16_t) : sizeof(uint32_t)) + sizeof(VT)) + uint64_t(v.n / 64 + 2) * 4;
    }
};


}  // namespace
}  // namespace mlp

// ---------------------------------------------------------------------------
// handle types
// ---------------------------------------------------------------------------
struct mlp_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = true;           // false for a sub-context on its parent's stream
    mlp::DevMem mem;
    double* partials = nullptr;        // kPartials doubles
    unsigned* counters = nullptr;      // 8
    double* scal = nullptr;            // 4 result slots for standalone BLAS-1
    mlp::SolveState* st = nullptr;
    double* host_pinned = nullptr;     // 64 doubles
    // batched right-hand sides (mlp_pcg_multi): R vectors of each kind, R states, R partial buffers
    mlp::DevMem multi_mem;
    uint32_t multi_R = 0, multi_n = 0;
    double *mr = nullptr, *mz = nullptr, *mp = nullptr, *mAp = nullptr, *mzp = nullptr, *mpart = nullptr;
    mlp::SolveState* mst = nullptr;
    mlp::SolveState* mst_host = nullptr;   // pinned, R states
    // ragged ensembles (mlp_pcg_ensemble): vectors of the block-diagonal system, per segment a state,
    // a partial-sum buffer and a result slot, the segment offsets on the device
    mlp::DevMem ens_mem;
    uint32_t ens_S = 0, ens_n = 0;
    double *er = nullptr, *ez = nullptr, *ep = nullptr, *eAp = nullptr, *ezp = nullptr, *epart = nullptr;
    double* eout = nullptr;
    uint32_t* esoff = nullptr;
    mlp::SolveState* est = nullptr;
    mlp::SolveState* est_host = nullptr;   // pinned, S states
    double* eout_host = nullptr;           // pinned, S doubles
    // solver workspace, grown on demand and kept for the next solve
    uint32_t ws_n = 0;
    mlp::DevMem ws_mem;
    double *r = nullptr, *z = nullptr, *p = nullptr, *Ap = nullptr, *zp = nullptr, *V = nullptr;
    double* mdot = nullptr;            // 32 results of the Gram-Schmidt dots
    // instantiated PCG iteration graphs, reused while the same matrix,
    // preconditioner and vectors are solved again
    struct GraphEntry {
        int solver = 0;                // 0 PCG, 1 Chebyshev, 2 simple, 3 batched PCG, 4 ensemble PCG
        uint32_t R = 1;                // right-hand sides (segments) per launch
        uint32_t gmax = 0;             // ensemble: the largest segment grid
        uint64_t A, B;
        const void *b, *x;
        int flex;
        uint32_t tre, ws_n;
        hipGraph_t g[3] = {nullptr, nullptr, nullptr};
        hipGraphExec_t e[3] = {nullptr, nullptr, nullptr};
    };
    // a list, so that an entry stays where it is while others are added or dropped
    std::list<GraphEntry> graphs;
    static constexpr size_t kMaxGraphs = 16;
    // Chebyshev outer-iteration coefficients alpha_k, beta_k (k = iteration
    // count) for the bounds (coef_lmin, coef_lmax), filled up to coef_filled
    double* coef = nullptr;
    uint32_t coef_filled = 0;
    double coef_lmin = 0.0, coef_lmax = 0.0;
    static void destroy_graph(GraphEntry& g) {
        for (int k = 0; k < 3; k++) {
            if (g.e[k]) (void)hipGraphExecDestroy(g.e[k]);
            if (g.g[k]) (void)hipGraphDestroy(g.g[k]);
            g.e[k] = nullptr;
            g.g[k] = nullptr;
        }
    }
    // drops the graphs of matrix or preconditioner `id` (0: all); the stream must be idle
    void drop_graphs(uint64_t id) {
        for (auto it = graphs.begin(); it != graphs.end();) {
            if (id == 0 || it->A == id || it->B == id) {
                destroy_graph(*it);
                it = graphs.erase(it);
            } else {
                ++it;
            }
        }
    }
    // caches an instantiated entry, dropping the oldest beyond kMaxGraphs; the stream must be idle
    GraphEntry* add_graph(const GraphEntry& e) {
        while (graphs.size() >= kMaxGraphs) {
            destroy_graph(graphs.front());
            graphs.pop_front();
        }
        graphs.push_back(e);
        return &graphs.back();
    }
};

struct mlp_mat_s {
    mlp_ctx ctx = nullptr;
    uint64_t id = 0;                   // unique per context lifetime (graph cache key)
    uint32_t n = 0, nnz = 0;
    mlp::DevMem mem;
    mlp::DevSell<double> A;
    mlp::HostVec<uint32_t> ia, ja;     // host copy for DiagPrec, GS, Composite and SymPrec
    mlp::HostVec<double> a;
    uint64_t bytes = 0;
    // a matrix assembled on the device keeps its Skyline CSR there (the device
    // construction reads it) and fetches the host copy when something needs it
    uint32_t* d_ia = nullptr;
    uint32_t* d_ja = nullptr;
    double* d_a = nullptr;
    bool host_ready = true;
};

struct mlp_hier_s {
    mlp::HostCSR A0;
    mlp::Hierarchy H;
};

// the host copy of a device-assembled matrix, fetched on first use
static void ensure_host(mlp_mat m) {
    if (m->host_ready) return;
    hipStream_t s = m->ctx->stream;
    m->ia.resize(size_t(m->n) + 1);
    m->ja.resize(m->nnz);
    m->a.resize(m->nnz);
    MLP_HIP_CHECK(hipMemcpyAsync(m->ia.data(), m->d_ia, (size_t(m->n) + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MLP_HIP_CHECK(hipMemcpyAsync(m->ja.data(), m->d_ja, size_t(m->nnz) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MLP_HIP_CHECK(hipMemcpyAsync(m->a.data(), m->d_a, size_t(m->nnz) * sizeof(double), hipMemcpyDeviceToHost, s));
    MLP_HIP_CHECK(hipStreamSynchronize(s));
    m->host_ready = true;
}

namespace mlp {
namespace {

// ---------------------------------------------------------------------------
// device hierarchy
// ---------------------------------------------------------------------------
template <class VT> struct DevLevel {
    uint32_t N = 0, M = 0, Md = 0, n = 0, nit = 0;
    double eps = 0.0;
    // Chebyshev coefficients for the iteration on the next level (cheb_solve.cc:53-101)
    double a1 = 0; int s1_skip = 0;
    double ma2 = 0; int s2_skip = 0; double opb2 = 0; int opb2_one = 0;
    std::vector<double> ak, bk;               // steps 3..ncheb
    uint32_t* map = nullptr;
    ProgView<VT> prog;
    uint32_t lds_down = 0, lds_up = 0, lds_solve = 0;
    uint32_t lbs = kLBS;                       // threads per block of the level kernels (64 or kLBS)
    VT* dval = nullptr;
    DevSell<VT> A;                             // this level's matrix (levels >= 1); uh_cheb stores its rows
                                               // in the order the level kernels read them
    VT *wt = nullptr, *F = nullptr, *u1 = nullptr, *u0 = nullptr, *x2 = nullptr;
    VT* r = nullptr;                           // multi_split inner residual
    // level 0 of the BF16 variant: F, u1, a1 u1 and x2 of the Chebyshev iteration on level 1 in BF16
    hip_bfloat16 *bF = nullptr, *bu1 = nullptr, *bu0 = nullptr, *bx2 = nullptr;
    // bookkeeping for the byte model
    uint32_t nnzL = 0, nnzU = 0;
    // the level run fused with the coarsest level (two launches per visit); the coarsest level's values
    // and its matrix (read by step 2's residual) are FP64 in every hierarchy
    bool fused = false;
    PairView<VT, double> pair;
    DevSell<double> A64;
};
static_assert(kPairHdr == kPairFields, "pair block header size");
static_assert(kChainHdrW == kChainHdr, "chain block header size");

struct DevPrecBase {
    virtual ~DevPrecBase() {}
    virtual void apply(const double* f, double* x, const uint32_t* guard) = 0;
    virtual double bytes_per_apply() const = 0;
    virtual uint32_t launches_per_apply() const = 0;
    virtual bool graph_safe() const { return true; }
    // a copy on another context that shares the device data (nullptr: not supported)
    virtual std::unique_ptr<DevPrecBase> clone_on(mlp_ctx) const { return nullptr; }
    // a copy that applies to rhs right-hand sides per call, vectors stored one
    // after another (nullptr: not supported)
    virtual std::unique_ptr<DevPrecBase> widen(mlp_ctx, uint32_t) const { return nullptr; }
    // the launch schedule in use (MLP_SCHED_* bits, level block size << 8)
    virtual uint32_t schedule() const { return 0; }
};

constexpr uint32_t kLdsBudget = 60 * 1024;

// The pinned staging buffer of one device, shared by its contexts (grown on
// demand, kept until the process ends; a flush holds the lock from packing to
// the end of its copy).
struct PinnedStage {
    std::mutex m;
    char* p = nullptr;
    size_t n = 0;
};
constexpr int kMaxStageDevices = 64;
inline PinnedStage& pinned_stage(int device) {
    static PinnedStage st[kMaxStageDevices];
    if (device < 0 || device >= kMaxStageDevices) throw Error("device index out of range", MLP_ERR_INVALID);
    return st[device];
}

// Uploads a set of host arrays as one device block: the host threads pack
// them (converting element types where asked) into the device's pinned
// staging buffer (a block over max_pinned() goes through a pageable buffer
// instead), one copy moves the block, and flush() then points every
// registered field at its array and runs the actions queued with after().
// The host arrays must live until flush(); own() keeps a temporary alive
// that long.  Packing runs as one parallel loop over pieces of at most 1 MB
// of every array.
struct Stager {
    mlp_ctx ctx;
    DevMem& mem;
    struct Item {
        size_t off, bytes;
        std::function<void(char*, size_t, size_t)> write;   // elements [b, e) into the block at this item
        std::function<void(char*)> set;
        size_t elem;
    };
    std::vector<Item> items;
    std::vector<std::shared_ptr<void>> owned;
    std::vector<std::function<void()>> actions;
    size_t total = 0;
    Stager(mlp_ctx c, DevMem& m) : ctx(c), mem(m) {}
    template <class P, class S> void add(const S* src, size_t n, P& field) {
        using T = std::remove_const_t<std::remove_pointer_t<P>>;
        const size_t off = total;
        total += (std::max<size_t>(n, 1) * sizeof(T) + 255) & ~size_t(255);
        items.push_back({off, n * sizeof(T),
                         [src](char* dst, size_t b, size_t e) {
                             T* d = reinterpret_cast<T*>(dst);
                             if constexpr (std::is_same_v<T, S>) std::memcpy(d + b, src + b, (e - b) * sizeof(T));
                             else for (size_t i = b; i < e; i++) d[i] = T(src[i]);
                         },
                         [&field](char* base) { field = reinterpret_cast<P>(base); }, sizeof(T)});
    }
    template <class P, class S, class Al> void add(const std::vector<S, Al>& v, P& field) {
        add(v.data(), v.size(), field);
    }
    // a temporary that lives until flush()
    template <class V> V& own(V&& v) {
        auto p = std::make_shared<std::decay_t<V>>(std::move(v));
        owned.push_back(p);
        return *p;
    }
    void after(std::function<void()> f) { actions.push_back(std::move(f)); }
    // blocks up to this size use the pinned buffer: 256 MB, or MLP_STAGE_MAX bytes
    static size_t max_pinned() {
        const char* e = getenv("MLP_STAGE_MAX");
        return e ? size_t(std::strtoull(e, nullptr, 10)) : size_t(256) << 20;
    }
    void pack(char* host) {
        struct Piece {
            const Item* it;
            size_t b, e;
        };
        std::vector<Piece> pieces;
        for (const Item& it : items) {
            const size_t n = it.bytes / it.elem, step = std::max<size_t>(1, (size_t(1) << 20) / it.elem);
            for (size_t b = 0; b < n; b += step) pieces.push_back({&it, b, std::min(n, b + step)});
        }
        parallel_for(pieces.size(), 1, [&](uint64_t p0, uint64_t p1) {
            for (uint64_t p = p0; p < p1; p++) pieces[p].it->write(host + pieces[p].it->off, pieces[p].b, pieces[p].e);
        });
    }
    void flush() {
        if (items.empty()) {
            for (auto& f : actions) f();
            actions.clear();
            owned.clear();
            return;
        }
        char* dev = mem.alloc<char>(total);
        if (total <= max_pinned()) {
            PinnedStage& st = pinned_stage(ctx->device);
            std::lock_guard<std::mutex> hold(st.m);
            if (st.n < total) {
                if (st.p) MLP_HIP_CHECK(hipHostFree(st.p));
                st.p = nullptr;
                st.n = 0;
                MLP_HIP_CHECK(hipHostMalloc((void**)&st.p, total, hipHostMallocDefault));
                st.n = total;
            }
            pack(st.p);
            MLP_HIP_CHECK(hipMemcpyAsync(dev, st.p, total, hipMemcpyHostToDevice, ctx->stream));
            MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        } else {
            HostVec<char> buf(total);
            pack(buf.data());
            MLP_HIP_CHECK(hipMemcpyAsync(dev, buf.data(), total, hipMemcpyHostToDevice, ctx->stream));
            MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        }
        for (Item& it : items) it.set(dev + it.off);
        items.clear();
        total = 0;
        for (auto& f : actions) f();
        actions.clear();
        owned.clear();
    }
};

// A matrix in its own row order is stored in the seven-point layout when it is
// a symmetric seven-point grid operator (MLP_NO_STENCIL=1 keeps the slices),
// else as sliced ELL.
// (staged on `up`; its device pointers are set at up.flush())
template <class VT>
void upload_sell(Stager& up, uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a, DevSell<VT>& out,
                 const uint32_t* rows = nullptr) {
    const char* off = getenv("MLP_NO_STENCIL");
    if (!rows && !(off && off[0] == '1')) {
        StencilLayout T;
        if (plan_stencil(n, ia, ja, a, T)) {
            StencilLayout& K = up.own(std::move(T));
            out.v.n = n;
            up.add(K.pres, out.v.spres);
            up.add(K.d, out.v.sd);
            up.add(K.lx, out.v.slx);
            up.add(K.ly, out.v.sly);
            up.add(K.lz, out.v.slz);
            out.v.ssy = K.sy;
            out.v.ssz = K.sz;
            out.stencil = true;
            out.nent = 0;
            out.nnz = ia[n];
            return;
        }
    }
    SellLayout& L = up.own(SellLayout());
    plan_sell(n, ia, ja, a, L, rows);
    out.v.n = n;
    up.add(L.off64, out.v.off64);
    if (!L.dcol.empty()) up.add(L.dcol, out.v.dcol);
    else up.add(L.col, out.v.col);
    up.add(L.val, out.v.val);
    out.nent = L.nent;
    out.nnz = n ? ia[n] : 0;
}
template <class VT>
void upload_sell(mlp_ctx ctx, DevMem& mem, uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                 DevSell<VT>& out, const uint32_t* rows = nullptr) {
    Stager up(ctx, mem);
    upload_sell<VT>(up, n, ia, ja, a, out, rows);
    up.flush();
}

template <class VT> struct DevHier : DevPrecBase {
    mlp_ctx ctx;
    int kind;
    uint32_t nlevels = 0;
    uint32_t R = 1;                          // right-hand sides per apply (level vectors R times over)
    std::vector<DevLevel<VT>> lv;
    std::shared_ptr<DevMem> data = std::make_shared<DevMem>();   // programs, matrices: shared by clones
    DevMem scratch;                                              // per-instance level vectors
    mlp_mat A0 = nullptr;
    mutable uint32_t launches = 0;
    uint32_t last_launches = 0;
    bool dynamic = false;
    int pair_l = -1;                         // level fused with the coarsest level (-1: none)
    bool bf16 = false;                       // level 0 keeps its Chebyshev vectors in BF16 (FP32 uh_cheb)
    // chain of the levels chain_a..nlevels-1 (plan.h: ChainProgram); -1: none
    int chain_a = -1;
    uint32_t chain_bs = kChainBS;            // threads per chain block: kChainBS, or 2 kChainBS (the wide chain)
    uint32_t want_chain_bs = 0;              // a tuner candidate's chain block size (0: as the build chooses)
    bool no_chain = false;                   // build without the chain (the batched copy, see widen)
    bool no_pair = false;                    // keep the last two levels separate (a tuner candidate)
    uint32_t sched_lbs = 0;                  // level blocks of this many threads, every level (a tuner candidate)
    // a chain hierarchy's level arrays, kept on the host for a batched copy built without the chain
    struct LevelCopy {
        std::vector<uint32_t> map, L_ia, L_ja, U_ia, U_ja, A_ia, A_ja;
        std::vector<double> L_a, U_a, dval, A_a;
    };
    std::shared_ptr<std::vector<LevelCopy>> host_levels;
    std::shared_ptr<std::vector<mlp_level_desc>> host_desc;
    ChainView<VT, double> cv;                // the chain program and the chain levels' vectors (device pointers)
    std::vector<uint32_t> chain_blk;         // per chain level: blocks with rows on it
    std::vector<uint32_t> chain_npt;         // per chain level and the coarsest: passthrough rows

    hipStream_t stream() const { return ctx->stream; }

    // Launch shape of the component blocks of a level: `lbs` threads per block
    // and at most `cap` nodes per block (a larger component gets a block of its
    // own).  By default a level whose elimination components all fit one
    // wavefront runs 64-thread blocks of at most 64 nodes (no workgroup
    // barriers in its sweeps), any other level 256-thread blocks of at most
    // 256 nodes.  MLP_LEVEL_BLOCK / MLP_LEVEL_CAP fix the shape of every level
    // for tuning (the cap defaults to 256 then), and MLP_LEVEL_LDS lowers the LDS budget of a block (smaller
    // budgets send more blocks to the global-memory sweeps).
    uint32_t env_lbs = 0, env_cap = 0, lds_budget = kLdsBudget;

    void read_launch_shape() {
        if (const char* e = getenv("MLP_LEVEL_BLOCK")) {
            const uint32_t v = uint32_t(atoi(e));
            if (v != 64 && v != 128 && v != 256) throw Error("MLP_LEVEL_BLOCK must be 64, 128 or 256", 1);
            env_lbs = v;
        }
        if (const char* e = getenv("MLP_LEVEL_CAP")) {
            const uint32_t v = uint32_t(atoi(e));
            if (v < 8 || v > 4096) throw Error("MLP_LEVEL_CAP must lie in [8, 4096]", 1);
            env_cap = v;
        }
        if (const char* e = getenv("MLP_LEVEL_LDS")) {
            const uint32_t v = uint32_t(atoi(e));
            if (v < 1024 || v > kLdsBudget) throw Error("MLP_LEVEL_LDS must lie in [1024, 61440]", 1);
            lds_budget = v;
        }
    }

    // LDS bytes of the kernels' worst case over the LDS-resident blocks of P
    static uint32_t lds_need(uint32_t lbs, uint32_t nodes, uint32_t lent, uint32_t uent, uint32_t unodes,
                             uint32_t levels) {
        if (nodes <= lbs) return 2 * ((nodes * uint32_t(sizeof(VT)) + 7u) & ~7u);   // owned-node path: w and x
        return lds_plan<VT>(nodes, lent, uent, unodes, levels, true, true).total;
    }

    // Plan a level for blocks of `lbs` threads with the largest packing (at
    // most cap0 nodes) whose LDS needs fit the budget.  A component that does
    // not fit alone runs from global memory; if the LDS-resident blocks fit one
    // at a time but not together (the kernels size LDS for the maxima), every
    // block beyond the owned-node path does.
    void plan_fitting(const LevelArrays& la, LevelProgram& P, uint32_t lbs, uint32_t cap0) const {
        const uint32_t budget = lds_budget;
        const BlockFits alone = [lbs, budget](uint32_t nodes, uint32_t lent, uint32_t uent, uint32_t unodes,
                                              uint32_t lev) {
            return lds_need(lbs, nodes, lent, uent, unodes, lev) <= budget;
        };
        auto together = [&]() {
            return lds_need(lbs, P.max_nodes, P.max_Lent, P.max_Uent, P.max_Unodes, P.max_levels) <= budget;
        };
        for (uint32_t cap = cap0;; cap /= 2) {
            plan_level(la, cap, P, alone);
            if (together() || cap <= 8) break;
        }
        if (!together()) {
            plan_level(la, std::min(cap0, lbs), P,
                       [lbs, &alone](uint32_t nodes, uint32_t lent, uint32_t uent, uint32_t unodes, uint32_t lev) {
                           return nodes <= lbs && alone(nodes, lent, uent, unodes, lev);
                       });
        }
        if (!together()) throw Error("level blocks do not fit LDS", 1);
    }

    // the launch shape of one level (see env_lbs) and its program
    // small64: every component has at most 64 nodes (else the 64-thread plan is not tried)
    uint32_t plan_shaped(const LevelArrays& la, LevelProgram& P, bool small64 = true) const {
        if (env_lbs || env_cap) {
            const uint32_t lbs = env_lbs ? env_lbs : kLBS;
            plan_fitting(la, P, lbs, env_cap ? env_cap : kLBS);
            return lbs;
        }
        if (small64) {
            plan_fitting(la, P, 64, 64);
            if (P.max_nodes <= 64 && !P.has_global) return 64;
        }
        plan_fitting(la, P, kLBS, kLBS);
        return kLBS;
    }

    // the device arrays of the fused pair program on level L (see plan.h: PairProgram)
    void upload_pair(Stager& up, uint32_t l, const mlp_level_desc& D, const PairProgram& PP, uint32_t pbs,
                     DevLevel<VT>& L) {
        L.fused = true;
        L.lbs = pbs;
        L.nnzL = D.L_ia ? D.L_ia[D.N - D.Md] : 0;
        L.nnzU = (D.U_ia && D.M) ? D.U_ia[D.M] : 0;
        PairView<VT, double>& V = L.pair;
        V.nblk = PP.nblk; V.nodes = PP.nodes; V.nodes2 = PP.nodes2; V.nflat2 = PP.nflat2; V.Md = D.Md;
        up.add(D.map, D.N, L.map);
        up.add(PP.blk, V.blk);
        up.add(PP.qrow, V.qrow);
        up.add(PP.Ldep, V.Ldep); up.add(PP.Lptr, V.Lptr); up.add(PP.Lloc, V.Lloc); up.add(PP.kept2, V.kept2);
        up.add(PP.Lval, V.Lval);
        up.add(PP.Udep, V.Udep); up.add(PP.uloc, V.uloc); up.add(PP.uxidx, V.uxidx); up.add(PP.Uptr, V.Uptr);
        up.add(PP.Ucol, V.Ucol); up.add(PP.Uval, V.Uval);
        up.add(PP.src2, V.src2); up.add(PP.row2, V.row2); up.add(PP.xl2, V.xl2); up.add(PP.Ldep2, V.Ldep2);
        up.add(PP.Lptr2, V.Lptr2); up.add(PP.Lloc2, V.Lloc2); up.add(PP.Lval2, V.Lval2); up.add(PP.d2, V.d2);
        up.add(PP.Udep2, V.Udep2); up.add(PP.uloc2, V.uloc2); up.add(PP.Uptr2, V.Uptr2); up.add(PP.Ucol2, V.Ucol2);
        up.add(PP.Uval2, V.Uval2);
        up.add(PP.frow2, V.frow2); up.add(PP.fxl, V.fxl); up.add(PP.fd2, V.fd2);
        if constexpr (std::is_same_v<VT, float>) {             // udiv: FP32 multiplies by 1 / u_ii
            std::vector<double>& r1 = up.own(std::vector<double>(PP.Udiag.size()));
            for (size_t i = 0; i < r1.size(); i++) r1[i] = 1.0 / PP.Udiag[i];
            up.add(r1, V.Udiag);
        } else {
            up.add(PP.Udiag, V.Udiag);
        }
        up.add(PP.Udiag2, V.Udiag2);                            // FP64: the coarsest level divides
        up.add(D.dval, D.Md, L.dval);
        up.after([&V, &L] { V.dval = L.dval; });
        // clones size the per-instance w of the block nodes from prog.nodes
        L.prog = ProgView<VT>();
        L.prog.nodes = PP.nodes;
        L.wt = scratch.alloc<VT>(std::max<uint32_t>(PP.nodes, 1));
        // level l's own matrix (l >= 1), read by its parent's fused Chebyshev residual in its consumption order
        if (l > 0) upload_sell<VT>(up, D.N, D.A_ia, D.A_ja, D.A_a, L.A, PP.qrow.data());
        if (L.n) {
            L.F = scratch.alloc<VT>(L.n);
            L.u1 = scratch.alloc<VT>(L.n);
            L.u0 = scratch.alloc<VT>(L.n);
            L.x2 = scratch.alloc<VT>(L.n);
        }
    }

    static LevelArrays arrays_of(const mlp_level_desc& D) {
        LevelArrays la;
        la.N = D.N; la.M = D.M; la.Md = D.Md; la.map = D.map;
        la.L_ia = D.L_ia; la.L_ja = D.L_ja; la.L_a = D.L_a;
        la.U_ia = D.U_ia; la.U_ja = D.U_ja; la.U_a = D.U_a; la.dval = D.dval;
        return la;
    }

    // the device arrays of chain level l (plan.h: ChainLevelProgram) and its vectors
    template <class T>
    void upload_chain_level(Stager& up, uint32_t l, const mlp_level_desc& D, const ChainLevelProgram& C,
                            ChainLv<T>& V, DevLevel<VT>& L) {
        const uint32_t c = nlevels - 1;
        L.nnzL = D.L_ia ? D.L_ia[D.N - D.Md] : 0;
        L.nnzU = (D.U_ia && D.M) ? D.U_ia[D.M] : 0;
        up.add(D.map, D.N, L.map);
        up.add(C.hdr, V.hdr); up.add(C.qrow, V.qrow); up.add(C.src, V.src); up.add(C.xl, V.xl);
        up.add(C.Lmeta, V.Lmeta); up.add(C.Lv2, V.Lv2); up.add(C.Lptr, V.Lptr); up.add(C.Lloc, V.Lloc);
        up.add(C.Lval, V.Lval);
        up.add(C.Umeta, V.Umeta); up.add(C.Uv1, V.Uv1); up.add(C.uxidx, V.uxidx); up.add(C.Uptr, V.Uptr);
        up.add(C.Ucol, V.Ucol); up.add(C.Uval, V.Uval); up.add(C.dval, V.dval);
        if constexpr (std::is_same_v<T, float>) {               // udiv: FP32 multiplies by 1 / u_ii
            std::vector<double>& urcp = up.own(std::vector<double>(C.Udiag.size()));
            for (size_t i = 0; i < urcp.size(); i++) urcp[i] = 1.0 / C.Udiag[i];
            up.add(urcp, V.Udiag);
        } else {
            up.add(C.Udiag, V.Udiag);
        }
        V.nslots = C.nslots;
        V.npt = C.npt;
        chain_npt[l - chain_a] = C.npt;
        // this level's matrix in its chain order: read by the Chebyshev residual of its second visits (the
        // coarsest level's in FP64, by step 2)
        if (l == c) {
            upload_sell<double>(up, D.N, D.A_ia, D.A_ja, D.A_a, L.A64, C.qrow.data());
            L.A.nnz = L.A64.nnz;
        } else if (l > 0) {
            upload_sell<VT>(up, D.N, D.A_ia, D.A_ja, D.A_a, L.A, C.qrow.data());
        }
        if (l < c) {
            L.prog = ProgView<VT>();
            L.prog.nodes = L.N;                                  // wt: one value per chain position
            L.wt = scratch.alloc<VT>(std::max<uint32_t>(L.N, 1));
            if (L.n) {                                           // (no u0: the chain forms a1 u1 where it reads it)
                L.F = scratch.alloc<VT>(L.n);
                L.u1 = scratch.alloc<VT>(L.n);
                L.x2 = scratch.alloc<VT>(L.n);
            }
        }
    }

    // the chain levels' vectors and combine coefficients in cv (after the coefficients are set, and again in
    // a clone, whose vectors are its own)
    void chain_vectors() {
        if (chain_a < 0) return;
        const uint32_t c = nlevels - 1;
        cv.nv = c - uint32_t(chain_a);
        cv.st = uint32_t(sizeof(SolveState) / sizeof(uint32_t));
        for (uint32_t l = uint32_t(chain_a); l < c; l++) {
            const DevLevel<VT>& L = lv[l];
            ChainVec<VT>& V = cv.V[l - chain_a];
            V.F = L.F; V.u1 = L.u1; V.x2 = L.x2; V.wt = L.wt;
            V.a1 = L.a1; V.s1_skip = L.s1_skip;
            V.ma2 = L.ma2; V.s2_skip = L.s2_skip; V.opb2 = L.opb2; V.opb2_one = L.opb2_one;
            V.n = L.n;
            V.N = L.N;
        }
        cv.a1c = lv[c - 1].a1;
        cv.A2 = lv[c].A64.v;
    }

    void build(const std::vector<mlp_level_desc>& d, int kind_, mlp_mat A0_) {
        kind = kind_;
        A0 = A0_;
        nlevels = uint32_t(d.size());
        lv.resize(nlevels);
        read_launch_shape();
        if (sched_lbs) env_lbs = env_cap = sched_lbs;
        hipStream_t s = stream();
        // plan every level first: uh_cheb links each level's F to the next
        // level's consumption order
        std::vector<LevelProgram> progs(nlevels);
        for (uint32_t l = 0; l < nlevels; l++) {
            const mlp_level_desc& D = d[l];
            if (D.M + D.Md > D.N) throw Error("level sizes inconsistent", 1);
            const uint32_t n = D.N - D.M - D.Md;
            if (l + 1 == nlevels && n != 0) throw Error("the coarsest level must have no kept rows", 1);
            if (l + 1 < nlevels && d[l + 1].N != n) throw Error("level sizes do not chain", 1);
        }
        if (bf16 && (kind != MLP_PREC_UH_CHEB || nlevels < 2 || d[0].nit != 2))
            throw Error("the BF16 variant needs a uh_cheb hierarchy of two or more levels with ncheb = 2 on level 0", 1);
        // uh_cheb runs its last levels as a chain (descent and ascent launches) from the highest level a such that
        // levels a..c-1 have ncheb = 2, at most kChainMax of them, and every merged component fits a wavefront
        // (MLP_NO_CHAIN=1: level kernels and the fused last pair instead).  The BF16 variant keeps level 0 apart.
        // every level's largest component (plan_chain, plan_pair and plan_level cannot fit a 64-thread block
        // past 64 nodes, so those attempts are skipped)
        std::vector<uint32_t> big(nlevels, 0);
        if (kind == MLP_PREC_UH_CHEB || kind == MLP_PREC_MULTI_SPLIT)
            parallel_for(nlevels, 1, [&](uint64_t k0, uint64_t k1) {
                for (uint64_t k = k0; k < k1; k++) big[k] = max_level_component(arrays_of(d[k]));
            });
        // The chain runs in one-wavefront blocks (components of at most kChainBS nodes).  Two-wavefront blocks
        // (at most 2 kChainBS nodes: most 60x220x85 SPE10 fields have a larger component on some level) are
        // used when asked for, by MLP_CHAIN_BS=128 or a tuner candidate: their plan costs about as much host
        // time as the level programs and the pair together, which pays only over many solves.
        ChainProgram CP;
        const char* nochain = getenv("MLP_NO_CHAIN");
        if (kind == MLP_PREC_UH_CHEB && nlevels >= 2 && !no_chain && !(nochain && nochain[0] == '1')) {
            const uint32_t c = nlevels - 1;
            uint32_t want = want_chain_bs;
            if (const char* e = getenv("MLP_CHAIN_BS")) want = uint32_t(std::atoi(e));
            if (want && want != kChainBS && want != 2 * kChainBS) throw Error("MLP_CHAIN_BS must be 64 or 128", 1);
            uint32_t top = c > uint32_t(kChainMax) ? c - uint32_t(kChainMax) : 0u;
            if (bf16) top = std::max(top, 1u);
            for (uint32_t l = top; l < c; l++)
                if (d[l].nit != 2) top = l + 1;
            for (const uint32_t bs : {kChainBS, 2 * kChainBS}) {
                if (bs != (want ? want : kChainBS)) continue;
                // a level whose own largest component exceeds a block rules out every chain through it
                uint32_t a0 = top;
                for (uint32_t l = a0; l <= c && a0 < c; l++)
                    if (big[l] > bs) a0 = l + 1;
                const uint32_t a1 = chain_a >= 0 ? uint32_t(chain_a) : c;   // only a longer chain replaces one
                for (uint32_t a = a0; a < a1; a++) {
                    std::vector<LevelArrays> la;
                    for (uint32_t l = a; l <= c; l++) la.push_back(arrays_of(d[l]));
                    ChainProgram P;
                    if (plan_chain(la, a, bs, 2 * bs, P)) {
                        CP = std::move(P);
                        chain_a = int(a);
                        chain_bs = bs;
                        chain_blk = CP.blk_from;
                        chain_npt.assign(CP.nlev, 0);
                        break;
                    }
                }
            }
        }
        const uint32_t nplan = chain_a >= 0 ? uint32_t(chain_a) : nlevels;
        // The last two uh_cheb levels run fused when every merged component fits a block (MLP_NO_PAIR=1 keeps
        // them separate; the BF16 variant runs level 0 on its own kernels).  The pair and the levels it does
        // not cover are planned independently, one thread each; the pair's levels get level programs only if
        // the pair does not fit.
        const char* nopair = getenv("MLP_NO_PAIR");
        const bool try_pair = kind == MLP_PREC_UH_CHEB && chain_a < 0 && nlevels >= 2 && d[nlevels - 2].nit == 2 &&
                              !(nopair && nopair[0] == '1') && !no_pair && !(bf16 && nlevels == 2);
        PairProgram PP;
        uint32_t pbs = 0;
        auto arrays = [&](uint32_t l) {
            const mlp_level_desc& D = d[l];
            LevelArrays la;
            la.N = D.N; la.M = D.M; la.Md = D.Md; la.map = D.map;
            la.L_ia = D.L_ia; la.L_ja = D.L_ja; la.L_a = D.L_a;
            la.U_ia = D.U_ia; la.U_ja = D.U_ja; la.U_a = D.U_a; la.dval = D.dval;
            return la;
        };
        constexpr uint32_t kPairTask = 0xffffffffu;
        auto plan_tasks = [&](const std::vector<uint32_t>& tasks) {
            std::vector<std::exception_ptr> err(tasks.size());
            parallel_for(tasks.size(), 1, [&](uint64_t t0, uint64_t t1) {
                for (uint64_t t = t0; t < t1; t++) {
                    try {
                        const uint32_t l = tasks[t];
                        if (l == kPairTask) {
                            const LevelArrays la = arrays(nlevels - 2), lb = arrays(nlevels - 1);
                            if (big[nlevels - 2] <= 64 && big[nlevels - 1] <= 64 && plan_pair(la, lb, 64, PP)) pbs = 64;
                            else if (plan_pair(la, lb, kLBS, PP)) pbs = kLBS;
                        } else {
                            lv[l].lbs = plan_shaped(arrays(l), progs[l], big[l] <= 64);
                        }
                    } catch (...) {
                        err[t] = std::current_exception();
                    }
                }
            });
            for (auto& e : err)
                if (e) std::rethrow_exception(e);
        };
        const uint32_t lead = try_pair ? nlevels - 2 : nplan;     // levels planned alongside the pair
        std::vector<uint32_t> tasks;
        if (try_pair) tasks.push_back(kPairTask);
        for (uint32_t l = 0; l < lead; l++) tasks.push_back(l);
        plan_tasks(tasks);
        if (try_pair) {
            if (pbs) {
                pair_l = int(nlevels - 2);
            } else {
                tasks.clear();
                for (uint32_t l = lead; l < nplan; l++) tasks.push_back(l);
                plan_tasks(tasks);
            }
        }
        if (kind == MLP_PREC_UH_CHEB)
            for (uint32_t l = 0; l < nplan && l + 1 < nlevels; l++) {
                if (int(l) == pair_l) continue;
                if (int(l + 1) == pair_l) link_levels(progs[l], PP.qrow.data(), PP.qrow.size());
                else if (int(l + 1) == chain_a) link_levels(progs[l], CP.lv[0].qrow.data(), CP.lv[0].qrow.size());
                else link_levels(progs[l], progs[l + 1]);
            }
        // every level's arrays go up in one staged copy (device pointers are set when it is flushed)
        Stager up(ctx, *data);
        for (uint32_t l = 0; l < nlevels; l++) {
            const mlp_level_desc& D = d[l];
            const LevelProgram& P = progs[l];
            DevLevel<VT>& L = lv[l];
            L.N = D.N; L.M = D.M; L.Md = D.Md; L.nit = D.nit; L.eps = D.eps;
            L.n = L.N - L.M - L.Md;
            if (L.eps != 0.0) dynamic = true;
            if (chain_a >= 0 && l >= uint32_t(chain_a)) {
                if (l + 1 == nlevels) upload_chain_level(up, l, D, CP.lv[l - chain_a], cv.C, L);
                else upload_chain_level(up, l, D, CP.lv[l - chain_a], cv.L[l - chain_a], L);
                continue;
            }
            if (int(l) == pair_l) {
                upload_pair(up, l, D, PP, pbs, L);
                continue;
            }
            if (pair_l >= 0 && int(l) == pair_l + 1) {
                // the coarsest level runs inside the pair kernels; only its matrix is read here (FP64),
                // by step 2's residual, rows in the pair's level-(l+1) consumption order
                upload_sell<double>(up, D.N, D.A_ia, D.A_ja, D.A_a, L.A64, PP.qrow2.data());
                L.A.nnz = L.A64.nnz;                            // for the byte model
                L.nnzL = D.L_ia ? D.L_ia[D.N - D.Md] : 0;
                L.nnzU = (D.U_ia && D.M) ? D.U_ia[D.M] : 0;
                continue;
            }
            up.add(D.map, D.N, L.map);
            L.nnzL = D.L_ia ? D.L_ia[D.N - D.Md] : 0;
            L.nnzU = (D.U_ia && D.M) ? D.U_ia[D.M] : 0;
            ProgView<VT>& G = L.prog;
            G.nblk = P.nblk;
            G.nodes = P.nodes;
            G.nflat = uint32_t(P.flat.size());
            G.max_nodes = P.max_nodes; G.max_Lent = P.max_Lent; G.max_Uent = P.max_Uent; G.max_Unodes = P.max_Unodes;
            G.max_levels = P.max_levels;
            up.add(P.qrow, G.qrow);
            up.add(P.flatq, G.flatq);
            if (P.nblk) {
                up.add(P.blk, G.blk);
                up.add(P.lvlL, G.lvlL);
                up.add(P.lvlU, G.lvlU);
                up.add(P.out, G.out);
                up.add(P.Lptr, G.Lptr);
                up.add(P.Lloc, G.Lloc);
                up.add(P.Lval, G.Lval);
                up.add(P.Ldep, G.Ldep);
                up.add(P.Udep, G.Udep);
                up.add(P.uloc, G.uloc);
                up.add(P.uxidx, G.uxidx);
                up.add(P.Uptr, G.Uptr);
                up.add(P.Ucol, G.Ucol);
                up.add(P.Uval, G.Uval);
                if constexpr (std::is_same_v<VT, float>) {         // udiv: FP32 multiplies by 1 / u_ii
                    std::vector<double>& urcp = up.own(std::vector<double>(P.Udiag.size()));
                    for (size_t i = 0; i < urcp.size(); i++) urcp[i] = 1.0 / P.Udiag[i];
                    up.add(urcp, G.Udiag);
                } else {
                    up.add(P.Udiag, G.Udiag);
                }
                L.wt = scratch.alloc<VT>(P.nodes);
                if (P.has_global) G.gx = scratch.alloc<VT>(P.nodes);
            }
            if (P.max_nodes <= L.lbs && P.max_Unodes <= L.lbs) {
                // every block takes the owned-node path, which keeps only w (sw) and x (sx) in LDS
                const uint32_t v = (P.max_nodes * uint32_t(sizeof(VT)) + 7u) & ~7u;
                L.lds_down = v;
                L.lds_up = L.lds_solve = 2 * v;
            } else {
                L.lds_down = lds_plan<VT>(P.max_nodes, P.max_Lent, P.max_Uent, P.max_Unodes, P.max_levels, true, false).total;
                L.lds_up = lds_plan<VT>(P.max_nodes, P.max_Lent, P.max_Uent, P.max_Unodes, P.max_levels, false, true).total;
                L.lds_solve = lds_plan<VT>(P.max_nodes, P.max_Lent, P.max_Uent, P.max_Unodes, P.max_levels, true, true).total;
            }
            up.add(D.dval, D.Md, L.dval);
            if (l > 0) {
                // uh_cheb reads A_l only through the fused Chebyshev residual, row by row in consumption order
                if (kind == MLP_PREC_UH_CHEB)
                    upload_sell<VT>(up, D.N, D.A_ia, D.A_ja, D.A_a, L.A, P.qrow.data());
                else
                    upload_sell<VT>(up, D.N, D.A_ia, D.A_ja, D.A_a, L.A);
            }
            if (L.n && bf16 && l == 0) {
                L.bF = scratch.alloc<hip_bfloat16>(L.n);
                L.bu1 = scratch.alloc<hip_bfloat16>(L.n);
                L.bu0 = scratch.alloc<hip_bfloat16>(L.n);
                L.bx2 = scratch.alloc<hip_bfloat16>(L.n);
            } else if (L.n) {
                L.F = scratch.alloc<VT>(L.n);
                L.u1 = scratch.alloc<VT>(L.n);
                L.u0 = scratch.alloc<VT>(L.n);
                L.x2 = scratch.alloc<VT>(L.n);
            }
            if (kind == MLP_PREC_MULTI_SPLIT && l > 0) L.r = scratch.alloc<VT>(L.N);
        }
        up.flush();
        if (kind == MLP_PREC_UH_CHEB) {
            for (uint32_t l = 0; l + 1 < nlevels; l++) {
                DevLevel<VT>& L = lv[l];
                if (L.nit == 0) throw Error("ncheb must be positive", 1);
                const double lmin = d[l + 1].lmin, lmax = d[l + 1].lmax;
                const double eta = (lmax + lmin) / (lmax - lmin);
                L.a1 = 2 / (lmax + lmin);
                L.s1_skip = host_is_one(L.a1);
                const double alpha2 = 4 / (lmax - lmin) * cheb_poly(eta, 1) / cheb_poly(eta, 2);
                const double beta2 = cheb_poly(eta, 0) / cheb_poly(eta, 2);
                L.ma2 = -alpha2;
                L.s2_skip = host_is_one(-alpha2);
                L.opb2 = 1 + beta2;
                L.opb2_one = host_is_one(1 + beta2);
                for (uint32_t i = 3; i <= L.nit; i++) {
                    L.ak.push_back(4 / (lmax - lmin) * cheb_poly(eta, i - 1) / cheb_poly(eta, i));
                    L.bk.push_back(cheb_poly(eta, i - 2) / cheb_poly(eta, i));
                }
            }
        }
        if (kind == MLP_PREC_MULTI_SPLIT && nlevels && lv[0].nit > 1)
            throw Error("multi_split level 0 runs one iteration", 1);
        chain_vectors();
        if (chain_a >= 0) keep_levels(d);
        MLP_HIP_CHECK(hipStreamSynchronize(s));
    }

    // a host copy of the level arrays (one thread per level), for widen
    void keep_levels(const std::vector<mlp_level_desc>& d) {
        auto copy = std::make_shared<std::vector<LevelCopy>>(d.size());
        auto desc = std::make_shared<std::vector<mlp_level_desc>>(d);
        parallel_for(d.size(), 1, [&](uint64_t l0, uint64_t l1) {
            for (uint64_t l = l0; l < l1; l++) {
                const mlp_level_desc& D = d[l];
                LevelCopy& C = (*copy)[l];
                mlp_level_desc& E = (*desc)[l];
                const uint32_t K = D.N - D.Md;
                auto take = [](auto& v, const auto* p, size_t n) {
                    if (p) v.assign(p, p + n);
                    return p ? v.data() : nullptr;
                };
                E.map = take(C.map, D.map, D.N);
                E.L_ia = take(C.L_ia, D.L_ia, K + 1);
                const size_t nl = D.L_ia ? D.L_ia[K] : 0, nu = (D.U_ia && D.M) ? D.U_ia[D.M] : 0;
                E.L_ja = take(C.L_ja, D.L_ja, nl);
                E.L_a = take(C.L_a, D.L_a, nl);
                E.U_ia = take(C.U_ia, D.U_ia, D.U_ia ? size_t(D.M) + 1 : 0);
                E.U_ja = take(C.U_ja, D.U_ja, nu);
                E.U_a = take(C.U_a, D.U_a, nu);
                E.dval = take(C.dval, D.dval, D.Md);
                const size_t na = D.A_ia ? D.A_ia[D.N] : 0;
                E.A_ia = take(C.A_ia, D.A_ia, D.A_ia ? size_t(D.N) + 1 : 0);
                E.A_ja = take(C.A_ja, D.A_ja, na);
                E.A_a = take(C.A_a, D.A_a, na);
            }
        });
        host_levels = copy;
        host_desc = desc;
    }

    // A preconditioner on context c that shares this one's device data and has
    // level vectors of its own, so both can be applied at the same time.
    std::unique_ptr<DevHier<VT>> clone(mlp_ctx c, uint32_t rhs = 1) const {
        auto h = std::make_unique<DevHier<VT>>();
        h->ctx = c;
        h->kind = kind;
        h->nlevels = nlevels;
        h->R = rhs;
        h->lv = lv;
        h->data = data;
        h->A0 = A0;
        h->dynamic = dynamic;
        h->pair_l = pair_l;
        h->bf16 = bf16;
        h->chain_a = chain_a;
        h->chain_bs = chain_bs;
        h->want_chain_bs = want_chain_bs;
        h->no_chain = no_chain;
        h->cv = cv;
        h->chain_blk = chain_blk;
        h->chain_npt = chain_npt;
        h->host_levels = host_levels;
        h->host_desc = host_desc;
        h->env_lbs = env_lbs;
        h->env_cap = env_cap;
        h->lds_budget = lds_budget;
        DevMem& sc = h->scratch;
        const size_t k = rhs;
        for (auto& L : h->lv) {
            if (L.wt) L.wt = sc.template alloc<VT>(k * std::max<uint32_t>(L.prog.nodes, 1));
            if (L.prog.gx) L.prog.gx = sc.template alloc<VT>(k * L.prog.nodes);
            if (L.F) {
                L.F = sc.template alloc<VT>(k * L.n);
                L.u1 = sc.template alloc<VT>(k * L.n);
                L.u0 = sc.template alloc<VT>(k * L.n);
                L.x2 = sc.template alloc<VT>(k * L.n);
            }
            if (L.bF) {
                L.bF = sc.template alloc<hip_bfloat16>(k * L.n);
                L.bu1 = sc.template alloc<hip_bfloat16>(k * L.n);
                L.bu0 = sc.template alloc<hip_bfloat16>(k * L.n);
                L.bx2 = sc.template alloc<hip_bfloat16>(k * L.n);
            }
            if (L.r) L.r = sc.template alloc<VT>(k * L.N);
        }
        h->chain_vectors();
        return h;
    }

    // a copy on context c that applies to `rhs` right-hand sides at once
    // (uh_cheb; multi_split's dynamic inner loops are per right-hand side)
    // With many right-hand sides per launch the apply is bound by throughput, where the level kernels (short,
    // light waves) beat the chain's long ones: a chain hierarchy builds its batched copy without the chain.
    std::unique_ptr<DevPrecBase> widen(mlp_ctx c, uint32_t rhs) const override {
        if (kind != MLP_PREC_UH_CHEB) return nullptr;
        if (chain_a >= 0 && rhs > 1 && host_desc) {
            auto h = std::make_unique<DevHier<VT>>();
            h->ctx = c;
            h->bf16 = bf16;
            h->no_chain = true;
            h->env_lbs = env_lbs;
            h->env_cap = env_cap;
            h->lds_budget = lds_budget;
            h->build(*host_desc, kind, A0);
            return h->clone(c, rhs);
        }
        return clone(c, rhs);
    }

    std::unique_ptr<DevPrecBase> clone_on(mlp_ctx c) const override { return clone(c); }

    // --- launch helpers ---
    // every launch covers the R right-hand sides of a batched apply (grid.y)
    template <class K, class... Args>
    void launch(K kern, uint32_t grid, uint32_t block, uint32_t smem, Args... args) const {
        if (grid == 0) return;
        hipLaunchKernelGGL(kern, dim3(grid, R), dim3(block), smem, ctx->stream, args...);
        launches++;
    }
    // per-right-hand-side strides of level L's vectors (R independent problems)
    static LevelBatch level_batch(const DevLevel<VT>& L) {
        LevelBatch b;
        b.N = L.N;
        b.nodes = L.prog.nodes;
        b.n = L.n;
        b.st = uint32_t(sizeof(SolveState) / sizeof(uint32_t));
        return b;
    }

    // Each level kernel comes in four instances: one-wavefront (64) or kLBS
    // blocks, and with or without the batched right-hand-side offsets.
    // (F: where the kept rows' right-hand side goes, in type FV)
    template <class FT, int IN, class XT, int DMODE, class FV>
    void down(const DevLevel<VT>& L, const InSrc<VT, FT, IN>& f, XT* x, const uint32_t* g, FV* F) const {
        const uint32_t grid = L.prog.nblk + cdiv(uint64_t(L.prog.nflat) + L.Md, L.lbs);
        auto k = L.lbs == 64 ? (R > 1 ? k_level_down<VT, FT, IN, XT, DMODE, 64, true, FV>
                                      : k_level_down<VT, FT, IN, XT, DMODE, 64, false, FV>)
                             : (R > 1 ? k_level_down<VT, FT, IN, XT, DMODE, kLBS, true, FV>
                                      : k_level_down<VT, FT, IN, XT, DMODE, kLBS, false, FV>);
        launch(k, grid, L.lbs, L.lds_down, L.prog, f, L.wt, F, L.Md, (const VT*)L.dval, x, g, level_batch(L));
    }
    template <class FT, int IN, class XT, int DMODE>
    void down(const DevLevel<VT>& L, const InSrc<VT, FT, IN>& f, XT* x, const uint32_t* g) const {
        down<FT, IN, XT, DMODE, VT>(L, f, x, g, L.F);
    }
    template <class XT, int OP, int UMODE, class CV = VT>
    void up(const DevLevel<VT>& L, const Combine<VT, CV>& C, XT* x, const uint32_t* g) const {
        const uint32_t grid = L.prog.nblk + cdiv(L.n, L.lbs);
        auto k = L.lbs == 64 ? (R > 1 ? k_level_up<VT, XT, OP, UMODE, 64, true, CV>
                                      : k_level_up<VT, XT, OP, UMODE, 64, false, CV>)
                             : (R > 1 ? k_level_up<VT, XT, OP, UMODE, kLBS, true, CV>
                                      : k_level_up<VT, XT, OP, UMODE, kLBS, false, CV>);
        launch(k, grid, L.lbs, L.lds_up, L.prog, (const VT*)L.wt, C, (const uint32_t*)(L.map + L.M), L.n, x, g,
               level_batch(L));
    }
    template <class FT, int IN, class XT, int MODE>
    void solve_last(const DevLevel<VT>& L, const InSrc<VT, FT, IN>& f, XT* x, const uint32_t* g) const {
        const uint32_t grid = L.prog.nblk + cdiv(L.Md, L.lbs);
        auto k = L.lbs == 64 ? (R > 1 ? k_level_solve<VT, FT, IN, XT, MODE, 64, true>
                                      : k_level_solve<VT, FT, IN, XT, MODE, 64, false>)
                             : (R > 1 ? k_level_solve<VT, FT, IN, XT, MODE, kLBS, true>
                                      : k_level_solve<VT, FT, IN, XT, MODE, kLBS, false>);
        launch(k, grid, L.lbs, L.lds_solve, L.prog, f, L.Md, (const VT*)L.dval, x, g, level_batch(L));
    }

    static InSrc<VT, VT, IN_PLAIN> plain(const VT* f) {
        InSrc<VT, VT, IN_PLAIN> S;
        S.f = f;
        return S;
    }
    // F of the parent level, stored in this level's consumption order
    static InSrc<VT, VT, IN_PERM> perm(const VT* f) {
        InSrc<VT, VT, IN_PERM> S;
        S.f = f;
        return S;
    }
    // f = A_{l+1} (s u) - F_l, read row by row by the visit of level l+1
    template <int IN> InSrc<VT, VT, IN> residual_src(uint32_t l, const VT* u, double s, VT* us) const {
        InSrc<VT, VT, IN> S;
        S.A = lv[l + 1].A.v;
        S.u = u;
        S.F = lv[l].F;
        S.us = us;
        S.s = s;
        return S;
    }

    // Prec::solve(level, f, x)  (modules/prec/cheb/cheb_solve.cc:13-123).  The
    // Chebyshev residual tmp = A_{l+1} u - F of steps >= 2 is not stored: the
    // visit of level l+1 computes each row where it reads it.
    // a visit of the level fused with the coarsest level: K1 and K2 (kernels.h: k_pair_down, k_pair_up)
    template <class FT, int IN, class XT>
    void pair_visit(const DevLevel<VT>& L, const DevLevel<VT>& Lc, const InSrc<VT, FT, IN>& f, XT* x,
                    const uint32_t* g) const {
        const PairView<VT, double>& P = L.pair;
        const LevelBatch lb = level_batch(L);
        const uint32_t g1 = P.nblk + cdiv(uint64_t(P.nflat2) + P.Md, L.lbs);
        auto k1 = L.lbs == 64 ? (R > 1 ? k_pair_down<VT, double, FT, IN, XT, 64, true>
                                       : k_pair_down<VT, double, FT, IN, XT, 64, false>)
                              : (R > 1 ? k_pair_down<VT, double, FT, IN, XT, kLBS, true>
                                       : k_pair_down<VT, double, FT, IN, XT, kLBS, false>);
        launch(k1, g1, L.lbs, 0, P, f, L.wt, L.F, L.u1, x, g, lb);
        const uint32_t g2 = P.nblk + cdiv(P.nflat2, L.lbs);
        const VT* wt = L.wt;
        if (L.s1_skip) {
            auto k2 = L.lbs == 64 ? (R > 1 ? k_pair_up<VT, double, XT, false, 64, true>
                                           : k_pair_up<VT, double, XT, false, 64, false>)
                                  : (R > 1 ? k_pair_up<VT, double, XT, false, kLBS, true>
                                           : k_pair_up<VT, double, XT, false, kLBS, false>);
            launch(k2, g2, L.lbs, 0, P, Lc.A64.v, (const VT*)L.u1, 1.0, (const VT*)L.F, wt, L.ma2, L.s2_skip,
                   L.opb2, L.opb2_one, x, g, lb);
        } else {
            auto k2 = L.lbs == 64 ? (R > 1 ? k_pair_up<VT, double, XT, true, 64, true>
                                           : k_pair_up<VT, double, XT, true, 64, false>)
                                  : (R > 1 ? k_pair_up<VT, double, XT, true, kLBS, true>
                                           : k_pair_up<VT, double, XT, true, kLBS, false>);
            launch(k2, g2, L.lbs, 0, P, Lc.A64.v, (const VT*)L.u1, L.a1, (const VT*)L.F, wt, L.ma2,
                   L.s2_skip, L.opb2, L.opb2_one, x, g, lb);
        }
    }

    // Level 0 of the BF16 variant (ncheb = 2): as cheb(0, ...), with the
    // vectors of the Chebyshev iteration on level 1 (F, u1, a1 u1, x2) stored
    // in BF16.  Every kernel still computes in the level's own type; only the
    // stores round to BF16 and the loads widen from it.
    void cheb_bf16(const InSrc<VT, double, IN_PLAIN>& f, double* x, const uint32_t* g) const {
        using BF = hip_bfloat16;
        const DevLevel<VT>& L = lv[0];
        down<double, IN_PLAIN, double, 0, BF>(L, f, x, g, L.bF);
        InSrc<VT, BF, IN_PERM> pf;
        pf.f = L.bF;
        cheb(1, pf, L.bu1, g);
        auto res = [&](auto S, BF* us, double s) {
            S.A = lv[1].A.v;
            S.u = L.bu1;
            S.F = L.bF;
            S.us = us;
            S.s = s;
            cheb(1, S, L.bx2, g);
        };
        if (L.s1_skip) res(InSrc<VT, BF, IN_RES>(), nullptr, 1.0);
        else res(InSrc<VT, BF, IN_RES_SCALE>(), L.bu0, L.a1);
        Combine<VT, BF> C;
        C.x2 = L.bx2; C.u1 = L.s1_skip ? L.bu1 : L.bu0;
        C.a = L.ma2; C.skip = L.s2_skip; C.b = L.opb2; C.one = L.opb2_one;
        up<double, OP_CHEB2, 0>(L, C, x, g);
    }

    // --- the chain (plan.h: ChainProgram) ---
    // The chain kernels come in two register budgets: NL = 1 for launches that run only the last chain level
    // (with the coarsest), NL = kChainMax for any other.
    template <class FT, int IN, class XT, int NL>
    static auto chain_down_kernel(bool batched, bool narrow) {
        constexpr int W = int(2 * kChainBS);
        return narrow ? (batched ? k_chain_down<VT, double, FT, IN, XT, true, kChainBS, NL>
                                 : k_chain_down<VT, double, FT, IN, XT, false, kChainBS, NL>)
                      : (batched ? k_chain_down<VT, double, FT, IN, XT, true, W, NL>
                                 : k_chain_down<VT, double, FT, IN, XT, false, W, NL>);
    }
    template <class FT, int IN, class XT>
    void chain_down(uint32_t j, const InSrc<VT, FT, IN>& f, XT* out, const uint32_t* g) const {
        const uint32_t i0 = j - uint32_t(chain_a);
        const uint32_t grid = chain_blk[i0] + cdiv(chain_npt[i0], chain_bs);
        const bool narrow = chain_bs == kChainBS;
        auto k = i0 + 1 == cv.nv ? chain_down_kernel<FT, IN, XT, 1>(R > 1, narrow)
                                 : chain_down_kernel<FT, IN, XT, kChainMax>(R > 1, narrow);
        launch(k, grid, chain_bs, 0, cv, i0, f, out, lv[j].N, chain_blk[i0], g);
    }
    template <class XT, bool SCALE, int NL>
    static auto chain_up_kernel(bool batched, bool narrow) {
        constexpr int W = int(2 * kChainBS);
        return narrow ? (batched ? k_chain_up<VT, double, XT, SCALE, true, kChainBS, NL>
                                 : k_chain_up<VT, double, XT, SCALE, false, kChainBS, NL>)
                      : (batched ? k_chain_up<VT, double, XT, SCALE, true, W, NL>
                                 : k_chain_up<VT, double, XT, SCALE, false, W, NL>);
    }
    template <class XT, bool SCALE>
    auto chain_up_pick(uint32_t ie) const {
        const bool narrow = chain_bs == kChainBS;
        return ie + 1 == cv.nv ? chain_up_kernel<XT, SCALE, 1>(R > 1, narrow)
                               : chain_up_kernel<XT, SCALE, kChainMax>(R > 1, narrow);
    }
    template <class XT>
    void chain_up(uint32_t ke, XT* out, const uint32_t* g) const {
        const uint32_t ie = ke - uint32_t(chain_a);
        const uint32_t grid = chain_blk[ie] + cdiv(chain_npt[ie + 1], chain_bs);
        if (lv[nlevels - 2].s1_skip)
            launch(chain_up_pick<XT, false>(ie), grid, chain_bs, 0, cv, ie, out, lv[ke].N, chain_blk[ie], g);
        else
            launch(chain_up_pick<XT, true>(ie), grid, chain_bs, 0, cv, ie, out, lv[ke].N, chain_blk[ie], g);
    }
    // The operations of Prec::solve on the chain levels, in order: D (L sweep of a level: its visit starts),
    // the coarsest level's first and second visits, U (U sweep: a visit ends).
    struct ChainOp {
        int kind;                            // 0 D, 1 coarsest first visit, 2 coarsest second visit, 3 U
        uint32_t k;
        bool top, first;                     // the chain's own visit; a first visit (input F of the level above)
    };
    void chain_ops(uint32_t k, bool first, bool top, std::vector<ChainOp>& ops) const {
        if (k + 1 == nlevels) {
            ops.push_back({first ? 1 : 2, k, false, first});
            return;
        }
        ops.push_back({0, k, top, first});
        chain_ops(k + 1, true, false, ops);
        chain_ops(k + 1, false, false, ops);
        ops.push_back({3, k, top, first});
    }
    // The chain's visit: a descent runs the D operations down to the coarsest level's first visit, an ascent the
    // coarsest level's second visit and the U operations that follow it.  A descent starts with the chain's own
    // input or with the Chebyshev residual of a second visit; an ascent ends with the chain's own output or with
    // a first visit's (into u1 of the level above).
    template <class FT, int IN, class XT>
    void run_chain(const InSrc<VT, FT, IN>& f, XT* x, const uint32_t* g) const {
        std::vector<ChainOp> ops;
        chain_ops(uint32_t(chain_a), true, true, ops);
        for (size_t i = 0; i < ops.size();) {
            const ChainOp& o = ops[i];
            if (o.kind == 0) {
                if (o.top) {
                    chain_down(o.k, f, x, g);
                } else {
                    const DevLevel<VT>& P = lv[o.k - 1];
                    // a1 u1 is formed again where the ascent's combine reads it (chain_u1s): not stored
                    if (P.s1_skip) chain_down(o.k, residual_src<IN_RES>(o.k - 1, P.u1, 1.0, nullptr), P.x2, g);
                    else chain_down(o.k, residual_src<IN_RES_SCALE>(o.k - 1, P.u1, P.a1, nullptr), P.x2, g);
                }
                while (i < ops.size() && ops[i].kind == 0) i++;
                if (i >= ops.size() || ops[i].kind != 1) throw Error("chain: unexpected operation order", 1);
                i++;
            } else {
                if (o.kind != 2) throw Error("chain: unexpected operation order", 1);
                size_t e = i + 1;
                while (e < ops.size() && ops[e].kind == 3) e++;
                if (e == i + 1) throw Error("chain: unexpected operation order", 1);
                const ChainOp& u = ops[e - 1];
                if (u.top) chain_up(u.k, x, g);
                else chain_up(u.k, u.first ? lv[u.k - 1].u1 : lv[u.k - 1].x2, g);
                i = e;
            }
        }
    }

    template <class FT, int IN, class XT>
    void cheb(uint32_t l, const InSrc<VT, FT, IN>& f, XT* x, const uint32_t* g) const {
        const DevLevel<VT>& L = lv[l];
        if (chain_a >= 0 && int(l) == chain_a) {
            run_chain(f, x, g);
            return;
        }
        if (int(l) == pair_l) {
            pair_visit(L, lv[l + 1], f, x, g);
            return;
        }
        if (l + 1 == nlevels) {
            solve_last<FT, IN, XT, 0>(L, f, x, g);
            return;
        }
        down<FT, IN, XT, 0>(L, f, x, g);
        const uint32_t n = L.n, gb = cdiv(n, kBS);
        Combine<VT> C;
        if (L.nit == 1) {
            cheb(l + 1, perm(L.F), L.x2, g);
            C.x2 = L.x2; C.a = L.a1; C.skip = L.s1_skip;
            up<XT, OP_SCALE, 0>(L, C, x, g);
            return;
        }
        cheb(l + 1, perm(L.F), L.u1, g);
        // step 2 on tmp = A (a1 u1) - F, with u1s = a1 u1 kept in u0
        if (L.s1_skip) cheb(l + 1, residual_src<IN_RES>(l, L.u1, 1.0, nullptr), L.x2, g);
        else cheb(l + 1, residual_src<IN_RES_SCALE>(l, L.u1, L.a1, L.u0), L.x2, g);
        VT* u1s = L.s1_skip ? L.u1 : L.u0;
        if (L.nit == 2) {
            C.x2 = L.x2; C.u1 = u1s; C.a = L.ma2; C.skip = L.s2_skip; C.b = L.opb2; C.one = L.opb2_one;
            up<XT, OP_CHEB2, 0>(L, C, x, g);
            return;
        }
        // three rotating buffers: prev (previous iterate), cur (latest), spare
        VT* prev = u1s;
        VT* spare = nullptr;
        for (VT* b : {L.u1, L.u0}) if (b != prev) spare = b;
        auto cheb2 = R > 1 ? k_cheb2<VT, XT, true, true> : k_cheb2<VT, XT, true, false>;
        launch(cheb2, gb, kBS, 0, n, (const VT*)L.x2, (const VT*)prev, L.ma2, L.s2_skip, L.opb2,
               L.opb2_one, (const uint32_t*)nullptr, (XT*)nullptr, spare, g, level_batch(L));
        VT* cur = spare;
        spare = L.x2;
        for (uint32_t i = 3; i <= L.nit; i++) {
            cheb(l + 1, residual_src<IN_RES>(l, cur, 1.0, nullptr), spare, g);
            const double a = L.ak[i - 3], b = L.bk[i - 3];
            if (i == L.nit) {
                C.x2 = spare; C.u1 = cur; C.u0 = prev; C.a = a; C.b = b;
                up<XT, OP_CHEBK, 0>(L, C, x, g);
            } else {
                auto step = R > 1 ? k_cheb_step<VT, XT, true, true> : k_cheb_step<VT, XT, true, false>;
                launch(step, gb, kBS, 0, n, (const VT*)cur, (const VT*)prev, (const VT*)spare, a, b,
                       (const uint32_t*)nullptr, (XT*)nullptr, spare, g, level_batch(L));
                VT* t = prev;
                prev = cur;
                cur = spare;
                spare = t;
            }
        }
    }

    // MultiSplitPrec::solve(level, f, x)  (multi_split_solve.cc:59-157)
    template <class FT, class XT>
    void msplit(uint32_t l, const FT* f, XT* x, const uint32_t* g) const {
        const DevLevel<VT>& L = lv[l];
        const bool last = l + 1 == nlevels;
        Combine<VT> C;
        C.x2 = L.x2;
        InSrc<VT, FT, IN_PLAIN> fs;
        fs.f = f;
        if (last) {
            solve_last<FT, IN_PLAIN, XT, 0>(L, fs, x, g);
        } else {
            down<FT, IN_PLAIN, XT, 0>(L, fs, x, g);
            msplit<VT, VT>(l + 1, L.F, L.x2, g);
            up<XT, OP_COPY, 0>(L, C, x, g);
        }
        if (L.nit <= 1 && L.eps == 0.0) return;
        if (l == 0) throw Error("multi_split level 0 runs one iteration", 1);
        // steps 2.. (levels >= 1, so FT == XT == VT here): x += solve(f - A x)
        const VT* fv = reinterpret_cast<const VT*>(f);
        VT* xv = reinterpret_cast<VT*>(x);
        double init_norm = 0.0;
        if (L.eps != 0.0) init_norm = host_nrm2(fv, L.N);
        for (uint32_t it = 2; L.eps != 0.0 || it <= L.nit; it++) {
            launch(k_residual<VT, VT>, cdiv(L.N, kBS), kBS, 0, L.A.v, fv, (const VT*)xv, L.r, g);
            if (L.eps != 0.0) {
                double nr = host_nrm2(L.r, L.N);
                if (nr / init_norm < L.eps) break;
            }
            if (last) {
                solve_last<VT, IN_PLAIN, VT, 1>(L, plain(L.r), xv, g);
            } else {
                down<VT, IN_PLAIN, VT, 1>(L, plain(L.r), xv, g);
                msplit<VT, VT>(l + 1, L.F, L.x2, g);
                up<VT, OP_COPY, 1>(L, C, xv, g);
            }
        }
    }

    // norm of a level vector for the dynamic inner stopping test (host sync)
    template <class T> double host_nrm2(const T* d, uint32_t n) const {
        std::vector<T> h(n);
        MLP_HIP_CHECK(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
        MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        double s = 0.0;
        for (uint32_t i = 0; i < n; i++) s += double(h[i]) * double(h[i]);
        return std::sqrt(s);
    }

    void apply(const double* f, double* x, const uint32_t* guard) override {
        const uint32_t l0 = launches;
        if (kind == MLP_PREC_UH_CHEB) {
            InSrc<VT, double, IN_PLAIN> fs;
            fs.f = f;
            bool done = false;
            if constexpr (std::is_same_v<VT, float>)     // the BF16 variant exists for FP32 hierarchies only
                if (bf16) {
                    cheb_bf16(fs, x, guard);
                    done = true;
                }
            if (!done) cheb(0, fs, x, guard);
        } else {
            msplit<double, double>(0, f, x, guard);
        }
        last_launches = launches - l0;
    }
    bool graph_safe() const override { return !dynamic; }

    // Algorithmic bytes of one apply, summed over level visits:
    // per sweep (w+4)(nnz L + nnz U) + 3w N + 4N; uh_cheb adds (ncheb-1)
    // SpMVs on A_{l+1} plus 6w n of vector traffic; multi_split adds a
    // residual SpMV on A_l per extra inner step.
    double bytes_per_apply() const override {
        const double w = sizeof(VT);
        double tot = 0.0, visits = 1.0;
        for (uint32_t l = 0; l < nlevels; l++) {
            const DevLevel<VT>& L = lv[l];
            const bool cheb_kind = kind == MLP_PREC_UH_CHEB;
            const double steps = cheb_kind ? 1.0 : double(std::max<uint32_t>(1, L.nit));
            // BF16 variant: level 1 reads its input and writes its output in BF16 (2 bytes), and so do
            // level 0's Chebyshev vectors
            const double wio = bf16 && l == 1 ? 2.0 : w;
            const double sweep = (w + 4) * (double(L.nnzL) + double(L.nnzU)) + (w + 2 * wio) * L.N + 4.0 * L.N;
            double b = steps * sweep;
            if (cheb_kind && l + 1 < nlevels) {
                const DevLevel<VT>& Ln = lv[l + 1];
                const double wv = bf16 && l == 0 ? 2.0 : w;
                const double bsp = (w + 4) * Ln.A.nnz + 4.0 * (Ln.N + 1) + 2 * wv * Ln.N;
                b += (double(L.nit) - 1) * (bsp + 6 * wv * Ln.N);
            }
            if (!cheb_kind && steps > 1) {
                const double bsp = (w + 4) * L.A.nnz + 4.0 * (L.N + 1) + 2 * w * L.N;
                b += (steps - 1) * (bsp + w * L.N);
            }
            tot += visits * b;
            visits *= cheb_kind ? double(L.nit) : steps;
        }
        return tot;
    }
    // kernels the last apply launched (0 before the first apply)
    uint32_t schedule() const override {
        uint32_t b = (chain_a >= 0 ? uint32_t(MLP_SCHED_CHAIN) : 0u) | (pair_l >= 0 ? uint32_t(MLP_SCHED_PAIR) : 0u);
        if (chain_a >= 0 && chain_bs != kChainBS) b |= MLP_SCHED_WIDE_CHAIN;
        return b | (sched_lbs << 8);
    }
    uint32_t launches_per_apply() const override { return last_launches; }
};

struct SimplePrec : DevPrecBase {
    mlp_ctx ctx;
    uint32_t n = 0;
    double* d = nullptr;      // nullptr: identity
    DevMem mem;
    void apply(const double* f, double* x, const uint32_t* guard) override {
        hipLaunchKernelGGL(k_diag_apply, dim3(cdiv(n, kBS)), dim3(kBS), 0, ctx->stream, n, (const double*)d, f, x, guard);
    }
    double bytes_per_apply() const override { return (d ? 24.0 : 16.0) * n; }
    uint32_t launches_per_apply() const override { return 1; }
};

// GSPrec (modules/prec/gs/gs_prec.cc:9-65): one forward Gauss-Seidel sweep over
// the lower triangle of A.  The rows are grouped into wavefront levels (a row's
// level is one more than the largest level of its lower columns) and each level
// is one launch; rows keep the reference's entry order.
struct DevGS : DevPrecBase {
    mlp_ctx ctx;
    uint32_t n = 0, nnzl = 0;
    DevMem mem;
    uint32_t *rows = nullptr, *ptr = nullptr, *col = nullptr;
    uint8_t* cnt = nullptr;
    double *val = nullptr, *diag = nullptr;
    uint32_t W = 0;                        // fixed lower-entry slots per row (0: row pointers)
    std::vector<uint32_t> lvl;             // level k: level-order rows [lvl[k], lvl[k+1])
    // brick sweep of a seven-point lower triangle (k_gs_brick): grid sizes, natural-order values, presence
    // bits, brick codes grouped by brick wavefront [blev[w], blev[w+1])
    uint32_t gx = 0, gy = 0, gz = 0;
    double *bvz = nullptr, *bvy = nullptr, *bvx = nullptr;
    uint8_t* bpres = nullptr;
    uint32_t* bricks = nullptr;
    std::vector<uint32_t> blev;

    // The lower entries of every row are at offsets {nx ny, nx, 1} of a grid
    // numbered x fastest (each only where that neighbour exists): set up the
    // brick sweep and return true.
    template <class VI, class VD>
    bool plan_bricks(const VI& ia, const VI& ja, const VD& a, const std::vector<uint32_t>& lo_end) {
        uint32_t off[3] = {0, 0, 0};
        int no = 0;
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t e = ia[i] + 1; e < lo_end[i]; e++) {
                const uint32_t d = i - ja[e];
                if (d == off[0] || d == off[1] || d == off[2]) continue;
                if (no == 3) return false;
                off[no++] = d;
            }
        if (no != 3) return false;
        std::sort(off, off + 3);
        if (off[0] != 1 || off[1] < 2 || off[2] % off[1] || n % off[2]) return false;
        const uint32_t nx = off[1], ny = off[2] / off[1], nz = n / off[2];
        const uint32_t BX = (nx + kBrick - 1) / kBrick, BY = (ny + kBrick - 1) / kBrick, BZ = (nz + kBrick - 1) / kBrick;
        if (ny < 2 || BX > 1024 || BY > 1024 || BZ > 4096) return false;
        const uint32_t nxy = nx * ny;
        std::vector<double> hz(n, 0.0), hy(n, 0.0), hx(n, 0.0), hd(n);
        std::vector<uint8_t> hp(n, 0);
        for (uint32_t i = 0; i < n; i++) {
            hd[i] = a[ia[i]];
            for (uint32_t e = ia[i] + 1; e < lo_end[i]; e++) {
                const uint32_t d = i - ja[e];
                if (d == nxy && i >= nxy) { hz[i] = a[e]; hp[i] |= 1; }
                else if (d == nx && (i / nx) % ny != 0) { hy[i] = a[e]; hp[i] |= 2; }
                else if (d == 1 && i % nx != 0) { hx[i] = a[e]; hp[i] |= 4; }
                else return false;                       // an offset that wraps around the grid
            }
        }
        blev.assign(BX + BY + BZ - 1, 0);
        for (uint32_t z = 0; z < BZ; z++)
            for (uint32_t y = 0; y < BY; y++)
                for (uint32_t x = 0; x < BX; x++) blev[x + y + z + 1]++;
        for (size_t w = 0; w + 1 < blev.size(); w++) blev[w + 1] += blev[w];
        std::vector<uint32_t> codes(size_t(BX) * BY * BZ), fill(blev.begin(), blev.end() - 1);
        for (uint32_t z = 0; z < BZ; z++)
            for (uint32_t y = 0; y < BY; y++)
                for (uint32_t x = 0; x < BX; x++) codes[fill[x + y + z]++] = x | (y << 10) | (z << 20);
        gx = nx; gy = ny; gz = nz;
        hipStream_t s = ctx->stream;
        diag = mem.upload(hd, s);
        bvz = mem.upload(hz, s);
        bvy = mem.upload(hy, s);
        bvx = mem.upload(hx, s);
        bpres = mem.upload(hp, s);
        bricks = mem.upload(codes, s);
        return true;
    }

    template <class VI, class VD> void build(const VI& ia, const VI& ja, const VD& a) {
        n = uint32_t(ia.size()) - 1;
        std::vector<uint32_t> lev(n, 0), lo_end(n);
        uint32_t nlev = 0;
        for (uint32_t i = 0; i < n; i++) {
            if (!(a[ia[i]] > 0))
                throw Error("Something is very strange: (" + std::to_string(i) + ") has " + std::to_string(a[ia[i]]) +
                                " on diagonal, aborting...",
                            MLP_ERR_INVALID);
            uint32_t e = ia[i] + 1;
            while (e < ia[i + 1] && ja[e] < i) {       // the Skyline row is sorted after its diagonal
                lev[i] = std::max(lev[i], lev[ja[e]] + 1);
                e++;
            }
            lo_end[i] = e;
            nlev = std::max(nlev, lev[i] + 1);
        }
        for (uint32_t i = 0; i < n; i++) nnzl += lo_end[i] - ia[i] - 1;
        if (plan_bricks(ia, ja, a, lo_end)) return;
        nnzl = 0;
        lvl.assign(nlev + 1, 0);
        for (uint32_t i = 0; i < n; i++) lvl[lev[i] + 1]++;
        for (uint32_t k = 0; k < nlev; k++) lvl[k + 1] += lvl[k];
        std::vector<uint32_t> order(n), fill(lvl.begin(), lvl.end() - 1);
        for (uint32_t i = 0; i < n; i++) order[fill[lev[i]]++] = i;
        std::vector<uint32_t> hptr(n + 1, 0), hcol;
        std::vector<double> hval, hdiag(n);
        for (uint32_t k = 0; k < n; k++) {
            const uint32_t i = order[k];
            hdiag[k] = a[ia[i]];
            for (uint32_t e = ia[i] + 1; e < lo_end[i]; e++) {
                hcol.push_back(ja[e]);
                hval.push_back(a[e]);
            }
            hptr[k + 1] = uint32_t(hcol.size());
        }
        nnzl = uint32_t(hcol.size());
        uint32_t wmax = 0;
        for (uint32_t k = 0; k < n; k++) wmax = std::max(wmax, hptr[k + 1] - hptr[k]);
        hipStream_t s = ctx->stream;
        if (wmax <= 8) {
            // fixed slots: the common case (a seven-point matrix has at most three lower entries)
            W = wmax == 0 ? 1 : wmax <= 4 ? wmax : 8;
            std::vector<uint8_t> hcnt(n);
            std::vector<uint32_t> ecol(size_t(W) * n, 0);
            std::vector<double> eval(size_t(W) * n, 0.0);
            for (uint32_t k = 0; k < n; k++) {
                hcnt[k] = uint8_t(hptr[k + 1] - hptr[k]);
                for (uint32_t q = hptr[k]; q < hptr[k + 1]; q++) {
                    ecol[size_t(q - hptr[k]) * n + k] = hcol[q];
                    eval[size_t(q - hptr[k]) * n + k] = hval[q];
                }
            }
            cnt = mem.upload(hcnt, s);
            col = mem.upload(ecol, s);
            val = mem.upload(eval, s);
        } else {
            W = 0;
            if (hcol.empty()) { hcol.push_back(0); hval.push_back(0.0); }
            ptr = mem.upload(hptr, s);
            col = mem.upload(hcol, s);
            val = mem.upload(hval, s);
        }
        rows = mem.upload(order, s);
        diag = mem.upload(hdiag, s);
        MLP_HIP_CHECK(hipStreamSynchronize(s));
    }
    void apply(const double* f, double* x, const uint32_t* guard) override {
        if (bricks) {
            for (size_t w = 0; w + 1 < blev.size(); w++)
                hipLaunchKernelGGL(k_gs_brick, dim3(blev[w + 1] - blev[w]), dim3(kBrick * kBrick * kBrick), 0,
                                   ctx->stream, gx, gy, gz, (const uint32_t*)bricks, blev[w], (const double*)diag,
                                   (const double*)bvz, (const double*)bvy, (const double*)bvx,
                                   (const uint8_t*)bpres, f, x, guard);
            return;
        }
        auto ell = W == 1 ? k_gs_level_ell<1> : W == 2 ? k_gs_level_ell<2> : W == 3 ? k_gs_level_ell<3>
                 : W == 4 ? k_gs_level_ell<4> : k_gs_level_ell<8>;
        for (size_t k = 0; k + 1 < lvl.size(); k++) {
            const uint32_t r0 = lvl[k], r1 = lvl[k + 1];
            if (W)
                hipLaunchKernelGGL(ell, dim3(cdiv(r1 - r0, kBS)), dim3(kBS), 0, ctx->stream, r0, r1, n,
                                   (const uint32_t*)rows, (const uint8_t*)cnt, (const uint32_t*)col,
                                   (const double*)val, (const double*)diag, f, x, guard);
            else
                hipLaunchKernelGGL(k_gs_level, dim3(cdiv(r1 - r0, kBS)), dim3(kBS), 0, ctx->stream, r0, r1,
                                   (const uint32_t*)rows, (const uint32_t*)ptr, (const uint32_t*)col,
                                   (const double*)val, (const double*)diag, f, x, guard);
        }
    }
    double bytes_per_apply() const override { return 12.0 * nnzl + 28.0 * n; }
    uint32_t launches_per_apply() const override {
        return bricks ? uint32_t(blev.size()) - 1 : uint32_t(lvl.size()) - 1;
    }
};

// CompositePrec (modules/prec/composite/composite_prec.cc:9-36): x = B2(B1(f)).
struct DevComposite : DevPrecBase {
    std::unique_ptr<DevPrecBase> first, second;
    DevMem mem;
    double* tmp = nullptr;
    void apply(const double* f, double* x, const uint32_t* guard) override {
        first->apply(f, tmp, guard);
        second->apply(tmp, x, guard);
    }
    double bytes_per_apply() const override { return first->bytes_per_apply() + second->bytes_per_apply(); }
    uint32_t launches_per_apply() const override {
        return first->launches_per_apply() + second->launches_per_apply();
    }
    bool graph_safe() const override { return first->graph_safe() && second->graph_safe(); }
};

}  // namespace
}  // namespace mlp

struct mlp_prec_s {
    mlp_ctx ctx = nullptr;
    uint64_t id = 0;
    int kind = 0, precision = 0;
    uint32_t n = 0;
    std::unique_ptr<mlp::DevPrecBase> impl;
    std::unique_ptr<mlp::DevPrecBase> wide;   // impl widened to wide_R right-hand sides (mlp_pcg_multi)
    uint32_t wide_R = 0;
};

namespace mlp {
namespace {
// scratch of a context: reduction partials, counters, scalars, solve state, pinned staging
void context_setup(mlp_ctx c) {
    c->partials = c->mem.alloc<double>(kPartials);
    c->counters = c->mem.alloc<unsigned>(8);
    c->scal = c->mem.alloc<double>(4);
    c->st = c->mem.alloc<SolveState>(1);
    MLP_HIP_CHECK(hipMemsetAsync(c->counters, 0, 8 * sizeof(unsigned), c->stream));
    MLP_HIP_CHECK(hipHostMalloc((void**)&c->host_pinned, 64 * sizeof(double), hipHostMallocDefault));
    MLP_HIP_CHECK(hipStreamSynchronize(c->stream));
}

void context_release(mlp_ctx ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->drop_graphs(0);
    ctx->ws_mem.release();
    ctx->multi_mem.release();
    ctx->ens_mem.release();
    ctx->mem.release();
    if (ctx->est_host) (void)hipHostFree(ctx->est_host);
    ctx->est_host = nullptr;
    if (ctx->eout_host) (void)hipHostFree(ctx->eout_host);
    ctx->eout_host = nullptr;
    if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
    ctx->host_pinned = nullptr;
    if (ctx->mst_host) (void)hipHostFree(ctx->mst_host);
    ctx->mst_host = nullptr;
    if (ctx->owns_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    ctx->stream = nullptr;
}

// A context with its own workspace and device state that runs on `parent`'s
// stream (used by preconditioners that run a solver inside their apply).
std::unique_ptr<mlp_ctx_s, void (*)(mlp_ctx)> make_sub_context(mlp_ctx parent) {
    std::unique_ptr<mlp_ctx_s, void (*)(mlp_ctx)> c(new mlp_ctx_s(), [](mlp_ctx p) {
        context_release(p);
        delete p;
    });
    c->device = parent->device;
    c->stream = parent->stream;
    c->owns_stream = false;
    context_setup(c.get());
    return c;
}
}  // namespace
}  // namespace mlp

namespace mlp {
namespace {
// SymPrec (modules/prec/sym/sym_prec.cc:8-38): Asym = A with each off-diagonal
// a_ij raised to a_ji when that is larger (max(A, A^T) off the diagonal), the
// uh_cheb hierarchy built on Asym, and an apply that is a PCG solve of Asym to
// 1e-14 from generate_x0.  The inner solve runs on a sub-context that shares
// the caller's stream; it synchronises with the host, so the apply is not
// captured in graphs.  The inner PCG always solves from the instance's own
// right-hand side and solution buffers (f is copied in, the solution out), so
// its iteration graph is captured once and replayed by every apply, whatever
// vectors the caller passes (GMRES passes a different Krylov vector each step).
struct DevSym : DevPrecBase {
    mlp_ctx parent;
    std::unique_ptr<mlp_ctx_s, void (*)(mlp_ctx)> sub;
    mlp_mat Asym = nullptr;
    mlp_prec B = nullptr;
    DevMem mem;
    double* x0 = nullptr;                  // generate_x0's vector, made once
    double* fin = nullptr;                 // the inner solve's right-hand side and solution
    double* xin = nullptr;
    uint32_t n = 0;
    double eps = 1e-14;                    // sym_prec.cc:37, SymPrec::set_eps (sym_prec.h:26-28)
    explicit DevSym(mlp_ctx c) : parent(c), sub(make_sub_context(c)) {}
    ~DevSym() override {
        if (B) mlp_prec_destroy(B);
        if (Asym) mlp_mat_destroy(Asym);
    }
    void build(mlp_mat A, const double* sigmas, uint32_t nsig, const uint32_t* niters, uint32_t nnit, int use_tails) {
        n = A->n;
        ensure_host(A);
        std::vector<double> a(A->a.begin(), A->a.end());
        const HostVec<uint32_t>&ia = A->ia, &ja = A->ja;
        auto at = [&](uint32_t r, uint32_t c) -> double {      // Asym(r, c) with the updates made so far
            if (ja[ia[r]] == c) return a[ia[r]];
            auto b = ja.begin() + ia[r] + 1, e = ja.begin() + ia[r + 1];
            auto it = std::lower_bound(b, e, c);
            return (it != e && *it == c) ? a[size_t(it - ja.begin())] : 0.0;
        };
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t j = ia[i] + 1; j < ia[i + 1]; j++) {
                const double d = at(ja[j], i);
                if (d > a[j]) a[j] = d;
            }
        int rc = mlp_csr_upload(sub.get(), n, ia.data(), ja.data(), a.data(), &Asym);
        if (rc) throw Error(mlp_last_error(), rc);
        mlp_hier h = nullptr;
        rc = mlp_hier_build_uh_cheb(n, ia.data(), ja.data(), a.data(), sigmas, nsig, niters, nnit, use_tails, &h);
        if (rc) throw Error(mlp_last_error(), rc);
        std::unique_ptr<mlp_hier_s> hold(h);
        rc = mlp_prec_from_hier(sub.get(), h, Asym, MLP_FP64, &B);
        if (rc) throw Error(mlp_last_error(), rc);
        std::vector<double> hx(n);
        srandom(3);                                            // generate_x0, modules/solvers/solvers.cc:9-21
        for (uint32_t i = 0; i < n; i++) hx[i] = 20. * (random() - 0.5 * RAND_MAX) / RAND_MAX + 100;
        x0 = mem.upload(hx, parent->stream);
        fin = mem.alloc<double>(n);
        xin = mem.alloc<double>(n);
        MLP_HIP_CHECK(hipStreamSynchronize(parent->stream));
    }
    void apply(const double* f, double* x, const uint32_t* guard) override {
        if (guard) {                                           // the enclosing solve has already converged
            MLP_HIP_CHECK(hipMemcpyAsync(sub->host_pinned, guard, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                         parent->stream));
            MLP_HIP_CHECK(hipStreamSynchronize(parent->stream));
            uint32_t g = 0;
            std::memcpy(&g, sub->host_pinned, sizeof(g));
            if (g) return;
        }
        const size_t bytes = size_t(n) * sizeof(double);
        MLP_HIP_CHECK(hipMemcpyAsync(fin, f, bytes, hipMemcpyDeviceToDevice, parent->stream));
        MLP_HIP_CHECK(hipMemcpyAsync(xin, x0, bytes, hipMemcpyDeviceToDevice, parent->stream));
        mlp_solve_opts o;
        std::memset(&o, 0, sizeof(o));
        o.eps = eps;
        o.maxit = 100000;
        o.x0_mode = MLP_X0_GIVEN;
        o.check_every = 8;
        o.use_graph = 1;
        mlp_solve_stats st;
        const int rc = mlp_pcg(sub.get(), Asym, B, fin, xin, &o, &st);
        if (rc) throw Error(mlp_last_error(), rc);
        MLP_HIP_CHECK(hipMemcpyAsync(x, xin, bytes, hipMemcpyDeviceToDevice, parent->stream));
    }
    double bytes_per_apply() const override { return 0.0; }
    uint32_t launches_per_apply() const override { return 0; }
    bool graph_safe() const override { return false; }
};
}  // namespace
}  // namespace mlp

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace mlp;

#define MLP_API_BEGIN try {
#define MLP_API_END                                   \
    return MLP_OK;                                    \
    }                                                 \
    catch (const mlp::Error& e) {                     \
        mlp::set_last_error(e.what());                \
        return e.code();                              \
    }                                                 \
    catch (const std::exception& e) {                 \
        mlp::set_last_error(e.what());                \
        return MLP_ERR_INVALID;                       \
    }

// Calls on a context run on its device, whichever device the calling thread
// had current (allocations and copies follow the current device).
static void use_device(mlp_ctx c) {
    int d = -1;
    if (c && (hipGetDevice(&d) != hipSuccess || d != c->device)) MLP_HIP_CHECK(hipSetDevice(c->device));
}

static void require(bool c, const char* msg) {
    if (!c) throw Error(msg, MLP_ERR_INVALID);
}

// y = A x (b null) or y = b - A x on a top-level FP64 matrix: two rows per
// thread on the seven-point layout, one on slices.
static void launch_rows(const SellView<double>& A, const double* b, const double* x, double* y, hipStream_t s) {
    if (!A.n) return;
    if (A.sd) {
        const dim3 g(cdiv(A.n, 2 * kBS)), t(kBS);
        const bool vec = pair_vec(A, uintptr_t(x) | uintptr_t(y) | uintptr_t(b));
        if (b && vec) hipLaunchKernelGGL((k_spmv_pairs<true, true>), g, t, 0, s, A, b, x, y);
        else if (b) hipLaunchKernelGGL((k_spmv_pairs<true, false>), g, t, 0, s, A, b, x, y);
        else if (vec) hipLaunchKernelGGL((k_spmv_pairs<false, true>), g, t, 0, s, A, b, x, y);
        else hipLaunchKernelGGL((k_spmv_pairs<false, false>), g, t, 0, s, A, b, x, y);
        return;
    }
    const uint32_t* g = nullptr;
    if (b) hipLaunchKernelGGL((k_residual<double, double>), dim3(cdiv(A.n, kBS)), dim3(kBS), 0, s, A, b, x, y, g);
    else hipLaunchKernelGGL((k_spmv<double, double, double>), dim3(cdiv(A.n, kBS)), dim3(kBS), 0, s, A, x, y, g);
}

extern "C" {

const char* mlp_last_error(void) { return mlp::g_last_error.c_str(); }

int mlp_ctx_create(int device, mlp_ctx* out) {
    MLP_API_BEGIN
    require(out, "null output");
    int ndev = 0;
    MLP_HIP_CHECK(hipGetDeviceCount(&ndev));
    require(device >= 0 && device < ndev, "no such HIP device");
    MLP_HIP_CHECK(hipSetDevice(device));
    std::unique_ptr<mlp_ctx_s, void (*)(mlp_ctx)> c(new mlp_ctx_s(), [](mlp_ctx p) {
        context_release(p);
        delete p;
    });
    c->device = device;
    MLP_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    context_setup(c.get());
    *out = c.release();
    MLP_API_END
}

int mlp_ctx_destroy(mlp_ctx ctx) {
    MLP_API_BEGIN
    if (!ctx) return MLP_OK;
    context_release(ctx);
    delete ctx;
    MLP_API_END
}

int mlp_ctx_sync(mlp_ctx ctx) {
    MLP_API_BEGIN
    require(ctx, "null context");
    use_device(ctx);
    MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    MLP_API_END
}

int mlp_ctx_stream(mlp_ctx ctx, void** s) {
    MLP_API_BEGIN
    require(ctx && s, "null argument");
    use_device(ctx);
    *s = (void*)ctx->stream;
    MLP_API_END
}

int mlp_dev_alloc(mlp_ctx ctx, size_t bytes, void** d_out) {
    MLP_API_BEGIN
    require(ctx && d_out, "null argument");
    use_device(ctx);
    size_t cls = 0;
    void* p = cache_alloc(std::max<size_t>(bytes, 1), &cls);
    {
        MemCache& c = mem_cache();
        std::lock_guard<std::mutex> hold(c.m);
        c.live[p] = {cls, ctx->device};
    }
    *d_out = p;
    MLP_API_END
}

int mlp_dev_free(mlp_ctx ctx, void* p) {
    MLP_API_BEGIN
    require(ctx, "null context");
    use_device(ctx);
    if (p) {
        MemCache& c = mem_cache();
        std::vector<CachedBlock> one;
        {
            std::lock_guard<std::mutex> hold(c.m);
            auto it = c.live.find(p);
            require(it != c.live.end(), "not a block of mlp_dev_alloc");
            one.push_back({p, it->second.first, it->second.second});
            c.live.erase(it);
        }
        cache_release(one);
    }
    MLP_API_END
}

}  // extern "C"

namespace mlp {
namespace {
// A copy between pageable host memory and the device: small ones directly, large ones in chunks through the
// device's pinned staging buffer, the host side of each chunk spread over the host threads.
void staged_copy(mlp_ctx ctx, void* dst, const void* src, size_t bytes, bool to_device) {
    constexpr size_t kDirect = size_t(1) << 20, kChunk = size_t(32) << 20;
    hipStream_t s = ctx->stream;
    if (bytes < kDirect) {
        MLP_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost, s));
        MLP_HIP_CHECK(hipStreamSynchronize(s));
        return;
    }
    PinnedStage& st = pinned_stage(ctx->device);
    std::lock_guard<std::mutex> hold(st.m);
    const size_t want = std::min(bytes, kChunk);
    if (st.n < want) {
        if (st.p) MLP_HIP_CHECK(hipHostFree(st.p));
        st.p = nullptr;
        st.n = 0;
        MLP_HIP_CHECK(hipHostMalloc((void**)&st.p, want, hipHostMallocDefault));
        st.n = want;
    }
    char* pin = st.p;
    for (size_t off = 0; off < bytes; off += kChunk) {
        const size_t len = std::min(kChunk, bytes - off);
        if (to_device) {
            const char* from = static_cast<const char*>(src) + off;
            parallel_for(len, size_t(1) << 20, [&](uint64_t a, uint64_t b) { std::memcpy(pin + a, from + a, b - a); });
            MLP_HIP_CHECK(hipMemcpyAsync(static_cast<char*>(dst) + off, pin, len, hipMemcpyHostToDevice, s));
            MLP_HIP_CHECK(hipStreamSynchronize(s));
        } else {
            MLP_HIP_CHECK(hipMemcpyAsync(pin, static_cast<const char*>(src) + off, len, hipMemcpyDeviceToHost, s));
            MLP_HIP_CHECK(hipStreamSynchronize(s));
            char* to = static_cast<char*>(dst) + off;
            parallel_for(len, size_t(1) << 20, [&](uint64_t a, uint64_t b) { std::memcpy(to + a, pin + a, b - a); });
        }
    }
}
}  // namespace
}  // namespace mlp

extern "C" {

int mlp_copy_h2d(mlp_ctx ctx, void* dst, const void* src, size_t bytes) {
    MLP_API_BEGIN
    require(ctx, "null context");
    use_device(ctx);
    if (bytes) staged_copy(ctx, dst, src, bytes, true);
    MLP_API_END
}

int mlp_copy_d2h(mlp_ctx ctx, void* dst, const void* src, size_t bytes) {
    MLP_API_BEGIN
    require(ctx, "null context");
    use_device(ctx);
    if (bytes) staged_copy(ctx, dst, src, bytes, false);
    MLP_API_END
}

// ---- matrices ----
static void check_skyline(uint32_t n, const uint32_t* ia, const uint32_t* ja) {
    require(ia && (n == 0 || ja), "null CSR arrays");
    require(ia[0] == 0, "ia[0] must be 0");
    for (uint32_t i = 0; i < n; i++) {
        require(ia[i + 1] > ia[i], "every row needs at least its diagonal");
        require(ja[ia[i]] == i, "Skyline rows must start with the diagonal");
        for (uint32_t p = ia[i]; p < ia[i + 1]; p++) require(ja[p] < n, "column index out of range");
    }
}

int mlp_csr_upload(mlp_ctx ctx, uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a, mlp_mat* out) {
    MLP_API_BEGIN
    require(ctx && out, "null argument");
    use_device(ctx);
    check_skyline(n, ia, ja);
    require(a || n == 0, "null values");
    auto* m = new mlp_mat_s();
    try {
        m->ctx = ctx;
        m->id = mlp::new_handle_id();
        m->n = n;
        m->nnz = ia[n];
        par_copy(m->ia, ia, size_t(n) + 1);
        par_copy(m->ja, ja, m->nnz);
        par_copy(m->a, a, m->nnz);
        upload_sell<double>(ctx, m->mem, n, ia, ja, a, m->A);
        m->bytes = m->A.stored_bytes();
    } catch (...) {
        delete m;
        throw;
    }
    *out = m;
    MLP_API_END
}

int mlp_mat_destroy(mlp_mat m) {
    MLP_API_BEGIN
    use_device(m ? m->ctx : nullptr);
    if (m) {
        (void)hipStreamSynchronize(m->ctx->stream);
        m->ctx->drop_graphs(m->id);
        delete m;
    }
    MLP_API_END
}

int mlp_mat_info(mlp_mat m, uint32_t* n, uint32_t* nnz, uint64_t* bytes) {
    MLP_API_BEGIN
    require(m, "null matrix");
    if (n) *n = m->n;
    if (nnz) *nnz = m->nnz;
    if (bytes) *bytes = m->bytes;
    MLP_API_END
}

// The device assembly behind mlp_spe_matrix_dev and mlp_spe_ensemble_matrix_dev: nsys systems of nx*ny*nzs[s]
// cells stacked along z (one system: nsys = 1).
static int spe_matrix_dev(mlp_ctx ctx, int kind, uint32_t nx, uint32_t ny, uint32_t nsys, const uint32_t* nzs,
                          const double* kx, double kz_ratio, double c, double hx, double hy, double hz, mlp_mat* out) {
    MLP_API_BEGIN
    require(ctx && kx && out && nzs, "null argument");
    require(kind == MLP_ASM_SPE || kind == MLP_ASM_TPFA, "kind must be MLP_ASM_SPE or MLP_ASM_TPFA");
    require(nx && ny && nsys, "grid dimensions must be positive");
    uint64_t nz64 = 0, nnz64 = 0;
    for (uint32_t k = 0; k < nsys; k++) {
        require(nzs[k] > 0, "grid dimensions must be positive");
        nz64 += nzs[k];
        nnz64 += spe_nnz(nx, ny, nzs[k]);
    }
    require(nz64 < (uint64_t(1) << 32), "grid too large for 32-bit indices");
    const uint32_t nz = uint32_t(nz64);
    const uint64_t n64 = uint64_t(nx) * ny * nz;
    require(n64 < (uint64_t(1) << 32) && nnz64 < (uint64_t(1) << 32), "grid too large for 32-bit indices");
    use_device(ctx);
    auto m = std::make_unique<mlp_mat_s>();
    m->ctx = ctx;
    m->id = mlp::new_handle_id();
    m->n = uint32_t(n64);
    m->nnz = uint32_t(nnz64);
    m->d_ia = m->mem.alloc<uint32_t>(size_t(m->n) + 1);
    m->d_ja = m->mem.alloc<uint32_t>(m->nnz);
    m->d_a = m->mem.alloc<double>(m->nnz);
    // the stacked systems' plane ranges
    uint32_t* zfirst = nullptr;
    uint32_t* zlast = nullptr;
    DevMem ztab;
    if (nsys > 1) {
        std::vector<uint32_t> zf(nz), zl(nz);
        for (uint32_t k = 0, z = 0; k < nsys; z += nzs[k], k++)
            for (uint32_t q = z; q < z + nzs[k]; q++) {
                zf[q] = z;
                zl[q] = z + nzs[k] - 1;
            }
        zfirst = ztab.upload(zf, ctx->stream);
        zlast = ztab.upload(zl, ctx->stream);
    }
    // a full grid (every dimension >= 2), or a stack of them, is stored in the seven-point layout (presence bits
    // per row), written by the assembly itself
    const char* off = getenv("MLP_NO_STENCIL");
    bool stencil = nx >= 2 && ny >= 2 && n64 >= 8 && !(off && off[0] == '1');
    for (uint32_t k = 0; k < nsys; k++) stencil = stencil && nzs[k] >= 2;
    if (stencil) {
        SpeStencilOut st;
        st.pres = m->mem.alloc<uint8_t>(m->n);
        st.d = m->mem.alloc<double>(m->n);
        st.lx = m->mem.alloc<double>(m->n);
        st.ly = m->mem.alloc<double>(m->n);
        st.lz = m->mem.alloc<double>(m->n);
        spe_assemble_device(ctx->stream, kind == MLP_ASM_SPE, nx, ny, nz, kx, kz_ratio, c, hx, hy, hz, m->d_ia,
                            m->d_ja, m->d_a, &st, zfirst, zlast);
        m->A.v.n = m->n;
        m->A.v.spres = st.pres;
        m->A.v.sd = st.d;
        m->A.v.slx = st.lx;
        m->A.v.sly = st.ly;
        m->A.v.slz = st.lz;
        m->A.v.ssy = nx;
        m->A.v.ssz = nx * ny;
        m->A.stencil = true;
        m->A.nent = 0;
        m->A.nnz = m->nnz;
        m->host_ready = false;
    } else {
        spe_assemble_device(ctx->stream, kind == MLP_ASM_SPE, nx, ny, nz, kx, kz_ratio, c, hx, hy, hz, m->d_ia,
                            m->d_ja, m->d_a, nullptr, zfirst, zlast);
        m->host_ready = false;
        ensure_host(m.get());
        upload_sell<double>(ctx, m->mem, m->n, m->ia.data(), m->ja.data(), m->a.data(), m->A);
    }
    m->bytes = m->A.stored_bytes();
    *out = m.release();
    MLP_API_END
}

int mlp_spe_matrix_dev(mlp_ctx ctx, int kind, uint32_t nx, uint32_t ny, uint32_t nz, const double* kx,
                       double kz_ratio, double c, double hx, double hy, double hz, mlp_mat* out) {
    return spe_matrix_dev(ctx, kind, nx, ny, 1, &nz, kx, kz_ratio, c, hx, hy, hz, out);
}

int mlp_spe_ensemble_matrix_dev(mlp_ctx ctx, int kind, uint32_t nx, uint32_t ny, uint32_t nsys, const uint32_t* nz,
                                const double* kx, double kz_ratio, double c, double hx, double hy, double hz,
                                mlp_mat* out) {
    return spe_matrix_dev(ctx, kind, nx, ny, nsys, nz, kx, kz_ratio, c, hx, hy, hz, out);
}

int mlp_mat_get_csr(mlp_mat A, uint32_t* ia, uint32_t* ja, double* a) {
    MLP_API_BEGIN
    require(A && ia && (A->n == 0 || (ja && a)), "null argument");
    use_device(A->ctx);
    ensure_host(A);
    std::memcpy(ia, A->ia.data(), (size_t(A->n) + 1) * sizeof(uint32_t));
    std::memcpy(ja, A->ja.data(), size_t(A->nnz) * sizeof(uint32_t));
    std::memcpy(a, A->a.data(), size_t(A->nnz) * sizeof(double));
    MLP_API_END
}

int mlp_spmv(mlp_ctx ctx, mlp_mat A, const double* x, double* y) {
    MLP_API_BEGIN
    require(ctx && A, "null argument");
    use_device(ctx);
    launch_rows(A->A.v, nullptr, x, y, ctx->stream);
    MLP_HIP_CHECK(hipGetLastError());
    MLP_API_END
}

int mlp_residual(mlp_ctx ctx, mlp_mat A, const double* b, const double* x, double* r) {
    MLP_API_BEGIN
    require(ctx && A, "null argument");
    use_device(ctx);
    launch_rows(A->A.v, b, x, r, ctx->stream);
    MLP_HIP_CHECK(hipGetLastError());
    MLP_API_END
}

// ---- BLAS-1 ----
// Streaming reductions run a grid-stride loop over at most kRedGrid blocks
// (8 per CU), which keeps the one-block finalize short.
constexpr uint32_t kMaxRhs = 64;          // right-hand sides of one mlp_pcg_multi call
constexpr uint32_t kMaxSegments = 256;    // systems of one mlp_pcg_ensemble call
static_assert(kRedGrid <= kRedBlocks, "partials buffer too small");
static_assert(2 * kRedGrid <= kPartials / 2, "each half of the partials buffer must hold two values per block");
static uint32_t red_grid(uint32_t n) { return std::max<uint32_t>(1, std::min<uint32_t>(cdiv(n, kBS), kRedGrid)); }
static uint32_t mdot_grid(uint32_t n) { return std::max<uint32_t>(1, std::min<uint32_t>(cdiv(n, kBS), kMdotBlocks)); }

static double dev_dot(mlp_ctx ctx, uint32_t n, const double* x, const double* y) {
    const uint32_t g = red_grid(n);
    hipLaunchKernelGGL(k_dot, dim3(g), dim3(kBS), 0, ctx->stream, n, x, y, ctx->partials);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBS), 0, ctx->stream, int(FIN_DOT), g, (const double*)ctx->partials,
                       ctx->st, ctx->scal, Batch{});
    MLP_HIP_CHECK(hipMemcpyAsync(ctx->host_pinned, ctx->scal, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return ctx->host_pinned[0];
}

int mlp_ddot(mlp_ctx ctx, uint32_t n, const double* x, const double* y, double* out) {
    MLP_API_BEGIN
    require(ctx && out, "null argument");
    use_device(ctx);
    *out = dev_dot(ctx, n, x, y);
    MLP_API_END
}

int mlp_dnrm2(mlp_ctx ctx, uint32_t n, const double* x, double* out) {
    MLP_API_BEGIN
    require(ctx && out, "null argument");
    use_device(ctx);
    *out = std::sqrt(dev_dot(ctx, n, x, x));
    MLP_API_END
}

int mlp_daxpy(mlp_ctx ctx, uint32_t n, double alpha, const double* x, double* y) {
    MLP_API_BEGIN
    require(ctx, "null context");
    use_device(ctx);
    if (n)
        hipLaunchKernelGGL(k_axpy, dim3(red_grid(n)), dim3(kBS), 0, ctx->stream, n, alpha, (int)host_is_one(alpha), x, y);
    MLP_HIP_CHECK(hipGetLastError());
    MLP_API_END
}

int mlp_dscal(mlp_ctx ctx, uint32_t n, double alpha, double* x) {
    MLP_API_BEGIN
    require(ctx, "null context");
    use_device(ctx);
    if (n && !host_is_one(alpha))
        hipLaunchKernelGGL(k_scal, dim3(red_grid(n)), dim3(kBS), 0, ctx->stream, n, alpha, x);
    MLP_HIP_CHECK(hipGetLastError());
    MLP_API_END
}

// ---- host hierarchies ----
static int build_uh_cheb_entry(mlp_ctx ctx, uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                               const double* sigmas, uint32_t nsigmas, const uint32_t* niters, uint32_t nniters,
                               int use_tails, mlp_hier* out) {
    MLP_API_BEGIN
    require(out, "null output");
    check_skyline(n, ia, ja);
    require(sigmas && nsigmas && niters && nniters, "sigmas and niters are required");
    auto h = std::make_unique<mlp_hier_s>();
    h->A0.nrow = h->A0.ncol = n;
    h->A0.ia.assign(ia, ia + n + 1);
    h->A0.ja.assign(ja, ja + ia[n]);
    h->A0.a.assign(a, a + ia[n]);
    ChebParams p;
    p.sigmas.assign(sigmas, sigmas + nsigmas);
    p.niters.assign(niters, niters + nniters);
    p.use_tails = use_tails != 0;
    if (ctx) {
        use_device(ctx);
        auto dev = device_cheb_builder(ctx->stream);
        build_uh_cheb(h->A0, p, h->H, dev.get());
    } else {
        build_uh_cheb(h->A0, p, h->H);
    }
    *out = h.release();
    MLP_API_END
}

int mlp_hier_build_uh_cheb(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a, const double* sigmas,
                           uint32_t nsigmas, const uint32_t* niters, uint32_t nniters, int use_tails, mlp_hier* out) {
    return build_uh_cheb_entry(nullptr, n, ia, ja, a, sigmas, nsigmas, niters, nniters, use_tails, out);
}

int mlp_hier_build_uh_cheb_dev(mlp_ctx ctx, uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                               const double* sigmas, uint32_t nsigmas, const uint32_t* niters, uint32_t nniters,
                               int use_tails, mlp_hier* out) {
    if (!ctx) {
        mlp::set_last_error("null context");
        return MLP_ERR_INVALID;
    }
    return build_uh_cheb_entry(ctx, n, ia, ja, a, sigmas, nsigmas, niters, nniters, use_tails, out);
}

int mlp_hier_build_uh_cheb_mat(mlp_ctx ctx, mlp_mat A, const double* sigmas, uint32_t nsigmas, const uint32_t* niters,
                               uint32_t nniters, int use_tails, mlp_hier* out) {
    MLP_API_BEGIN
    require(ctx && A && out, "null argument");
    require(sigmas && nsigmas && niters && nniters, "sigmas and niters are required");
    use_device(ctx);
    if (!A->d_ia) {                                      // an uploaded matrix: its host copy goes up again
        ensure_host(A);
        return build_uh_cheb_entry(ctx, A->n, A->ia.data(), A->ja.data(), A->a.data(), sigmas, nsigmas, niters,
                                   nniters, use_tails, out);
    }
    auto h = std::make_unique<mlp_hier_s>();
    h->A0.nrow = h->A0.ncol = A->n;                     // level 0 is read on the device
    ChebParams p;
    p.sigmas.assign(sigmas, sigmas + nsigmas);
    p.niters.assign(niters, niters + nniters);
    p.use_tails = use_tails != 0;
    DevCSRView v;
    v.ia = A->d_ia;
    v.ja = A->d_ja;
    v.a = A->d_a;
    v.nnz = A->nnz;
    auto dev = device_cheb_builder(ctx->stream, v);
    build_uh_cheb(h->A0, p, h->H, dev.get());
    *out = h.release();
    MLP_API_END
}

int mlp_hier_build_multi_split(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a,
                               const char* config_path, const mlp_msplit_params* params, mlp_hier* out) {
    MLP_API_BEGIN
    require(out, "null output");
    check_skyline(n, ia, ja);
    MultiSplitParams p;
    if (params) {
        if (params->q && params->nq) p.q.assign(params->q, params->q + params->nq);
        if (params->niters && params->nniters) p.niters.assign(params->niters, params->niters + params->nniters);
        if (params->eps && params->neps) p.eps.assign(params->eps, params->eps + params->neps);
        p.inner_dynamic = params->inner_dynamic != 0;
        require(params->order >= 0 && params->order <= 3, "unknown elimination order");
        p.order = ElimOrder(params->order);
        p.use_tails = params->use_tails != 0;
        p.degree_max = params->degree_max;
        p.max_levels = params->max_levels;
        p.coarse_n = params->coarse_n;
    }
    if (config_path && *config_path) {
        try {
            read_multi_split_config(config_path, p);
        } catch (const Error& e) {
            throw Error(e.what(), MLP_ERR_CONFIG);
        }
    }
    auto h = std::make_unique<mlp_hier_s>();
    h->A0.nrow = h->A0.ncol = n;
    h->A0.ia.assign(ia, ia + n + 1);
    h->A0.ja.assign(ja, ja + ia[n]);
    h->A0.a.assign(a, a + ia[n]);
    build_multi_split(h->A0, p, h->H);
    *out = h.release();
    MLP_API_END
}

int mlp_hier_destroy(mlp_hier h) {
    MLP_API_BEGIN
    delete h;
    MLP_API_END
}

int mlp_hier_nlevels(mlp_hier h, uint32_t* out) {
    MLP_API_BEGIN
    require(h && out, "null argument");
    *out = h->H.nlevels;
    MLP_API_END
}

int mlp_hier_level_info(mlp_hier h, uint32_t l, mlp_level_info* o) {
    MLP_API_BEGIN
    require(h && o, "null argument");
    require(l < h->H.nlevels, "level out of range");
    const HostLevel& L = h->H.levels[l];
    const bool next = l + 1 < h->H.nlevels;
    o->N = L.N; o->nnz = L.nnz; o->M = L.M; o->Md = L.Md;
    o->nit = h->H.kind == PrecKind::UH_CHEB ? L.ncheb : L.niter;
    o->nnz_L = L.L.nnz(); o->nnz_U = L.U.nnz();
    o->n_next = next ? h->H.levels[l + 1].A.nrow : 0;
    o->nnz_next = next ? h->H.levels[l + 1].A.nnz() : 0;
    o->alpha = L.alpha;
    o->beta_or_q = h->H.kind == PrecKind::UH_CHEB ? L.beta : L.q;
    o->lmin = L.lmin; o->lmax = L.lmax; o->eps = L.eps;
    MLP_API_END
}

int mlp_hier_level_array(mlp_hier h, uint32_t l, int which, void* out) {
    MLP_API_BEGIN
    require(h && out, "null argument");
    require(l < h->H.nlevels, "level out of range");
    const HostLevel& L = h->H.levels[l];
    const bool next = l + 1 < h->H.nlevels;
    auto cp = [&](const auto& v) {
        if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(v[0]));
    };
    switch (which) {
        case 0: cp(L.map); break;
        case 1: cp(L.L.ia); break;
        case 2: cp(L.L.ja); break;
        case 3: cp(L.L.a); break;
        case 4: cp(L.U.ia); break;
        case 5: cp(L.U.ja); break;
        case 6: cp(L.U.a); break;
        case 7: cp(L.dval); break;
        case 8: if (next) cp(h->H.levels[l + 1].A.ia); break;
        case 9: if (next) cp(h->H.levels[l + 1].A.ja); break;
        case 10: if (next) cp(h->H.levels[l + 1].A.a); break;
        case 11: cp(L.rmap); break;
        case 12: cp(L.aux); break;
        default: throw Error("unknown array id", MLP_ERR_INVALID);
    }
    MLP_API_END
}

// ---- device preconditioners ----
}  // extern "C"
// A launch schedule to build with (the autotuner's candidates); the default picks per hierarchy.
struct Schedule {
    bool no_chain = false, no_pair = false;
    uint32_t lbs = 0;
    uint32_t chain_bs = 0;
};

template <class VT>
static DevPrecBase* make_hier(mlp_ctx ctx, int kind, uint32_t nlevels, const mlp_level_desc* levels, mlp_mat A0,
                              bool bf16 = false, const Schedule& sc = Schedule()) {
    auto* h = new DevHier<VT>();
    h->ctx = ctx;
    h->bf16 = bf16;
    h->no_chain = sc.no_chain;
    h->no_pair = sc.no_pair;
    h->sched_lbs = sc.lbs;
    h->want_chain_bs = sc.chain_bs;
    try {
        std::vector<mlp_level_desc> d(levels, levels + nlevels);
        h->build(d, kind, A0);
    } catch (...) {
        delete h;
        throw;
    }
    return h;
}
extern "C" {

static int prec_upload(mlp_ctx ctx, int kind, uint32_t nlevels, const mlp_level_desc* levels, mlp_mat A0,
                       int precision, mlp_prec* out, const Schedule& sc) {
    MLP_API_BEGIN
    require(ctx && out && levels && nlevels && A0, "null argument");
    require(kind == MLP_PREC_UH_CHEB || kind == MLP_PREC_MULTI_SPLIT, "unknown hierarchy kind");
    require(precision == MLP_FP64 || precision == MLP_FP32 || precision == MLP_FP32_BF16, "unknown precision");
    require(levels[0].N == A0->n, "level 0 size differs from A0");
    use_device(ctx);
    auto p = std::make_unique<mlp_prec_s>();
    p->ctx = ctx;
    p->id = mlp::new_handle_id();
    p->kind = kind;
    p->precision = precision;
    p->n = A0->n;
    if (precision == MLP_FP64) p->impl.reset(make_hier<double>(ctx, kind, nlevels, levels, A0, false, sc));
    else p->impl.reset(make_hier<float>(ctx, kind, nlevels, levels, A0, precision == MLP_FP32_BF16, sc));
    *out = p.release();
    MLP_API_END
}

int mlp_prec_upload(mlp_ctx ctx, int kind, uint32_t nlevels, const mlp_level_desc* levels, mlp_mat A0, int precision,
                    mlp_prec* out) {
    return prec_upload(ctx, kind, nlevels, levels, A0, precision, out, Schedule());
}

static std::vector<mlp_level_desc> level_descs(const Hierarchy& H) {
    std::vector<mlp_level_desc> d(H.nlevels);
    for (uint32_t l = 0; l < H.nlevels; l++) {
        const HostLevel& L = H.levels[l];
        mlp_level_desc& D = d[l];
        std::memset(&D, 0, sizeof(D));
        D.N = L.N; D.M = L.M; D.Md = L.Md;
        D.nit = H.kind == PrecKind::UH_CHEB ? L.ncheb : L.niter;
        D.lmin = L.lmin; D.lmax = L.lmax; D.eps = L.eps;
        D.map = L.map.data();
        D.L_ia = L.L.ia.data(); D.L_ja = L.L.ja.data(); D.L_a = L.L.a.data();
        D.U_ia = L.U.ia.data(); D.U_ja = L.U.ja.data(); D.U_a = L.U.a.data();
        D.dval = L.dval.data();
        if (l > 0) { D.A_ia = L.A.ia.data(); D.A_ja = L.A.ja.data(); D.A_a = L.A.a.data(); }
    }
    return d;
}

int mlp_prec_from_hier(mlp_ctx ctx, mlp_hier h, mlp_mat A0, int precision, mlp_prec* out) {
    MLP_API_BEGIN
    require(ctx && h && A0 && out, "null argument");
    use_device(ctx);
    const Hierarchy& H = h->H;
    const std::vector<mlp_level_desc> d = level_descs(H);
    int kind = H.kind == PrecKind::UH_CHEB ? MLP_PREC_UH_CHEB : MLP_PREC_MULTI_SPLIT;
    int rc = mlp_prec_upload(ctx, kind, H.nlevels, d.data(), A0, precision, out);
    if (rc) return rc;
    MLP_API_END
}

int mlp_prec_from_hier_tuned(mlp_ctx ctx, mlp_hier h, mlp_mat A0, int precision, uint32_t reps,
                             mlp_tune_report* report, mlp_prec* out) {
    MLP_API_BEGIN
    require(ctx && h && A0 && out, "null argument");
    require(reps >= 1 && reps <= 1000, "reps must lie in [1, 1000]");
    use_device(ctx);
    const Hierarchy& H = h->H;
    const std::vector<mlp_level_desc> d = level_descs(H);
    const int kind = H.kind == PrecKind::UH_CHEB ? MLP_PREC_UH_CHEB : MLP_PREC_MULTI_SPLIT;
    std::vector<Schedule> cand;
    cand.push_back(Schedule());                                  // the default: chain where it fits
    if (kind == MLP_PREC_UH_CHEB) {
        cand.push_back({false, false, 0, 2 * kChainBS});         // the chain in two-wavefront blocks
        cand.push_back({true, false, 0});                        // level kernels, fused last pair
        cand.push_back({true, true, 0});                         // level kernels only
    }
    cand.push_back({true, false, 64});                           // every level program in 64-thread blocks
    cand.push_back({true, false, 256});                          // ... in 256-thread blocks
    // time `reps` applies of each candidate on one random right-hand side (after one untimed apply)
    const uint32_t n = A0->n;
    DevMem tmp;
    double* f = tmp.alloc<double>(n);
    double* x = tmp.alloc<double>(n);
    {
        std::vector<double> hf(n);
        uint64_t z = 0x9e3779b97f4a7c15ull;
        for (uint32_t i = 0; i < n; i++) {
            z = z * 6364136223846793005ull + 1442695040888963407ull;
            hf[i] = double(int64_t(z >> 11) - (int64_t(1) << 52)) / double(int64_t(1) << 52);
        }
        MLP_HIP_CHECK(hipMemcpyAsync(f, hf.data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
    mlp_prec best = nullptr;
    double best_ms = 0.0;
    mlp_tune_report rp;
    std::memset(&rp, 0, sizeof(rp));
    std::vector<uint32_t> seen;
    for (const Schedule& sc : cand) {
        mlp_prec p = nullptr;
        int rc = prec_upload(ctx, kind, H.nlevels, d.data(), A0, precision, &p, sc);
        if (rc) {
            if (best) mlp_prec_destroy(best);
            return rc;
        }
        const uint32_t sb = p->impl->schedule();
        if (std::find(seen.begin(), seen.end(), sb) != seen.end()) {   // the same schedule as an earlier one
            mlp_prec_destroy(p);
            continue;
        }
        seen.push_back(sb);
        double ms = 0.0;
        p->impl->apply(f, x, nullptr);
        rc = mlp_time_prec_apply(ctx, p, f, x, reps, &ms);
        if (rc) {
            mlp_prec_destroy(p);
            if (best) mlp_prec_destroy(best);
            return rc;
        }
        if (rp.ncand < 8) {
            rp.sched[rp.ncand] = sb;
            rp.us[rp.ncand] = 1e3 * ms;
        }
        if (!best || ms < best_ms) {
            if (best) mlp_prec_destroy(best);
            best = p;
            best_ms = ms;
            rp.chosen = rp.ncand;
        } else {
            mlp_prec_destroy(p);
        }
        rp.ncand++;
    }
    if (report) *report = rp;
    *out = best;
    MLP_API_END
}

int mlp_prec_simple(mlp_ctx ctx, int kind, mlp_mat A0, mlp_prec* out) {
    MLP_API_BEGIN
    require(ctx && A0 && out, "null argument");
    require(kind == MLP_PREC_DIAG || kind == MLP_PREC_ID, "kind must be DIAG or ID");
    use_device(ctx);
    auto sp = std::make_unique<SimplePrec>();
    sp->ctx = ctx;
    sp->n = A0->n;
    if (kind == MLP_PREC_DIAG) {
        ensure_host(A0);
        std::vector<double> d(A0->n);
        for (uint32_t i = 0; i < A0->n; i++) {
            require(A0->a[A0->ia[i]] > 0, "DiagPrec needs a positive diagonal");
            d[i] = 1. / A0->a[A0->ia[i]];
        }
        sp->d = sp->mem.upload(d, ctx->stream);
        MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
    auto p = std::make_unique<mlp_prec_s>();
    p->ctx = ctx;
    p->id = mlp::new_handle_id();
    p->kind = kind;
    p->n = A0->n;
    p->impl = std::move(sp);
    *out = p.release();
    MLP_API_END
}

int mlp_prec_gs(mlp_ctx ctx, mlp_mat A, mlp_prec* out) {
    MLP_API_BEGIN
    require(ctx && A && out, "null argument");
    use_device(ctx);
    ensure_host(A);
    auto g = std::make_unique<DevGS>();
    g->ctx = ctx;
    g->build(A->ia, A->ja, A->a);
    auto p = std::make_unique<mlp_prec_s>();
    p->ctx = ctx;
    p->id = mlp::new_handle_id();
    p->kind = MLP_PREC_GS;
    p->n = A->n;
    p->impl = std::move(g);
    *out = p.release();
    MLP_API_END
}

int mlp_prec_composite(mlp_ctx ctx, mlp_mat A, mlp_hier ms, mlp_prec* out) {
    MLP_API_BEGIN
    require(ctx && A && ms && out, "null argument");
    require(ms->H.kind == PrecKind::MULTI_SPLIT, "the second stage must be a multi_split hierarchy");
    ensure_host(A);
    require(ms->A0.nrow == A->n && std::equal(A->ia.begin(), A->ia.end(), ms->A0.ia.begin(), ms->A0.ia.end()) &&
                std::equal(A->ja.begin(), A->ja.end(), ms->A0.ja.begin(), ms->A0.ja.end()) &&
                std::equal(A->a.begin(), A->a.end(), ms->A0.a.begin(), ms->A0.a.end()),
            "the multi_split hierarchy must be built on A");
    use_device(ctx);
    ensure_host(A);
    auto g = std::make_unique<DevGS>();
    g->ctx = ctx;
    g->build(A->ia, A->ja, A->a);
    mlp_prec second = nullptr;
    int rc = mlp_prec_from_hier(ctx, ms, A, MLP_FP64, &second);
    if (rc) return rc;
    std::unique_ptr<mlp_prec_s> holder(second);
    auto c = std::make_unique<DevComposite>();
    c->first = std::move(g);
    c->second = std::move(holder->impl);
    c->tmp = c->mem.alloc<double>(std::max<uint32_t>(A->n, 1));
    auto p = std::make_unique<mlp_prec_s>();
    p->ctx = ctx;
    p->id = mlp::new_handle_id();
    p->kind = MLP_PREC_COMPOSITE;
    p->n = A->n;
    p->impl = std::move(c);
    *out = p.release();
    MLP_API_END
}

int mlp_prec_sym(mlp_ctx ctx, mlp_mat A, const double* sigmas, uint32_t nsigmas, const uint32_t* niters,
                 uint32_t nniters, int use_tails, mlp_prec* out) {
    MLP_API_BEGIN
    require(ctx && A && out, "null argument");
    require(sigmas && nsigmas && niters && nniters, "sigmas and niters are required");
    use_device(ctx);
    auto y = std::make_unique<DevSym>(ctx);
    y->build(A, sigmas, nsigmas, niters, nniters, use_tails);
    auto p = std::make_unique<mlp_prec_s>();
    p->ctx = ctx;
    p->id = mlp::new_handle_id();
    p->kind = MLP_PREC_SYM;
    p->n = A->n;
    p->impl = std::move(y);
    *out = p.release();
    MLP_API_END
}

int mlp_prec_clone(mlp_ctx ctx, mlp_prec src, mlp_prec* out) {
    MLP_API_BEGIN
    require(ctx && src && out, "null argument");
    require(ctx->device == src->ctx->device, "the clone must be on the device of the source");
    use_device(ctx);
    std::unique_ptr<DevPrecBase> impl = src->impl->clone_on(ctx);
    require(impl != nullptr, "only uh_cheb and multi_split preconditioners can be cloned");
    auto p = std::make_unique<mlp_prec_s>();
    p->ctx = ctx;
    p->id = mlp::new_handle_id();
    p->kind = src->kind;
    p->precision = src->precision;
    p->n = src->n;
    p->impl = std::move(impl);
    *out = p.release();
    MLP_API_END
}

int mlp_prec_sym_set_eps(mlp_prec p, double eps) {
    MLP_API_BEGIN
    require(p && p->kind == MLP_PREC_SYM, "not a SymPrec");
    require(eps > 0, "eps must be positive");
    static_cast<DevSym*>(p->impl.get())->eps = eps;
    MLP_API_END
}

int mlp_prec_destroy(mlp_prec p) {
    MLP_API_BEGIN
    use_device(p ? p->ctx : nullptr);
    if (p) {
        (void)hipStreamSynchronize(p->ctx->stream);
        p->ctx->drop_graphs(p->id);
        delete p;
    }
    MLP_API_END
}

int mlp_prec_apply(mlp_ctx ctx, mlp_prec p, const double* f, double* x) {
    MLP_API_BEGIN
    require(ctx && p, "null argument");
    use_device(ctx);
    p->impl->apply(f, x, nullptr);
    MLP_HIP_CHECK(hipGetLastError());
    MLP_API_END
}

int mlp_time_spmv(mlp_ctx ctx, mlp_mat A, const double* x, double* y, uint32_t reps, size_t flush_bytes,
                  double* avg_ms) {
    MLP_API_BEGIN
    require(ctx && A && avg_ms && reps, "null argument");
    use_device(ctx);
    // flush_bytes > 0: launches rotate over copies of the matrix totalling at
    // least flush_bytes, so each launch finds its matrix evicted from L2 and
    // the Infinity Cache by the launches before it; 0: the same matrix every
    // launch (cache-warm).  Launches run back to back between two events.
    std::vector<SellView<double>> views{A->A.v};
    std::vector<void*> extra;
    const size_t mbytes = A->A.stored_bytes();
    // a device copy of n elements of T (freed with the other copies)
    auto copy_of = [&](const auto* src, size_t n) {
        using T = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
        void* p = nullptr;
        MLP_HIP_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
        extra.push_back(p);
        MLP_HIP_CHECK(hipMemcpyAsync(p, src, n * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
        return static_cast<const T*>(p);
    };
    if (flush_bytes && A->A.stencil) {
        const size_t nc = std::max<size_t>(2, (flush_bytes + mbytes - 1) / mbytes);
        for (size_t c = 1; c < nc; c++) {
            SellView<double> v = A->A.v;
            v.spres = copy_of(A->A.v.spres, A->n);
            v.sd = copy_of(A->A.v.sd, A->n);
            v.slx = copy_of(A->A.v.slx, A->n);
            v.sly = copy_of(A->A.v.sly, A->n);
            v.slz = copy_of(A->A.v.slz, A->n);
            views.push_back(v);
        }
    } else if (flush_bytes && mbytes) {
        const size_t nc = std::max<size_t>(2, (flush_bytes + mbytes - 1) / mbytes);
        for (size_t c = 1; c < nc; c++) {
            SellView<double> v = A->A.v;
            const size_t ib = A->A.v.dcol ? sizeof(int16_t) : sizeof(uint32_t);
            void* idx = nullptr;
            double* val = nullptr;
            MLP_HIP_CHECK(hipMalloc(&idx, A->A.nent * ib));
            extra.push_back(idx);
            MLP_HIP_CHECK(hipMalloc(&val, A->A.nent * sizeof(double)));
            extra.push_back(val);
            MLP_HIP_CHECK(hipMemcpyAsync(idx, A->A.v.dcol ? (const void*)A->A.v.dcol : (const void*)A->A.v.col,
                                         A->A.nent * ib, hipMemcpyDeviceToDevice, ctx->stream));
            MLP_HIP_CHECK(hipMemcpyAsync(val, A->A.v.val, A->A.nent * sizeof(double), hipMemcpyDeviceToDevice,
                                         ctx->stream));
            if (A->A.v.dcol) v.dcol = static_cast<const int16_t*>(idx);
            else v.col = static_cast<const uint32_t*>(idx);
            v.val = val;
            views.push_back(v);
        }
    }
    auto run = [&](uint32_t r) { launch_rows(views[r % views.size()], nullptr, x, y, ctx->stream); };
    for (uint32_t r = 0; r < views.size(); r++) run(r);   // untimed pass over every copy
    hipEvent_t e0, e1;
    MLP_HIP_CHECK(hipEventCreate(&e0));
    MLP_HIP_CHECK(hipEventCreate(&e1));
    MLP_HIP_CHECK(hipEventRecord(e0, ctx->stream));
    for (uint32_t r = 0; r < reps; r++) run(r);
    MLP_HIP_CHECK(hipEventRecord(e1, ctx->stream));
    MLP_HIP_CHECK(hipEventSynchronize(e1));
    MLP_HIP_CHECK(hipGetLastError());
    float ms = 0.f;
    MLP_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (void* p : extra) (void)hipFree(p);
    *avg_ms = double(ms) / reps;
    MLP_API_END
}

int mlp_time_prec_apply(mlp_ctx ctx, mlp_prec p, const double* f, double* x, uint32_t reps, double* avg_ms) {
    MLP_API_BEGIN
    require(ctx && p && avg_ms && reps, "null argument");
    use_device(ctx);
    hipEvent_t e0, e1;
    MLP_HIP_CHECK(hipEventCreate(&e0));
    MLP_HIP_CHECK(hipEventCreate(&e1));
    MLP_HIP_CHECK(hipEventRecord(e0, ctx->stream));
    for (uint32_t r = 0; r < reps; r++) p->impl->apply(f, x, nullptr);
    MLP_HIP_CHECK(hipEventRecord(e1, ctx->stream));
    MLP_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    MLP_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = double(ms) / reps;
    MLP_API_END
}

uint64_t mlp_spe_nnz(uint32_t nx, uint32_t ny, uint32_t nz) { return spe_nnz(nx, ny, nz); }

int mlp_spe_assemble(int kind, uint32_t nx, uint32_t ny, uint32_t nz, const double* kx, double kz_ratio,
                     double c, double hx, double hy, double hz, uint32_t* ia, uint32_t* ja, double* a) {
    MLP_API_BEGIN
    require(kx && ia && ja && a, "null argument");
    require(kind == MLP_ASM_SPE || kind == MLP_ASM_TPFA, "kind must be MLP_ASM_SPE or MLP_ASM_TPFA");
    spe_assemble(SpeKind(kind), nx, ny, nz, kx, kz_ratio, c, hx, hy, hz, ia, ja, a);
    MLP_API_END
}

int mlp_prec_bytes(mlp_prec p, double* bytes, uint32_t* launches) {
    MLP_API_BEGIN
    require(p, "null argument");
    if (bytes) *bytes = p->impl->bytes_per_apply();
    if (launches) *launches = p->impl->launches_per_apply();
    MLP_API_END
}

int mlp_prec_schedule(mlp_prec p, uint32_t* sched) {
    MLP_API_BEGIN
    require(p && sched, "null argument");
    *sched = p->impl->schedule();
    MLP_API_END
}

}  // extern "C"

// ---------------------------------------------------------------------------
// solvers
// ---------------------------------------------------------------------------
namespace mlp {
namespace {

constexpr uint32_t kGmresRestart = 30;     // gmres.cc:21
// Flexible PCG: PCG phases restarted from the FP64 true residual b - A x after the
// recurrence residual met eps, at most this many times (include/mlp.h).
constexpr uint32_t kFlexRestarts = 8;

void ensure_workspace(mlp_ctx ctx, uint32_t n, bool gmres) {
    if (ctx->ws_n >= n && (!gmres || ctx->V)) return;
    MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->drop_graphs(0);
    ctx->ws_mem.release();
    ctx->V = nullptr;
    const uint32_t m = std::max(n, ctx->ws_n);
    ctx->r = ctx->ws_mem.alloc<double>(m);
    ctx->z = ctx->ws_mem.alloc<double>(m);
    ctx->p = ctx->ws_mem.alloc<double>(m);
    ctx->Ap = ctx->ws_mem.alloc<double>(m);
    ctx->zp = ctx->ws_mem.alloc<double>(m);
    ctx->mdot = ctx->ws_mem.alloc<double>(32);
    if (gmres) ctx->V = ctx->ws_mem.alloc<double>(size_t(m) * kGmresRestart);
    ctx->ws_n = m;
}

void fill_x0(mlp_ctx ctx, uint32_t n, int mode, double* x) {
    if (mode == MLP_X0_GIVEN) return;
    if (mode != MLP_X0_ZERO && mode != MLP_X0_REFERENCE) throw Error("unknown x0 mode", MLP_ERR_INVALID);
    if (mode == MLP_X0_ZERO) {
        MLP_HIP_CHECK(hipMemsetAsync(x, 0, size_t(n) * sizeof(double), ctx->stream));
        return;
    }
    std::vector<double> h(n, 0.0);
    if (mode == MLP_X0_REFERENCE) {        // generate_x0, modules/solvers/solvers.cc:9-21
        srandom(3);
        for (uint32_t i = 0; i < n; i++) h[i] = 20. * (random() - 0.5 * RAND_MAX) / RAND_MAX + 100;
    }
    MLP_HIP_CHECK(hipMemcpyAsync(x, h.data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

// the device state `st` through the context's pinned buffer
SolveState read_state(mlp_ctx ctx, const SolveState* st) {
    SolveState s;
    MLP_HIP_CHECK(hipMemcpyAsync(ctx->host_pinned, st, sizeof(SolveState), hipMemcpyDeviceToHost, ctx->stream));
    MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    std::memcpy(&s, ctx->host_pinned, sizeof(SolveState));
    return s;
}
SolveState read_state(mlp_ctx ctx) { return read_state(ctx, ctx->st); }

void write_state(mlp_ctx ctx, SolveState* st, const SolveState& s) {
    std::memcpy(ctx->host_pinned, &s, sizeof(SolveState));
    MLP_HIP_CHECK(hipMemcpyAsync(st, ctx->host_pinned, sizeof(SolveState), hipMemcpyHostToDevice, ctx->stream));
    MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}
void write_state(mlp_ctx ctx, const SolveState& s) { write_state(ctx, ctx->st, s); }

void launch_residual(mlp_ctx ctx, mlp_mat A, const double* b, const double* x, double* r) {
    launch_rows(A->A.v, b, x, r, ctx->stream);
}

double true_rel_res(mlp_ctx ctx, mlp_mat A, const double* b, const double* x, double* r, double init_norm) {
    launch_residual(ctx, A, b, x, r);
    const double nr = std::sqrt(dev_dot(ctx, A->n, r, r));
    return init_norm > 0 ? nr / init_norm : nr;
}

struct GraphPair {
    hipGraph_t g[2] = {nullptr, nullptr};
    hipGraphExec_t e[2] = {nullptr, nullptr};
    ~GraphPair() {
        for (int i = 0; i < 2; i++) {
            if (e[i]) (void)hipGraphExecDestroy(e[i]);
            if (g[i]) (void)hipGraphDestroy(g[i]);
        }
    }
};

// The vectors, states and partial sums of R PCG solves run together (R = 1:
// the context's own); vectors and buffers of solve y sit y strides on.
struct PcgVecs {
    uint32_t R = 1;
    double *r = nullptr, *z = nullptr, *p = nullptr, *Ap = nullptr, *zp = nullptr;
    SolveState* st = nullptr;
    double* partials = nullptr;          // kPartials doubles per solve
    double* slot = nullptr;              // result of a standalone dot (R = 1)
    Batch bt;
};

PcgVecs ctx_vecs(mlp_ctx ctx, uint32_t n) {
    PcgVecs V;
    V.r = ctx->r; V.z = ctx->z; V.p = ctx->p; V.Ap = ctx->Ap; V.zp = ctx->zp;
    V.st = ctx->st;
    V.partials = ctx->partials;
    V.slot = ctx->scal;
    V.bt.n = n;
    V.bt.part = uint32_t(kPartials);
    return V;
}

// (x, y) with the partial sums and result slot of workspace V (dev_dot's kernels and grid)
double dev_dot(mlp_ctx ctx, uint32_t n, const double* x, const double* y, const PcgVecs& V) {
    const uint32_t g = red_grid(n);
    hipLaunchKernelGGL(k_dot, dim3(g), dim3(kBS), 0, ctx->stream, n, x, y, V.partials);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBS), 0, ctx->stream, int(FIN_DOT), g, (const double*)V.partials,
                       V.st, V.slot, Batch{});
    MLP_HIP_CHECK(hipMemcpyAsync(ctx->host_pinned, V.slot, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return ctx->host_pinned[0];
}

// ||b - A x|| / init_norm, the residual left in V.r
double true_rel_res(mlp_ctx ctx, mlp_mat A, const double* b, const double* x, const PcgVecs& V, double init_norm) {
    launch_residual(ctx, A, b, x, V.r);
    const double nr = std::sqrt(dev_dot(ctx, A->n, V.r, V.r, V));
    return init_norm > 0 ? nr / init_norm : nr;
}

// One PCG iteration (pcg.cc:46-76) on the context's stream for the solves of
// V; every kernel is a no-op for a solve whose device state says `done`.
// replace = true: the recurrence residual is replaced by b - A x before the
// preconditioner is applied, and the stop test uses the replaced residual.
void enqueue_pcg_iteration(mlp_ctx ctx, mlp_mat A, DevPrecBase* B, const double* b, double* x, bool flex,
                           bool replace, const PcgVecs& V) {
    const uint32_t n = A->n, rg = red_grid(n);
    hipStream_t s = ctx->stream;
    const dim3 grid(rg, V.R), one(1, V.R);
    // two partial-sum buffers: k_pcg_xr reads the (Ap, p) partials while it writes the (r, r) ones
    double* P0 = V.partials;
    double* P1 = V.partials + kPartials / 2;
    auto fin = [&](int op, const double* part) {
        hipLaunchKernelGGL(k_finalize, one, dim3(kBS), 0, s, op, rg, part, V.st, (double*)nullptr, V.bt);
    };
    hipLaunchKernelGGL(k_pcg_spmv_dot<false>, grid, dim3(kBS), 0, s, A->A.v, (const double*)V.p, V.Ap,
                       (const SolveState*)V.st, P0, V.bt);
    hipLaunchKernelGGL(k_pcg_xr<false>, grid, dim3(kBS), 0, s, n, (const double*)V.p, (const double*)V.Ap, x, V.r, V.st,
                       (const double*)P0, rg, P1, V.bt);
    if (replace) {
        fin(FIN_XR_NOTEST, P1);
        hipLaunchKernelGGL(k_pcg_true_res<false>, grid, dim3(kBS), 0, s, A->A.v, b, (const double*)x, V.r,
                           (const SolveState*)V.st, P0, V.bt);
        fin(FIN_TRUE_RES, P0);
    } else {
        fin(FIN_XR, P1);
    }
    B->apply(V.r, V.z, &V.st->done);
    if (flex) {
        hipLaunchKernelGGL(k_pcg_rz<true>, grid, dim3(kBS), 0, s, n, (const double*)V.r, (const double*)V.z,
                           (const double*)V.zp, (const SolveState*)V.st, P0, V.bt);
        hipLaunchKernelGGL(k_pcg_p<true>, grid, dim3(kBS), 0, s, n, (const double*)V.z, V.p, V.zp, V.st,
                           (const double*)P0, rg, V.bt);
    } else {
        hipLaunchKernelGGL(k_pcg_rz<false>, grid, dim3(kBS), 0, s, n, (const double*)V.r, (const double*)V.z,
                           (const double*)V.zp, (const SolveState*)V.st, P0, V.bt);
        hipLaunchKernelGGL(k_pcg_p<false>, grid, dim3(kBS), 0, s, n, (const double*)V.z, V.p, V.zp, V.st,
                           (const double*)P0, rg, V.bt);
    }
}
void enqueue_pcg_iteration(mlp_ctx ctx, mlp_mat A, mlp_prec B, const double* b, double* x, bool flex, bool replace) {
    enqueue_pcg_iteration(ctx, A, B->impl.get(), b, x, flex, replace, ctx_vecs(ctx, A->n));
}

// p = z, rz = (r, z) after z = B r (pcg.cc:32-33); no-ops when `done`
void enqueue_pcg_start(mlp_ctx ctx, DevPrecBase* B, uint32_t n, bool flex, const PcgVecs& V) {
    const uint32_t rg = red_grid(n);
    hipStream_t s = ctx->stream;
    const dim3 grid(rg, V.R), one(1, V.R);
    B->apply(V.r, V.z, &V.st->done);
    if (flex)
        hipLaunchKernelGGL(k_pcg_init<true>, grid, dim3(kBS), 0, s, n, (const double*)V.r, (const double*)V.z, V.p,
                           V.zp, (const SolveState*)V.st, V.partials, V.bt);
    else
        hipLaunchKernelGGL(k_pcg_init<false>, grid, dim3(kBS), 0, s, n, (const double*)V.r, (const double*)V.z, V.p,
                           V.zp, (const SolveState*)V.st, V.partials, V.bt);
    hipLaunchKernelGGL(k_finalize, one, dim3(kBS), 0, s, int(FIN_INIT), rg, (const double*)V.partials, V.st,
                       (double*)nullptr, V.bt);
}
void enqueue_pcg_start(mlp_ctx ctx, mlp_prec B, uint32_t n, bool flex) {
    enqueue_pcg_start(ctx, B->impl.get(), n, flex, ctx_vecs(ctx, n));
}

// MLP_GRAPHS=0 turns hipGraph capture off in every solver, whatever the
// options ask (for tracing tools that follow plain launches only).
bool graphs_allowed() {
    static const bool on = [] {
        const char* e = getenv("MLP_GRAPHS");
        return !(e && e[0] == '0');
    }();
    return on;
}

// The cached graphs of one PCG iteration (plain and, with true_res_every,
// residual-replacing) for these operands; nullptr when graphs are off.
mlp_ctx_s::GraphEntry* pcg_graphs(mlp_ctx ctx, mlp_mat A, mlp_prec B, const double* b, double* x, bool flex,
                                  uint32_t tre, bool use_graph) {
    if (!use_graph || !graphs_allowed()) return nullptr;
    for (auto& e : ctx->graphs)
        if (e.solver == 0 && e.A == A->id && e.B == B->id && e.b == b && e.x == x && e.flex == int(flex) &&
            e.tre == tre && e.ws_n == ctx->ws_n)
            return &e;
    hipStream_t s = ctx->stream;
    MLP_HIP_CHECK(hipStreamSynchronize(s));
    mlp_ctx_s::GraphEntry e;
    e.A = A->id; e.B = B->id; e.b = b; e.x = x; e.flex = int(flex); e.tre = tre; e.ws_n = ctx->ws_n;
    for (int v = 0; v < (tre ? 2 : 1); v++) {
        MLP_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        enqueue_pcg_iteration(ctx, A, B, b, x, flex, v == 1);
        MLP_HIP_CHECK(hipStreamEndCapture(s, &e.g[v]));
        MLP_HIP_CHECK(hipGraphInstantiate(&e.e[v], e.g[v], nullptr, nullptr, 0));
    }
    return ctx->add_graph(e);
}

// Workspace of R batched solves of size n on the context (kept for the next
// call; growing it drops the batched graphs, which hold its addresses).
PcgVecs ensure_multi(mlp_ctx ctx, uint32_t n, uint32_t R) {
    if (ctx->multi_R < R || ctx->multi_n < n) {
        MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (auto it = ctx->graphs.begin(); it != ctx->graphs.end();) {
            if (it->solver != 3) { ++it; continue; }
            mlp_ctx_s::destroy_graph(*it);
            it = ctx->graphs.erase(it);
        }
        ctx->multi_mem.release();
        if (ctx->mst_host) (void)hipHostFree(ctx->mst_host);
        ctx->mst_host = nullptr;
        const uint32_t r = std::max(R, ctx->multi_R), m = std::max(n, ctx->multi_n);
        const size_t v = size_t(r) * m;
        ctx->mr = ctx->multi_mem.alloc<double>(v);
        ctx->mz = ctx->multi_mem.alloc<double>(v);
        ctx->mp = ctx->multi_mem.alloc<double>(v);
        ctx->mAp = ctx->multi_mem.alloc<double>(v);
        ctx->mzp = ctx->multi_mem.alloc<double>(v);
        ctx->mpart = ctx->multi_mem.alloc<double>(size_t(r) * kPartials);
        ctx->mst = ctx->multi_mem.alloc<SolveState>(r);
        MLP_HIP_CHECK(hipHostMalloc((void**)&ctx->mst_host, size_t(r) * sizeof(SolveState), hipHostMallocDefault));
        ctx->multi_R = r;
        ctx->multi_n = m;
    }
    PcgVecs V;
    V.R = R;
    V.r = ctx->mr; V.z = ctx->mz; V.p = ctx->mp; V.Ap = ctx->mAp; V.zp = ctx->mzp;
    V.st = ctx->mst;
    V.partials = ctx->mpart;
    V.bt.n = n;
    V.bt.part = uint32_t(kPartials);
    return V;
}

// The cached graphs of one batched PCG iteration (plain and, with
// true_res_every, residual-replacing) for these operands.
mlp_ctx_s::GraphEntry* multi_graph(mlp_ctx ctx, mlp_mat A, mlp_prec B, DevPrecBase* W, const double* b, double* x,
                                   bool flex, uint32_t tre, const PcgVecs& V) {
    for (auto& e : ctx->graphs)
        if (e.solver == 3 && e.A == A->id && e.B == B->id && e.b == b && e.x == x && e.R == V.R &&
            e.flex == int(flex) && e.tre == tre && e.ws_n == ctx->multi_n)
            return &e;
    hipStream_t s = ctx->stream;
    MLP_HIP_CHECK(hipStreamSynchronize(s));
    mlp_ctx_s::GraphEntry e;
    e.solver = 3;
    e.A = A->id; e.B = B->id; e.b = b; e.x = x; e.flex = int(flex); e.tre = tre; e.ws_n = ctx->multi_n; e.R = V.R;
    for (int v = 0; v < (tre ? 2 : 1); v++) {
        MLP_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        enqueue_pcg_iteration(ctx, A, W, b, x, flex, v == 1, V);
        MLP_HIP_CHECK(hipStreamEndCapture(s, &e.g[v]));
        MLP_HIP_CHECK(hipGraphInstantiate(&e.e[v], e.g[v], nullptr, nullptr, 0));
    }
    return ctx->add_graph(e);
}

// Workspace of an ensemble solve on the context: the block-diagonal system's n rows, S segments with the
// offsets seg_off (uploaded); growing it drops the ensemble graphs, which hold its addresses.
PcgVecs ensure_ens(mlp_ctx ctx, uint32_t n, uint32_t S, const uint32_t* seg_off) {
    if (ctx->ens_S < S || ctx->ens_n < n) {
        MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (auto it = ctx->graphs.begin(); it != ctx->graphs.end();) {
            if (it->solver != 4) { ++it; continue; }
            mlp_ctx_s::destroy_graph(*it);
            it = ctx->graphs.erase(it);
        }
        ctx->ens_mem.release();
        if (ctx->est_host) (void)hipHostFree(ctx->est_host);
        if (ctx->eout_host) (void)hipHostFree(ctx->eout_host);
        ctx->est_host = nullptr;
        ctx->eout_host = nullptr;
        const uint32_t r = std::max(S, ctx->ens_S), m = std::max(n, ctx->ens_n);
        ctx->er = ctx->ens_mem.alloc<double>(m);
        ctx->ez = ctx->ens_mem.alloc<double>(m);
        ctx->ep = ctx->ens_mem.alloc<double>(m);
        ctx->eAp = ctx->ens_mem.alloc<double>(m);
        ctx->ezp = ctx->ens_mem.alloc<double>(m);
        ctx->epart = ctx->ens_mem.alloc<double>(size_t(r) * kPartials);
        ctx->eout = ctx->ens_mem.alloc<double>(r);
        ctx->esoff = ctx->ens_mem.alloc<uint32_t>(size_t(r) + 1);
        ctx->est = ctx->ens_mem.alloc<SolveState>(r);
        MLP_HIP_CHECK(hipHostMalloc((void**)&ctx->est_host, size_t(r) * sizeof(SolveState), hipHostMallocDefault));
        MLP_HIP_CHECK(hipHostMalloc((void**)&ctx->eout_host, size_t(r) * sizeof(double), hipHostMallocDefault));
        ctx->ens_S = r;
        ctx->ens_n = m;
    }
    MLP_HIP_CHECK(hipMemcpyAsync(ctx->esoff, seg_off, (size_t(S) + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                                 ctx->stream));
    MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    PcgVecs V;
    V.R = S;
    V.r = ctx->er; V.z = ctx->ez; V.p = ctx->ep; V.Ap = ctx->eAp; V.zp = ctx->ezp;
    V.st = ctx->est;
    V.partials = ctx->epart;
    V.bt.n = 0;
    V.bt.part = uint32_t(kPartials);
    V.bt.soff = ctx->esoff;
    return V;
}

// One PCG iteration of every segment of an ensemble (enqueue_pcg_iteration with the segments' own grids); the
// preconditioner applies once to the whole block-diagonal system.  gmax: the largest segment grid.
void enqueue_ens_iteration(mlp_ctx ctx, mlp_mat A, DevPrecBase* B, const double* b, double* x, bool flex,
                           bool replace, const PcgVecs& V, uint32_t gmax) {
    hipStream_t s = ctx->stream;
    const dim3 grid(gmax, V.R), one(1, V.R);
    double* P0 = V.partials;
    double* P1 = V.partials + kPartials / 2;
    auto fin = [&](int op, const double* part) {
        hipLaunchKernelGGL(k_finalize, one, dim3(kBS), 0, s, op, 0u, part, V.st, (double*)nullptr, V.bt);
    };
    hipLaunchKernelGGL(k_pcg_spmv_dot<true>, grid, dim3(kBS), 0, s, A->A.v, (const double*)V.p, V.Ap,
                       (const SolveState*)V.st, P0, V.bt);
    hipLaunchKernelGGL(k_pcg_xr<true>, grid, dim3(kBS), 0, s, 0u, (const double*)V.p, (const double*)V.Ap, x, V.r,
                       V.st, (const double*)P0, 0u, P1, V.bt);
    if (replace) {
        fin(FIN_XR_NOTEST, P1);
        hipLaunchKernelGGL(k_pcg_true_res<true>, grid, dim3(kBS), 0, s, A->A.v, b, (const double*)x, V.r,
                           (const SolveState*)V.st, P0, V.bt);
        fin(FIN_TRUE_RES, P0);
    } else {
        fin(FIN_XR, P1);
    }
    B->apply(V.r, V.z, nullptr);
    if (flex) {
        hipLaunchKernelGGL((k_pcg_rz<true, true>), grid, dim3(kBS), 0, s, 0u, (const double*)V.r, (const double*)V.z,
                           (const double*)V.zp, (const SolveState*)V.st, P0, V.bt);
        hipLaunchKernelGGL((k_pcg_p<true, true>), grid, dim3(kBS), 0, s, 0u, (const double*)V.z, V.p, V.zp, V.st,
                           (const double*)P0, 0u, V.bt);
    } else {
        hipLaunchKernelGGL((k_pcg_rz<false, true>), grid, dim3(kBS), 0, s, 0u, (const double*)V.r,
                           (const double*)V.z, (const double*)V.zp, (const SolveState*)V.st, P0, V.bt);
        hipLaunchKernelGGL((k_pcg_p<false, true>), grid, dim3(kBS), 0, s, 0u, (const double*)V.z, V.p, V.zp, V.st,
                           (const double*)P0, 0u, V.bt);
    }
}

// p = z, rz = (r, z) after z = B r for every segment not done
void enqueue_ens_start(mlp_ctx ctx, DevPrecBase* B, bool flex, const PcgVecs& V, uint32_t gmax) {
    hipStream_t s = ctx->stream;
    const dim3 grid(gmax, V.R), one(1, V.R);
    B->apply(V.r, V.z, nullptr);
    if (flex)
        hipLaunchKernelGGL((k_pcg_init<true, true>), grid, dim3(kBS), 0, s, 0u, (const double*)V.r, (const double*)V.z,
                           V.p, V.zp, (const SolveState*)V.st, V.partials, V.bt);
    else
        hipLaunchKernelGGL((k_pcg_init<false, true>), grid, dim3(kBS), 0, s, 0u, (const double*)V.r,
                           (const double*)V.z, V.p, V.zp, (const SolveState*)V.st, V.partials, V.bt);
    hipLaunchKernelGGL(k_finalize, one, dim3(kBS), 0, s, int(FIN_INIT), 0u, (const double*)V.partials, V.st,
                       (double*)nullptr, V.bt);
}

// (r, r) of every segment r = b - A x (k_dot_batch with the segments' grids) into the partials, then op
void enqueue_ens_norms(mlp_ctx ctx, mlp_mat A, const double* b, const double* x, const PcgVecs& V, uint32_t gmax,
                       int op, double* out) {
    hipStream_t s = ctx->stream;
    launch_residual(ctx, A, b, x, V.r);
    hipLaunchKernelGGL(k_dot_batch<true>, dim3(gmax, V.R), dim3(kBS), 0, s, 0u, (const double*)V.r, V.partials, V.bt);
    hipLaunchKernelGGL(k_finalize, dim3(1, V.R), dim3(kBS), 0, s, op, 0u, (const double*)V.partials, V.st, out, V.bt);
}

// the cached graphs of one ensemble iteration (plain and residual-replacing)
mlp_ctx_s::GraphEntry* ens_graph(mlp_ctx ctx, mlp_mat A, mlp_prec B, const double* b, double* x, bool flex,
                                 uint32_t tre, const PcgVecs& V, uint32_t gmax) {
    for (auto& e : ctx->graphs)
        if (e.solver == 4 && e.A == A->id && e.B == B->id && e.b == b && e.x == x && e.R == V.R &&
            e.flex == int(flex) && e.tre == tre && e.ws_n == ctx->ens_n && e.gmax == gmax)
            return &e;
    hipStream_t s = ctx->stream;
    MLP_HIP_CHECK(hipStreamSynchronize(s));
    mlp_ctx_s::GraphEntry e;
    e.solver = 4;
    e.A = A->id; e.B = B->id; e.b = b; e.x = x; e.flex = int(flex); e.tre = tre; e.ws_n = ctx->ens_n; e.R = V.R;
    e.gmax = gmax;
    for (int v = 0; v < (tre ? 2 : 1); v++) {
        MLP_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        enqueue_ens_iteration(ctx, A, B->impl.get(), b, x, flex, v == 1, V, gmax);
        MLP_HIP_CHECK(hipStreamEndCapture(s, &e.g[v]));
        MLP_HIP_CHECK(hipGraphInstantiate(&e.e[v], e.g[v], nullptr, nullptr, 0));
    }
    return ctx->add_graph(e);
}

// iterations k0+1 .. k0+chunk (graph replays or direct launches)
void enqueue_pcg_chunk(mlp_ctx ctx, mlp_mat A, mlp_prec B, const double* b, double* x, bool flex, uint32_t tre,
                       mlp_ctx_s::GraphEntry* ge, uint32_t k0, uint32_t chunk, const PcgVecs& V) {
    for (uint32_t k = k0 + 1; k <= k0 + chunk; k++) {
        const bool replace = tre && (k % tre == 0);
        if (ge) MLP_HIP_CHECK(hipGraphLaunch(ge->e[replace ? 1 : 0], ctx->stream));
        else enqueue_pcg_iteration(ctx, A, B->impl.get(), b, x, flex, replace, V);
    }
}

// Iterations to enqueue before the next look at the device state.  An
// iteration enqueued after convergence still launches its kernels (they return
// at once), and each look costs a host round trip, so after the first chunk
// the count is predicted from the mean reduction per iteration so far:
// ceil(log(eps / rel) / log(rate)), between 1 and max(check_every, 64).
uint32_t next_chunk(const SolveState& st, uint32_t check_every) {
    const double rel = st.init_norm > 0 ? st.norm / st.init_norm : 0.0;
    if (!(rel > 0 && rel < 1) || st.niter == 0 || !(st.eps > 0)) return check_every;
    const double rate = std::pow(rel, 1.0 / st.niter);
    if (!(rate > 0 && rate < 1)) return check_every;
    const double need = std::ceil(std::log(st.eps / rel) / std::log(rate));
    if (!(need >= 1)) return 1;
    return uint32_t(std::min<double>(need, std::max<uint32_t>(check_every, 64)));
}

// One PCG phase from the current x: r = b - A x, z = B r, p = z, then
// iterations until the device sets `done` (pcg.cc:24-76).  W: a workspace
// other than the context's (the iterations then run as plain launches).
SolveState pcg_phase(mlp_ctx ctx, mlp_mat A, mlp_prec B, const double* b, double* x, const mlp_solve_opts* o,
                     double init_norm, double eps, uint32_t niter0, bool* no_iter, const PcgVecs* W = nullptr) {
    const uint32_t n = A->n;
    const bool flex = o->flexible != 0;
    const PcgVecs V = W ? *W : ctx_vecs(ctx, n);
    launch_residual(ctx, A, b, x, V.r);
    const double norm = std::sqrt(dev_dot(ctx, n, V.r, V.r, V));
    SolveState st0;
    std::memset(&st0, 0, sizeof(st0));
    st0.norm = norm;
    st0.init_norm = init_norm;
    st0.eps = eps;
    st0.maxit = o->maxit;
    st0.niter = niter0;
    write_state(ctx, V.st, st0);
    *no_iter = !(norm / init_norm > eps);
    if (*no_iter) return st0;
    enqueue_pcg_start(ctx, B->impl.get(), n, flex, V);
    const uint32_t chunk = o->check_every ? o->check_every : 16;
    const uint32_t tre = o->true_res_every;
    mlp_ctx_s::GraphEntry* ge =
        W ? nullptr : pcg_graphs(ctx, A, B, b, x, flex, tre, o->use_graph && B->impl->graph_safe());
    SolveState st;
    uint32_t k = 0, c = chunk;
    while (true) {
        enqueue_pcg_chunk(ctx, A, B, b, x, flex, tre, ge, k, c, V);
        k += c;
        MLP_HIP_CHECK(hipGetLastError());
        st = read_state(ctx, V.st);
        if (st.done) break;
        c = next_chunk(st, chunk);
    }
    return st;
}

constexpr uint32_t kCoefCap = 1u << 20;   // Chebyshev outer iterations covered by the coefficient table

// alpha_k, beta_k of ChebSolver steps 2+ (cheb.cc:59-60), k = iteration count,
// computed with the reference's cheb() and operation order
void ensure_coef(mlp_ctx ctx, double lmin, double lmax, uint32_t upto) {
    if (upto > kCoefCap) throw Error("Chebyshev iteration count beyond the coefficient table", MLP_ERR_MAXIT);
    if (!ctx->coef) ctx->coef = ctx->mem.alloc<double>(2 * size_t(kCoefCap));
    if (lmin != ctx->coef_lmin || lmax != ctx->coef_lmax) {
        ctx->coef_filled = 0;
        ctx->coef_lmin = lmin;
        ctx->coef_lmax = lmax;
    }
    if (upto <= ctx->coef_filled) return;
    const uint32_t from = ctx->coef_filled, to = std::min(kCoefCap, std::max(upto, from + 4096));
    const double eta = (lmax + lmin) / (lmax - lmin);
    std::vector<double> h(2 * size_t(to - from), 0.0);
    for (uint32_t k = std::max(from, 2u); k < to; k++) {
        h[2 * size_t(k - from)] = 4 / (lmax - lmin) * cheb_poly(eta, k - 1) / cheb_poly(eta, k);
        h[2 * size_t(k - from) + 1] = cheb_poly(eta, k - 2) / cheb_poly(eta, k);
    }
    MLP_HIP_CHECK(hipMemcpyAsync(ctx->coef + 2 * size_t(from), h.data(), h.size() * sizeof(double),
                                 hipMemcpyHostToDevice, ctx->stream));
    MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->coef_filled = to;
}

// Replays one outer iteration (variant v = pick(k) for the k-th iteration,
// 1-based) in chunks until the device sets `done`.  Variants are captured as
// hipGraphs and cached on the context.  before_chunk(k_last) runs on the host
// before a chunk whose last iteration is k_last.
template <class Iter, class Pick, class Before>
SolveState run_outer(mlp_ctx ctx, int solver, mlp_mat A, mlp_prec B, const void* b, const void* x, int nvar,
                     const mlp_solve_opts* o, Iter iteration, Pick pick, Before before_chunk) {
    hipStream_t s = ctx->stream;
    const uint32_t chunk = o->check_every ? o->check_every : 16;
    const bool use_graph = o->use_graph && B->impl->graph_safe() && graphs_allowed();
    mlp_ctx_s::GraphEntry* ge = nullptr;
    if (use_graph) {
        for (auto& e : ctx->graphs)
            if (e.solver == solver && e.A == A->id && e.B == B->id && e.b == b && e.x == x && e.ws_n == ctx->ws_n)
                ge = &e;
        if (!ge) {
            MLP_HIP_CHECK(hipStreamSynchronize(s));
            mlp_ctx_s::GraphEntry e;
            e.solver = solver; e.A = A->id; e.B = B->id; e.b = b; e.x = x; e.flex = 0; e.tre = 0; e.ws_n = ctx->ws_n;
            for (int v = 0; v < nvar; v++) {
                MLP_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                iteration(v);
                MLP_HIP_CHECK(hipStreamEndCapture(s, &e.g[v]));
                MLP_HIP_CHECK(hipGraphInstantiate(&e.e[v], e.g[v], nullptr, nullptr, 0));
            }
            ge = ctx->add_graph(e);
        }
    }
    SolveState st;
    uint32_t k = 0, n = chunk;
    while (true) {
        before_chunk(k + n);
        for (uint32_t c = 0; c < n; c++) {
            k++;
            const int v = pick(k);
            if (use_graph) MLP_HIP_CHECK(hipGraphLaunch(ge->e[v], s));
            else iteration(v);
        }
        MLP_HIP_CHECK(hipGetLastError());
        st = read_state(ctx);
        if (st.done) break;
        n = next_chunk(st, chunk);
    }
    return st;
}

void init_state(mlp_ctx ctx, double norm, double init_norm, double eps, uint32_t maxit, uint32_t niter,
                uint32_t maxit_offset) {
    SolveState st0;
    std::memset(&st0, 0, sizeof(st0));
    st0.norm = norm;
    st0.init_norm = init_norm;
    st0.eps = eps;
    st0.maxit = maxit;
    st0.niter = niter;
    st0.pad = maxit_offset;
    write_state(ctx, st0);
}

}  // namespace
}  // namespace mlp

extern "C" {

// PCGSolver (modules/solvers/pcg.cc:8-102) with device-resident scalars: a
// batch of one (mlp_pcg_batch), so the solve starts without host round trips.
// flexible: Polak-Ribiere beta and, after convergence of the recurrence
// residual, restarts from the FP64 true residual until it meets eps as well.
int mlp_pcg(mlp_ctx ctx, mlp_mat A, mlp_prec B, const double* b, double* x, const mlp_solve_opts* o,
            mlp_solve_stats* stats) {
    MLP_API_BEGIN
    require(ctx && A && B && b && x && o, "null argument");
    use_device(ctx);
    mlp_pcg_job J;
    std::memset(&J, 0, sizeof(J));
    J.ctx = ctx;
    J.A = A;
    J.B = B;
    J.d_b = b;
    J.d_x = x;
    const int rc = mlp_pcg_batch(1, &J, o);
    if (stats) *stats = J.stats;
    if (rc) return rc;
    MLP_API_END
}

// R PCG solves with one matrix and preconditioner (many right-hand sides):
// b and x hold R vectors of length n one after another.  Every launch of an
// iteration covers all R (grid.y), so the latency of the preconditioner's
// dependent launches is paid once per iteration for the batch.  Each solve
// runs mlp_pcg's kernels on its own vectors with its own state and stops on
// its own test; the batch iterates until all have stopped.  Flexible solves
// then take mlp_pcg's true-residual restarts one by one.  The preconditioner
// must be uh_cheb.
int mlp_pcg_multi(mlp_ctx ctx, mlp_mat A, mlp_prec B, uint32_t nrhs, const double* b, double* x,
                  const mlp_solve_opts* o, mlp_solve_stats* stats) {
    MLP_API_BEGIN
    require(ctx && A && B && b && x && o && stats, "null argument");
    require(nrhs >= 1 && nrhs <= kMaxRhs, "nrhs must lie in [1, 64]");
    require(B->n == A->n, "preconditioner size differs from A");
    require(B->ctx == ctx && A->ctx->device == ctx->device, "the preconditioner must live on the context");
    use_device(ctx);
    const bool flex = o->flexible != 0;
    const uint32_t tre = o->true_res_every;
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t n = A->n, rg = red_grid(n), R = nrhs;
    if (!B->wide || B->wide_R != R) {
        MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        ctx->drop_graphs(B->id);
        B->wide = B->impl->widen(ctx, R);
        B->wide_R = B->wide ? R : 0;
        if (!B->wide) throw Error("this preconditioner has no batched apply (uh_cheb has)", MLP_ERR_INVALID);
    }
    DevPrecBase* W = B->wide.get();
    ensure_workspace(ctx, n, false);
    const PcgVecs V = ensure_multi(ctx, n, R);
    hipStream_t s = ctx->stream;
    for (uint32_t k = 0; k < R; k++) fill_x0(ctx, n, o->x0_mode, x + size_t(k) * n);
    for (uint32_t k = 0; k < R; k++) {
        SolveState& st0 = ctx->mst_host[k];
        std::memset(&st0, 0, sizeof(st0));
        st0.eps = o->eps;
        st0.maxit = o->maxit;
    }
    MLP_HIP_CHECK(hipMemcpyAsync(V.st, ctx->mst_host, R * sizeof(SolveState), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_residual_batch, dim3(cdiv(n, kBS), R), dim3(kBS), 0, s, A->A.v, b, (const double*)x, V.r,
                       V.bt);
    hipLaunchKernelGGL(k_dot_batch<false>, dim3(rg, R), dim3(kBS), 0, s, n, (const double*)V.r, V.partials, V.bt);
    hipLaunchKernelGGL(k_finalize, dim3(1, R), dim3(kBS), 0, s, int(FIN_START), rg, (const double*)V.partials, V.st,
                       (double*)nullptr, V.bt);
    enqueue_pcg_start(ctx, W, n, flex, V);
    mlp_ctx_s::GraphEntry* ge =
        o->use_graph && W->graph_safe() && graphs_allowed() ? multi_graph(ctx, A, B, W, b, x, flex, tre, V)
                                                            : nullptr;
    const uint32_t chunk = o->check_every ? o->check_every : 16;
    uint32_t c = chunk, k = 0;
    while (true) {
        for (uint32_t i = 0; i < c; i++) {
            k++;
            const bool replace = tre && (k % tre == 0);
            if (ge) MLP_HIP_CHECK(hipGraphLaunch(ge->e[replace ? 1 : 0], s));
            else enqueue_pcg_iteration(ctx, A, W, b, x, flex, replace, V);
        }
        MLP_HIP_CHECK(hipGetLastError());
        MLP_HIP_CHECK(hipMemcpyAsync(ctx->mst_host, V.st, R * sizeof(SolveState), hipMemcpyDeviceToHost, s));
        MLP_HIP_CHECK(hipStreamSynchronize(s));
        bool all = true;
        uint32_t next = 1;
        for (uint32_t k = 0; k < R; k++) {
            const SolveState& st = ctx->mst_host[k];
            if (st.done) continue;
            all = false;
            next = std::max(next, next_chunk(st, chunk));
        }
        if (all) break;
        c = next;
    }
    std::vector<SolveState> fin(ctx->mst_host, ctx->mst_host + R);
    std::vector<double> trr(R, -1.0);          // a final x's true residual, when the restart test has it
    if (flex) {
        // mlp_pcg's restarts from the FP64 true residual, per right-hand side
        for (uint32_t q = 0; q < R; q++) {
            SolveState& st = fin[q];
            const double* bq = b + size_t(q) * n;
            double* xq = x + size_t(q) * n;
            for (uint32_t restart = 0; restart < kFlexRestarts && st.status == 0; restart++) {
                const double t = true_rel_res(ctx, A, bq, xq, ctx->r, st.init_norm);
                if (t <= st.eps) {
                    trr[q] = t;
                    break;
                }
                bool again = false;
                const double init_norm = st.init_norm, eps = st.eps;
                st = pcg_phase(ctx, A, B, bq, xq, o, init_norm, eps, st.niter, &again);
                st.init_norm = init_norm;
                st.eps = eps;
                if (again) break;
            }
        }
        MLP_HIP_CHECK(hipStreamSynchronize(s));
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    int first = MLP_OK;
    std::string msg;
    for (uint32_t k = 0; k < R; k++) {
        const SolveState& st = fin[k];
        mlp_solve_stats& S = stats[k];
        std::memset(&S, 0, sizeof(S));
        S.niter = st.niter;
        S.init_norm = st.init_norm;
        S.final_norm = st.norm;
        S.eps_used = st.eps;
        S.t_solve_ms = ms;
        S.true_rel_res = trr[k] >= 0.0 ? trr[k]
                                       : true_rel_res(ctx, A, b + size_t(k) * n, x + size_t(k) * n, ctx->r, st.init_norm);
        int code = MLP_OK;
        const char* what = "";
        if (st.status == 4) {
            code = MLP_ERR_NO_ITER;
            what = "No need to iterate";
        } else if (st.status == 3) {
            code = MLP_ERR_NOT_SPD;
            what = "<Ap, p> <= 0";
        } else if (st.status == 5) {
            code = MLP_ERR_MAXIT;
            what = "maximum number of iterations reached";
        } else if (flex && S.true_rel_res > st.eps) {
            code = MLP_ERR_NO_CONV;
            what = "flexible PCG: the FP64 true residual is above eps after the restarts";
        }
        S.status = code;
        if (code && first == MLP_OK) {
            first = code;
            msg = "right-hand side " + std::to_string(k) + ": " + what;
        }
    }
    if (first != MLP_OK) {
        set_last_error(msg.c_str());
        return first;
    }
    MLP_API_END
}

}  // extern "C"

namespace mlp {
namespace {

// One ensemble solve (mlp_pcg_ensemble) advanced in steps, so that the solves of several contexts can be driven
// round-robin from one host thread.
struct EnsRun {
    mlp_ctx ctx = nullptr;
    mlp_mat A = nullptr;
    mlp_prec B = nullptr;
    uint32_t S = 0;
    const uint32_t* seg_off = nullptr;
    const double* b = nullptr;
    double* x = nullptr;
    const mlp_solve_opts* o = nullptr;
    bool flex = false;
    uint32_t tre = 0, chunk = 16, gmax = 1, k = 0;
    PcgVecs V;
    mlp_ctx_s::GraphEntry* ge = nullptr;
    hipEvent_t ev = nullptr;
    std::chrono::steady_clock::time_point t0;

    ~EnsRun() {
        if (ev) (void)hipEventDestroy(ev);
    }
    void check() const {
        require(ctx && A && B && b && x && o && seg_off, "null argument");
        require(S >= 1 && S <= kMaxSegments, "nseg must lie in [1, 256]");
        require(seg_off[0] == 0 && seg_off[S] == A->n, "segments must cover the matrix rows");
        for (uint32_t y = 0; y < S; y++)
            require(seg_off[y + 1] > seg_off[y], "segments must be non-empty and ascending");
        require(B->n == A->n, "preconditioner size differs from A");
        require(B->ctx == ctx && A->ctx->device == ctx->device, "the preconditioner must live on the context");
    }
    // x0, the initial norms and eps floors, z = B r and p = z for every segment; the iteration graphs
    void start() {
        use_device(ctx);
        t0 = std::chrono::steady_clock::now();
        flex = o->flexible != 0;
        tre = o->true_res_every;
        chunk = o->check_every ? o->check_every : 16;
        hipStream_t s = ctx->stream;
        V = ensure_ens(ctx, A->n, S, seg_off);
        gmax = 1;
        for (uint32_t y = 0; y < S; y++) gmax = std::max(gmax, red_grid(seg_off[y + 1] - seg_off[y]));
        for (uint32_t y = 0; y < S; y++) fill_x0(ctx, seg_off[y + 1] - seg_off[y], o->x0_mode, x + seg_off[y]);
        for (uint32_t y = 0; y < S; y++) {
            SolveState& st0 = ctx->est_host[y];
            std::memset(&st0, 0, sizeof(st0));
            st0.eps = o->eps;
            st0.maxit = o->maxit;
        }
        MLP_HIP_CHECK(hipMemcpyAsync(V.st, ctx->est_host, S * sizeof(SolveState), hipMemcpyHostToDevice, s));
        enqueue_ens_norms(ctx, A, b, x, V, gmax, FIN_START, nullptr);
        enqueue_ens_start(ctx, B->impl.get(), flex, V, gmax);
        const bool graphs = o->use_graph && B->impl->graph_safe() && graphs_allowed();
        ge = graphs ? ens_graph(ctx, A, B, b, x, flex, tre, V, gmax) : nullptr;
        if (!ev) MLP_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        k = 0;
    }
    // c iterations, then the states copied to the host behind an event
    void enqueue(uint32_t c) {
        hipStream_t s = ctx->stream;
        for (uint32_t i = 0; i < c; i++) {
            k++;
            const bool replace = tre && (k % tre == 0);
            if (ge) MLP_HIP_CHECK(hipGraphLaunch(ge->e[replace ? 1 : 0], s));
            else enqueue_ens_iteration(ctx, A, B->impl.get(), b, x, flex, replace, V, gmax);
        }
        MLP_HIP_CHECK(hipGetLastError());
        MLP_HIP_CHECK(hipMemcpyAsync(ctx->est_host, V.st, S * sizeof(SolveState), hipMemcpyDeviceToHost, s));
        MLP_HIP_CHECK(hipEventRecord(ev, s));
    }
    // after the event: true when every segment is done, else the next chunk is enqueued
    bool advance() {
        bool all = true;
        uint32_t next = 1;
        for (uint32_t y = 0; y < S; y++) {
            const SolveState& st = ctx->est_host[y];
            if (st.done) continue;
            all = false;
            next = std::max(next, next_chunk(st, chunk));
        }
        if (!all) enqueue(next);
        return all;
    }
    // true when the enqueued work has finished and every segment is done
    bool poll() {
        const hipError_t q = hipEventQuery(ev);
        if (q == hipErrorNotReady) return false;
        MLP_HIP_CHECK(q);
        return advance();
    }
    void run_sync() {
        enqueue(chunk);
        while (true) {
            MLP_HIP_CHECK(hipEventSynchronize(ev));
            if (advance()) return;
        }
    }
    // FP64 true relative residuals ||b - A x|| / init_norm of every segment
    void true_rel(const std::vector<SolveState>& fin, std::vector<double>& tr) {
        hipStream_t s = ctx->stream;
        enqueue_ens_norms(ctx, A, b, x, V, gmax, FIN_DOT, ctx->eout);
        MLP_HIP_CHECK(hipMemcpyAsync(ctx->eout_host, ctx->eout, S * sizeof(double), hipMemcpyDeviceToHost, s));
        MLP_HIP_CHECK(hipStreamSynchronize(s));
        tr.resize(S);
        for (uint32_t y = 0; y < S; y++) {
            const double nr = std::sqrt(ctx->eout_host[y]);
            tr[y] = fin[y].init_norm > 0 ? nr / fin[y].init_norm : nr;
        }
    }
    // flexible restarts, the stats; returns the first failing segment's code (message in *msg)
    int finish(mlp_solve_stats* stats, std::string* msg) {
        use_device(ctx);
        hipStream_t s = ctx->stream;
        std::vector<SolveState> fin(ctx->est_host, ctx->est_host + S);
        std::vector<double> tr;
        if (flex) {
            // mlp_pcg's restarts from the FP64 true residual (pcg_phase), every segment that needs one at once
            std::vector<char> stop(S, 0);
            for (uint32_t y = 0; y < S; y++) stop[y] = fin[y].status != 0;
            for (uint32_t restart = 0; restart < kFlexRestarts; restart++) {
                true_rel(fin, tr);
                bool need = false;
                for (uint32_t y = 0; y < S; y++) {
                    if (!stop[y] && tr[y] <= fin[y].eps) stop[y] = 1;
                    SolveState& st0 = ctx->est_host[y];              // pcg_phase's starting state
                    std::memset(&st0, 0, sizeof(st0));
                    st0.init_norm = fin[y].init_norm;
                    st0.eps = fin[y].eps;
                    st0.maxit = o->maxit;
                    st0.niter = fin[y].niter;
                    st0.done = stop[y] ? 1u : 0u;
                    need = need || !stop[y];
                }
                if (!need) break;
                MLP_HIP_CHECK(hipMemcpyAsync(V.st, ctx->est_host, S * sizeof(SolveState), hipMemcpyHostToDevice, s));
                enqueue_ens_norms(ctx, A, b, x, V, gmax, FIN_RESTART, nullptr);
                // a segment whose restarted residual already meets eps stops there (no iteration, no more restarts)
                MLP_HIP_CHECK(hipMemcpyAsync(ctx->est_host, V.st, S * sizeof(SolveState), hipMemcpyDeviceToHost, s));
                MLP_HIP_CHECK(hipStreamSynchronize(s));
                bool go = false;
                for (uint32_t y = 0; y < S; y++) {
                    if (stop[y]) continue;
                    if (ctx->est_host[y].done) {
                        fin[y] = ctx->est_host[y];
                        fin[y].done = 1;
                        stop[y] = 1;
                    } else {
                        go = true;
                    }
                }
                if (!go) break;
                enqueue_ens_start(ctx, B->impl.get(), flex, V, gmax);
                run_sync();
                for (uint32_t y = 0; y < S; y++)
                    if (!stop[y]) {
                        fin[y] = ctx->est_host[y];
                        if (fin[y].status != 0) stop[y] = 1;
                    }
            }
        }
        true_rel(fin, tr);
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        int first = MLP_OK;
        for (uint32_t y = 0; y < S; y++) {
            const SolveState& st = fin[y];
            mlp_solve_stats& St = stats[y];
            std::memset(&St, 0, sizeof(St));
            St.niter = st.niter;
            St.init_norm = st.init_norm;
            St.final_norm = st.norm;
            St.eps_used = st.eps;
            St.t_solve_ms = ms;
            St.true_rel_res = tr[y];
            int code = MLP_OK;
            const char* what = "";
            if (st.status == 4) {
                code = MLP_ERR_NO_ITER;
                what = "No need to iterate";
            } else if (st.status == 3) {
                code = MLP_ERR_NOT_SPD;
                what = "<Ap, p> <= 0";
            } else if (st.status == 5) {
                code = MLP_ERR_MAXIT;
                what = "maximum number of iterations reached";
            } else if (flex && St.true_rel_res > st.eps) {
                code = MLP_ERR_NO_CONV;
                what = "flexible PCG: the FP64 true residual is above eps after the restarts";
            }
            St.status = code;
            if (code && first == MLP_OK) {
                first = code;
                *msg = "segment " + std::to_string(y) + ": " + what;
            }
        }
        return first;
    }
};

}  // namespace
}  // namespace mlp

extern "C" {

// An ensemble of independent systems given as one block-diagonal system (segment y = rows
// [seg_off[y], seg_off[y+1]), no entries coupling two segments), with one preconditioner for the whole of it.
// Every segment runs PCGSolver's steps (pcg.cc:8-102) on its own rows with its own state, reductions of its
// own grid and its own stop test, as mlp_pcg on that system alone would; every launch covers all segments,
// the preconditioner's launches included.  Flexible solves then take mlp_pcg's true-residual restarts, all
// segments that need one in one batched phase per round.
int mlp_pcg_ensemble(mlp_ctx ctx, mlp_mat A, mlp_prec B, uint32_t nseg, const uint32_t* seg_off, const double* b,
                     double* x, const mlp_solve_opts* o, mlp_solve_stats* stats) {
    MLP_API_BEGIN
    require(stats, "null argument");
    EnsRun r;
    r.ctx = ctx; r.A = A; r.B = B; r.S = nseg; r.seg_off = seg_off; r.b = b; r.x = x; r.o = o;
    r.check();
    r.start();
    r.run_sync();
    std::string msg;
    const int rc = r.finish(stats, &msg);
    if (rc != MLP_OK) {
        set_last_error(msg);
        return rc;
    }
    MLP_API_END
}

// Several ensembles (each its own context and stream) driven from this thread: their iteration chunks are
// enqueued round-robin and each one's states polled through an event.  Every job's results are those of
// mlp_pcg_ensemble on it alone.
int mlp_pcg_ensemble_batch(uint32_t count, mlp_ens_job* jobs, const mlp_solve_opts* o) {
    MLP_API_BEGIN
    require(o && (jobs || !count), "null argument");
    std::vector<EnsRun> runs(count);
    std::vector<mlp_ctx> seen;
    for (uint32_t i = 0; i < count; i++) {
        mlp_ens_job& J = jobs[i];
        require(J.stats, "null job argument");
        EnsRun& r = runs[i];
        r.ctx = J.ctx; r.A = J.A; r.B = J.B; r.S = J.nseg; r.seg_off = J.seg_off; r.b = J.d_b; r.x = J.d_x; r.o = o;
        r.check();
        for (mlp_ctx c : seen) require(c != J.ctx, "two jobs share a context");
        seen.push_back(J.ctx);
        J.status = MLP_OK;
    }
    for (EnsRun& r : runs) {
        r.start();
        r.enqueue(r.chunk);
    }
    std::vector<char> active(count, 1);
    uint32_t left = count;
    while (left) {
        bool progressed = false;
        for (uint32_t i = 0; i < count; i++) {
            if (!active[i]) continue;
            use_device(runs[i].ctx);
            const hipError_t q = hipEventQuery(runs[i].ev);
            if (q == hipErrorNotReady) continue;
            MLP_HIP_CHECK(q);
            progressed = true;
            if (runs[i].advance()) {
                active[i] = 0;
                left--;
            }
        }
        if (!progressed) std::this_thread::yield();
    }
    int first = MLP_OK;
    for (uint32_t i = 0; i < count; i++) {
        std::string msg;
        const int rc = runs[i].finish(jobs[i].stats, &msg);
        jobs[i].status = rc;
        if (rc != MLP_OK && first == MLP_OK) {
            first = rc;
            set_last_error("job " + std::to_string(i) + ", " + msg);
        }
    }
    return first;
    MLP_API_END
}

// Many independent PCG solves at once (ensembles of systems).  Each
// job has its own context, so its own stream, workspace and device state; the
// solves start without host synchronisation (FIN_START computes the initial
// norm and the eps floor on the device) and their iteration chunks are
// enqueued round-robin, each job's state being polled through an event.  Per
// job the kernels and their order are those of mlp_pcg, so every result is
// bitwise the one a single mlp_pcg call on that job gives.
int mlp_pcg_batch(uint32_t count, mlp_pcg_job* jobs, const mlp_solve_opts* o) {
    MLP_API_BEGIN
    require(o && (jobs || !count), "null argument");
    struct Run {
        mlp_pcg_job* j;
        hipEvent_t ev = nullptr;
        mlp_ctx_s::GraphEntry* ge = nullptr;
        uint32_t k = 0, c = 0;
        bool active = false;
        std::chrono::steady_clock::time_point t0;
    };
    std::vector<Run> runs(count);
    std::vector<mlp_ctx> seen;
    const bool flex = o->flexible != 0;
    const uint32_t chunk = o->check_every ? o->check_every : 16;
    const uint32_t tre = o->true_res_every;
    for (uint32_t i = 0; i < count; i++) {
        mlp_pcg_job& J = jobs[i];
        require(J.ctx && J.A && J.B && J.d_b && J.d_x, "null job argument");
        require(J.B->n == J.A->n, "preconditioner size differs from A");
        require(J.B->ctx == J.ctx, "a job's preconditioner must live on its context (mlp_prec_clone)");
        require(J.A->ctx->device == J.ctx->device, "a job's matrix must be on the device of its context");
        for (mlp_ctx c : seen) require(c != J.ctx, "two jobs share a context");
        seen.push_back(J.ctx);
        std::memset(&J.stats, 0, sizeof(J.stats));
        J.status = MLP_OK;
    }
    for (uint32_t i = 0; i < count; i++) {
        Run& R = runs[i];
        R.j = &jobs[i];
        mlp_ctx ctx = R.j->ctx;
        use_device(ctx);
        mlp_mat A = R.j->A;
        const uint32_t n = A->n, rg = red_grid(n);
        R.t0 = std::chrono::steady_clock::now();
        ensure_workspace(ctx, n, false);
        fill_x0(ctx, n, o->x0_mode, R.j->d_x);
        SolveState st0;
        std::memset(&st0, 0, sizeof(st0));
        st0.eps = o->eps;
        st0.maxit = o->maxit;
        std::memcpy(ctx->host_pinned, &st0, sizeof(st0));
        MLP_HIP_CHECK(hipMemcpyAsync(ctx->st, ctx->host_pinned, sizeof(SolveState), hipMemcpyHostToDevice, ctx->stream));
        launch_residual(ctx, A, R.j->d_b, R.j->d_x, ctx->r);
        hipLaunchKernelGGL(k_dot, dim3(rg), dim3(kBS), 0, ctx->stream, n, (const double*)ctx->r, (const double*)ctx->r,
                           ctx->partials);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBS), 0, ctx->stream, int(FIN_START), rg,
                           (const double*)ctx->partials, ctx->st, (double*)nullptr, Batch{});
        enqueue_pcg_start(ctx, R.j->B, n, flex);
        R.ge = pcg_graphs(ctx, A, R.j->B, R.j->d_b, R.j->d_x, flex, tre, o->use_graph && R.j->B->impl->graph_safe());
        MLP_HIP_CHECK(hipEventCreateWithFlags(&R.ev, hipEventDisableTiming));
        R.active = true;
    }
    auto enqueue = [&](Run& R) {
        mlp_ctx ctx = R.j->ctx;
        const uint32_t c = R.c ? R.c : chunk;
        enqueue_pcg_chunk(ctx, R.j->A, R.j->B, R.j->d_b, R.j->d_x, flex, tre, R.ge, R.k, c, ctx_vecs(ctx, R.j->A->n));
        R.k += c;
        MLP_HIP_CHECK(hipMemcpyAsync(ctx->host_pinned, ctx->st, sizeof(SolveState), hipMemcpyDeviceToHost,
                                     ctx->stream));
        MLP_HIP_CHECK(hipEventRecord(R.ev, ctx->stream));
    };
    for (Run& R : runs) enqueue(R);
    uint32_t left = count;
    while (left) {
        bool progressed = false;
        for (Run& R : runs) {
            if (!R.active) continue;
            const hipError_t q = hipEventQuery(R.ev);
            if (q == hipErrorNotReady) continue;
            MLP_HIP_CHECK(q);
            SolveState st;
            std::memcpy(&st, R.j->ctx->host_pinned, sizeof(st));
            progressed = true;
            if (st.done) {
                R.active = false;
                left--;
            } else {
                R.c = next_chunk(st, chunk);
                enqueue(R);
            }
        }
        if (!progressed) std::this_thread::yield();
    }
    int first = MLP_OK;
    for (Run& R : runs) {
        (void)hipEventDestroy(R.ev);
        mlp_pcg_job& J = *R.j;
        mlp_ctx ctx = J.ctx;
        try {
            SolveState st = read_state(ctx);
            const double init_norm = st.init_norm, eps = st.eps;
            if (st.status == 4) {
                J.stats.init_norm = init_norm;
                J.stats.status = MLP_ERR_NO_ITER;
                throw Error("No need to iterate", MLP_ERR_NO_ITER);
            }
            double trr = -1.0;                    // the final x's true residual, when the restart test has it
            if (flex) {
                for (uint32_t restart = 0; restart < kFlexRestarts && st.status == 0; restart++) {
                    const double t = true_rel_res(ctx, J.A, J.d_b, J.d_x, ctx->r, init_norm);
                    if (t <= eps) {
                        trr = t;
                        break;
                    }
                    bool again = false;
                    st = pcg_phase(ctx, J.A, J.B, J.d_b, J.d_x, o, init_norm, eps, st.niter, &again);
                    if (again) break;
                }
            }
            MLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            J.stats.niter = st.niter;
            J.stats.init_norm = init_norm;
            J.stats.final_norm = st.norm;
            J.stats.eps_used = eps;
            J.stats.t_solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - R.t0).count();
            J.stats.status = int(st.status);
            J.stats.true_rel_res = trr >= 0.0 ? trr : true_rel_res(ctx, J.A, J.d_b, J.d_x, ctx->r, init_norm);
            if (st.status == 3) {
                char buf[160];
                std::snprintf(buf, sizeof(buf), "<Ap, p> = %g", st.app);
                throw Error(buf, MLP_ERR_NOT_SPD);
            }
            if (st.status == 5) throw Error("maximum number of iterations reached", MLP_ERR_MAXIT);
            if (flex && J.stats.true_rel_res > eps) {
                J.stats.status = MLP_ERR_NO_CONV;
                char buf[200];
                std::snprintf(buf, sizeof(buf),
                              "flexible PCG: FP64 true relative residual %g is above eps %g after %u restarts",
                              J.stats.true_rel_res, eps, kFlexRestarts);
                throw Error(buf, MLP_ERR_NO_CONV);
            }
        } catch (const Error& e) {
            J.status = e.code();
            if (first == MLP_OK) {
                first = e.code();
                set_last_error(e.what());
            }
        }
    }
    return first;
    MLP_API_END
}

// Gradient of lam^T A(theta) p with respect to theta = ln kx for the
// two-point-flux assembly of modules/mesh/spe.cc:69-131 (see k_spe_grad):
// with lam the adjoint solution of A lam = dL/dp this is -dL/dtheta.
int mlp_spe_adjoint_grad(mlp_ctx ctx, uint32_t nx, uint32_t ny, uint32_t nz, double fx, double fy, double fz,
                         double kz_ratio, const double* kx, const double* p, const double* lam, double* grad) {
    MLP_API_BEGIN
    require(ctx && kx && p && lam && grad, "null argument");
    use_device(ctx);
    const uint64_t n = uint64_t(nx) * ny * nz;
    require(n > 0 && n < (1ull << 32), "grid size out of range");
    require(fx > 0 && fy > 0 && fz > 0 && kz_ratio > 0, "face factors must be positive");
    hipLaunchKernelGGL(k_spe_grad, dim3(cdiv(uint32_t(n), kBS)), dim3(kBS), 0, ctx->stream, nx, ny, nz, fx, fy, fz,
                       kz_ratio, kx, p, lam, grad);
    MLP_HIP_CHECK(hipGetLastError());
    MLP_API_END
}

// GMRESSolver (modules/solvers/gmres.cc:13-168): restart 30, classical
// Gram-Schmidt on the device, Givens rotations and the triangular solve on the
// host in FP64 exactly as the reference orders them (including its
// back-substitution bound j < n in update_x, gmres.cc:156-159).
int mlp_gmres(mlp_ctx ctx, mlp_mat A, mlp_prec B, const double* b, double* x, const mlp_solve_opts* o,
              mlp_solve_stats* stats) {
    MLP_API_BEGIN
    require(ctx && A && B && b && x && o, "null argument");
    require(B->n == A->n, "preconditioner size differs from A");
    use_device(ctx);
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t n = A->n, m = kGmresRestart;
    ensure_workspace(ctx, n, true);
    hipStream_t s = ctx->stream;
    const size_t ld = ctx->ws_n;
    auto v = [&](uint32_t k) { return ctx->V + k * ld; };
    double cs[kGmresRestart], sn[kGmresRestart], rs[kGmresRestart];
    std::vector<double> H((m + 1) * m, 0.0);
    auto hh = [&](uint32_t r, uint32_t c) -> double& { return H[size_t(r) * m + c]; };
    double* z = ctx->z;
    double* tmp = ctx->p;
    fill_x0(ctx, n, o->x0_mode, x);
    launch_residual(ctx, A, b, x, v(0));
    double norm = std::sqrt(dev_dot(ctx, n, v(0), v(0)));
    const double init_norm = norm;
    double eps = o->eps;
    if (init_norm * eps < 1e-14) eps = 1e-14 / init_norm;
    uint32_t niter = 0;
    int status = 0;
    const uint32_t rg = red_grid(n);
    auto scal = [&](double a, double* d) {
        if (!host_is_one(a)) hipLaunchKernelGGL(k_scal, dim3(rg), dim3(kBS), 0, s, n, a, d);
    };
    // z = B * sum_{j<=nn} rs_j v_j ; x += z   (update_x, gmres.cc:144-168)
    auto update_x = [&](uint32_t nn) {
        if (nn == 0) return;
        nn--;
        rs[nn] /= hh(nn, nn);
        for (uint32_t i = 1; i <= nn; i++) {
            const int k = int(nn - i);
            for (uint32_t j = uint32_t(k) + 1; j < nn; j++) rs[k] -= hh(k, j) * rs[j];
            rs[k] /= hh(k, k);
        }
        MaxpyCoef cf;
        cf.nk = nn + 1;
        for (uint32_t j = 0; j <= nn; j++) { cf.c[j] = rs[j]; cf.one[j] = host_is_one(rs[j]); }
        hipLaunchKernelGGL(k_maxpy, dim3(rg), dim3(kBS), 0, s, n, (const double*)ctx->V, ld, cf, tmp, 1);
        B->impl->apply(tmp, z, nullptr);
        hipLaunchKernelGGL(k_axpy, dim3(rg), dim3(kBS), 0, s, n, 1.0, 1, (const double*)z, x);
    };
    while (true) {
        if (norm / init_norm <= eps) break;
        scal(1. / norm, v(0));
        rs[0] = norm;
        uint32_t i = 0;
        for (; i < m - 1;) {
            B->impl->apply(v(i), z, nullptr);
            launch_rows(A->A.v, nullptr, z, v(i + 1), s);
            // Gram-Schmidt column and the new vector's norm stay on the device until
            // one copy brings h_0..h_i and (v_{i+1}, v_{i+1}) to the host
            hipLaunchKernelGGL(k_mdot, dim3(mdot_grid(n)), dim3(kBS), 0, s, n, i + 1, (const double*)ctx->V, ld,
                               (const double*)v(i + 1), ctx->partials);
            hipLaunchKernelGGL(k_mdot_finalize, dim3(i + 1), dim3(kBS), 0, s, mdot_grid(n), i + 1,
                               (const double*)ctx->partials, ctx->mdot);
            hipLaunchKernelGGL(k_cgs_update, dim3(rg), dim3(kBS), 0, s, n, (const double*)ctx->V, ld,
                               (const double*)ctx->mdot, i + 1, v(i + 1), ctx->partials);
            hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBS), 0, s, int(FIN_DOT), rg, (const double*)ctx->partials,
                               ctx->st, ctx->mdot + i + 1, Batch{});
            MLP_HIP_CHECK(hipMemcpyAsync(ctx->host_pinned, ctx->mdot, (i + 2) * sizeof(double), hipMemcpyDeviceToHost, s));
            MLP_HIP_CHECK(hipStreamSynchronize(s));
            for (uint32_t k = 0; k <= i; k++) hh(k, i) = ctx->host_pinned[k];
            hh(i + 1, i) = std::sqrt(ctx->host_pinned[i + 1]);
            scal(1. / hh(i + 1, i), v(i + 1));
            for (uint32_t k = 1; k <= i; k++) {
                const uint32_t k1 = k - 1;
                const double d = hh(k1, i);
                hh(k1, i) = cs[k1] * d + sn[k1] * hh(k, i);
                hh(k, i) = -sn[k1] * d + cs[k1] * hh(k, i);
            }
            double d = std::sqrt(hh(i, i) * hh(i, i) + hh(i + 1, i) * hh(i + 1, i));
            d = 1.0 / d;
            cs[i] = hh(i, i) * d;
            sn[i] = hh(i + 1, i) * d;
            hh(i, i) = cs[i] * hh(i, i) + sn[i] * hh(i + 1, i);
            rs[i + 1] = -sn[i] * rs[i];
            rs[i] *= cs[i];
            norm = std::fabs(rs[i + 1]);
            niter++;
            i++;
            if (norm / init_norm <= eps) break;
        }
        update_x(i);
        if (norm / init_norm <= eps) break;
        if (o->maxit && niter >= o->maxit) { status = MLP_ERR_MAXIT; break; }
        launch_residual(ctx, A, b, x, v(0));
    }
    MLP_HIP_CHECK(hipStreamSynchronize(s));
    const auto t1 = std::chrono::steady_clock::now();
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->niter = niter;
        stats->init_norm = init_norm;
        stats->final_norm = norm;
        stats->eps_used = eps;
        stats->t_solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        stats->status = status;
        stats->true_rel_res = true_rel_res(ctx, A, b, x, ctx->r, init_norm);
    }
    if (status == MLP_ERR_MAXIT) throw Error("maximum number of iterations reached", MLP_ERR_MAXIT);
    MLP_API_END
}


// ChebSolver (modules/solvers/cheb.cc:9-100): Chebyshev outer iteration on
// [lmin, lmax] (Prec::lmin()/lmax(), cheb_prec.h:122-123) with the
// preconditioner; stops on ||r|| / ||r0|| <= eps.  The reported niter is the
// reference's (loop count + 1, cheb.cc:28, 47, 77, 81).
int mlp_cheb_solver(mlp_ctx ctx, mlp_mat A, mlp_prec B, double lmin, double lmax, const double* b, double* x,
                    const mlp_solve_opts* o, mlp_solve_stats* stats) {
    MLP_API_BEGIN
    require(ctx && A && B && b && x && o, "null argument");
    require(B->n == A->n, "preconditioner size differs from A");
    require(lmax > lmin && lmin > 0, "Chebyshev bounds need 0 < lmin < lmax");
    use_device(ctx);
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t n = A->n;
    hipStream_t s = ctx->stream;
    ensure_workspace(ctx, n, false);
    fill_x0(ctx, n, o->x0_mode, x);
    launch_residual(ctx, A, b, x, ctx->r);
    const double init_norm = std::sqrt(dev_dot(ctx, n, ctx->r, ctx->r));
    double eps = o->eps;
    if (init_norm * eps < 1e-14) eps = 1e-14 / init_norm;      // check_and_replace_eps, solvers.cc:49-58
    if (stats) std::memset(stats, 0, sizeof(*stats));
    // step 1 (cheb.cc:37-47): u1 = x; x = alpha B r + u1; r = b - A x
    double* bufs[3] = {x, ctx->z, ctx->p};   // x, then the two rotating work vectors
    const double alpha = 2 / (lmax + lmin);
    const uint32_t rg = red_grid(n);
    MLP_HIP_CHECK(hipMemcpyAsync(bufs[1], x, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, s));
    B->impl->apply(ctx->r, x, nullptr);
    hipLaunchKernelGGL(k_cheb_first, dim3(rg), dim3(kBS), 0, s, n, x, (const double*)bufs[1], alpha,
                       int(host_is_one(alpha)));
    launch_residual(ctx, A, b, x, ctx->r);
    const double norm1 = std::sqrt(dev_dot(ctx, n, ctx->r, ctx->r));
    // SolveState.niter holds the reference's niter (2 after step 1); maxit counts niter - 1
    init_state(ctx, norm1, init_norm, eps, o->maxit, 2, 1);
    SolveState st;
    if (!(norm1 / init_norm > eps)) {
        st = read_state(ctx);
        st.done = 1;
    } else if (o->maxit && 1 >= o->maxit) {
        st = read_state(ctx);
        st.status = 5;
    } else {
        // loop iteration k (1-based) uses roles (X, U0, U1) = rotation (k-1) mod 3 of (bufs[2], bufs[1], bufs[0])
        auto iteration = [&](int v) {
            double* X = bufs[(2 + 2 * v) % 3];
            double* U0 = bufs[(1 + 2 * v) % 3];
            double* U1 = bufs[(0 + 2 * v) % 3];
            B->impl->apply(ctx->r, X, &ctx->st->done);
            hipLaunchKernelGGL(k_cheb_outer, dim3(rg), dim3(kBS), 0, s, n, X, (const double*)U1, (const double*)U0,
                               (const double*)ctx->coef, (const SolveState*)ctx->st);
            hipLaunchKernelGGL(k_pcg_true_res<false>, dim3(rg), dim3(kBS), 0, s, A->A.v, b, (const double*)X, ctx->r,
                               (const SolveState*)ctx->st, ctx->partials, Batch{});
            hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBS), 0, s, int(FIN_STEP), rg, (const double*)ctx->partials,
                               ctx->st, (double*)nullptr, Batch{});
        };
        ensure_coef(ctx, lmin, lmax, 2);
        st = run_outer(ctx, 1, A, B, b, x, 3, o, iteration, [](uint32_t k) { return int((k - 1) % 3); },
                       [&](uint32_t k_last) { ensure_coef(ctx, lmin, lmax, k_last + 3); });
        const uint32_t loops = st.niter - 2;
        if (loops) {
            double* X = bufs[(2 + 2 * ((loops - 1) % 3)) % 3];
            if (X != x) MLP_HIP_CHECK(hipMemcpyAsync(x, X, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, s));
        }
    }
    MLP_HIP_CHECK(hipStreamSynchronize(s));
    const auto t1 = std::chrono::steady_clock::now();
    if (stats) {
        stats->niter = st.niter - 1;                  // cheb.cc:81
        stats->init_norm = init_norm;
        stats->final_norm = st.norm;
        stats->eps_used = eps;
        stats->t_solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        stats->status = int(st.status);
        stats->true_rel_res = true_rel_res(ctx, A, b, x, ctx->r, init_norm);
    }
    if (st.status == 5) throw Error("maximum number of iterations reached", MLP_ERR_MAXIT);
    MLP_API_END
}

// SimpleSolver (modules/solvers/simple_solver.cc:8-88): x += B (b - A x)
// until ||r|| / ||r0|| <= eps; like the reference it reports "No need to
// iterate" when fewer than two iterations ran (x is updated all the same).
int mlp_simple_solver(mlp_ctx ctx, mlp_mat A, mlp_prec B, const double* b, double* x, const mlp_solve_opts* o,
                      mlp_solve_stats* stats) {
    MLP_API_BEGIN
    require(ctx && A && B && b && x && o, "null argument");
    require(B->n == A->n, "preconditioner size differs from A");
    use_device(ctx);
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t n = A->n;
    hipStream_t s = ctx->stream;
    ensure_workspace(ctx, n, false);
    fill_x0(ctx, n, o->x0_mode, x);
    launch_residual(ctx, A, b, x, ctx->r);
    const double init_norm = std::sqrt(dev_dot(ctx, n, ctx->r, ctx->r));
    double eps = o->eps;
    if (init_norm * eps < 1e-14) eps = 1e-14 / init_norm;      // check_and_replace_eps, solvers.cc:49-58
    if (stats) std::memset(stats, 0, sizeof(*stats));
    init_state(ctx, init_norm, init_norm, eps, o->maxit, 0, 0);
    SolveState st;
    if (!(init_norm / init_norm > eps)) {
        st = read_state(ctx);
    } else {
        const uint32_t rg = red_grid(n);
        auto iteration = [&](int) {
            B->impl->apply(ctx->r, ctx->z, &ctx->st->done);
            hipLaunchKernelGGL(k_add, dim3(rg), dim3(kBS), 0, s, n, (const double*)ctx->z, x, (const SolveState*)ctx->st);
            hipLaunchKernelGGL(k_pcg_true_res<false>, dim3(rg), dim3(kBS), 0, s, A->A.v, b, (const double*)x, ctx->r,
                               (const SolveState*)ctx->st, ctx->partials, Batch{});
            hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBS), 0, s, int(FIN_STEP), rg, (const double*)ctx->partials,
                               ctx->st, (double*)nullptr, Batch{});
        };
        st = run_outer(ctx, 2, A, B, b, x, 1, o, iteration, [](uint32_t) { return 0; }, [](uint32_t) {});
    }
    MLP_HIP_CHECK(hipStreamSynchronize(s));
    const auto t1 = std::chrono::steady_clock::now();
    if (stats) {
        stats->niter = st.niter;
        stats->init_norm = init_norm;
        stats->final_norm = st.norm;
        stats->eps_used = eps;
        stats->t_solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        stats->status = int(st.status);
        stats->true_rel_res = true_rel_res(ctx, A, b, x, ctx->r, init_norm);
    }
    if (st.status == 5) throw Error("maximum number of iterations reached", MLP_ERR_MAXIT);
    if (st.niter < 2) {
        if (stats) stats->status = MLP_ERR_NO_ITER;
        throw Error("No need to iterate", MLP_ERR_NO_ITER);
    }
    MLP_API_END
}

}  // extern "C"

// ---------------------------------------------------------------------------
// allocation-free PCG (no counterpart in the reference): mlp_pcg with its
// workspace in a static device array instead of context allocations
// ---------------------------------------------------------------------------
namespace mlp {
namespace {
constexpr uint32_t kNoAllocRows = (1u << 20) + (1u << 17);     // 1,179,648 rows
constexpr size_t kNoAllocHead = 32;                            // doubles (256 bytes) per small slot
constexpr size_t kNoAllocDoubles = kNoAllocHead               // SolveState
                                   + kPartials                // reduction partial sums
                                   + kNoAllocHead             // result of a standalone dot
                                   + 5 * size_t(kNoAllocRows); // r, z, p, Ap, zp
static_assert(sizeof(SolveState) <= kNoAllocHead * sizeof(double), "solve state slot too small");
static_assert((kNoAllocRows * sizeof(double)) % 256 == 0, "vectors keep 256-byte alignment");
}  // namespace
}  // namespace mlp

// The workspace of mlp_pcg_noalloc: 47.4 MB per GPU, held from the library's
// load until the process ends.  Host code never uses this name as an address:
// it takes the device address from hipGetSymbolAddress once per device.
static __device__ __align__(256) double scratch[mlp::kNoAllocDoubles];

namespace mlp {
namespace {
struct NoAllocDevice {
    std::once_flag once;
    hipError_t rc = hipSuccess;
    double* ptr = nullptr;
};
NoAllocDevice g_noalloc[kMaxStageDevices];

// The device address of `scratch` on `device` (which the caller has made
// current), looked up on the first call for that device and kept.
double* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxStageDevices) throw Error("device index out of range", MLP_ERR_INVALID);
    NoAllocDevice& d = g_noalloc[device];
    std::call_once(d.once, [&d] { d.rc = hipGetSymbolAddress((void**)&d.ptr, HIP_SYMBOL(scratch)); });
    if (d.rc != hipSuccess || !d.ptr)
        throw Error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(d.rc), MLP_ERR_HIP);
    return d.ptr;
}

// The PCG workspace carved out of the device's scratch for an n-row solve.
PcgVecs noalloc_vecs(double* s, uint32_t n) {
    PcgVecs V;
    V.st = reinterpret_cast<SolveState*>(s);
    V.partials = s + kNoAllocHead;
    V.slot = V.partials + kPartials;
    double* v = V.slot + kNoAllocHead;
    V.r = v;
    V.z = v + kNoAllocRows;
    V.p = v + 2 * size_t(kNoAllocRows);
    V.Ap = v + 3 * size_t(kNoAllocRows);
    V.zp = v + 4 * size_t(kNoAllocRows);
    V.bt.n = n;
    V.bt.part = uint32_t(kPartials);
    return V;
}
}  // namespace
}  // namespace mlp

extern "C" {

// mlp_pcg without allocations: the same kernels in the same order as mlp_pcg
// (so the same x, iteration count and stats), as plain launches (no graph is
// instantiated), with the solve state, partial sums and vectors in `scratch`.
int mlp_pcg_noalloc(mlp_ctx ctx, mlp_mat A, mlp_prec B, const double* b, double* x, const mlp_solve_opts* o,
                    mlp_solve_stats* stats) {
    MLP_API_BEGIN
    require(ctx && A && B && b && x && o, "null argument");
    require(B->n == A->n, "preconditioner size differs from A");
    require(B->ctx == ctx && A->ctx->device == ctx->device, "the matrix and preconditioner must live on ctx's device");
    if (A->n > kNoAllocRows)
        throw Error("mlp_pcg_noalloc: " + std::to_string(A->n) + " rows exceed its static workspace (at most " +
                        std::to_string(kNoAllocRows) + ")",
                    MLP_ERR_INVALID);
    require(o->x0_mode == MLP_X0_ZERO || o->x0_mode == MLP_X0_GIVEN,
            "mlp_pcg_noalloc: x0_mode must be MLP_X0_ZERO or MLP_X0_GIVEN");
    require(B->impl->graph_safe(),
            "mlp_pcg_noalloc: this preconditioner's apply synchronises with the host (SymPrec, dynamic multi_split)");
    use_device(ctx);
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t n = A->n, rg = red_grid(n);
    const PcgVecs V = noalloc_vecs(noalloc_scratch(ctx->device), n);
    hipStream_t s = ctx->stream;
    const bool flex = o->flexible != 0;
    const uint32_t chunk = o->check_every ? o->check_every : 16;
    const uint32_t tre = o->true_res_every;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    fill_x0(ctx, n, o->x0_mode, x);
    SolveState st0;                                   // every field of the shared state starts from zero
    std::memset(&st0, 0, sizeof(st0));
    st0.eps = o->eps;
    st0.maxit = o->maxit;
    write_state(ctx, V.st, st0);
    launch_residual(ctx, A, b, x, V.r);
    hipLaunchKernelGGL(k_dot, dim3(rg), dim3(kBS), 0, s, n, (const double*)V.r, (const double*)V.r, V.partials);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBS), 0, s, int(FIN_START), rg, (const double*)V.partials, V.st,
                       (double*)nullptr, Batch{});
    enqueue_pcg_start(ctx, B->impl.get(), n, flex, V);
    SolveState st;
    uint32_t k = 0, c = chunk;
    while (true) {
        enqueue_pcg_chunk(ctx, A, B, b, x, flex, tre, nullptr, k, c, V);
        k += c;
        MLP_HIP_CHECK(hipGetLastError());
        st = read_state(ctx, V.st);
        if (st.done) break;
        c = next_chunk(st, chunk);
    }
    const double init_norm = st.init_norm, eps = st.eps;
    int code = MLP_OK;
    std::string msg;
    double known = -1.0;                           // the final x's true residual, when the restart test has it
    if (st.status == 4) {
        code = MLP_ERR_NO_ITER;
        msg = "No need to iterate";
    } else {
        if (flex) {
            for (uint32_t restart = 0; restart < kFlexRestarts && st.status == 0; restart++) {
                const double t = true_rel_res(ctx, A, b, x, V, init_norm);
                if (t <= eps) {
                    known = t;
                    break;
                }
                bool again = false;
                st = pcg_phase(ctx, A, B, b, x, o, init_norm, eps, st.niter, &again, &V);
                if (again) break;
            }
        }
        MLP_HIP_CHECK(hipStreamSynchronize(s));
    }
    const double trr = code != MLP_OK ? 0.0 : known >= 0.0 ? known : true_rel_res(ctx, A, b, x, V, init_norm);
    if (code == MLP_OK) {
        if (st.status == 3) {
            code = MLP_ERR_NOT_SPD;
            msg = "<Ap, p> = " + std::to_string(st.app);
        } else if (st.status == 5) {
            code = MLP_ERR_MAXIT;
            msg = "maximum number of iterations reached";
        } else if (flex && trr > eps) {
            code = MLP_ERR_NO_CONV;
            msg = "flexible PCG: the FP64 true residual is above eps after the restarts";
        }
    }
    if (stats) {
        stats->niter = st.niter;
        stats->init_norm = init_norm;
        stats->final_norm = st.norm;
        stats->eps_used = eps;
        stats->true_rel_res = trr;
        stats->t_solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->status = code;
    }
    if (code != MLP_OK) throw Error(msg, code);
    MLP_API_END
}

}  // extern "C"

// Host check of the chain program (plan.cc: plan_chain) and of the launch
// sequence the device runs with it.  For a uh_cheb hierarchy with ncheb = 2
// (the arrays Prec::Level holds, cheb_prec.h:17-40, read from a file that the
// test writes) the apply of Prec::solve (cheb_solve.cc:13-123) is computed
//   (1) by plain recursion over the level arrays, in the reference's order of
//       operations (FP64), and
//   (2) as the device runs it: the recursion's operations grouped into
//       descent and ascent launches, each executed block by block and slot by
//       slot over the chain program, with the level vectors in chain order,
// and the two results must be equal bit for bit, for chains starting at every
// level.  The Chebyshev coefficients are fixed test values (both sides use the
// same ones).
//   chain_check <levels.bin>
// Built and run by tests/test_host_checks.py.
#include "../../ml_prec_amd/csrc/plan.cc"

#include <cstdio>
#include <cstring>
#include <random>

using namespace mlp;

struct FileLevel {
    uint32_t N = 0, M = 0, Md = 0, nA = 0;
    std::vector<uint32_t> map, L_ia, L_ja, U_ia, U_ja, A_ia, A_ja;   // A: the next level's matrix
    std::vector<double> L_a, U_a, dval, A_a;
    LevelArrays arrays() const {
        LevelArrays la;
        la.N = N; la.M = M; la.Md = Md; la.map = map.data();
        la.L_ia = L_ia.data(); la.L_ja = L_ja.data(); la.L_a = L_a.data();
        la.U_ia = U_ia.data(); la.U_ja = U_ja.data(); la.U_a = U_a.data(); la.dval = dval.data();
        return la;
    }
};

template <class T> static bool rd(FILE* f, std::vector<T>& v, size_t n) {
    v.resize(n);
    return n == 0 || std::fread(v.data(), sizeof(T), n, f) == n;
}

static bool is_one(double x) { return std::fabs(x - 1.0) < 1e-15; }   // both sides use the same test

struct Coef {
    double a1, ma2, opb2;
    int s1_skip, s2_skip, opb2_one;
};

static std::vector<FileLevel> lv;
static std::vector<Coef> co;           // per level l: the iteration on level l+1
static uint32_t nlev;

static int packed_bad = 0;              // packed node words that disagree with the full arrays

// A node's L entries as the kernels read them: the first two from its packed word and inline values
// (ChainLevelProgram::Lmeta, Lv2), the rest through Lptr; the word's depth and count must be Ldep's and Lptr's.
template <class Vec> static double l_sub(const ChainLevelProgram& C, uint32_t q, double acc, const Vec& x) {
    const uint32_t m = C.Lmeta[q], c = (m >> 8) & 0xffu, p0 = C.Lptr[q];
    if ((m & 0xffu) != C.Ldep[q] || c != C.Lptr[q + 1] - p0 ||
        (c > 0 && (((m >> 16) & 0xffu) != C.Lloc[p0] || C.Lv2[2 * size_t(q)] != C.Lval[p0])) ||
        (c > 1 && ((m >> 24) != C.Lloc[p0 + 1] || C.Lv2[2 * size_t(q) + 1] != C.Lval[p0 + 1])))
        packed_bad++;
    if (c > 0) acc -= C.Lv2[2 * size_t(q)] * x[(m >> 16) & 0xffu];
    if (c > 1) acc -= C.Lv2[2 * size_t(q) + 1] * x[m >> 24];
    for (uint32_t p = p0 + 2; p < p0 + c; p++) acc -= C.Lval[p] * x[C.Lloc[p]];
    return acc;
}
// A U-list entry's solve as the kernels read it: its first off-diagonal entry from the packed word and Uv1,
// the rest through Uptr; writes x[loc] and returns loc.
template <class Vec> static uint32_t u_solve(const ChainLevelProgram& C, uint32_t u, double z, Vec& x) {
    const uint32_t m = C.Umeta[u], c = (m >> 8) & 0xffu, p0 = C.Uptr[u], loc = (m >> 16) & 0xffu;
    if ((m & 0xffu) != C.Udep[u] || c != C.Uptr[u + 1] - p0 || loc != C.uloc[u] ||
        (c > 0 && ((m >> 24) != C.Ucol[p0] || C.Uv1[u] != C.Uval[p0])))
        packed_bad++;
    if (c > 0) z -= C.Uv1[u] * x[m >> 24];
    for (uint32_t p = p0 + 1; p < p0 + c; p++) z -= C.Uval[p] * x[C.Ucol[p]];
    x[loc] = z / C.Udiag[u];
    return loc;
}

// A_{l+1} row r times (s u), entries in storage order (sym_multiply's full-row order)
static double arow(uint32_t l, uint32_t r, const std::vector<double>& u, double s, bool scale) {
    const FileLevel& L = lv[l];
    double acc = 0.0;
    for (uint32_t e = L.A_ia[r]; e < L.A_ia[r + 1]; e++) {
        double t = u[L.A_ja[e]];
        if (scale) t = t * s;
        acc += L.A_a[e] * t;
    }
    return acc;
}

static double cheb2(double x2, double us, const Coef& c) {
    const double v = c.s2_skip ? x2 : x2 * c.ma2;
    return c.opb2_one ? v + us : v + c.opb2 * us;
}

// (1) plain recursion: f, x natural vectors of level l
static void solve_ref(uint32_t l, const std::vector<double>& f, std::vector<double>& x) {
    const FileLevel& L = lv[l];
    const uint32_t N = L.N, M = L.M, K = N - L.Md, n = K - M;
    std::vector<double> w(K);
    for (uint32_t i = 0; i < K; i++) {
        double a = f[L.map[i]];
        for (uint32_t p = L.L_ia[i]; p < L.L_ia[i + 1]; p++) a -= L.L_a[p] * w[L.L_ja[p]];
        w[i] = a;
    }
    x.assign(N, 0.0);
    if (l + 1 < nlev) {
        const Coef& c = co[l];
        std::vector<double> F(w.begin() + M, w.begin() + K), u1, x2, tmp(n);
        solve_ref(l + 1, F, u1);
        std::vector<double> us = u1;
        if (!c.s1_skip)
            for (auto& v : us) v = v * c.a1;
        for (uint32_t r = 0; r < n; r++) tmp[r] = arow(l, r, u1, c.a1, !c.s1_skip) - F[r];
        solve_ref(l + 1, tmp, x2);
        for (uint32_t r = 0; r < n; r++) x[L.map[M + r]] = cheb2(x2[r], us[r], c);
    }
    for (uint32_t i = M; i-- > 0;) {
        double z = w[i];
        for (uint32_t p = L.U_ia[i] + 1; p < L.U_ia[i + 1]; p++) z -= L.U_a[p] * x[L.map[L.U_ja[p]]];
        x[L.map[i]] = z / L.U_a[L.U_ia[i]];
    }
    for (uint32_t t = 0; t < L.Md; t++) x[L.map[K + t]] = L.dval[t] * f[L.map[K + t]];
}

// (2) the chain launches.  Level vectors (per level l < c, as on the device):
// F[l] (level l+1 input, chain order of l+1), u1/u0/x2[l] (level l+1 natural
// order), wt[l] (chain positions of level l).
struct Bufs {
    std::vector<std::vector<double>> F, u1, u0, x2, wt;
};
static ChainProgram CP;
static Bufs B;

enum { IN_PLAIN_ = 0, IN_PERM_ = 1, IN_RES_ = 2 };
struct Input {
    int kind;
    const std::vector<double>* f;       // plain (natural) or perm (chain order)
};

// the input of level k (chain level index i = k - a) at chain position q
static double fetch_in(uint32_t k, const Input& in, uint32_t q) {
    const ChainLevelProgram& C = CP.lv[k - CP.a];
    if (in.kind == IN_PLAIN_) return (*in.f)[C.qrow[q]];
    if (in.kind == IN_PERM_) return (*in.f)[q];
    // residual of the second visit of k: A_k (a1 u1) - F, u1 = u1[k-1]; us written at the natural row
    const uint32_t p = k - 1, r = C.qrow[q];
    const Coef& c = co[p];
    const double v = arow(p, r, B.u1[p], c.a1, !c.s1_skip) - B.F[p][q];
    B.u0[p][r] = c.s1_skip ? B.u1[p][r] : B.u1[p][r] * c.a1;
    return v;
}

static void descent(uint32_t j, const Input& in, std::vector<double>& out_j) {
    const uint32_t a = CP.a, c = nlev - 1;
    auto out_of = [&](uint32_t k) -> std::vector<double>& { return k == j ? out_j : B.u1[k - 1]; };
    for (uint32_t b = 0; b < CP.blk_from[j - a]; b++) {
        std::vector<std::vector<double>> V(nlev);
        for (uint32_t k = j; k <= c; k++) {
            const ChainLevelProgram& C = CP.lv[k - a];
            const FileLevel& L = lv[k];
            const uint32_t* H = &C.hdr[size_t(kChainHdr) * b];
            const uint32_t base = H[0], S = H[1], nn = H[2], nkeep = H[2] + H[3];
            V[k].assign(S, 0.0);
            for (uint32_t s = 0; s < S; s++) {
                const double v = k == j ? fetch_in(k, in, base + s) : V[k - 1][C.src[base + s]];
                if (k > j) B.F[k - 1][base + s] = v;
                V[k][s] = v;
                if (s >= nkeep) out_of(k)[C.qrow[base + s]] = C.dval[base + s] * v;
            }
            for (uint32_t s = 0; s < nn; s++) {          // slot order is depth order
                const double acc = l_sub(C, base + s, V[k][s], V[k]);
                V[k][s] = acc;
                if (k < c) B.wt[k][base + s] = acc;
            }
            if (k == c) {                                // step 1 on the coarsest level
                std::vector<double> X(S, 0.0);
                for (uint32_t u = H[5]; u < H[5] + H[6]; u++) {
                    const uint32_t loc = u_solve(C, u, V[k][C.uloc[u]], X);
                    B.u1[c - 1][C.uxidx[u]] = X[loc];
                }
            }
            (void)L;
        }
    }
    // passthrough rows of level j
    const ChainLevelProgram& Cj = CP.lv[j - a];
    for (uint32_t t = 0; t < Cj.npt; t++) {
        const double v = fetch_in(j, in, Cj.nslots + t);
        uint32_t k = j;
        while (true) {
            const ChainLevelProgram& C = CP.lv[k - a];
            const uint32_t q = C.nslots + t;
            if (k > j) B.F[k - 1][q] = v;
            const bool deeper = k < c && t < CP.lv[k + 1 - a].npt;
            if (!deeper) {
                out_of(k)[C.qrow[q]] = C.dval[q] * v;
                break;
            }
            k++;
        }
    }
}

static void ascent(uint32_t ke, std::vector<double>& out_top) {
    const uint32_t a = CP.a, c = nlev - 1;
    const Coef& cc = co[c - 1];
    // u1s of level k's combine at natural row r of level k+1
    auto us_of = [&](uint32_t k, uint32_t r) {
        if (k == c - 1) return cc.s1_skip ? B.u1[k][r] : B.u1[k][r] * cc.a1;
        return co[k].s1_skip ? B.u1[k][r] : B.u0[k][r];
    };
    for (uint32_t b = 0; b < CP.blk_from[ke - a]; b++) {
        std::vector<std::vector<double>> X(nlev);
        {   // step 2 on the coarsest level
            const ChainLevelProgram& C = CP.lv[c - a];
            const uint32_t* H = &C.hdr[size_t(kChainHdr) * b];
            const uint32_t base = H[0], S = H[1], nn = H[2];
            std::vector<double> w(S);
            for (uint32_t s = 0; s < S; s++)
                w[s] = arow(c - 1, C.qrow[base + s], B.u1[c - 1], cc.a1, !cc.s1_skip) - B.F[c - 1][base + s];
            X[c].assign(S, 0.0);
            for (uint32_t s = nn; s < S; s++) X[c][s] = C.dval[base + s] * w[s];
            for (uint32_t s = 0; s < nn; s++) w[s] = l_sub(C, base + s, w[s], w);
            for (uint32_t u = H[5]; u < H[5] + H[6]; u++) u_solve(C, u, w[C.uloc[u]], X[c]);
        }
        for (uint32_t k = c; k-- > ke;) {
            const ChainLevelProgram& C = CP.lv[k - a];
            const ChainLevelProgram& Cn = CP.lv[k + 1 - a];
            const uint32_t* H = &C.hdr[size_t(kChainHdr) * b];
            const uint32_t* Hn = &Cn.hdr[size_t(kChainHdr) * b];
            const uint32_t base = H[0], S = H[1], nkeep = H[2] + H[3];
            std::vector<double>& out = k == ke ? out_top : B.x2[k - 1];
            X[k].assign(S, 0.0);
            for (uint32_t s = 0; s < Hn[1]; s++) {        // kept rows: the combine of level k
                const uint32_t q = Hn[0] + s;
                const double v = cheb2(X[k + 1][s], us_of(k, Cn.qrow[q]), co[k]);
                X[k][Cn.src[q]] = v;
                if (k == ke) out[Cn.xl[q]] = v;
            }
            if (k > ke)
                for (uint32_t s = nkeep; s < S; s++) X[k][s] = B.x2[k - 1][C.qrow[base + s]];
            for (uint32_t u = H[5]; u < H[5] + H[6]; u++) {
                const uint32_t loc = u_solve(C, u, B.wt[k][base + C.uloc[u]], X[k]);
                if (k == ke) out[C.uxidx[u]] = X[k][loc];
            }
        }
    }
    // passthrough rows that reach level ke + 1
    if (ke + 1 - a < CP.nlev)
        for (uint32_t t = 0; t < CP.lv[ke + 1 - a].npt; t++) {
            uint32_t m = ke + 1;
            while (m < c && t < CP.lv[m + 1 - a].npt) m++;
            double v;
            const ChainLevelProgram& Cm = CP.lv[m - a];
            const uint32_t qm = Cm.nslots + t;
            if (m == c) v = Cm.dval[qm] * (arow(c - 1, Cm.qrow[qm], B.u1[c - 1], cc.a1, !cc.s1_skip) - B.F[c - 1][qm]);
            else v = B.x2[m - 1][Cm.qrow[qm]];
            for (uint32_t k = m; k-- > ke;) {
                const ChainLevelProgram& Cn = CP.lv[k + 1 - a];
                v = cheb2(v, us_of(k, Cn.qrow[Cn.nslots + t]), co[k]);
            }
            const ChainLevelProgram& Ce = CP.lv[ke - a];
            out_top[Ce.qrow[Ce.nslots + t]] = v;
        }
}

// The recursion's operations, grouped into launches as the device does.
struct Op {
    int kind;        // 0 D, 1 coarse first visit, 2 coarse second visit, 3 U
    uint32_t k;
    Input in;
    std::vector<double>* out;
};
static void gen(uint32_t k, Input in, std::vector<double>* out, std::vector<Op>& ops) {
    const uint32_t c = nlev - 1;
    if (k == c) {
        ops.push_back({in.kind == IN_RES_ ? 2 : 1, k, in, out});
        return;
    }
    ops.push_back({0, k, in, out});
    gen(k + 1, {IN_PERM_, &B.F[k]}, &B.u1[k], ops);
    gen(k + 1, {IN_RES_, nullptr}, &B.x2[k], ops);
    ops.push_back({3, k, in, out});
}
static uint32_t run_chain(const Input& in, std::vector<double>& out) {
    std::vector<Op> ops;
    gen(CP.a, in, &out, ops);
    uint32_t launches = 0;
    for (size_t i = 0; i < ops.size();) {
        if (ops[i].kind == 0 || ops[i].kind == 1) {          // descent: D ... D, coarse first visit
            descent(ops[i].k, ops[i].in, *ops[i].out);
            while (ops[i].kind == 0) i++;
            if (ops[i].kind != 1) throw Error("op sequence", 1);
            i++;
        } else {                                              // ascent: coarse second visit, U ... U
            if (ops[i].kind != 2) throw Error("op sequence", 1);
            i++;
            size_t e = i;
            while (e < ops.size() && ops[e].kind == 3) e++;
            if (e == i) throw Error("op sequence", 1);
            ascent(ops[e - 1].k, *ops[e - 1].out);
            i = e;
        }
        launches++;
    }
    return launches;
}

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    FILE* fp = std::fopen(argv[1], "rb");
    if (!fp) return 2;
    if (std::fread(&nlev, 4, 1, fp) != 1) return 2;
    lv.resize(nlev);
    for (auto& L : lv) {
        uint32_t h[7];
        if (std::fread(h, 4, 7, fp) != 7) return 2;
        L.N = h[0]; L.M = h[1]; L.Md = h[2]; L.nA = h[5];
        const uint32_t K = L.N - L.Md;
        if (!rd(fp, L.map, L.N) || !rd(fp, L.L_ia, K + 1) || !rd(fp, L.L_ja, h[3]) || !rd(fp, L.L_a, h[3]) ||
            !rd(fp, L.U_ia, L.M + 1) || !rd(fp, L.U_ja, h[4]) || !rd(fp, L.U_a, h[4]) || !rd(fp, L.dval, L.Md) ||
            !rd(fp, L.A_ia, h[5] ? h[5] + 1 : 0) || !rd(fp, L.A_ja, h[6]) || !rd(fp, L.A_a, h[6]))
            return 2;
    }
    std::fclose(fp);
    // test coefficients (one level skips the first scaling, one the second, to cover both forms)
    co.resize(nlev);
    for (uint32_t l = 0; l < nlev; l++) {
        co[l] = {0.6 + 0.01 * l, -1.3 - 0.02 * l, 1.15 + 0.03 * l, 0, 0, 0};
        if (l == 1) { co[l].a1 = 1.0; co[l].s1_skip = 1; }
        if (l == 2) { co[l].ma2 = 1.0; co[l].s2_skip = 1; }
    }
    std::mt19937_64 rng(11);
    std::normal_distribution<double> nd;
    const uint32_t c = nlev - 1;
    int bad = 0;
    for (const uint32_t bs : {64u, 128u})       // one- and two-wavefront blocks
    for (uint32_t a = 0; a < c; a++) {
        std::vector<LevelArrays> la;
        for (uint32_t l = a; l < nlev; l++) la.push_back(lv[l].arrays());
        ChainProgram P;
        if (!plan_chain(la, a, bs, 2 * bs, P)) {
            std::printf("chain %u..%u, %u-thread blocks: does not fit\n", a, c, bs);
            continue;
        }
        CP = std::move(P);
        // the level a input: a natural vector (as the solver hands level 0 its input)
        std::vector<double> f(lv[a].N), xr, xc(lv[a].N, 0.0);
        for (auto& v : f) v = nd(rng);
        solve_ref(a, f, xr);
        B = Bufs();
        B.F.resize(nlev); B.u1.resize(nlev); B.u0.resize(nlev); B.x2.resize(nlev); B.wt.resize(nlev);
        for (uint32_t l = a; l < c; l++) {
            const uint32_t n = lv[l + 1].N;
            B.F[l].assign(n, 0.0); B.u1[l].assign(n, 0.0); B.u0[l].assign(n, 0.0); B.x2[l].assign(n, 0.0);
            B.wt[l].assign(lv[l].N, 0.0);
        }
        const uint32_t launches = run_chain({IN_PLAIN_, &f}, xc);
        size_t diff = 0;
        for (uint32_t i = 0; i < lv[a].N; i++)
            if (std::memcmp(&xr[i], &xc[i], sizeof(double)) != 0) diff++;
        std::printf("chain %u..%u, %u-thread blocks: %u blocks, max slots %u, max nodes %u, %u launches, %zu of %u rows "
                    "differ\n", a, c, bs, CP.nblk, CP.max_slots, CP.max_nodes, launches, diff, lv[a].N);
        if (diff) bad++;
    }
    if (packed_bad) {
        std::printf("%d packed node words disagree with the full arrays\n", packed_bad);
        bad++;
    }
    std::printf(bad ? "FAIL %d\n" : "OK\n", bad);
    return bad ? 1 : 0;
}

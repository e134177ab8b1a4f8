#include "plan.h"

#include "par.h"

#include "mlp_error.h"

#include <algorithm>
#include <atomic>
#include <exception>
#include <cstring>
#include <numeric>
#include <string>

namespace mlp {

namespace {

struct DSU {
    std::vector<uint32_t> p;
    explicit DSU(uint32_t n) : p(n) { std::iota(p.begin(), p.end(), 0u); }
    uint32_t find(uint32_t x) {
        while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
        return x;
    }
    void unite(uint32_t a, uint32_t b) {
        a = find(a); b = find(b);
        if (a != b) p[std::max(a, b)] = std::min(a, b);
    }
};

constexpr uint32_t kNoNode = 0xffffffffu;

// appends the flat and isolated rows to the consumption order and checks that
// it covers every level row once; flat rows write F slot r until linked
void finish_order(const LevelArrays& lv, LevelProgram& P) {
    for (uint32_t r : P.flat) P.qrow.push_back(lv.map[P.M + r]);
    for (uint32_t k = 0; k < P.Md; k++) P.qrow.push_back(lv.map[P.N - P.Md + k]);
    if (P.qrow.size() != P.N) throw Error("level rows are not covered exactly once", 1);
    std::vector<char> seen(P.N, 0);
    for (uint32_t j : P.qrow) {
        if (j >= P.N || seen[j]) throw Error("level rows are not covered exactly once", 1);
        seen[j] = 1;
    }
    P.flatq = P.flat;
}

}  // namespace

void plan_level(const LevelArrays& lv, uint32_t cap, LevelProgram& P, const BlockFits& fits) {
    const uint32_t N = lv.N, M = lv.M, Md = lv.Md, K = N - Md, n = N - M - Md;
    P = LevelProgram();
    P.M = M; P.n = n; P.Md = Md; P.N = N;

    // the structure the kernels rely on
    for (uint32_t i = 0; i < K; i++)
        for (uint32_t p = lv.L_ia[i]; p < lv.L_ia[i + 1]; p++) {
            uint32_t k = lv.L_ja[p];
            if (k >= i || k >= M)
                throw Error("L row " + std::to_string(i) + " has a column outside the eliminated block", 1);
        }
    for (uint32_t i = 0; i < M; i++) {
        if (lv.U_ia[i + 1] <= lv.U_ia[i] || lv.U_ja[lv.U_ia[i]] != i)
            throw Error("U row " + std::to_string(i) + " does not start with its pivot", 1);
        for (uint32_t p = lv.U_ia[i] + 1; p < lv.U_ia[i + 1]; p++)
            if (lv.U_ja[p] <= i || lv.U_ja[p] >= K)
                throw Error("U row " + std::to_string(i) + " has a column outside (i, N-Md)", 1);
    }

    // nodes: eliminated rows and kept rows fed by L entries
    std::vector<char> is_node(K, 0);
    for (uint32_t i = 0; i < M; i++) is_node[i] = 1;
    for (uint32_t i = M; i < K; i++) {
        if (lv.L_ia[i + 1] > lv.L_ia[i]) is_node[i] = 1;
        else P.flat.push_back(i - M);
    }

    // dependency components and depths
    DSU dsu(std::max<uint32_t>(K, 1));
    std::vector<uint32_t> dL(K, 0), dU(M, 0);
    for (uint32_t i = 0; i < K; i++) {
        if (!is_node[i]) continue;
        for (uint32_t p = lv.L_ia[i]; p < lv.L_ia[i + 1]; p++) {
            dsu.unite(i, lv.L_ja[p]);
            dL[i] = std::max(dL[i], dL[lv.L_ja[p]] + 1);
        }
    }
    for (uint32_t i = M; i-- > 0;)
        for (uint32_t p = lv.U_ia[i] + 1; p < lv.U_ia[i + 1]; p++) {
            uint32_t j = lv.U_ja[p];
            if (j < M) {
                dsu.unite(i, j);
                dU[i] = std::max(dU[i], dU[j] + 1);
            }
        }
    for (uint32_t i = 0; i < K; i++) if (is_node[i]) P.max_depth = std::max(P.max_depth, dL[i]);
    for (uint32_t i = 0; i < M; i++) P.max_depth = std::max(P.max_depth, dU[i]);

    std::vector<uint32_t> comp_of(K, kNoNode), root_comp(std::max<uint32_t>(K, 1), kNoNode);
    uint32_t ncomp = 0;
    for (uint32_t i = 0; i < K; i++) {
        if (!is_node[i]) continue;
        uint32_t r = dsu.find(i);
        if (root_comp[r] == kNoNode) root_comp[r] = ncomp++;
        comp_of[i] = root_comp[r];
    }
    if (ncomp == 0) {
        finish_order(lv, P);
        return;
    }
    std::vector<uint32_t> csize(ncomp, 0);
    for (uint32_t i = 0; i < K; i++) if (is_node[i]) csize[comp_of[i]]++;
    P.max_comp_size = *std::max_element(csize.begin(), csize.end());
    // components in order, a new block whenever the next one would overflow
    // cap (so a component larger than cap sits alone)
    std::vector<uint32_t> comp_blk(ncomp);
    uint32_t nblk = 0, fill = 0;
    for (uint32_t c = 0; c < ncomp; c++) {
        if (c == 0 || fill + csize[c] > cap) { nblk++; fill = 0; }
        comp_blk[c] = nblk - 1;
        fill += csize[c];
    }
    // the blocks' rows in ascending order, as one array with block offsets
    std::vector<uint32_t> roff(size_t(nblk) + 1, 0), rows_all;
    for (uint32_t i = 0; i < K; i++) if (is_node[i]) roff[comp_blk[comp_of[i]] + 1]++;
    for (uint32_t b = 0; b < nblk; b++) roff[b + 1] += roff[b];
    rows_all.resize(roff[nblk]);
    {
        std::vector<uint32_t> at(roff.begin(), roff.end() - 1);
        for (uint32_t i = 0; i < K; i++) if (is_node[i]) rows_all[at[comp_blk[comp_of[i]]]++] = i;
    }

    P.nblk = nblk;
    P.blk.resize(kBlkFields * size_t(nblk));
    const uint32_t total = roff[nblk];
    P.nodes = total;
    P.qrow.resize(total);
    P.out.resize(total);
    P.Ldep.resize(total);
    P.Lptr.assign(total + 1, 0);
    P.uloc.resize(M);
    P.uxidx.resize(M);
    P.Udep.resize(M);
    P.Udiag.resize(M);
    P.Uptr.assign(M + 1, 0);
    std::vector<uint32_t> local(K, kNoNode);
    // rows [b, e) below `below` ordered by key, equal keys in row order (a counting sort into out)
    auto by_depth = [](const uint32_t* b, const uint32_t* e, const std::vector<uint32_t>& key, uint32_t below,
                       std::vector<uint32_t>& out, std::vector<uint32_t>& count) {
        uint32_t kmax = 0, m = 0;
        for (const uint32_t* r = b; r < e; r++)
            if (*r < below) { kmax = std::max(kmax, key[*r]); m++; }
        count.assign(size_t(kmax) + 2, 0);
        for (const uint32_t* r = b; r < e; r++)
            if (*r < below) count[key[*r] + 1]++;
        for (uint32_t k = 0; k <= kmax; k++) count[k + 1] += count[k];
        out.resize(m);
        for (const uint32_t* r = b; r < e; r++)
            if (*r < below) out[count[key[*r]]++] = *r;
    };
    // The blocks are independent: pass 1 counts each block's depth levels and entries, pass 2 writes every
    // block's ranges at the offsets those counts give (the arrays the sequential append would build).
    struct Counts { uint32_t lvL = 0, entL = 0, nU = 0, lvU = 0, entU = 0; };
    std::vector<Counts> bc(nblk);
    parallel_for(nblk, 256, [&](uint64_t b0, uint64_t b1) {
        std::vector<uint32_t> lo, uo, count;
        for (uint64_t b = b0; b < b1; b++) {
            const uint32_t* Rb = rows_all.data() + roff[b];
            const uint32_t* Re = rows_all.data() + roff[b + 1];
            by_depth(Rb, Re, dL, K, lo, count);
            by_depth(Rb, Re, dU, M, uo, count);
            Counts& c = bc[b];
            for (size_t k = 0; k < lo.size(); k++) {
                if (k + 1 == lo.size() || dL[lo[k + 1]] != dL[lo[k]]) c.lvL++;
                c.entL += lv.L_ia[lo[k] + 1] - lv.L_ia[lo[k]];
            }
            c.nU = uint32_t(uo.size());
            for (size_t k = 0; k < uo.size(); k++) {
                if (k + 1 == uo.size() || dU[uo[k + 1]] != dU[uo[k]]) c.lvU++;
                c.entU += lv.U_ia[uo[k] + 1] - lv.U_ia[uo[k]] - 1;
            }
        }
    });
    std::vector<uint32_t> oLv(size_t(nblk) + 1, 0), oLe(size_t(nblk) + 1, 0), oU(size_t(nblk) + 1, 0),
        oUv(size_t(nblk) + 1, 0), oUe(size_t(nblk) + 1, 0);
    for (uint32_t b = 0; b < nblk; b++) {
        oLv[b + 1] = oLv[b] + bc[b].lvL;
        oLe[b + 1] = oLe[b] + bc[b].entL;
        oU[b + 1] = oU[b] + bc[b].nU;
        oUv[b + 1] = oUv[b] + bc[b].lvU;
        oUe[b + 1] = oUe[b] + bc[b].entU;
    }
    P.lvlL.resize(oLv[nblk]);
    P.Lloc.resize(oLe[nblk]);
    P.Lval.resize(oLe[nblk]);
    P.lvlU.resize(oUv[nblk]);
    P.Ucol.resize(oUe[nblk]);
    P.Uval.resize(oUe[nblk]);
    std::vector<char> glob(nblk, 0);
    std::vector<uint32_t> nlev(nblk, 0);
    parallel_for(nblk, 256, [&](uint64_t b0, uint64_t b1) {
        std::vector<uint32_t> lo, uo, count;
        for (uint64_t b = b0; b < b1; b++) {
            const uint32_t* Rb = rows_all.data() + roff[b];
            const uint32_t* Re = rows_all.data() + roff[b + 1];
            const uint32_t base = roff[b];
            by_depth(Rb, Re, dL, K, lo, count);
            by_depth(Rb, Re, dU, M, uo, count);
            for (uint32_t k = 0; k < lo.size(); k++) local[lo[k]] = k;
            uint32_t* B = &P.blk[kBlkFields * b];
            B[0] = base;
            B[1] = uint32_t(lo.size());
            B[2] = oLv[b];
            B[6] = oLe[b];
            uint32_t lvl = oLv[b], le = oLe[b];
            for (uint32_t k = 0; k < lo.size(); k++) {
                const uint32_t i = lo[k];
                P.Ldep[base + k] = lvl - B[2];
                if (k + 1 == lo.size() || dL[lo[k + 1]] != dL[i]) P.lvlL[lvl++] = k + 1;
                P.qrow[base + k] = lv.map[i];
                P.out[base + k] = i >= M ? (kGlobalRef | (i - M)) : k;
                for (uint32_t p = lv.L_ia[i]; p < lv.L_ia[i + 1]; p++) {
                    P.Lloc[le] = local[lv.L_ja[p]];
                    P.Lval[le] = lv.L_a[p];
                    le++;
                }
                P.Lptr[base + k + 1] = le;
            }
            B[3] = lvl;
            B[7] = le;
            B[4] = oUv[b];
            B[8] = oUe[b];
            B[10] = oU[b];
            B[11] = uint32_t(uo.size());
            uint32_t ulv = oUv[b], ue = oUe[b];
            const uint32_t upos = oU[b];
            for (uint32_t k = 0; k < uo.size(); k++) {
                const uint32_t i = uo[k];
                P.Udep[upos + k] = ulv - B[4];
                if (k + 1 == uo.size() || dU[uo[k + 1]] != dU[i]) P.lvlU[ulv++] = k + 1;
                P.uloc[upos + k] = local[i];
                P.uxidx[upos + k] = lv.map[i];
                P.Udiag[upos + k] = lv.U_a[lv.U_ia[i]];
                for (uint32_t p = lv.U_ia[i] + 1; p < lv.U_ia[i + 1]; p++) {
                    const uint32_t j = lv.U_ja[p];
                    P.Ucol[ue] = j < M ? local[j] : (kGlobalRef | (j - M));
                    P.Uval[ue] = lv.U_a[p];
                    ue++;
                }
                P.Uptr[upos + k + 1] = ue;
            }
            B[5] = ulv;
            B[9] = ue;
            nlev[b] = std::max(B[3] - B[2], B[5] - B[4]);
            if (fits && !fits(uint32_t(lo.size()), B[7] - B[6], B[9] - B[8], uint32_t(uo.size()), nlev[b])) {
                B[1] |= kGlobalBlock;
                glob[b] = 1;
            }
        }
    });
    for (uint32_t b = 0; b < nblk; b++) {
        const uint32_t* B = &P.blk[kBlkFields * b];
        if (glob[b]) {
            P.has_global = true;
            continue;
        }
        P.max_nodes = std::max<uint32_t>(P.max_nodes, B[1]);
        P.max_Unodes = std::max<uint32_t>(P.max_Unodes, B[11]);
        P.max_Lent = std::max<uint32_t>(P.max_Lent, B[7] - B[6]);
        P.max_Uent = std::max<uint32_t>(P.max_Uent, B[9] - B[8]);
        P.max_levels = std::max<uint32_t>(P.max_levels, nlev[b]);
    }
    finish_order(lv, P);
    if (P.Lloc.empty()) { P.Lloc.push_back(0); P.Lval.push_back(0.0); }
    if (P.Ucol.empty()) { P.Ucol.push_back(0); P.Uval.push_back(0.0); }
    if (P.lvlU.empty()) P.lvlU.push_back(0);
}

void link_levels(LevelProgram& P, const uint32_t* next_qrow, size_t n) {
    if (n != P.n) throw Error("level sizes do not chain", 1);
    std::vector<uint32_t> qpos(P.n);
    for (uint32_t q = 0; q < P.n; q++) qpos[next_qrow[q]] = q;
    for (uint32_t& o : P.out)
        if (o & kGlobalRef) o = kGlobalRef | qpos[o & ~kGlobalRef];
    for (size_t i = 0; i < P.flat.size(); i++) P.flatq[i] = qpos[P.flat[i]];
}

void link_levels(LevelProgram& P, const LevelProgram& next) {
    if (next.N != P.n) throw Error("level sizes do not chain", 1);
    link_levels(P, next.qrow.data(), next.qrow.size());
}

bool plan_pair(const LevelArrays& la, const LevelArrays& lb, uint32_t cap, PairProgram& out) {
    const uint32_t N = la.N, M = la.M, Md = la.Md, K = N - Md, n = N - M - Md;
    const uint32_t M2 = lb.M, Md2 = lb.Md;
    if (lb.N != n || M2 + Md2 != lb.N) return false;          // level l+1 must keep no rows
    // the structure the kernels rely on (as plan_level checks it)
    for (uint32_t i = 0; i < K; i++)
        for (uint32_t p = la.L_ia[i]; p < la.L_ia[i + 1]; p++)
            if (la.L_ja[p] >= i || la.L_ja[p] >= M) throw Error("L row has a column outside the eliminated block", 1);
    for (uint32_t i = 0; i < M; i++) {
        if (la.U_ia[i + 1] <= la.U_ia[i] || la.U_ja[la.U_ia[i]] != i) throw Error("U row does not start with its pivot", 1);
        for (uint32_t p = la.U_ia[i] + 1; p < la.U_ia[i + 1]; p++)
            if (la.U_ja[p] <= i || la.U_ja[p] >= K) throw Error("U row has a column outside (i, N-Md)", 1);
    }
    for (uint32_t i = 0; i < M2; i++) {
        for (uint32_t p = lb.L_ia[i]; p < lb.L_ia[i + 1]; p++)
            if (lb.L_ja[p] >= i) throw Error("L row has a column outside the eliminated block", 1);
        if (lb.U_ia[i + 1] <= lb.U_ia[i] || lb.U_ja[lb.U_ia[i]] != i) throw Error("U row does not start with its pivot", 1);
        for (uint32_t p = lb.U_ia[i] + 1; p < lb.U_ia[i + 1]; p++)
            if (lb.U_ja[p] <= i || lb.U_ja[p] >= M2) throw Error("U row has a column outside (i, N-Md)", 1);
    }
    std::vector<uint32_t> pos2(n);                          // level-(l+1) permuted index of row r
    for (uint32_t i = 0; i < n; i++) pos2[lb.map[i]] = i;
    // nodes: eliminated rows, kept rows with L entries or U references, kept rows eliminated on level l+1
    std::vector<char> is_node(K, 0);
    for (uint32_t i = 0; i < M; i++) is_node[i] = 1;
    for (uint32_t i = M; i < K; i++)
        if (la.L_ia[i + 1] > la.L_ia[i] || pos2[i - M] < M2) is_node[i] = 1;
    for (uint32_t i = 0; i < M; i++)
        for (uint32_t p = la.U_ia[i] + 1; p < la.U_ia[i + 1]; p++) is_node[la.U_ja[p]] = 1;
    DSU dsu(std::max<uint32_t>(K, 1));
    std::vector<uint32_t> dL(K, 0), dU(M, 0), dL2(M2, 0), dU2(M2, 0);
    for (uint32_t i = 0; i < K; i++)
        for (uint32_t p = la.L_ia[i]; p < la.L_ia[i + 1]; p++) {
            dsu.unite(i, la.L_ja[p]);
            dL[i] = std::max(dL[i], dL[la.L_ja[p]] + 1);
        }
    for (uint32_t i = M; i-- > 0;)
        for (uint32_t p = la.U_ia[i] + 1; p < la.U_ia[i + 1]; p++) {
            const uint32_t j = la.U_ja[p];
            dsu.unite(i, j);
            if (j < M) dU[i] = std::max(dU[i], dU[j] + 1);
        }
    for (uint32_t i = 0; i < M2; i++) {
        const uint32_t a = M + lb.map[i];
        for (uint32_t p = lb.L_ia[i]; p < lb.L_ia[i + 1]; p++) {
            dsu.unite(a, M + lb.map[lb.L_ja[p]]);
            dL2[i] = std::max(dL2[i], dL2[lb.L_ja[p]] + 1);
        }
    }
    for (uint32_t i = M2; i-- > 0;)
        for (uint32_t p = lb.U_ia[i] + 1; p < lb.U_ia[i + 1]; p++) {
            const uint32_t j = lb.U_ja[p];
            dsu.unite(M + lb.map[i], M + lb.map[j]);
            dU2[i] = std::max(dU2[i], dU2[j] + 1);
        }
    // components in order of their first row; per component its four counts
    std::vector<uint32_t> comp_of(K, kNoNode), root_comp(std::max<uint32_t>(K, 1), kNoNode);
    struct Count { uint32_t S = 0, nk = 0, nu = 0, nu2 = 0; };
    std::vector<Count> cnt;
    for (uint32_t i = 0; i < K; i++) {
        if (!is_node[i]) continue;
        const uint32_t r = dsu.find(i);
        if (root_comp[r] == kNoNode) { root_comp[r] = uint32_t(cnt.size()); cnt.emplace_back(); }
        const uint32_t c = comp_of[i] = root_comp[r];
        cnt[c].S++;
        if (i < M) cnt[c].nu++;
        else {
            cnt[c].nk++;
            if (pos2[i - M] < M2) cnt[c].nu2++;
        }
    }
    for (const Count& c : cnt)
        if (c.S > cap || c.nk > cap || c.nu > cap || c.nu2 > cap) return false;
    std::vector<uint32_t> comp_blk(cnt.size());
    uint32_t nblk = 0;
    Count fill;
    for (size_t c = 0; c < cnt.size(); c++) {
        const Count& k = cnt[c];
        if (c == 0 || fill.S + k.S > cap || fill.nk + k.nk > cap || fill.nu + k.nu > cap || fill.nu2 + k.nu2 > cap) {
            nblk++;
            fill = Count();
        }
        comp_blk[c] = nblk - 1;
        fill.S += k.S; fill.nk += k.nk; fill.nu += k.nu; fill.nu2 += k.nu2;
    }
    // the blocks' rows in ascending order, as one array with block offsets
    std::vector<uint32_t> roff(size_t(nblk) + 1, 0), rows_all;
    for (uint32_t i = 0; i < K; i++)
        if (is_node[i]) roff[comp_blk[comp_of[i]] + 1]++;
    for (uint32_t b = 0; b < nblk; b++) roff[b + 1] += roff[b];
    rows_all.resize(roff[nblk]);
    {
        std::vector<uint32_t> at(roff.begin(), roff.end() - 1);
        for (uint32_t i = 0; i < K; i++)
            if (is_node[i]) rows_all[at[comp_blk[comp_of[i]]]++] = i;
    }

    PairProgram P;
    P.N = N; P.M = M; P.Md = Md; P.n = n; P.nblk = nblk;
    P.blk.assign(size_t(kPairFields) * nblk, 0);
    std::vector<uint32_t> local(K, kNoNode), local2(K, kNoNode);
    // rows of [b, e) passing `keep`, ordered by key, equal keys in their order (a counting sort into o)
    auto by_key = [](const uint32_t* b, const uint32_t* e, auto key, auto keep, std::vector<uint32_t>& o,
                     std::vector<uint32_t>& kcount) {
        uint32_t kmax = 0, m = 0;
        for (const uint32_t* r = b; r < e; r++)
            if (keep(*r)) { kmax = std::max(kmax, key(*r)); m++; }
        kcount.assign(size_t(kmax) + 2, 0);
        for (const uint32_t* r = b; r < e; r++)
            if (keep(*r)) kcount[key(*r) + 1]++;
        for (uint32_t k = 0; k <= kmax; k++) kcount[k + 1] += kcount[k];
        o.resize(m);
        for (const uint32_t* r = b; r < e; r++)
            if (keep(*r)) o[kcount[key(*r)]++] = *r;
    };
    auto levels_of = [](const std::vector<uint32_t>& o, auto key) {
        uint32_t nl = 0;
        for (size_t k = 0; k < o.size(); k++) nl = std::max(nl, key(o[k]) + 1);
        return nl;
    };
    auto kL = [&](uint32_t i) { return dL[i]; };
    auto kU = [&](uint32_t i) { return dU[i]; };
    // kept nodes, by level-(l+1) L depth (rows isolated there at depth 0)
    auto key2 = [&](uint32_t i) { const uint32_t q = pos2[i - M]; return q < M2 ? dL2[q] : 0u; };
    auto keyU2 = [&](uint32_t i) { return dU2[pos2[i - M]]; };
    auto all = [](uint32_t) { return true; };
    auto elim = [&](uint32_t i) { return i < M; };
    auto kept = [&](uint32_t i) { return i >= M; };
    auto kept_elim2 = [&](uint32_t i) { return i >= M && pos2[i - M] < M2; };
    // The blocks are independent: pass 1 counts each block's rows and entries of every list, pass 2 writes every
    // block's ranges at the offsets those counts give (the arrays the sequential append would build).
    struct Counts { uint32_t nU = 0, nK = 0, nU2 = 0, eL = 0, eU = 0, eL2 = 0, eU2 = 0; };
    std::vector<Counts> bc(nblk);
    parallel_for(nblk, 256, [&](uint64_t b0, uint64_t b1) {
        std::vector<uint32_t> uo, ko, u2o, kc;
        for (uint64_t b = b0; b < b1; b++) {
            const uint32_t* Rb = rows_all.data() + roff[b];
            const uint32_t* Re = rows_all.data() + roff[b + 1];
            Counts& c = bc[b];
            for (const uint32_t* r = Rb; r < Re; r++) c.eL += la.L_ia[*r + 1] - la.L_ia[*r];
            by_key(Rb, Re, kU, elim, uo, kc);
            by_key(Rb, Re, key2, kept, ko, kc);
            by_key(Rb, Re, keyU2, kept_elim2, u2o, kc);
            c.nU = uint32_t(uo.size());
            c.nK = uint32_t(ko.size());
            c.nU2 = uint32_t(u2o.size());
            for (uint32_t i : uo) c.eU += la.U_ia[i + 1] - la.U_ia[i] - 1;
            for (uint32_t i : ko) {
                const uint32_t q = pos2[i - M];
                if (q < M2) c.eL2 += lb.L_ia[q + 1] - lb.L_ia[q];
            }
            for (uint32_t i : u2o) {
                const uint32_t q = pos2[i - M];
                c.eU2 += lb.U_ia[q + 1] - lb.U_ia[q] - 1;
            }
        }
    });
    std::vector<uint32_t> oU(size_t(nblk) + 1, 0), oK(size_t(nblk) + 1, 0), oU2(size_t(nblk) + 1, 0),
        oeL(size_t(nblk) + 1, 0), oeU(size_t(nblk) + 1, 0), oeL2(size_t(nblk) + 1, 0), oeU2(size_t(nblk) + 1, 0);
    for (uint32_t b = 0; b < nblk; b++) {
        oU[b + 1] = oU[b] + bc[b].nU;
        oK[b + 1] = oK[b] + bc[b].nK;
        oU2[b + 1] = oU2[b] + bc[b].nU2;
        oeL[b + 1] = oeL[b] + bc[b].eL;
        oeU[b + 1] = oeU[b] + bc[b].eU;
        oeL2[b + 1] = oeL2[b] + bc[b].eL2;
        oeU2[b + 1] = oeU2[b] + bc[b].eU2;
    }
    const uint32_t nodes = roff[nblk];
    P.qrow.resize(nodes);
    P.Ldep.resize(nodes);
    P.kept2.resize(nodes);
    P.Lptr.resize(size_t(nodes) + 1);
    P.Lloc.resize(oeL[nblk]);
    P.Lval.resize(oeL[nblk]);
    P.Udep.resize(oU[nblk]);
    P.uloc.resize(oU[nblk]);
    P.uxidx.resize(oU[nblk]);
    P.Udiag.resize(oU[nblk]);
    P.Uptr.resize(size_t(oU[nblk]) + 1);
    P.Ucol.resize(oeU[nblk]);
    P.Uval.resize(oeU[nblk]);
    P.qrow2.resize(oK[nblk]);
    P.src2.resize(oK[nblk]);
    P.row2.resize(oK[nblk]);
    P.xl2.resize(oK[nblk]);
    P.Ldep2.resize(oK[nblk]);
    P.d2.resize(oK[nblk]);
    P.Lptr2.resize(size_t(oK[nblk]) + 1);
    P.Lloc2.resize(oeL2[nblk]);
    P.Lval2.resize(oeL2[nblk]);
    P.Udep2.resize(oU2[nblk]);
    P.uloc2.resize(oU2[nblk]);
    P.Udiag2.resize(oU2[nblk]);
    P.Uptr2.resize(size_t(oU2[nblk]) + 1);
    P.Ucol2.resize(oeU2[nblk]);
    P.Uval2.resize(oeU2[nblk]);
    parallel_for(nblk, 256, [&](uint64_t b0, uint64_t b1) {
        std::vector<uint32_t> lo, uo, ko, u2o, kc;
        for (uint64_t b = b0; b < b1; b++) {
            const uint32_t* Rb = rows_all.data() + roff[b];
            const uint32_t* Re = rows_all.data() + roff[b + 1];
            by_key(Rb, Re, kL, all, lo, kc);
            by_key(Rb, Re, kU, elim, uo, kc);
            by_key(Rb, Re, key2, kept, ko, kc);
            by_key(Rb, Re, keyU2, kept_elim2, u2o, kc);
            for (uint32_t k = 0; k < lo.size(); k++) local[lo[k]] = k;
            for (uint32_t k = 0; k < ko.size(); k++) local2[ko[k]] = k;
            uint32_t* B = &P.blk[size_t(kPairFields) * b];
            B[0] = roff[b];
            B[1] = uint32_t(lo.size());
            B[2] = levels_of(lo, kL);
            B[3] = oU[b];
            B[4] = uint32_t(uo.size());
            B[5] = levels_of(uo, kU);
            B[6] = oK[b];
            B[7] = uint32_t(ko.size());
            B[8] = levels_of(ko, key2);
            B[9] = oU2[b];
            B[10] = uint32_t(u2o.size());
            B[11] = levels_of(u2o, keyU2);
            uint32_t q0 = roff[b], e = oeL[b];
            for (uint32_t i : lo) {
                P.qrow[q0] = la.map[i];
                P.Ldep[q0] = dL[i];
                P.kept2[q0] = i >= M ? local2[i] : kNoRef;
                for (uint32_t p = la.L_ia[i]; p < la.L_ia[i + 1]; p++) {
                    P.Lloc[e] = local[la.L_ja[p]];
                    P.Lval[e] = la.L_a[p];
                    e++;
                }
                P.Lptr[++q0] = e;
            }
            q0 = oU[b];
            e = oeU[b];
            for (uint32_t i : uo) {
                P.Udep[q0] = dU[i];
                P.uloc[q0] = local[i];
                P.uxidx[q0] = la.map[i];
                P.Udiag[q0] = la.U_a[la.U_ia[i]];
                for (uint32_t p = la.U_ia[i] + 1; p < la.U_ia[i + 1]; p++) {
                    const uint32_t j = la.U_ja[p];
                    P.Ucol[e] = j < M ? local[j] : (kGlobalRef | local2[j]);
                    P.Uval[e] = la.U_a[p];
                    e++;
                }
                P.Uptr[++q0] = e;
            }
            q0 = oK[b];
            e = oeL2[b];
            for (uint32_t i : ko) {
                const uint32_t r = i - M, q = pos2[r];
                P.qrow2[q0] = r;
                P.src2[q0] = local[i];
                P.row2[q0] = q < M2 ? r : (r | kIso2);
                P.xl2[q0] = la.map[i];
                P.Ldep2[q0] = q < M2 ? dL2[q] : 0u;
                P.d2[q0] = q < M2 ? 0.0 : lb.dval[q - M2];
                if (q < M2)
                    for (uint32_t p = lb.L_ia[q]; p < lb.L_ia[q + 1]; p++) {
                        P.Lloc2[e] = local2[M + lb.map[lb.L_ja[p]]];
                        P.Lval2[e] = lb.L_a[p];
                        e++;
                    }
                P.Lptr2[++q0] = e;
            }
            q0 = oU2[b];
            e = oeU2[b];
            for (uint32_t i : u2o) {
                const uint32_t q = pos2[i - M];
                P.Udep2[q0] = dU2[q];
                P.uloc2[q0] = local2[i];
                P.Udiag2[q0] = lb.U_a[lb.U_ia[q]];
                for (uint32_t p = lb.U_ia[q] + 1; p < lb.U_ia[q + 1]; p++) {
                    P.Ucol2[e] = local2[M + lb.map[lb.U_ja[p]]];
                    P.Uval2[e] = lb.U_a[p];
                    e++;
                }
                P.Uptr2[++q0] = e;
            }
        }
    });
    for (uint32_t b = 0; b < nblk; b++) {
        const uint32_t* B = &P.blk[size_t(kPairFields) * b];
        P.max_nodes = std::max<uint32_t>(P.max_nodes, B[1]);
        P.max_U = std::max<uint32_t>(P.max_U, B[4]);
        P.max_nodes2 = std::max<uint32_t>(P.max_nodes2, B[7]);
        P.max_U2 = std::max<uint32_t>(P.max_U2, B[10]);
    }
    P.nodes = uint32_t(P.Ldep.size());
    P.nU = uint32_t(P.Udep.size());
    P.nodes2 = uint32_t(P.src2.size());
    P.nU2 = uint32_t(P.Udep2.size());
    // flat kept rows, then isolated rows
    for (uint32_t i = M; i < K; i++) {
        if (is_node[i]) continue;
        const uint32_t r = i - M;
        P.qrow.push_back(la.map[i]);
        P.qrow2.push_back(r);
        P.frow2.push_back(r);
        P.fxl.push_back(la.map[i]);
        P.fd2.push_back(lb.dval[pos2[r] - M2]);
    }
    P.nflat2 = uint32_t(P.frow2.size());
    for (uint32_t k = 0; k < Md; k++) P.qrow.push_back(la.map[K + k]);
    if (P.qrow.size() != N || P.qrow2.size() != n) throw Error("pair program does not cover the levels' rows", 1);
    {
        std::vector<char> seen(N, 0);
        for (uint32_t j : P.qrow) {
            if (j >= N || seen[j]) throw Error("pair program does not cover the level's rows once", 1);
            seen[j] = 1;
        }
    }
    // no empty device arrays
    for (auto* v : {&P.Lloc, &P.Ucol, &P.Lloc2, &P.Ucol2, &P.kept2, &P.Ldep, &P.Udep, &P.uloc, &P.uxidx, &P.src2,
                    &P.row2, &P.xl2, &P.Ldep2, &P.Udep2, &P.uloc2, &P.frow2, &P.fxl, &P.blk})
        if (v->empty()) v->push_back(0);
    for (auto* v : {&P.Lval, &P.Uval, &P.Udiag, &P.Lval2, &P.d2, &P.Uval2, &P.Udiag2, &P.fd2})
        if (v->empty()) v->push_back(0.0);
    out = std::move(P);
    return true;
}

bool plan_stencil(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a, StencilLayout& S) {
    S = StencilLayout();
    if (n < 8 || ia[1] - ia[0] != 4 || ja[ia[0]] != 0) return false;
    // row 0 is a grid corner: its neighbours are +1, +nx, +nx*ny
    const uint32_t c1 = ja[1], sy = ja[2], sz = ja[3];
    if (c1 != 1 || sy < 2 || sz <= sy || sz % sy || n % sz) return false;
    const uint32_t nx = sy, ny = sz / sy, nz = n / sz;
    if (ny < 2 || nz < 2) return false;
    S.n = n; S.sy = sy; S.sz = sz;
    S.pres.resize(n); S.d.resize(n); S.lx.resize(n); S.ly.resize(n); S.lz.resize(n);
    std::atomic<bool> bad{false};
    parallel_for(n, 1 << 14, [&](uint64_t r0, uint64_t r1) {
        for (uint64_t rr = r0; rr < r1 && !bad.load(std::memory_order_relaxed); rr++) {
            const uint32_t r = uint32_t(rr), i = r % nx, j = (r / nx) % ny, k = r / sz;
            const uint32_t b = ia[r], e = ia[r + 1];
            uint32_t want[7], m = 0, cnt = 0;
            want[cnt++] = r;
            if (k > 0) { want[cnt++] = r - sz; m |= 1; }
            if (j > 0) { want[cnt++] = r - sy; m |= 2; }
            if (i > 0) { want[cnt++] = r - 1; m |= 4; }
            if (i + 1 < nx) { want[cnt++] = r + 1; m |= 8; }
            if (j + 1 < ny) { want[cnt++] = r + sy; m |= 16; }
            if (k + 1 < nz) { want[cnt++] = r + sz; m |= 32; }
            if (e - b != cnt) { bad = true; break; }
            for (uint32_t t = 0; t < cnt; t++)
                if (ja[b + t] != want[t]) { bad = true; break; }
            S.pres[r] = uint8_t(m);
            S.d[r] = a[b];
            uint32_t t = b + 1;
            S.lz[r] = (m & 1) ? a[t++] : 0.0;
            S.ly[r] = (m & 2) ? a[t++] : 0.0;
            S.lx[r] = (m & 4) ? a[t++] : 0.0;
        }
    });
    if (bad) return false;
    // symmetry: each upper value equals the neighbour's lower value, bit for bit
    auto same = [](double x, double y) { return std::memcmp(&x, &y, sizeof(double)) == 0; };
    parallel_for(n, 1 << 14, [&](uint64_t r0, uint64_t r1) {
        for (uint64_t rr = r0; rr < r1 && !bad.load(std::memory_order_relaxed); rr++) {
            const uint32_t r = uint32_t(rr), m = S.pres[r];
            uint32_t t = ia[r] + 1 + ((m & 1) != 0) + ((m & 2) != 0) + ((m & 4) != 0);
            if ((m & 8) && !same(a[t++], S.lx[r + 1])) bad = true;
            if ((m & 16) && !same(a[t++], S.ly[r + sy])) bad = true;
            if ((m & 32) && !same(a[t++], S.lz[r + sz])) bad = true;
        }
    });
    if (bad) {
        S = StencilLayout();
        return false;
    }
    return true;
}

void plan_sell(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a, SellLayout& S,
               const uint32_t* rows) {
    S = SellLayout();
    S.n = n;
    S.nslices = (n + 63) / 64;
    S.off64.assign(S.nslices + 1, 0);
    auto row = [&](uint32_t q) { return rows ? rows[q] : q; };
    std::vector<uint32_t> width(S.nslices);
    parallel_for(S.nslices, 1024, [&](uint64_t s0, uint64_t s1) {
        for (uint64_t s = s0; s < s1; s++) {
            uint32_t w = 0;
            for (uint32_t q = uint32_t(64 * s); q < std::min<uint64_t>(n, 64 * s + 64); q++)
                w = std::max(w, ia[row(q) + 1] - ia[row(q)]);
            width[s] = w;
        }
    });
    for (uint32_t s = 0; s < S.nslices; s++) S.off64[s + 1] = S.off64[s] + width[s];
    S.nent = uint64_t(S.off64[S.nslices]) * 64;
    // 16-bit column deltas when the rows are in their own order and every |j - i| fits
    bool use16 = false;
    if (!rows) {
        std::atomic<bool> wide{false};
        parallel_for(n, 1 << 15, [&](uint64_t r0, uint64_t r1) {
            for (uint64_t r = r0; r < r1 && !wide.load(std::memory_order_relaxed); r++)
                for (uint32_t p = ia[r]; p < ia[r + 1]; p++) {
                    const int64_t d = int64_t(ja[p]) - int64_t(r);
                    if (d < -32767 || d > 32767) { wide = true; break; }
                }
        });
        use16 = !wide;
    }
    const uint64_t cap = S.nent ? S.nent : 1;
    if (use16) S.dcol.resize(cap);
    else S.col.resize(cap);
    S.val.resize(cap);
    if (!S.nent) {
        if (use16) S.dcol[0] = int16_t(-32768);
        else S.col[0] = 0xffffffffu;
        S.val[0] = 0.0;
        return;
    }
    // every slot of a slice written once: the row's entries, then padding
    parallel_for(S.nslices, 256, [&](uint64_t s0, uint64_t s1) {
        for (uint64_t s = s0; s < s1; s++) {
            const uint32_t w = width[s];
            for (uint32_t t = 0; t < 64; t++) {
                const uint64_t q = 64 * s + t;
                const uint32_t r = q < n ? row(uint32_t(q)) : 0u;
                const uint32_t len = q < n ? ia[r + 1] - ia[r] : 0u;
                for (uint32_t k = 0; k < w; k++) {
                    const uint64_t e = (uint64_t(S.off64[s]) + k) * 64 + t;
                    if (k < len) {
                        const uint32_t j = ja[ia[r] + k];
                        if (use16) S.dcol[e] = int16_t(int64_t(j) - int64_t(q));
                        else S.col[e] = j;
                        S.val[e] = a[ia[r] + k];
                    } else {
                        if (use16) S.dcol[e] = int16_t(-32768);
                        else S.col[e] = 0xffffffffu;
                        S.val[e] = 0.0;
                    }
                }
            }
        }
    });
}

uint32_t max_level_component(const LevelArrays& lv) {
    const uint32_t K = lv.N - lv.Md, M = lv.M;
    DSU dsu(K);
    for (uint32_t i = 0; i < K; i++)
        for (uint32_t p = lv.L_ia[i]; p < lv.L_ia[i + 1]; p++)
            if (lv.L_ja[p] < K) dsu.unite(i, lv.L_ja[p]);
    for (uint32_t i = 0; i < M; i++)
        for (uint32_t p = lv.U_ia[i] + 1; p < lv.U_ia[i + 1]; p++)
            if (lv.U_ja[p] < K) dsu.unite(i, lv.U_ja[p]);
    std::vector<uint32_t> cnt(K, 0);
    uint32_t best = 0;
    for (uint32_t i = 0; i < K; i++)
        if (i < M || lv.L_ia[i + 1] > lv.L_ia[i]) best = std::max(best, ++cnt[dsu.find(i)]);
    return best;
}

namespace {
// The first error of a parallel loop: the one a sequential loop over the same
// positions would have met first (smallest position), raised after the loop.
struct FirstError {
    std::mutex m;
    uint64_t at = ~uint64_t(0);
    const char* msg = nullptr;
    void put(uint64_t i, const char* s) {
        std::lock_guard<std::mutex> g(m);
        if (i < at) {
            at = i;
            msg = s;
        }
    }
    void raise() const {
        if (msg) throw Error(msg, 1);
    }
};

// The x in [0, n) with keep(x), in ascending order within buckets key(x) in [0, nb), bucket 0 first (a stable
// counting sort, chunks of x counted and placed in parallel).
template <class Keep, class Key>
std::vector<uint32_t> bucket_ids(uint32_t n, uint32_t nb, const Keep& keep, const Key& key) {
    const uint64_t chunks = std::max<uint64_t>(1, std::min<uint64_t>(uint64_t(host_threads()) * 4, (n + 65535) / 65536));
    const uint64_t step = (n + chunks - 1) / chunks;
    std::vector<uint32_t> cnt(chunks * nb, 0);
    parallel_for(chunks, 1, [&](uint64_t c0, uint64_t c1) {
        for (uint64_t c = c0; c < c1; c++)
            for (uint64_t x = c * step; x < std::min<uint64_t>(n, (c + 1) * step); x++)
                if (keep(uint32_t(x))) cnt[c * nb + key(uint32_t(x))]++;
    });
    uint32_t tot = 0;
    for (uint32_t b = 0; b < nb; b++)
        for (uint64_t c = 0; c < chunks; c++) {
            const uint32_t v = cnt[c * nb + b];
            cnt[c * nb + b] = tot;
            tot += v;
        }
    std::vector<uint32_t> out(tot);
    parallel_for(chunks, 1, [&](uint64_t c0, uint64_t c1) {
        for (uint64_t c = c0; c < c1; c++) {
            uint32_t* o = &cnt[c * nb];
            for (uint64_t x = c * step; x < std::min<uint64_t>(n, (c + 1) * step); x++)
                if (keep(uint32_t(x))) out[o[key(uint32_t(x))]++] = uint32_t(x);
        }
    });
    return out;
}
}  // namespace

bool plan_chain(const std::vector<LevelArrays>& L, uint32_t a, uint32_t max_nodes, uint32_t max_slots,
                ChainProgram& out) {
    const uint32_t nl = uint32_t(L.size());
    if (nl < 2) return false;
    if (max_slots > 256 || max_nodes > max_slots) throw Error("chain blocks hold at most 256 slots", 1);
    constexpr uint64_t kGrain = 1 << 14;
    // the structure the kernels rely on (as plan_level checks it)
    for (uint32_t k = 0; k < nl; k++) {
        const LevelArrays& lv = L[k];
        const uint32_t M = lv.M, K = lv.N - lv.Md;
        if (lv.M + lv.Md > lv.N) throw Error("level sizes inconsistent", 1);
        FirstError fl, fu;
        parallel_for(K, kGrain, [&](uint64_t i0, uint64_t i1) {
            for (uint64_t i = i0; i < i1; i++)
                for (uint32_t p = lv.L_ia[i]; p < lv.L_ia[i + 1]; p++)
                    if (lv.L_ja[p] >= i || lv.L_ja[p] >= M) {
                        fl.put(i, "L row has a column outside the eliminated block");
                        return;
                    }
        });
        fl.raise();
        parallel_for(M, kGrain, [&](uint64_t i0, uint64_t i1) {
            for (uint64_t i = i0; i < i1; i++) {
                if (lv.U_ia[i + 1] <= lv.U_ia[i] || lv.U_ja[lv.U_ia[i]] != i) {
                    fu.put(i, "U row does not start with its pivot");
                    return;
                }
                for (uint32_t p = lv.U_ia[i] + 1; p < lv.U_ia[i + 1]; p++)
                    if (lv.U_ja[p] <= i || lv.U_ja[p] >= K) {
                        fu.put(i, "U row has a column outside (i, N-Md)");
                        return;
                    }
            }
        });
        fu.raise();
        if (k + 1 < nl && L[k + 1].N != lv.N - lv.M - lv.Md) throw Error("level sizes do not chain", 1);
    }
    if (L[nl - 1].M + L[nl - 1].Md != L[nl - 1].N) return false;       // the last level keeps no rows
    const uint32_t NA = L[0].N;
    // identity of every row: its natural row on level a.  perm[k][i]: identity of permuted row i of level k
    std::vector<std::vector<uint32_t>> perm(nl);
    for (uint32_t k = 0; k < nl; k++) {
        const LevelArrays& lv = L[k];
        const uint32_t* nat = k ? perm[k - 1].data() + L[k - 1].M : nullptr;   // identity of level-k row r
        perm[k].resize(lv.N);
        uint32_t* pk = perm[k].data();
        FirstError fe;
        parallel_for(lv.N, kGrain, [&](uint64_t i0, uint64_t i1) {
            for (uint64_t i = i0; i < i1; i++) {
                const uint32_t m = lv.map[i];
                if (m >= lv.N) {
                    fe.put(i, "level map out of range");
                    return;
                }
                pk[i] = nat ? nat[m] : m;
            }
        });
        fe.raise();
    }
    // deepest level of every identity; components of the couplings of all levels
    std::vector<uint8_t> deep(NA, 0);
    std::vector<char> has_node(NA, 0);
    for (uint32_t k = 0; k < nl; k++) {
        const LevelArrays& lv = L[k];
        const uint32_t K = lv.N - lv.Md;
        const uint32_t* pk = perm[k].data();
        parallel_for(lv.N, kGrain, [&](uint64_t i0, uint64_t i1) {
            for (uint64_t i = i0; i < i1; i++) {
                if (k) deep[pk[i]] = uint8_t(k);
                if (i < lv.M || (i < K && lv.L_ia[i + 1] > lv.L_ia[i])) has_node[pk[i]] = 1;
            }
        });
    }
    DSU dsu(NA);
    for (uint32_t k = 0; k < nl; k++) {
        const LevelArrays& lv = L[k];
        const uint32_t K = lv.N - lv.Md;
        const uint32_t* pk = perm[k].data();
        for (uint32_t i = 0; i < K; i++)
            for (uint32_t p = lv.L_ia[i]; p < lv.L_ia[i + 1]; p++) dsu.unite(pk[i], pk[lv.L_ja[p]]);
        for (uint32_t i = 0; i < lv.M; i++)
            for (uint32_t p = lv.U_ia[i] + 1; p < lv.U_ia[i + 1]; p++) dsu.unite(pk[i], pk[lv.U_ja[p]]);
    }
    // components with a node: their rows by identity; the others are passthrough rows.  Every parent is a
    // smaller identity, so one ascending pass resolves every root (the smallest identity of its component).
    std::vector<uint32_t> root(NA);
    for (uint32_t x = 0; x < NA; x++) root[x] = dsu.p[x] == x ? x : root[dsu.p[x]];
    std::vector<char> comp_node(NA, 0);
    for (uint32_t x = 0; x < NA; x++)
        if (has_node[x]) comp_node[root[x]] = 1;
    std::vector<uint32_t> comp_start(NA + 1, 0), comp_rows;
    for (uint32_t x = 0; x < NA; x++)
        if (comp_node[root[x]]) comp_start[root[x] + 1]++;
    for (uint32_t x = 0; x < NA; x++) comp_start[x + 1] += comp_start[x];
    comp_rows.resize(comp_start[NA]);
    {
        std::vector<uint32_t> fill(comp_start.begin(), comp_start.end() - 1);
        for (uint32_t x = 0; x < NA; x++)
            if (comp_node[root[x]]) comp_rows[fill[root[x]]++] = x;   // ascending identity within a component
    }
    // each component's deepest level (its rows are comp_rows[comp_start[r] ..])
    std::vector<uint8_t> comp_deep(NA, 0);
    parallel_for(NA, kGrain, [&](uint64_t r0, uint64_t r1) {
        for (uint64_t r = r0; r < r1; r++) {
            uint8_t d = 0;
            for (uint32_t e = comp_start[r]; e < comp_start[r + 1]; e++) d = std::max(d, deep[comp_rows[e]]);
            comp_deep[r] = d;
        }
    });
    // component roots and passthrough identities: deepest level first, then by smallest identity (the root
    // is the smallest identity of its component)
    const std::vector<uint32_t> comps = bucket_ids(
        NA, nl, [&](uint32_t x) { return comp_node[x] && root[x] == x; },
        [&](uint32_t x) { return nl - 1 - uint32_t(comp_deep[x]); });
    const std::vector<uint32_t> pt = bucket_ids(
        NA, nl, [&](uint32_t x) { return !comp_node[root[x]]; }, [&](uint32_t x) { return nl - 1 - uint32_t(deep[x]); });
    auto is_node = [&](uint32_t k, uint32_t i) {
        const LevelArrays& lv = L[k];
        return i < lv.M || (i < lv.N - lv.Md && lv.L_ia[i + 1] > lv.L_ia[i]);
    };
    const uint64_t NC = comps.size();
    // per-level node / slot counts of each component, then the blocks
    std::vector<std::vector<uint32_t>> cnodes(nl, std::vector<uint32_t>(NC, 0)), cslots(nl, std::vector<uint32_t>(NC, 0));
    {
        std::vector<uint32_t> at(NA, kNoNode);                          // identity -> permuted row (level k)
        for (uint32_t k = 0; k < nl; k++) {
            const uint32_t* pk = perm[k].data();
            parallel_for(L[k].N, kGrain, [&](uint64_t i0, uint64_t i1) {
                for (uint64_t i = i0; i < i1; i++) at[pk[i]] = uint32_t(i);
            });
            parallel_for(NC, 1024, [&](uint64_t c0, uint64_t c1) {
                for (uint64_t c = c0; c < c1; c++) {
                    const uint32_t r = comps[c];
                    uint32_t nn = 0, ns = 0;
                    for (uint32_t e = comp_start[r]; e < comp_start[r + 1]; e++) {
                        const uint32_t x = comp_rows[e];
                        if (deep[x] < k) continue;
                        ns++;
                        if (is_node(k, at[x])) nn++;
                    }
                    cnodes[k][c] = nn;
                    cslots[k][c] = ns;
                }
            });
            parallel_for(L[k].N, kGrain, [&](uint64_t i0, uint64_t i1) {
                for (uint64_t i = i0; i < i1; i++) at[pk[i]] = kNoNode;
            });
        }
    }
    std::vector<uint32_t> bstart;                                       // first component of each block
    {
        std::vector<uint32_t> bn(nl, 0), bsl(nl, 0);
        for (size_t c = 0; c < NC; c++) {
            // components are packed while every level keeps at most max_nodes nodes and max_nodes slots (one
            // slot per thread); a component with more slots than that gets a block of its own
            bool fits = !bstart.empty();
            for (uint32_t k = 0; k < nl && fits; k++)
                fits = bn[k] + cnodes[k][c] <= max_nodes && bsl[k] + cslots[k][c] <= max_nodes;
            if (!fits) {
                for (uint32_t k = 0; k < nl; k++) {
                    if (cnodes[k][c] > max_nodes || cslots[k][c] > max_slots) return false;
                    bn[k] = bsl[k] = 0;
                }
                bstart.push_back(uint32_t(c));
            }
            for (uint32_t k = 0; k < nl; k++) { bn[k] += cnodes[k][c]; bsl[k] += cslots[k][c]; }
        }
        bstart.push_back(uint32_t(NC));
    }
    const uint32_t nblk = uint32_t(bstart.size()) - 1;
    ChainProgram P;
    P.a = a;
    P.nlev = nl;
    P.nblk = nblk;
    P.lv.resize(nl);
    P.blk_from.assign(nl, 0);
    for (uint32_t b = 0; b < nblk; b++) {
        const uint32_t d = comp_deep[comps[bstart[b]]];                // the block's deepest level
        for (uint32_t k = 0; k <= d; k++) P.blk_from[k] = b + 1;
    }
    // Per level: slots of every block, then the passthrough rows.  Pass 1 orders each block's slots and U list
    // (blocks in parallel), places the blocks one after another and records every identity's slot; pass 2
    // writes the arrays (the source slots of a level are the slots of the level above), blocks in parallel
    // from their entry offsets.  The levels run in turn.
    // slot[k][i]: the block slot of level-k row i (kNoNode: a passthrough row).  Level k's rows are the kept
    // rows of level k-1: level-k row i is row M_{k-1} + map_k[i] of level k-1.
    std::vector<std::vector<uint32_t>> slot(nl), order(nl), uorder(nl), dLk(nl), dUk(nl);
    HostVec<uint32_t> at(NA);                                           // identity -> row of the current level
    for (uint32_t k = 0; k < nl; k++) {
        const LevelArrays& lv = L[k];
        ChainLevelProgram& C = P.lv[k];
        const uint32_t N = lv.N, M = lv.M, K = N - lv.Md;
        const uint32_t* pk = perm[k].data();
        C.N = N; C.M = M; C.Md = lv.Md;
        // (only identities on level k are looked up)
        parallel_for(N, kGrain, [&](uint64_t i0, uint64_t i1) {
            for (uint64_t i = i0; i < i1; i++) at[pk[i]] = uint32_t(i);
        });
        // depths (L: over eliminated columns; U: over eliminated columns, kept columns are known beforehand)
        std::vector<uint32_t>& dL = dLk[k];
        std::vector<uint32_t>& dU = dUk[k];
        dL.assign(K, 0);
        for (uint32_t i = 0; i < K; i++)
            for (uint32_t p = lv.L_ia[i]; p < lv.L_ia[i + 1]; p++) dL[i] = std::max(dL[i], dL[lv.L_ja[p]] + 1);
        dU.assign(M, 0);
        for (uint32_t i = M; i-- > 0;)
            for (uint32_t p = lv.U_ia[i] + 1; p < lv.U_ia[i + 1]; p++)
                if (lv.U_ja[p] < M) dU[i] = std::max(dU[i], dU[lv.U_ja[p]] + 1);
        C.hdr.assign(size_t(kChainHdr) * nblk, 0);
        slot[k].assign(N, kNoNode);
        std::vector<std::vector<uint32_t>> bsl(nblk), bul(nblk);     // each block's slot order and U list
        parallel_for(nblk, 16, [&](uint64_t b0, uint64_t b1) {
            std::vector<uint32_t> nodes, flat, iso, ul;
            for (uint64_t b = b0; b < b1; b++) {
                nodes.clear(); flat.clear(); iso.clear(); ul.clear();
                for (uint32_t c = bstart[b]; c < bstart[b + 1]; c++) {
                    const uint32_t r = comps[c];
                    for (uint32_t e = comp_start[r]; e < comp_start[r + 1]; e++) {
                        const uint32_t x = comp_rows[e];
                        if (deep[x] < k) continue;
                        const uint32_t i = at[x];
                        if (is_node(k, i)) nodes.push_back(i);
                        else if (i < K) flat.push_back(i);
                        else iso.push_back(i);
                        if (i < M) ul.push_back(i);
                    }
                }
                // (total orders: by depth, then by row)
                std::sort(nodes.begin(), nodes.end(), [&](uint32_t u, uint32_t v) {
                    return dL[u] != dL[v] ? dL[u] < dL[v] : u < v;
                });
                std::sort(flat.begin(), flat.end());
                std::sort(iso.begin(), iso.end());
                std::sort(ul.begin(), ul.end(), [&](uint32_t u, uint32_t v) {
                    return dU[u] != dU[v] ? dU[u] < dU[v] : u < v;
                });
                uint32_t* H = &C.hdr[size_t(kChainHdr) * b];
                H[1] = uint32_t(nodes.size() + flat.size() + iso.size());
                H[2] = uint32_t(nodes.size());
                H[3] = uint32_t(flat.size());
                uint32_t nlv = 0, nuv = 0;
                for (uint32_t i : nodes) nlv = std::max(nlv, dL[i] + 1);
                for (uint32_t i : ul) nuv = std::max(nuv, dU[i] + 1);
                H[4] = nlv;
                H[6] = uint32_t(ul.size());
                H[7] = nuv;
                std::vector<uint32_t>& o = bsl[b];
                o.reserve(H[1]);
                o.insert(o.end(), nodes.begin(), nodes.end());
                o.insert(o.end(), flat.begin(), flat.end());
                o.insert(o.end(), iso.begin(), iso.end());
                bul[b] = ul;
            }
        });
        uint32_t pos = 0, upos = 0;
        for (uint32_t b = 0; b < nblk; b++) {
            uint32_t* H = &C.hdr[size_t(kChainHdr) * b];
            H[0] = pos;
            H[5] = upos;
            pos += H[1];
            upos += H[6];
        }
        order[k].resize(pos);
        uorder[k].resize(upos);
        parallel_for(nblk, 16, [&](uint64_t b0, uint64_t b1) {
            for (uint64_t b = b0; b < b1; b++) {
                const uint32_t* H = &C.hdr[size_t(kChainHdr) * b];
                for (uint32_t sl = 0; sl < H[1]; sl++) {
                    const uint32_t i = bsl[b][sl];
                    slot[k][i] = sl;
                    order[k][H[0] + sl] = i;
                }
                std::copy(bul[b].begin(), bul[b].end(), uorder[k].begin() + H[5]);
            }
        });
        C.nslots = pos;
    }
    for (uint32_t k = 0; k < nl; k++) {
        const LevelArrays& lv = L[k];
        ChainLevelProgram& C = P.lv[k];
        const uint32_t N = lv.N, M = lv.M, K = N - lv.Md;
        const uint32_t* pk = perm[k].data();
        parallel_for(N, kGrain, [&](uint64_t i0, uint64_t i1) {
            for (uint64_t i = i0; i < i1; i++) at[pk[i]] = uint32_t(i);
        });
        const std::vector<uint32_t>& sk = slot[k];
        const std::vector<uint32_t>& dL = dLk[k];
        const std::vector<uint32_t>& dU = dUk[k];
        // passthrough rows on this level: a prefix of pt (deepest level descending)
        uint32_t npt = 0;
        while (npt < pt.size() && deep[pt[npt]] >= k) npt++;
        {
            FirstError fe;
            parallel_for(npt, kGrain, [&](uint64_t t0, uint64_t t1) {
                for (uint64_t t = t0; t < t1; t++) {
                    const uint32_t x = pt[t], i = at[x];
                    if (i < M || (i < K && lv.L_ia[i + 1] > lv.L_ia[i])) {
                        fe.put(t, "chain: a passthrough row is a node");
                        return;
                    }
                    if ((i >= K) != (deep[x] == k)) {
                        fe.put(t, "chain: a passthrough row ends without being isolated");
                        return;
                    }
                }
            });
            if (C.nslots + uint64_t(npt) != N) {
                // the blocks' rows are checked in pass 2 of a consistent plan; an inconsistent one stops here
                fe.raise();
                throw Error("chain: level rows are not covered exactly once", 1);
            }
            fe.raise();
        }
        // every block's L and U entry counts, then their offsets
        const uint32_t NU = uint32_t(uorder[k].size());
        std::vector<uint32_t> offL(nblk + 1, 0), offU(nblk + 1, 0);
        parallel_for(nblk, 64, [&](uint64_t b0, uint64_t b1) {
            for (uint64_t b = b0; b < b1; b++) {
                const uint32_t* H = &C.hdr[size_t(kChainHdr) * b];
                uint32_t nLe = 0, nUe = 0;
                for (uint32_t s = 0; s < H[2]; s++) {
                    const uint32_t i = order[k][H[0] + s];
                    nLe += lv.L_ia[i + 1] - lv.L_ia[i];
                }
                for (uint32_t u = H[5]; u < H[5] + H[6]; u++) {
                    const uint32_t i = uorder[k][u];
                    nUe += lv.U_ia[i + 1] - lv.U_ia[i] - 1;
                }
                offL[b + 1] = nLe;
                offU[b + 1] = nUe;
            }
        });
        for (uint32_t b = 0; b < nblk; b++) {
            offL[b + 1] += offL[b];
            offU[b + 1] += offU[b];
        }
        const uint32_t ns = C.nslots;
        C.qrow.resize(N);
        C.dval.resize(N);                                               // every position written below
        if (k > 0) {
            C.src.resize(ns);
            C.xl.resize(ns);
        }
        C.Ldep.resize(ns);
        C.Lptr.resize(size_t(ns) + 1);
        C.Lptr[0] = 0;
        C.Lloc.resize(offL[nblk]);
        C.Lval.resize(offL[nblk]);
        C.Udep.resize(NU);
        C.uloc.resize(NU);
        C.uxidx.resize(NU);
        C.Udiag.resize(NU);
        C.Uptr.resize(size_t(NU) + 1);
        C.Uptr[0] = 0;
        C.Ucol.resize(offU[nblk]);
        C.Uval.resize(offU[nblk]);
        C.Lmeta.resize(ns);
        C.Lv2.resize(size_t(2) * ns);
        C.Umeta.resize(NU);
        C.Uv1.resize(NU);
        FirstError fe;
        parallel_for(nblk, 16, [&](uint64_t b0, uint64_t b1) {
            for (uint64_t b = b0; b < b1; b++) {
                const uint32_t* H = &C.hdr[size_t(kChainHdr) * b];
                uint32_t cl = offL[b];
                for (uint32_t s = 0; s < H[1]; s++) {
                    const uint32_t q = H[0] + s, i = order[k][q];
                    C.qrow[q] = lv.map[i];
                    if (k > 0) {
                        const uint32_t iu = L[k - 1].M + lv.map[i];        // the row on the level above
                        const uint32_t ps = slot[k - 1][iu];
                        if (ps == kNoNode) {
                            fe.put((b << 32) | s, "chain: a row has no slot on the level above");
                            return;
                        }
                        C.src[q] = ps;
                        C.xl[q] = L[k - 1].map[iu];
                    }
                    const bool node = s < H[2];
                    C.Ldep[q] = node ? dL[i] : 0u;
                    const uint32_t l0 = cl;
                    if (node)
                        for (uint32_t p = lv.L_ia[i]; p < lv.L_ia[i + 1]; p++) {
                            C.Lloc[cl] = sk[lv.L_ja[p]];
                            C.Lval[cl] = lv.L_a[p];
                            cl++;
                        }
                    C.Lptr[q + 1] = cl;
                    const uint32_t ne = cl - l0;
                    if (C.Ldep[q] > 0xffu || ne > 0xffu) {
                        fe.put((b << 32) | s, "chain: an L depth or entry count does not fit 8 bits");
                        return;
                    }
                    C.Lmeta[q] = C.Ldep[q] | ne << 8 | (ne > 0 ? C.Lloc[l0] : 0u) << 16 | (ne > 1 ? C.Lloc[l0 + 1] : 0u) << 24;
                    C.Lv2[2 * size_t(q)] = ne > 0 ? C.Lval[l0] : 0.0;
                    C.Lv2[2 * size_t(q) + 1] = ne > 1 ? C.Lval[l0 + 1] : 0.0;
                    C.dval[q] = i >= K ? lv.dval[i - K] : 0.0;
                }
                uint32_t cu = offU[b];
                for (uint32_t u = H[5]; u < H[5] + H[6]; u++) {
                    const uint32_t i = uorder[k][u];
                    C.Udep[u] = dU[i];
                    C.uloc[u] = sk[i];
                    C.uxidx[u] = lv.map[i];
                    C.Udiag[u] = lv.U_a[lv.U_ia[i]];
                    const uint32_t u0 = cu;
                    for (uint32_t p = lv.U_ia[i] + 1; p < lv.U_ia[i + 1]; p++) {
                        C.Ucol[cu] = sk[lv.U_ja[p]];
                        C.Uval[cu] = lv.U_a[p];
                        cu++;
                    }
                    C.Uptr[u + 1] = cu;
                    const uint32_t ne = cu - u0;
                    if (C.Udep[u] > 0xffu || ne > 0xffu || C.uloc[u] > 0xffu) {
                        fe.put((b << 32) | 0xffffffffu, "chain: a U depth, entry count or slot does not fit 8 bits");
                        return;
                    }
                    C.Umeta[u] = C.Udep[u] | ne << 8 | C.uloc[u] << 16 | (ne > 0 ? C.Ucol[u0] : 0u) << 24;
                    C.Uv1[u] = ne > 0 ? C.Uval[u0] : 0.0;
                }
            }
        });
        fe.raise();
        parallel_for(npt, kGrain, [&](uint64_t t0, uint64_t t1) {
            for (uint64_t t = t0; t < t1; t++) {
                const uint32_t q = ns + uint32_t(t), i = at[pt[t]];
                C.qrow[q] = lv.map[i];
                C.dval[q] = i >= K ? lv.dval[i - K] : 0.0;
            }
        });
        C.npt = npt;
        {
            HostVec<uint8_t> seen(N);
            parallel_for(N, kGrain, [&](uint64_t q0, uint64_t q1) { std::fill(&seen[q0], &seen[q0] + (q1 - q0), 0); });
            std::atomic<bool> bad{false};
            parallel_for(N, kGrain, [&](uint64_t q0, uint64_t q1) {
                for (uint64_t q = q0; q < q1; q++) {
                    const uint32_t j = C.qrow[q];
                    if (j >= N || __atomic_exchange_n(&seen[j], uint8_t(1), __ATOMIC_RELAXED)) {
                        bad = true;
                        return;
                    }
                }
            });
            if (bad) throw Error("chain: level rows are not covered exactly once", 1);
        }
        // no empty device arrays
        for (auto* v : {&C.src, &C.xl, &C.Lloc, &C.Udep, &C.uloc, &C.uxidx, &C.Ucol, &C.hdr, &C.Lmeta, &C.Umeta})
            if (v->empty()) v->push_back(0);
        for (auto* v : {&C.Lval, &C.Uval, &C.Udiag, &C.Lv2, &C.Uv1})
            if (v->empty()) v->push_back(0.0);
    }
    for (uint32_t k = 0; k < nl; k++)
        for (uint32_t b = 0; b < nblk; b++) {
            const uint32_t* H = &P.lv[k].hdr[size_t(kChainHdr) * b];
            P.max_slots = std::max(P.max_slots, H[1]);
            P.max_nodes = std::max(P.max_nodes, H[2]);
        }
    // index checks: every local index inside its block, every position inside its level
    for (uint32_t k = 0; k < nl; k++) {
        const ChainLevelProgram& C = P.lv[k];
        FirstError fe;
        parallel_for(nblk, 64, [&](uint64_t b0, uint64_t b1) {
            for (uint64_t b = b0; b < b1; b++) {
                auto bad = [&](const char* m) { fe.put(b, m); };
                const uint32_t* H = &C.hdr[size_t(kChainHdr) * b];
                const uint32_t base = H[0], S = H[1], nn = H[2];
                if (nn > S || H[2] + H[3] > S || base + S > C.nslots) return bad("chain: block header out of range");
                for (uint32_t s = 0; s < nn; s++) {
                    if (C.Ldep[base + s] >= H[4]) return bad("chain: L depth out of range");
                    for (uint32_t p = C.Lptr[base + s]; p < C.Lptr[base + s + 1]; p++)
                        if (C.Lloc[p] >= s) return bad("chain: L column not earlier in its block");
                }
                for (uint32_t s = nn; s < S; s++)
                    if (C.Lptr[base + s + 1] != C.Lptr[base + s]) return bad("chain: a non-node slot has L entries");
                if (k > 0) {
                    const uint32_t* Hp = &P.lv[k - 1].hdr[size_t(kChainHdr) * b];
                    for (uint32_t s = 0; s < S; s++)
                        if (C.src[base + s] >= Hp[1] || C.src[base + s] >= Hp[2] + Hp[3] ||
                            C.xl[base + s] >= P.lv[k - 1].N)
                            return bad("chain: a row's source on the level above is not a kept slot");
                }
                for (uint32_t u = H[5]; u < H[5] + H[6]; u++) {
                    if (C.uloc[u] >= nn || C.Udep[u] >= H[7]) return bad("chain: U-list entry out of range");
                    for (uint32_t p = C.Uptr[u]; p < C.Uptr[u + 1]; p++)
                        if (C.Ucol[p] >= H[2] + H[3] || C.Ucol[p] == C.uloc[u]) return bad("chain: U column out of range");
                }
                if (S > max_slots || nn > max_nodes || H[6] > nn) return bad("chain: block too large");
            }
        });
        fe.raise();
        if (k > 0 && C.npt > P.lv[k - 1].npt) throw Error("chain: passthrough rows do not nest", 1);
    }
    out = std::move(P);
    return true;
}

}  // namespace mlp

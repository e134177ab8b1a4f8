#include "plan.h"

#include "mlp_error.h"

#include <algorithm>
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

}  // namespace

void plan_tail(const LevelArrays& lv, uint32_t cap, TailLayout& T) {
    const uint32_t M = lv.M;
    T = TailLayout();
    T.M = M;
    T.pos_of_row.assign(M, 0);
    T.Lptr.assign(M + 1, 0);
    T.Uptr.assign(M + 1, 0);
    if (M == 0) return;

    // validate the structure the kernels rely on
    for (uint32_t i = 0; i < M; i++) {
        for (uint32_t p = lv.L_ia[i]; p < lv.L_ia[i + 1]; p++)
            if (lv.L_ja[p] >= i) throw Error("L row " + std::to_string(i) + " is not strictly lower", 1);
        if (lv.U_ia[i + 1] <= lv.U_ia[i] || lv.U_ja[lv.U_ia[i]] != i)
            throw Error("U row " + std::to_string(i) + " does not start with its pivot", 1);
        for (uint32_t p = lv.U_ia[i] + 1; p < lv.U_ia[i + 1]; p++)
            if (lv.U_ja[p] <= i || lv.U_ja[p] >= lv.N - lv.Md)
                throw Error("U row " + std::to_string(i) + " has a column outside (i, N-Md)", 1);
    }

    // dependency components among the eliminated rows
    DSU dsu(M);
    std::vector<uint32_t> dL(M, 0), dU(M, 0);
    for (uint32_t i = 0; i < M; i++)
        for (uint32_t p = lv.L_ia[i]; p < lv.L_ia[i + 1]; p++) {
            dsu.unite(i, lv.L_ja[p]);
            dL[i] = std::max(dL[i], dL[lv.L_ja[p]] + 1);
        }
    for (uint32_t i = M; i-- > 0;)
        for (uint32_t p = lv.U_ia[i] + 1; p < lv.U_ia[i + 1]; p++) {
            uint32_t j = lv.U_ja[p];
            if (j < M) {
                dsu.unite(i, j);
                dU[i] = std::max(dU[i], dU[j] + 1);
            }
        }
    // components in order of their first row
    std::vector<uint32_t> comp_id(M), root_to_comp(M, 0xffffffffu);
    uint32_t ncomp = 0;
    for (uint32_t i = 0; i < M; i++) {
        uint32_t r = dsu.find(i);
        if (root_to_comp[r] == 0xffffffffu) root_to_comp[r] = ncomp++;
        comp_id[i] = root_to_comp[r];
    }
    std::vector<uint32_t> csize(ncomp, 0);
    for (uint32_t i = 0; i < M; i++) csize[comp_id[i]]++;
    uint32_t maxc = *std::max_element(csize.begin(), csize.end());
    if (maxc > cap) cap = maxc;
    // greedy packing of consecutive components into blocks of at most cap rows
    std::vector<uint32_t> comp_blk(ncomp);
    uint32_t nblk = 0, fill = 0;
    for (uint32_t c = 0; c < ncomp; c++) {
        if (c == 0 || fill + csize[c] > cap) { nblk++; fill = 0; }
        comp_blk[c] = nblk - 1;
        fill += csize[c];
    }
    std::vector<std::vector<uint32_t>> rows(nblk);
    for (uint32_t i = 0; i < M; i++) rows[comp_blk[comp_id[i]]].push_back(i);

    T.nblk = nblk;
    T.blk.resize(6 * size_t(nblk));
    T.fidx.resize(M);
    T.uloc.resize(M);
    T.uxidx.resize(M);
    T.Udiag.resize(M);
    std::vector<uint32_t> local(M);
    uint32_t pos = 0, upos = 0;
    for (uint32_t b = 0; b < nblk; b++) {
        auto& R = rows[b];
        const uint32_t base = pos;
        std::vector<uint32_t> lo = R, uo = R;
        std::stable_sort(lo.begin(), lo.end(), [&](uint32_t x, uint32_t y) { return dL[x] < dL[y]; });
        std::stable_sort(uo.begin(), uo.end(), [&](uint32_t x, uint32_t y) { return dU[x] < dU[y]; });
        for (uint32_t k = 0; k < lo.size(); k++) {
            local[lo[k]] = k;
            T.pos_of_row[lo[k]] = base + k;
        }
        T.blk[6 * b + 0] = base;
        T.blk[6 * b + 2] = uint32_t(T.lvlL.size());
        for (uint32_t k = 0; k < lo.size(); k++) {
            uint32_t i = lo[k];
            if (k + 1 == lo.size() || dL[lo[k + 1]] != dL[i]) T.lvlL.push_back(base + k + 1);
            T.fidx[base + k] = lv.map[i];
            for (uint32_t p = lv.L_ia[i]; p < lv.L_ia[i + 1]; p++) {
                T.Lloc.push_back(local[lv.L_ja[p]]);
                T.Lval.push_back(lv.L_a[p]);
            }
            T.Lptr[base + k + 1] = uint32_t(T.Lloc.size());
        }
        T.blk[6 * b + 3] = uint32_t(T.lvlL.size());
        T.blk[6 * b + 4] = uint32_t(T.lvlU.size());
        for (uint32_t k = 0; k < uo.size(); k++) {
            uint32_t i = uo[k];
            if (k + 1 == uo.size() || dU[uo[k + 1]] != dU[i]) T.lvlU.push_back(upos + k + 1);
            T.uloc[upos + k] = local[i];
            T.uxidx[upos + k] = lv.map[i];
            T.Udiag[upos + k] = lv.U_a[lv.U_ia[i]];
            for (uint32_t p = lv.U_ia[i] + 1; p < lv.U_ia[i + 1]; p++) {
                uint32_t j = lv.U_ja[p];
                T.Ucol.push_back(j < M ? local[j] : (kGlobalRef | lv.map[j]));
                T.Uval.push_back(lv.U_a[p]);
            }
            T.Uptr[upos + k + 1] = uint32_t(T.Ucol.size());
        }
        T.blk[6 * b + 5] = uint32_t(T.lvlU.size());
        pos += uint32_t(lo.size());
        upos += uint32_t(uo.size());
        T.blk[6 * b + 1] = pos;
        T.max_nodes = std::max<uint32_t>(T.max_nodes, uint32_t(lo.size()));
        T.max_levels = std::max<uint32_t>(T.max_levels, std::max(T.blk[6 * b + 3] - T.blk[6 * b + 2],
                                                                  T.blk[6 * b + 5] - T.blk[6 * b + 4]));
    }
    if (T.Lloc.empty()) { T.Lloc.push_back(0); T.Lval.push_back(0.0); }
    if (T.Ucol.empty()) { T.Ucol.push_back(0); T.Uval.push_back(0.0); }
}

void plan_schur(const LevelArrays& lv, const TailLayout& T, SchurLayout& S) {
    const uint32_t n = lv.N - lv.M - lv.Md;
    S = SchurLayout();
    S.ptr.assign(n + 1, 0);
    for (uint32_t r = 0; r < n; r++) {
        uint32_t i = lv.M + r;
        for (uint32_t p = lv.L_ia[i]; p < lv.L_ia[i + 1]; p++) {
            uint32_t k = lv.L_ja[p];
            if (k >= lv.M) throw Error("Schur row " + std::to_string(i) + " references a kept row", 1);
            S.col.push_back(T.pos_of_row[k]);
            S.val.push_back(lv.L_a[p]);
        }
        S.ptr[r + 1] = uint32_t(S.col.size());
    }
    if (S.col.empty()) { S.col.push_back(0); S.val.push_back(0.0); }
}

void plan_sell(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a, SellLayout& S) {
    S = SellLayout();
    S.n = n;
    S.nslices = (n + 63) / 64;
    S.off64.assign(S.nslices + 1, 0);
    for (uint32_t s = 0; s < S.nslices; s++) {
        uint32_t w = 0;
        for (uint32_t r = 64 * s; r < std::min(n, 64 * s + 64); r++) w = std::max(w, ia[r + 1] - ia[r]);
        S.off64[s + 1] = S.off64[s] + w;
    }
    S.nent = uint64_t(S.off64[S.nslices]) * 64;
    S.col.assign(S.nent ? S.nent : 1, 0xffffffffu);
    S.val.assign(S.nent ? S.nent : 1, 0.0);
    for (uint32_t r = 0; r < n; r++) {
        uint32_t s = r / 64, t = r % 64;
        for (uint32_t k = 0; k < ia[r + 1] - ia[r]; k++) {
            uint64_t e = (uint64_t(S.off64[s]) + k) * 64 + t;
            S.col[e] = ja[ia[r] + k];
            S.val[e] = a[ia[r] + k];
        }
    }
}

}  // namespace mlp

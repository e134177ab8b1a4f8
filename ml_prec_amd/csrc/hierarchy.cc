// Host construction of the multilevel hierarchies (FP64, reference-exact).
//
// uh_cheb:      modules/prec/cheb/cheb_constr.cc:38-248, cheb_mark.cc:8-92, cheb_misc.h:15-96
// multi_split:  modules/prec/multi_split/multi_split_constr.cc:39-299, multi_split_mark.cc:8-252
// sparse LU:    modules/prec/misc/sparse_lu.cc:6-133
//
// The integer decisions (which links drop, the tail chains, the permutation,
// the sparsity of L, U and the next level) and the floating-point values are
// the reference's: the same comparisons in the same order and the same
// arithmetic sequence per entry (this file is compiled without FMA contraction).
#include "hierarchy.h"

#include "mlp_error.h"
#include "par.h"
#include "../cpp/config_file.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <numeric>
#include <sstream>

namespace mlp {

namespace {

constexpr uint32_t kNone = 0xffffffffu;

// Position of column j in Skyline row i, or kNone (modules/matrix/skyline.cc:15-28).
inline uint32_t skyline_find(const HostCSR& A, uint32_t i, uint32_t j) {
    if (i == j) return A.ia[i];
    auto b = A.ja.begin() + A.ia[i] + 1, e = A.ja.begin() + A.ia[i + 1];
    auto it = std::lower_bound(b, e, j);
    return (it != e && *it == j) ? uint32_t(it - A.ja.begin()) : kNone;
}

// Off-diagonal positions of one row ordered by |a|, ties in storage order
// (the stable insertion sort psort, cheb_constr.cc:26-36).
void order_by_magnitude(const double* a, uint32_t n, std::vector<uint32_t>& idx) {
    idx.resize(n);
    for (uint32_t t = 0; t < n; t++) {
        const double v = std::fabs(a[t]);
        uint32_t j = t;
        for (; j > 0 && std::fabs(a[idx[j - 1]]) > v; j--) idx[j] = idx[j - 1];
        idx[j] = t;
    }
}

// Link state for uh_cheb: a removal counter per upper-triangular entry that
// starts at 2 and reaches 0 once both end points agreed (cheb_misc.h:45-96).
// Lower entries point at their upper twin.
struct TwoSidedLinks {
    const HostCSR* A;
    HostVec<uint32_t> cnt;          // upper entries: counter; lower entries: position of the twin
    HostVec<uint32_t> utwin;        // upper entries: position of the twin (host construction only)
    explicit TwoSidedLinks(const HostCSR& M) : A(&M), cnt(M.nnz()), utwin(M.nnz()) {
        std::atomic<bool> asym{false};
        parallel_for(M.nrow, 4096, [&](uint64_t b, uint64_t e) {
            for (uint32_t i = uint32_t(b); i < e; i++) {
                cnt[M.ia[i]] = 2u;
                for (uint32_t p = M.ia[i] + 1; p < M.ia[i + 1]; p++) {
                    const uint32_t j = M.ja[p];
                    if (j > i) {                                       // its twin's row writes utwin[p]
                        cnt[p] = 2u;
                        continue;
                    }
                    const uint32_t t = skyline_find(M, j, i);           // the twin entry (j, i)
                    if (t == kNone) {
                        asym = true;
                        continue;
                    }
                    cnt[p] = t;
                    utwin[t] = p;
                }
            }
        });
        if (asym) throw Error("uh_cheb needs a structurally symmetric matrix", 1);
        // every upper entry must have been found as a twin: a structurally
        // symmetric matrix has as many lower as upper entries, each lower entry
        // found exactly one upper twin, so equal counts settle it
        std::atomic<uint64_t> lower{0}, upper{0};
        parallel_for(M.nrow, 4096, [&](uint64_t b, uint64_t e) {
            uint64_t lo = 0, up = 0;
            for (uint32_t i = uint32_t(b); i < e; i++)
                for (uint32_t p = M.ia[i] + 1; p < M.ia[i + 1]; p++) (M.ja[p] < i ? lo : up)++;
            lower += lo;
            upper += up;
        });
        if (lower != upper) throw Error("uh_cheb needs a structurally symmetric matrix", 1);
    }
    // counters and twin positions given (host checks of the tail walk)
    TwoSidedLinks(const HostCSR& M, HostVec<uint32_t>&& c) : A(&M), cnt(std::move(c)) {}
    uint32_t slot(uint32_t row, uint32_t p) const { return A->ja[p] < row ? cnt[p] : p; }
    bool present(uint32_t row, uint32_t p) const { return cnt[slot(row, p)] != 0; }
    bool mark(uint32_t row, uint32_t p) { return --cnt[slot(row, p)] == 0; }   // true: removed from both ends
    void remove(uint32_t i, uint32_t j) {
        if (i > j) std::swap(i, j);
        cnt[skyline_find(*A, i, j)] = 0;
    }
};

// Link state for multi_split: one flag per stored entry (multi_split_misc.h:10-73).
struct OneSidedLinks {
    const HostCSR* A;
    std::vector<unsigned char> on;
    explicit OneSidedLinks(const HostCSR& M) : A(&M), on(M.nnz(), 1) {}
    bool present(uint32_t /*row*/, uint32_t p) const { return on[p] != 0; }
    int present_ij(uint32_t i, uint32_t j) const {          // 1 present, 0 removed, -1 absent
        uint32_t p = skyline_find(*A, i, j);
        return p == kNone ? -1 : (on[p] ? 1 : 0);
    }
};

// Sorted index set used as the row buffer of the elimination (include/svector.h semantics).
struct SortedSet {
    std::vector<uint32_t> v;
    void clear() { v.clear(); }
    void insert(uint32_t x) { v.insert(std::lower_bound(v.begin(), v.end(), x), x); }
    size_t lower(uint32_t x) const { return size_t(std::lower_bound(v.begin(), v.end(), x) - v.begin()); }
    size_t upper(uint32_t x) const { return size_t(std::upper_bound(v.begin(), v.end(), x) - v.begin()); }
};

// Sparse elimination of the first M permuted rows, producing L, U and the
// Schur complement on the next M..N-Md-1 rows (sparse_lu.cc:6-133).  Row i
// only reads U rows k < min(i, M): the eliminated rows run in order, then the
// kept rows, which read the finished U only, run in parallel chunks whose L
// and Schur-complement rows are stitched together in row order.
template <class Links>
void eliminate(const HostCSR& A, const std::vector<uint32_t>& map, const std::vector<uint32_t>& rmap,
               uint32_t Md, uint32_t M, const Links& links, double beta, const std::vector<double>& aux,
               HostCSR& next, HostCSR& U, HostCSR& L) {
    const uint32_t N = A.nrow, nkeep = N - Md, n = N - M - Md;
    L = HostCSR();  L.nrow = L.ncol = nkeep;
    U = HostCSR();  U.nrow = M; U.ncol = nkeep;
    next = HostCSR(); next.nrow = next.ncol = n;
    U.ja.reserve(3 * size_t(M)); U.a.reserve(3 * size_t(M));

    // one row: its L entries (lja, la), then its U row (i < M) or its Schur-complement row
    struct Work {
        std::vector<int32_t> slot;
        SortedSet cols;
        std::vector<double> vals;
        std::vector<std::pair<uint32_t, double>> links;     // the direct path's (column, a/beta), sorted
    };
    // The direct path for the common row: every eliminated column k < min(i, M)
    // it links to has the U row {k, i}, as a tail's does, so elimination only
    // lowers the diagonal.  It performs the generic path's operations in the
    // same order (the diagonal minus the links' a/beta in storage order, then
    // minus l_ik u_ki in ascending k) and returns false, untouched, otherwise.
    auto row_direct = [&](uint32_t i, Work& w, auto& lja, auto& la, auto& oja, auto& oa) -> bool {
        const uint32_t r = map[i], lim = std::min(i, M);
        auto& lk = w.links;
        lk.clear();
        double d = aux[r];
        for (uint32_t p = A.ia[r] + 1; p < A.ia[r + 1]; p++) {
            if (!links.present(r, p)) continue;
            const uint32_t c = rmap[A.ja[p]];
            if (c < lim && !(U.ia[c + 1] - U.ia[c] == 2 && U.ja[U.ia[c] + 1] == i)) return false;
            const double z = A.a[p] / beta;
            d -= z;
            lk.emplace_back(c, z);
        }
        for (size_t t = 1; t < lk.size(); t++)                   // columns ascending (a handful of links)
            for (size_t u = t; u > 0 && lk[u - 1].first > lk[u].first; u--) std::swap(lk[u - 1], lk[u]);
        size_t t = 0;
        for (; t < lk.size() && lk[t].first < lim; t++) {
            const uint32_t k = lk[t].first;
            const double lik = lk[t].second / U.a[U.ia[k]];
            lja.push_back(k);
            la.push_back(lik);
            d -= lik * U.a[U.ia[k] + 1];
        }
        if (i < M) {
            oja.push_back(i);
            oa.push_back(d);
            for (; t < lk.size(); t++) {
                oja.push_back(lk[t].first);
                oa.push_back(lk[t].second);
            }
        } else {
            oja.push_back(i - M);
            oa.push_back(d);
            for (; t < lk.size(); t++)
                if (lk[t].first >= M) {
                    oja.push_back(lk[t].first - M);
                    oa.push_back(lk[t].second);
                }
        }
        return true;
    };
    auto row = [&](uint32_t i, Work& w, auto& lja, auto& la, auto& oja, auto& oa) {
        if (row_direct(i, w, lja, la, oja, oa)) return;
        if (w.slot.empty()) w.slot.assign(N, -1);         // only rows that need the generic path pay for it
        for (uint32_t c : w.cols.v) w.slot[c] = -1;
        w.cols.clear();
        w.vals.clear();
        const uint32_t r = map[i];
        w.slot[i] = 0;
        w.cols.insert(i);
        w.vals.push_back(aux[r]);                          // diagonal: c_i - sum(kept a_ij / beta)
        for (uint32_t p = A.ia[r] + 1; p < A.ia[r + 1]; p++) {
            if (!links.present(r, p)) continue;
            uint32_t c = rmap[A.ja[p]];
            w.slot[c] = int32_t(w.vals.size());
            w.cols.insert(c);
            double z = A.a[p] / beta;
            w.vals.push_back(z);
            w.vals[0] -= z;
        }
        const uint32_t lim = std::min(i, M);
        uint32_t k = w.cols.v.front();
        while (k < lim) {
            const double piv = U.a[U.ia[k]];
            const double lik = w.vals[w.slot[k]] / piv;
            lja.push_back(k);
            la.push_back(lik);
            for (uint32_t p = U.ia[k] + 1; p < U.ia[k + 1]; p++) {
                uint32_t c = U.ja[p];
                if (w.slot[c] != -1) {
                    w.vals[w.slot[c]] -= lik * U.a[p];
                } else {
                    w.slot[c] = int32_t(w.vals.size());
                    w.cols.insert(c);
                    w.vals.push_back(-lik * U.a[p]);
                }
            }
            k = w.cols.v[w.cols.upper(k)];
        }
        if (i < M) {
            for (size_t t = w.cols.lower(i); t < w.cols.v.size(); t++) {
                oja.push_back(w.cols.v[t]);
                oa.push_back(w.vals[w.slot[w.cols.v[t]]]);
            }
        } else {
            oja.push_back(i - M);
            oa.push_back(w.vals[0]);
            for (size_t t = w.cols.lower(M); t < w.cols.v.size(); t++) {
                uint32_t c = w.cols.v[t];
                if (c == i) continue;
                oja.push_back(c - M);
                oa.push_back(w.vals[w.slot[c]]);
            }
        }
    };

    // eliminated rows, in order (each reads the U rows before it)
    {
        Work w;
        for (uint32_t i = 0; i < M; i++) {
            row(i, w, L.ja, L.a, U.ja, U.a);
            L.ia.push_back(uint32_t(L.ja.size()));
            U.ia.push_back(uint32_t(U.ja.size()));
        }
    }
    // kept rows in parallel chunks
    struct Chunk {
        std::vector<uint32_t> lja, nja, lcnt, ncnt;
        std::vector<double> la, na;
    };
    const uint32_t nch = std::max<uint32_t>(1, std::min<uint32_t>(64, (n + 8191) / 8192));
    std::vector<Chunk> ch(nch);
    const uint64_t step = (uint64_t(n) + nch - 1) / nch;
    parallel_for(nch, 1, [&](uint64_t c0, uint64_t c1) {
        Work w;
        for (uint64_t c = c0; c < c1; c++) {
            Chunk& C = ch[c];
            const uint32_t b = M + uint32_t(std::min<uint64_t>(n, c * step)), e = M + uint32_t(std::min<uint64_t>(n, (c + 1) * step));
            for (uint32_t i = b; i < e; i++) {
                const size_t l0 = C.lja.size(), n0 = C.nja.size();
                row(i, w, C.lja, C.la, C.nja, C.na);
                C.lcnt.push_back(uint32_t(C.lja.size() - l0));
                C.ncnt.push_back(uint32_t(C.nja.size() - n0));
            }
        }
    });
    // stitch the chunks in row order: offsets first, then every chunk copies its
    // rows into place on its own (the destination arrays are not zero-filled)
    std::vector<size_t> roff(nch + 1, 0), loff(nch + 1, 0), noff(nch + 1, 0);
    for (uint32_t c = 0; c < nch; c++) {
        roff[c + 1] = roff[c] + ch[c].lcnt.size();
        loff[c + 1] = loff[c] + ch[c].lja.size();
        noff[c + 1] = noff[c] + ch[c].nja.size();
    }
    const size_t l0 = L.ja.size(), r0 = L.ia.size();
    L.ja.resize(l0 + loff[nch]);
    L.a.resize(l0 + loff[nch]);
    L.ia.resize(r0 + roff[nch]);
    next.ja.resize(noff[nch]);
    next.a.resize(noff[nch]);
    next.ia.resize(1 + roff[nch]);
    parallel_for(nch, 1, [&](uint64_t c0, uint64_t c1) {
        for (uint64_t c = c0; c < c1; c++) {
            const Chunk& C = ch[c];
            std::copy(C.lja.begin(), C.lja.end(), L.ja.begin() + (l0 + loff[c]));
            std::copy(C.la.begin(), C.la.end(), L.a.begin() + (l0 + loff[c]));
            std::copy(C.nja.begin(), C.nja.end(), next.ja.begin() + noff[c]);
            std::copy(C.na.begin(), C.na.end(), next.a.begin() + noff[c]);
            uint32_t lp = uint32_t(l0 + loff[c]), np = uint32_t(noff[c]);
            for (size_t t = 0; t < C.lcnt.size(); t++) L.ia[r0 + roff[c] + t] = lp += C.lcnt[t];
            for (size_t t = 0; t < C.ncnt.size(); t++) next.ia[1 + roff[c] + t] = np += C.ncnt[t];
        }
    });
}

std::vector<uint32_t> inverse(const std::vector<uint32_t>& map) {
    std::vector<uint32_t> r(map.size());
    parallel_for(map.size(), 1 << 16, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) r[map[i]] = uint32_t(i);
    });
    return r;
}

// ------------------------------- uh_cheb ---------------------------------

struct ChebBuilder {
    Hierarchy& H;
    bool use_tails;
    uint32_t max_levels;

    // Tails first (chains of nodes left with one link), isolated nodes last
    // (cheb_mark.cc:8-92).
    void permutation(const HostCSR& A, const TwoSidedLinks& links, std::vector<int>& deg, HostLevel& li) {
        const uint32_t N = A.nrow;
        // links the tail walks remove, flagged by upper-entry slot (a byte per
        // entry instead of a copy of the link counters)
        std::vector<unsigned char> gone(A.nnz(), 0);
        auto present = [&](uint32_t r, uint32_t p) {
            const uint32_t u = links.slot(r, p);
            return links.cnt[u] != 0 && !gone[u];
        };
        std::vector<uint32_t>& map = li.map;
        map.assign(N, 0);
        std::vector<char> done(N, 0);
        uint32_t pos = 0;
        if (use_tails) {
            uint32_t nb = kNone;
            for (uint32_t s = 0; s < N; s++) {
                if (deg[s] != 1) continue;
                uint32_t cur = s;
                do {
                    map[pos++] = cur;
                    done[cur] = 1;
                    uint32_t u = kNone;                     // the link's upper-entry slot
                    for (uint32_t p = A.ia[cur] + 1; p < A.ia[cur + 1]; p++)
                        if (present(cur, p)) { nb = A.ja[p]; u = links.slot(cur, p); break; }
                    if (u == kNone) {                       // degree and link counters disagree (cheb_mark.cc:45)
                        std::ostringstream os;
                        os << "Row " << cur << ": could not find remaining link";
                        throw Error(os.str());
                    }
                    gone[u] = 1;
                    deg[cur] = -1;
                    deg[nb]--;
                    cur = nb;
                } while (deg[cur] == 1);
                if (deg[cur] == 0) {
                    map[pos++] = cur;
                    done[cur] = 1;
                    deg[cur] = -1;
                }
            }
        }
        li.M = pos;
        uint32_t back = N - 1;
        for (uint32_t i = 0; i < N; i++) {
            if (done[i]) continue;
            if (deg[i] == 0) map[back--] = i;
            else map[pos++] = i;
        }
        li.Md = N - pos;
    }

    void level(uint32_t l, const HostCSR& A) {
        HostLevel& li = H.levels[l];
        li.N = A.nrow;
        li.nnz = A.nnz();
        const uint32_t N = li.N;
        std::vector<int> deg(N);
        li.aux.resize(N);
        TwoSidedLinks links(A);
        // drop weak links against the row budget (cheb_constr.cc:62-103, to_remove cheb_prec.h:71-99).
        // A row's choice depends only on its own values, so the rows choose in
        // parallel (marked[p]: row i asks to drop the link at its entry p); a
        // link is removed once both of its rows asked, which lowers both degrees.
        HostVec<unsigned char> marked(A.nnz());                    // every row clears its own entries
        parallel_for(N, 4096, [&](uint64_t r0, uint64_t r1) {
            std::vector<uint32_t> order;
            for (uint32_t i = uint32_t(r0); i < r1; i++) {
                const uint32_t b = A.ia[i], e = A.ia[i + 1], nrz = e - b - 1;
                double c = 0.0;
                for (uint32_t p = b; p < e; p++) c += A.a[p];
                li.aux[i] = c;
                for (uint32_t p = b; p < e; p++) marked[p] = 0;
                order_by_magnitude(A.a.data() + b + 1, nrz, order);
                double s = 0.0;
                for (uint32_t t = 0; t < nrz; t++) {
                    const uint32_t p = b + 1 + order[t];
                    const double v = A.a[p];
                    const double share = -2 * v / (c * (li.beta - 1));
                    if (s + share > 1) break;
                    s += share;
                    marked[p] = 1;
                }
            }
        });
        parallel_for(N, 4096, [&](uint64_t r0, uint64_t r1) {
            for (uint32_t i = uint32_t(r0); i < r1; i++) {
                int d = int(A.ia[i + 1] - A.ia[i] - 1);
                for (uint32_t p = A.ia[i] + 1; p < A.ia[i + 1]; p++) {
                    const uint32_t u = links.slot(i, p);          // the upper entry of the link
                    const uint32_t twin = u == p ? links.utwin[p] : u;
                    const uint32_t m = uint32_t(marked[p]) + uint32_t(marked[twin]);
                    if (m == 2) d--;
                    if (u == p) links.cnt[p] = 2u - m;
                }
                deg[i] = d;
            }
        });
        finish(l, A, links, deg);
    }

    // tails, permutation, elimination and the next level (cheb_constr.cc:104-161)
    void finish(uint32_t l, const HostCSR& A, TwoSidedLinks& links, std::vector<int>& deg) {
        HostLevel& li = H.levels[l];
        const uint32_t N = li.N;
        permutation(A, links, deg, li);
        li.rmap = inverse(li.map);
        HostLevel& ln = H.levels[l + 1];
        eliminate(A, li.map, li.rmap, li.Md, li.M, links, li.beta, li.aux, ln.A, li.U, li.L);
        li.dval.resize(li.Md);
        for (uint32_t i = 0; i < li.Md; i++) li.dval[i] = 1. / li.aux[li.map[i + (N - li.Md)]];
        if (ln.A.nrow) {
            if (l + 1 >= max_levels) {
                std::ostringstream os;
                os << "Too many levels: " << l + 1;
                throw Error(os.str());
            }
            level(l + 1, ln.A);
        } else {
            H.nlevels = l + 1;
        }
    }
};

// ----------------------------- multi_split --------------------------------

struct MultiSplitBuilder {
    Hierarchy& H;
    const MultiSplitParams& P;
    uint32_t coarse_n;

    void order_none(uint32_t N, const std::vector<int>& din, const std::vector<int>& dout, HostLevel& li) {
        li.map.assign(N, 0);
        uint32_t pos = 0, back = N - 1;
        li.M = 0;
        for (uint32_t i = 0; i < N; i++) {
            if (din[i] == 0 && dout[i] == 0) li.map[back--] = i;
            else li.map[pos++] = i;
        }
        li.Md = N - pos;
    }

    void order_original(const HostCSR& A, const OneSidedLinks& links_in, std::vector<int>& din,
                        std::vector<int>& dout, HostLevel& li) {
        const uint32_t N = A.nrow;
        OneSidedLinks links = links_in;
        li.map.assign(N, 0);
        std::vector<char> done(N, 0);
        uint32_t pos = 0;
        if (P.use_tails) {
            for (uint32_t s = 0; s < N; s++) {
                if (!(dout[s] == 1 && din[s] <= 1)) continue;
                uint32_t cur = s;
                while (dout[cur] == 1) {
                    uint32_t p = A.ia[cur] + 1;
                    while (p < A.ia[cur + 1] && !links.present(cur, p)) p++;
                    if (p == A.ia[cur + 1]) {
                        std::ostringstream os;
                        os << "Row " << cur << ": remaining link was not found";
                        throw Error(os.str());
                    }
                    const uint32_t nb = A.ja[p];
                    const bool back_link = links.present_ij(nb, cur) == 1;
                    if (!(din[cur] == 0 || (din[cur] == 1 && back_link))) break;
                    li.map[pos++] = cur;
                    done[cur] = 1;
                    links.on[p] = 0;
                    if (back_link) {
                        links.on[skyline_find(A, nb, cur)] = 0;
                        din[cur]--;
                        dout[nb]--;
                    }
                    dout[cur] = -1;
                    din[nb]--;
                    cur = nb;
                }
                if (dout[cur] == 0 && din[cur] == 0) {
                    li.map[pos++] = cur;
                    done[cur] = 1;
                    dout[cur] = -1;
                }
            }
        }
        for (uint32_t i = 0; i < N; i++)                   // group 2: nodes with <= 2 links
            if (!done[i] && (din[i] <= 2 || dout[i] <= 2)) {
                li.map[pos++] = i;
                din[i] = 0;
                dout[i] = -1;
                done[i] = 1;
            }
        li.M = pos;
        uint32_t back = N - 1;
        for (uint32_t i = 0; i < N; i++) {
            if (done[i]) continue;
            if (din[i] == 0 && dout[i] == 0) li.map[back--] = i;
            else li.map[pos++] = i;
        }
        li.Md = N - pos;
    }

    void order_simple_1(uint32_t N, const std::vector<int>& din, const std::vector<int>& dout, HostLevel& li) {
        std::vector<uint32_t> bucket[4];
        for (uint32_t i = 0; i < N; i++) bucket[dout[i] <= 2 ? dout[i] : 3].push_back(i);
        li.map.assign(N, 0);
        uint32_t pos = 0;
        for (uint32_t i : bucket[0]) if (din[i] != 0) li.map[pos++] = i;
        for (uint32_t d = 1; d < 4; d++) {
            for (uint32_t i : bucket[d]) li.map[pos++] = i;
            if (d == P.degree_max) li.M = pos;
        }
        const uint32_t kept = pos;
        for (uint32_t i : bucket[0]) if (din[i] == 0) li.map[pos++] = i;
        li.Md = N - kept;
    }

    void order_block(uint32_t N, const std::vector<int>& din, const std::vector<int>& dout, HostLevel& li) {
        const uint32_t nx = 60, ny = 220, nz = 85, bs = 5;
        if (N != nx * ny * nz) {
            std::ostringstream os;
            os << "Wrong mesh size: " << N << " (expected 1122000)";
            throw Error(os.str());
        }
        li.map.assign(N, 0);
        std::vector<char> done(N, 0);
        uint32_t pos = 0;
        for (uint32_t K = 0; K < nz; K += bs)
            for (uint32_t J = 0; J < ny; J += bs)
                for (uint32_t I = 0; I < nx; I += bs)
                    for (uint32_t k = K + 1; k < std::min(K + bs, nz) - 1; k++)
                        for (uint32_t j = J + 1; j < std::min(J + bs, ny) - 1; j++)
                            for (uint32_t i = I + 1; i < std::min(I + bs, nx) - 1; i++) {
                                uint32_t id = (k * ny + j) * nx + i;
                                if (din[id] || dout[id]) { li.map[pos++] = id; done[id] = 1; }
                            }
        li.M = pos;
        uint32_t back = N - 1;
        for (uint32_t i = 0; i < N; i++) {
            if (done[i]) continue;
            if (din[i] == 0 && dout[i] == 0) li.map[back--] = i;
            else li.map[pos++] = i;
        }
        li.Md = N - pos;
    }

    void level(uint32_t l, const HostCSR& A) {
        HostLevel& li = H.levels[l];
        li.N = A.nrow;
        li.nnz = A.nnz();
        if (l == H.nlevels - 1) {                           // built without a direct coarse solver
            std::ostringstream os;
            os << "Too many levels: " << H.nlevels;
            throw Error(os.str());
        }
        if (li.N <= coarse_n) {
            H.nlevels = l + 1;
            coarse_n = li.N;
            return;
        }
        const uint32_t N = li.N;
        std::vector<int> dout(N), din(N, 0);
        OneSidedLinks links(A);
        li.aux.assign(N, 0.0);
        // one-sided link removal against q/(1-q)*c (multi_split_constr.cc:70-113).
        // A row's removals depend only on its own values, so the rows decide in
        // parallel; the in-degrees count the links left in each column (integer
        // atomics: the same counts in any order).
        parallel_for(N, 4096, [&](uint64_t r0, uint64_t r1) {
            std::vector<uint32_t> order;
            for (uint32_t i = uint32_t(r0); i < r1; i++) {
                const uint32_t b = A.ia[i], e = A.ia[i + 1], nrz = e - b - 1;
                dout[i] = int(nrz);
                double c = A.a[b];
                for (uint32_t p = b + 1; p < e; p++) c += A.a[p];
                order_by_magnitude(A.a.data() + b + 1, nrz, order);
                double budget = li.q / (1 - li.q) * c;
                for (uint32_t t = 0; t < nrz; t++) {
                    const uint32_t p = b + 1 + order[t];
                    const double v = -A.a[p];
                    if (!(v <= budget)) break;
                    links.on[p] = 0;
                    dout[i]--;
                    budget -= v;
                }
                li.aux[i] = c * (1. / (1 - li.q));
                for (uint32_t p = b + 1; p < e; p++)
                    if (links.on[p]) __atomic_fetch_add(&din[A.ja[p]], 1, __ATOMIC_RELAXED);
            }
        });
        switch (P.order) {
            case ElimOrder::NONE: order_none(N, din, dout, li); break;
            case ElimOrder::ORIGINAL: order_original(A, links, din, dout, li); break;
            case ElimOrder::BLOCK: order_block(N, din, dout, li); break;
            case ElimOrder::SIMPLE_1: order_simple_1(N, din, dout, li); break;
        }
        li.rmap = inverse(li.map);
        HostLevel& ln = H.levels[l + 1];
        eliminate(A, li.map, li.rmap, li.Md, li.M, links, 1.0, li.aux, ln.A, li.U, li.L);
        li.dval.resize(li.Md);
        for (uint32_t i = 0; i < li.Md; i++) li.dval[i] = 1. / li.aux[li.map[i + (N - li.Md)]];
        if (ln.A.nrow) level(l + 1, ln.A);
        else H.nlevels = l + 1;
    }
};

}  // namespace

double cheb_poly(double x, uint32_t k) {
    switch (k) {
        case 0: return 1;
        case 1: return x;
        case 2: return 2 * x * x - 1;
        case 3: return x * (4 * x * x - 3);
        default: return std::cosh(k * std::acosh(x));
    }
}

void build_uh_cheb(const HostCSR& A, const ChebParams& p, Hierarchy& H, ChebLevelBuilder* dev) {
    if (p.sigmas.empty() || p.niters.empty()) throw Error("uh_cheb needs at least one sigma and one niter");
    for (double s : p.sigmas)
        if (!(s > 1)) throw Error("All sigmas must be > 1");
    const uint32_t kMaxLevels = 30;                        // cheb_constr.cc:175
    H = Hierarchy();
    H.kind = PrecKind::UH_CHEB;
    H.A0 = &A;
    H.nlevels = kMaxLevels;
    H.levels.assign(kMaxLevels + 1, HostLevel());
    for (uint32_t l = 0; l < kMaxLevels; l++) {
        H.levels[l].alpha = 1.0;
        H.levels[l].beta = l < p.sigmas.size() ? p.sigmas[l] : p.sigmas.back();
    }
    if (dev) {
        dev->run(A, p.use_tails, kMaxLevels, H);
    } else {
        ChebBuilder b{H, p.use_tails, kMaxLevels};
        b.level(0, A);
    }
    H.levels.resize(H.nlevels + 1);    // keep one slot: the last level's (empty) successor
    H.levels.back() = HostLevel();
    for (uint32_t l = 0; l < H.nlevels; l++)
        H.levels[l].ncheb = l < p.niters.size() ? p.niters[l] : p.niters.back();
    // spectral bounds, coarsest level upwards (cheb_constr.cc:227-236)
    HostLevel& last = H.levels[H.nlevels - 1];
    last.lmin = last.alpha;
    last.lmax = last.beta;
    for (int l = int(H.nlevels) - 2; l >= 0; l--) {
        HostLevel& li = H.levels[l];
        const HostLevel& ln = H.levels[l + 1];
        const double cs = cheb_poly((ln.lmax + ln.lmin) / (ln.lmax - ln.lmin), li.ncheb);
        li.lmin = li.alpha * (1 - 1 / cs);
        li.lmax = li.beta * (1 + 1 / cs);
    }
    H.levels.resize(H.nlevels);
}

void build_multi_split(const HostCSR& A, const MultiSplitParams& p, Hierarchy& H) {
    if (p.q.empty()) throw Error("multi_split needs at least one q");
    if (!p.inner_dynamic && p.niters.empty()) throw Error("multi_split needs level_niter");
    if (p.inner_dynamic && p.eps.empty()) throw Error("multi_split needs level_eps");
    if (p.max_levels == 0) throw Error("max_levels must be positive");
    H = Hierarchy();
    H.kind = PrecKind::MULTI_SPLIT;
    H.A0 = &A;
    H.nlevels = p.max_levels;
    H.levels.assign(p.max_levels + 1, HostLevel());
    for (uint32_t l = 0; l < p.max_levels; l++) H.levels[l].q = l < p.q.size() ? p.q[l] : p.q.back();
    H.levels[0].niter = 1;                                 // level 0 always does one step
    for (uint32_t l = 1; l < p.max_levels; l++) {
        if (!p.inner_dynamic) H.levels[l].niter = l < p.niters.size() ? p.niters[l] : p.niters.back();
        else H.levels[l].eps = l < p.eps.size() ? p.eps[l] : p.eps.back();
    }
    MultiSplitBuilder b{H, p, p.coarse_n};
    b.level(0, A);
    H.levels.resize(H.nlevels);
    HostLevel& last = H.levels.back();                     // multi_split_constr.cc:270-275
    if (last.M == 0) {
        last.niter = 0;
        last.eps = 0.0;
        last.q = 0.0;
    }
}

bool read_multi_split_config(const std::string& path, MultiSplitParams& p) {
    try {
        ConfigFile cfg(path);
        auto numbers = [](const std::string& s, auto sample) {
            std::vector<decltype(sample)> v;
            std::istringstream is(s);
            decltype(sample) x;
            while (is >> x) v.push_back(x);
            return v;
        };
        std::string qs = cfg.read<std::string>("q", "0.8");
        std::string inner = cfg.read<std::string>("inner_iter", "fixed");
        std::string lu = cfg.read<std::string>("lu_method", "exact");
        std::string order = cfg.read<std::string>("elim_order", "original");
        std::string stop = cfg.read<std::string>("elim_stop", "degree");
        p.coarse_n = cfg.read<uint32_t>("coarse_n", 0);
        p.max_levels = cfg.read<uint32_t>("max_levels", 15);
        p.q = numbers(qs, double());
        if (inner == "fixed") {
            p.inner_dynamic = false;
            p.niters = numbers(cfg.read<std::string>("level_niter", "5"), uint32_t());
        } else if (inner == "dynamic") {
            p.inner_dynamic = true;
            p.eps = numbers(cfg.read<std::string>("level_eps", "1e-2"), double());
        } else {
            throw Error("Unknown inner_iter type: " + inner);
        }
        if (lu != "exact" && lu != "ilut") throw Error("Unknown lu type: " + lu);
        if (order == "none") p.order = ElimOrder::NONE;
        else if (order == "original") p.order = ElimOrder::ORIGINAL;
        else if (order == "block") p.order = ElimOrder::BLOCK;
        else if (order == "simple_1") p.order = ElimOrder::SIMPLE_1;
        else throw Error("Unknown elim_order type: " + order);
        p.degree_max = 0;
        p.level_ratio = 0.0;
        if (stop == "degree") p.degree_max = cfg.read<uint32_t>("degree_max", 3);
        else if (stop == "goal") p.level_ratio = cfg.read<double>("level_ratio", 0.25);
        else throw Error("Unknown elim_stop type: " + stop);
        return true;
    } catch (ConfigFile::file_not_found&) {
        return false;
    }
}

}  // namespace mlp

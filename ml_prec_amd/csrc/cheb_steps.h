// One uh_cheb level built by data-parallel steps (link dropping, tails,
// permutation, sparse LU and the next level's matrix), with every integer
// decision and every floating-point value equal to the sequential
// construction (modules/prec/cheb/cheb_constr.cc:38-161, cheb_mark.cc:8-92,
// modules/prec/misc/sparse_lu.cc:6-133).
//
// The steps are functors over one index each, run by an executor X: the
// device executor launches them as kernels, a host executor runs them as loops
// (in any order: no step's result depends on the order of its indices).  The
// includer defines MLP_HD (the functors' call qualifier) and supplies the
// atomics type O (add, sub, max).
//
// Why the parallel steps give the sequential order.  After the link phase,
// every node i has deg[i] present links.  The sequential tail walk
// (cheb_mark.cc:24-64) scans i = 0..N-1; from each node still holding one
// link it walks along that link, removing it, and goes on while the node it
// reaches is left with one link.  The nodes it ever walks through are exactly
// the nodes outside the 2-core of the link graph (repeatedly removing nodes
// with at most one link), except the isolated ones (deg 0): they form a forest
// whose trees hang off the core by one link, or are free trees.  In a hanging
// tree every walk moves towards the core; in a free tree towards its largest
// leaf (no walk can start there before every smaller leaf has started, and a
// walk only ever ends on a leaf or the core).  Root each tree that way.  A walk
// reaching node v continues through v exactly when it is the last of v's
// children to arrive, and walks run to completion in the order of their start
// leaves, so the walk leaving v is the one started at the largest leaf below v:
// W(v) = ML(v), the largest leaf of v's subtree.  The walk that ends in a free
// tree's root is its child's.  The map therefore lists the walks by start leaf
// (a leaf that is not a root), each from its leaf upwards for as long as the
// next node's W is that leaf; then the kept (core) nodes in ascending order,
// then the isolated nodes in descending order.
//
// The peeling runs in rounds (round r removes the nodes left with one link,
// or none, after round r-1), which gives every node its height rnd, its parent
// (the neighbour left when it is removed) and ML by max-propagation to the
// parent.  Free trees end in one node left with no link, or two adjacent nodes
// removed in the same round that point at each other; such a tree is then
// re-rooted at its largest leaf by one thread walking from the centre down to
// that leaf, reversing the links on the way.
//
// Elimination.  Every tail's children come before it in the map and its parent
// after it, so row k of U is {k, parent(k)} and a row's elimination never
// fills: a row's pivot is its row budget minus its kept links' a/beta in
// storage order, minus l_ik u_k,parent over its children in ascending permuted
// order (exactly the sequential operations, sparse_lu.cc:50-104).  Pivots run
// in rounds of height; kept rows then form the next level's rows.
#pragma once

#include <cstdint>
#include <sstream>
#include <string>

#include "hierarchy.h"
#include "mlp_error.h"

#ifndef MLP_HD
#error "define MLP_HD before including cheb_steps.h"
#endif

namespace mlp {
namespace steps {

constexpr uint32_t kNone = 0xffffffffu;     // no node / not reached by the peeling

// Present after the link phase: cnt holds 2 on the diagonal, the counter on an
// upper entry (0: removed from both ends) and the twin's position on a lower
// entry (cheb_misc.h:45-96).
MLP_HD inline bool present(const uint32_t* cnt, const uint32_t* ja, uint32_t row, uint32_t p) {
    return cnt[ja[p] < row ? cnt[p] : p] != 0u;
}

// Row i: the row sum c (storage order) into aux[i]; then the off-diagonal
// links in ascending (|a|, position) order (the stable insertion sort psort,
// cheb_constr.cc:26-36) ask to be dropped while the shares -2 a / (c (beta-1))
// add up to at most 1 (to_remove, cheb_prec.h:71-99).
struct LinkMark {
    const uint32_t* ia;
    const double* a;
    double beta;
    double* aux;
    uint8_t* marked;
    MLP_HD void operator()(uint32_t i) const {
        const uint32_t b = ia[i], e = ia[i + 1];
        double c = 0.0;
        for (uint32_t p = b; p < e; p++) c += a[p];
        aux[i] = c;
        for (uint32_t p = b; p < e; p++) marked[p] = 0;
        double s = 0.0, pv = -1.0;
        uint32_t pp = 0;
        for (uint32_t t = b + 1; t < e; t++) {
            uint32_t best = kNone;
            double bv = 0.0;
            for (uint32_t p = b + 1; p < e; p++) {
                const double v = a[p] < 0 ? -a[p] : a[p];
                const bool after = v > pv || (v == pv && p > pp);
                if (after && (best == kNone || v < bv)) {     // p ascends: the first of equal keys wins
                    best = p;
                    bv = v;
                }
            }
            const double share = -2 * a[best] / (c * (beta - 1));
            if (s + share > 1) break;
            s += share;
            marked[best] = 1;
            pv = bv;
            pp = best;
        }
    }
};

// Row i: each off-diagonal entry's twin by binary search; a link asked to be
// dropped from both ends is removed.  cnt as for present(); deg[i] = links kept.
struct LinkCount {
    const uint32_t* ia;
    const uint32_t* ja;
    const uint8_t* marked;
    uint32_t* cnt;
    int* deg;
    uint32_t* asym;
    MLP_HD void operator()(uint32_t i) const {
        const uint32_t b = ia[i], e = ia[i + 1];
        cnt[b] = 2u;
        int d = int(e - b - 1);
        for (uint32_t p = b + 1; p < e; p++) {
            const uint32_t j = ja[p], je = ia[j + 1];
            uint32_t lo = ia[j] + 1, hi = je;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (ja[mid] < i) lo = mid + 1;
                else hi = mid;
            }
            uint32_t t = p;
            if (lo < je && ja[lo] == i) t = lo;
            else *asym = 1u;
            const uint32_t m = uint32_t(marked[p]) + uint32_t(marked[t]);
            if (m == 2) d--;
            cnt[p] = j > i ? 2u - m : t;
        }
        deg[i] = d;
    }
};

// Peeling state.  rem: links to nodes not yet removed; rnd: the round that
// removes the node (kNone: never, a core or isolated node); ml: largest leaf
// below the node (-1 none); par: the neighbour left at removal; queue: removed
// nodes by round, round r at queue[bnd[r] .. bnd[r+1]).
struct Peel {
    const uint32_t* ia;
    const uint32_t* ja;
    const uint32_t* cnt;
    const int* deg;
    int* rem;
    uint32_t* rnd;
    int* ml;
    uint32_t* par;
    uint32_t* queue;
    uint32_t* qcount;
    uint32_t* bnd;
    uint8_t* center;
    int* W;                 // walk id (start leaf) of the walk through the node
    uint32_t* hgt;          // height in the rooted forest
    uint32_t* wlen;         // walk length, at walk starts; 0 elsewhere
    uint32_t* hmax;         // largest height
    uint32_t* err;          // set when a walk does not end where the forest says it must
    uint32_t N;
};

// every node: the state before round 0; the leaves (one link) form round 0
template <class O> struct PeelInit {
    Peel P;
    bool tails;
    MLP_HD void operator()(uint32_t v) const {
        const int d = P.deg[v];
        P.rem[v] = d;
        P.par[v] = kNone;
        P.center[v] = 0;
        P.wlen[v] = 0;
        if (tails && d == 1) {
            P.ml[v] = int(v);
            P.rnd[v] = 0;
            P.queue[O::add(P.qcount, 1u)] = v;
        } else {
            P.ml[v] = -1;
            P.rnd[v] = kNone;
        }
    }
};

// one index: the end of round r's segment is the queue length before round r runs
struct PeelBound {
    Peel P;
    uint32_t r;
    MLP_HD void operator()(uint32_t) const { P.bnd[r + 1] = *P.qcount; }
};

// round r, one removed node v: its parent is the one present neighbour not
// removed in an earlier round (none: v is a free tree's centre).  A parent
// outside this round gets v's ML and loses a link; the link that leaves it with
// one makes it a node of round r + 1.  Two nodes of this round pointing at each
// other end a free tree together.
template <class O> struct PeelRound {
    Peel P;
    uint32_t r;
    MLP_HD void operator()(uint32_t q) const {
        const uint32_t v = P.queue[q];
        uint32_t u = kNone;
        for (uint32_t p = P.ia[v] + 1; p < P.ia[v + 1]; p++) {
            if (!present(P.cnt, P.ja, v, p)) continue;
            const uint32_t j = P.ja[p];
            if (P.rnd[j] >= r) {             // not removed before this round (kNone: not yet at all)
                u = j;
                break;
            }
        }
        P.par[v] = u;
        if (u == kNone || P.rnd[u] == r) return;
        O::max(P.ml + u, P.ml[v]);
        O::fence();
        if (O::sub(P.rem + u, 1) == 2) {
            P.rnd[u] = r + 1;
            P.queue[O::add(P.qcount, 1u)] = u;
        }
    }
};

// every removed node: W and height as the peeling left them; the free trees'
// centres (no parent, or the side of a mutual pair holding the larger ML)
template <class O> struct PeelCentres {
    Peel P;
    MLP_HD void operator()(uint32_t q) const {
        const uint32_t v = P.queue[q];
        P.W[v] = P.ml[v];
        P.hgt[v] = P.rnd[v];
        O::maxu(P.hmax, P.rnd[v]);
        const uint32_t u = P.par[v];
        if (u == kNone) P.center[v] = 1;
        else if (P.par[u] == v && P.ml[v] > P.ml[u]) P.center[v] = 2;
    }
};

// one free tree's centre c: re-root the tree at its largest leaf lm = ML(c).
// Walking from c towards lm, every node on the way takes the neighbour towards
// lm as its parent; its W becomes the largest leaf of all its other subtrees
// (including the one it came from) and its height one more than theirs.
template <class O> struct Reroot {
    Peel P;
    MLP_HD void operator()(uint32_t q) const {
        const uint32_t c = P.queue[q];
        if (!P.center[c]) return;
        const int lm = P.ml[c];
        uint32_t prev = kNone, cur = c;
        int pml = -1;
        uint32_t ph = 0;
        for (uint32_t step = 0;; step++) {
            if (cur == kNone || step > P.N) {       // lm below no child: not a tree
                *P.err = 1u;
                return;
            }
            if (int(cur) == lm) {                   // the root: its walk is its child's
                P.par[cur] = kNone;
                if (prev == kNone) {                // c itself, a leaf whose only neighbour is its child
                    for (uint32_t p = P.ia[cur] + 1; p < P.ia[cur + 1]; p++)
                        if (present(P.cnt, P.ja, cur, p)) {
                            const uint32_t y = P.ja[p];
                            pml = P.ml[y];
                            ph = P.rnd[y];
                        }
                }
                P.W[cur] = pml;
                P.hgt[cur] = ph + 1;
                O::maxu(P.hmax, ph + 1);
                return;
            }
            uint32_t x = kNone;
            int m = P.deg[cur] == 1 ? int(cur) : -1;
            uint32_t h = 0;
            bool any = false;
            for (uint32_t p = P.ia[cur] + 1; p < P.ia[cur + 1]; p++) {
                if (!present(P.cnt, P.ja, cur, p)) continue;
                const uint32_t y = P.ja[p];
                if (y == prev) continue;
                if (x == kNone && P.ml[y] == lm) {
                    x = y;
                    continue;
                }
                if (P.ml[y] > m) m = P.ml[y];
                if (!any || P.rnd[y] > h) h = P.rnd[y];
                any = true;
            }
            if (prev != kNone) {
                if (pml > m) m = pml;
                if (!any || ph > h) h = ph;
                any = true;
            }
            const uint32_t nh = any ? h + 1 : 0;
            P.par[cur] = x;
            P.W[cur] = m;
            P.hgt[cur] = nh;
            O::maxu(P.hmax, nh);
            prev = cur;
            pml = m;
            ph = nh;
            cur = x;
        }
    }
};

// a walk start s (a leaf with a parent): nodes s, par(s), ... while the next
// node is a tail whose W is s
struct WalkLen {
    Peel P;
    MLP_HD void operator()(uint32_t q) const {
        const uint32_t s = P.queue[q];
        if (P.deg[s] != 1 || P.par[s] == kNone) return;
        uint32_t len = 1, cur = s;
        for (;;) {
            const uint32_t nx = P.par[cur];
            if (nx == kNone || P.rnd[nx] == kNone || P.W[nx] != int(s)) break;
            cur = nx;
            if (++len > P.N) {                      // a cycle of parents
                *P.err = 1u;
                break;
            }
        }
        P.wlen[s] = len;
    }
};

struct WalkWrite {
    Peel P;
    const uint32_t* woff;        // exclusive scan of wlen
    uint32_t* rmap;
    uint32_t* map;
    MLP_HD void operator()(uint32_t q) const {
        const uint32_t s = P.queue[q];
        if (P.deg[s] != 1 || P.par[s] == kNone) return;
        uint32_t pos = woff[s], cur = s;
        for (;;) {
            if (pos >= P.N) {
                *P.err = 1u;
                return;
            }
            rmap[cur] = pos;
            map[pos] = cur;
            pos++;
            const uint32_t nx = P.par[cur];
            if (nx == kNone || P.rnd[nx] == kNone || P.W[nx] != int(s)) break;
            cur = nx;
        }
    }
};

// kept nodes (never removed, with links) and isolated nodes (no links)
struct GroupFlags {
    const int* deg;
    const uint32_t* rnd;
    uint32_t* kflag;
    uint32_t* iflag;
    MLP_HD void operator()(uint32_t v) const {
        kflag[v] = (rnd[v] == kNone && deg[v] > 0) ? 1u : 0u;
        iflag[v] = deg[v] == 0 ? 1u : 0u;
    }
};

// kept nodes at M.. in ascending order, isolated nodes from N-1 down
struct GroupPlace {
    const uint32_t* kflag;
    const uint32_t* koff;
    const uint32_t* iflag;
    const uint32_t* ioff;
    uint32_t M, N;
    uint32_t* rmap;
    uint32_t* map;
    MLP_HD void operator()(uint32_t v) const {
        uint32_t p;
        if (kflag[v]) p = M + koff[v];
        else if (iflag[v]) p = N - 1 - ioff[v];
        else return;
        rmap[v] = p;
        map[p] = v;
    }
};

// map and rmap are inverse permutations: every position holds a node placed
// there (a node out of range or placed elsewhere sets *err)
struct PermCheck {
    const uint32_t* map;
    const uint32_t* rmap;
    uint32_t N;
    uint32_t* err;
    MLP_HD void operator()(uint32_t p) const {
        const uint32_t v = map[p];
        if (v >= N || rmap[v] != p) *err = 1u;
    }
};

// Row sizes of L (children), U (tail rows: pivot and parent) and the next
// level (kept rows: diagonal and kept links), the U off-diagonal a/beta of
// every tail with a parent, and a check that every child's U row is
// {child, this row} (the forest guarantees it; *bad reports otherwise).
struct RowCounts {
    const uint32_t* ia;
    const uint32_t* ja;
    const double* a;
    const uint32_t* cnt;
    const uint32_t* rmap;
    const uint32_t* par;
    double beta;
    uint32_t M, nkeep;
    uint32_t* Lc;
    uint32_t* Uc;
    uint32_t* Nc;
    double* uoff;
    uint32_t* bad;
    MLP_HD void operator()(uint32_t v) const {
        const uint32_t i = rmap[v];
        if (i >= nkeep) return;
        const uint32_t lim = i < M ? i : M;
        uint32_t ch = 0, up = 0;
        for (uint32_t p = ia[v] + 1; p < ia[v + 1]; p++) {
            if (!present(cnt, ja, v, p)) continue;
            const uint32_t j = ja[p];
            if (rmap[j] < lim) {
                ch++;
                if (par[j] != v) *bad = 1u;
            } else {
                up++;
                if (i < M) uoff[v] = a[p] / beta;
            }
        }
        Lc[i] = ch;
        if (i < M) {
            if (up > 1) *bad = 1u;
            Uc[i] = 1 + up;
        } else {
            Nc[i - M] = 1 + up;
        }
    }
};

// The pivot of row i = rmap[v]: the row budget minus the present links' a/beta
// in storage order, then minus l_ik u_k,i over the children k in ascending
// permuted order, with l_ik = (a_ik/beta) / u_kk; the L row in that order.
struct RowElim {
    const uint32_t* ia;
    const uint32_t* ja;
    const double* a;
    const uint32_t* cnt;
    const uint32_t* rmap;
    const double* aux;
    double beta;
    uint32_t M;
    const double* piv;
    const double* uoff;
    const uint32_t* L_ia;
    uint32_t* L_ja;
    double* L_a;
    MLP_HD double operator()(uint32_t v, uint32_t i) const {
        const uint32_t b = ia[v] + 1, e = ia[v + 1];
        double d = aux[v];
        for (uint32_t p = b; p < e; p++)
            if (present(cnt, ja, v, p)) {
                const double z = a[p] / beta;
                d -= z;
            }
        const uint32_t lim = i < M ? i : M;
        uint32_t last = 0, out = L_ia[i];
        bool first = true;
        for (;;) {                                  // the next child in permuted order
            uint32_t best = kNone, bp = 0;
            for (uint32_t p = b; p < e; p++) {
                if (!present(cnt, ja, v, p)) continue;
                const uint32_t c = rmap[ja[p]];
                if (c >= lim || (!first && c <= last)) continue;
                if (best == kNone || c < best) {
                    best = c;
                    bp = p;
                }
            }
            if (best == kNone) break;
            const uint32_t k = ja[bp];
            const double lik = (a[bp] / beta) / piv[k];
            d -= lik * uoff[k];
            L_ja[out] = best;
            L_a[out] = lik;
            out++;
            last = best;
            first = false;
        }
        return d;
    }
};

// tails of height h: pivot, L row and U row {i, parent}
struct TailRows {
    RowElim E;
    const uint32_t* queue;
    const uint32_t* hgt;
    const uint32_t* par;
    uint32_t h;
    double* piv;
    const uint32_t* U_ia;
    uint32_t* U_ja;
    double* U_a;
    MLP_HD void operator()(uint32_t q) const {
        const uint32_t v = queue[q];
        if (hgt[v] != h) return;
        const uint32_t i = E.rmap[v];
        const double d = E(v, i);
        piv[v] = d;
        const uint32_t o = U_ia[i];
        U_ja[o] = i;
        U_a[o] = d;
        if (par[v] != kNone) {
            U_ja[o + 1] = E.rmap[par[v]];
            U_a[o + 1] = E.uoff[v];
        }
    }
};

// kept rows: pivot (the next level's diagonal), L row, and the next level's
// row: diagonal first, then the kept links in storage order (ascending
// permuted order, as kept nodes keep their relative order)
struct KeptRows {
    RowElim E;
    uint32_t nkeep;
    const uint32_t* N_ia;
    uint32_t* N_ja;
    double* N_a;
    MLP_HD void operator()(uint32_t v) const {
        const uint32_t i = E.rmap[v];
        if (i < E.M || i >= nkeep) return;
        const double d = E(v, i);
        uint32_t o = N_ia[i - E.M];
        N_ja[o] = i - E.M;
        N_a[o] = d;
        o++;
        for (uint32_t p = E.ia[v] + 1; p < E.ia[v + 1]; p++) {
            if (!present(E.cnt, E.ja, v, p)) continue;
            const uint32_t c = E.rmap[E.ja[p]];
            if (c < E.M) continue;
            N_ja[o] = c - E.M;
            N_a[o] = E.a[p] / E.beta;
            o++;
        }
    }
};

// reciprocals of the isolated nodes' diagonals (cheb_constr.cc:131-137)
struct IsoDiag {
    const uint32_t* map;
    const double* aux;
    uint32_t first;
    double* dval;
    MLP_HD void operator()(uint32_t t) const { dval[t] = 1. / aux[map[first + t]]; }
};

// ---------------------------------------------------------------------------
// The level driver: the steps in order on executor X with atomics O.  X
// provides alloc<T>(n), zero(p, n), each(n, f) (f(i) for i < n), each_seg(bnd,
// r, f) (f(q) for bnd[r] <= q < bnd[r+1], bounds read where f runs), scan(in,
// out, n) (exclusive, out[n] = total), get(dst, src, n) and get1(src) (copies
// to the host, waiting for the steps before them) and get_many(items) (several
// copies, one wait).
// ---------------------------------------------------------------------------
struct LevelMatrix {               // a level matrix on the executor (Skyline CSR)
    uint32_t n = 0, nnz = 0;
    const uint32_t* ia = nullptr;
    const uint32_t* ja = nullptr;
    const double* a = nullptr;
};

struct Fetch {                     // one copy of get_many: bytes from the executor's src to the host's dst
    void* dst;
    const void* src;
    size_t bytes;
};
template <class V, class T> void fetch(std::vector<Fetch>& f, V& dst, const T* src, size_t n) {
    dst.resize(n);
    if (n) f.push_back({dst.data(), src, n * sizeof(T)});
}

// One level: fills li (N, nnz, M, Md, map, rmap, aux, L, U, dval) and returns
// the next level's matrix, on the executor and in *next_host.
template <class X, class O>
LevelMatrix level_steps(X& x, const LevelMatrix& A, double beta, bool tails, HostLevel& li, HostCSR* next_host) {
    const uint32_t N = A.n, nnz = A.nnz;
    li.N = N;
    li.nnz = nnz;
    double* aux = x.template alloc<double>(N);
    uint8_t* marked = x.template alloc<uint8_t>(nnz);
    uint32_t* cnt = x.template alloc<uint32_t>(nnz);
    int* deg = x.template alloc<int>(N);
    uint32_t* flags = x.template alloc<uint32_t>(3);      // asymmetric structure; fill; inconsistent tails
    x.zero(flags, 3);
    x.each(N, LinkMark{A.ia, A.a, beta, aux, marked});
    x.each(N, LinkCount{A.ia, A.ja, marked, cnt, deg, flags});

    Peel P;
    P.ia = A.ia; P.ja = A.ja; P.cnt = cnt; P.deg = deg;
    P.rem = x.template alloc<int>(N);
    P.rnd = x.template alloc<uint32_t>(N);
    P.ml = x.template alloc<int>(N);
    P.par = x.template alloc<uint32_t>(N);
    P.queue = x.template alloc<uint32_t>(N);
    P.qcount = x.template alloc<uint32_t>(1);
    P.bnd = x.template alloc<uint32_t>(size_t(N) + 16);
    P.center = x.template alloc<uint8_t>(N);
    P.W = x.template alloc<int>(N);
    P.hgt = x.template alloc<uint32_t>(N);
    P.wlen = x.template alloc<uint32_t>(N);
    P.hmax = x.template alloc<uint32_t>(1);
    P.err = flags + 2;
    P.N = N;
    x.zero(P.qcount, 1);
    x.zero(P.hmax, 1);
    x.zero(P.bnd, 1);
    x.each(N, PeelInit<O>{P, tails});
    uint32_t removed = 0;
    if (tails) {
        // rounds in batches; a batch ends by reading whether its last round removed anything
        uint32_t r = 0;
        for (;;) {
            for (uint32_t k = 0; k < 8 && r < N + 8; k++, r++) {
                x.each(1, PeelBound{P, r});
                x.each_seg(P.bnd, r, PeelRound<O>{P, r});
            }
            x.each(1, PeelBound{P, r});
            uint32_t b2[2];
            x.get(b2, P.bnd + r, 2);
            if (b2[0] == b2[1] || r >= N + 8) {
                removed = b2[1];
                break;
            }
        }
        if (removed) {
            x.each(removed, PeelCentres<O>{P});
            x.each(removed, Reroot<O>{P});
            x.each(removed, WalkLen{P});
        }
    }
    uint32_t* woff = x.template alloc<uint32_t>(size_t(N) + 1);
    uint32_t* rmap = x.template alloc<uint32_t>(N);
    uint32_t* map = x.template alloc<uint32_t>(N);
    x.scan(P.wlen, woff, N);
    const uint32_t M = x.get1(woff + N);
    if (M != removed || x.get1(P.err))
        throw Error("device construction: the tail walks do not cover the removed nodes", 6);
    if (removed) x.each(removed, WalkWrite{P, woff, rmap, map});
    uint32_t* kflag = x.template alloc<uint32_t>(N);
    uint32_t* iflag = x.template alloc<uint32_t>(N);
    uint32_t* koff = x.template alloc<uint32_t>(size_t(N) + 1);
    uint32_t* ioff = x.template alloc<uint32_t>(size_t(N) + 1);
    x.each(N, GroupFlags{deg, P.rnd, kflag, iflag});
    x.scan(kflag, koff, N);
    x.scan(iflag, ioff, N);
    const uint32_t n = x.get1(koff + N), Md = x.get1(ioff + N);
    if (uint64_t(M) + n + Md != N) throw Error("device construction: the groups do not cover the level", 6);
    x.each(N, GroupPlace{kflag, koff, iflag, ioff, M, N, rmap, map});
    x.each(N, PermCheck{map, rmap, N, P.err});
    if (x.get1(P.err)) throw Error("device construction: the permutation is not one", 6);
    const uint32_t nkeep = N - Md;
    li.M = M;
    li.Md = Md;

    uint32_t* Lc = x.template alloc<uint32_t>(nkeep);
    uint32_t* Uc = x.template alloc<uint32_t>(M);
    uint32_t* Nc = x.template alloc<uint32_t>(n);
    double* uoff = x.template alloc<double>(N);
    x.each(N, RowCounts{A.ia, A.ja, A.a, cnt, rmap, P.par, beta, M, nkeep, Lc, Uc, Nc, uoff, flags + 1});
    uint32_t fl[2];
    x.get(fl, flags, 2);
    if (fl[0]) throw Error("uh_cheb needs a structurally symmetric matrix", 1);
    if (fl[1]) throw Error("device construction: a tail's elimination would fill", 6);
    uint32_t* L_ia = x.template alloc<uint32_t>(size_t(nkeep) + 1);
    uint32_t* U_ia = x.template alloc<uint32_t>(size_t(M) + 1);
    uint32_t* N_ia = x.template alloc<uint32_t>(size_t(n) + 1);
    x.scan(Lc, L_ia, nkeep);
    x.scan(Uc, U_ia, M);
    x.scan(Nc, N_ia, n);
    const uint32_t nnzL = x.get1(L_ia + nkeep), nnzU = x.get1(U_ia + M), nnzN = x.get1(N_ia + n);
    uint32_t* L_ja = x.template alloc<uint32_t>(nnzL);
    double* L_a = x.template alloc<double>(nnzL);
    uint32_t* U_ja = x.template alloc<uint32_t>(nnzU);
    double* U_a = x.template alloc<double>(nnzU);
    uint32_t* N_ja = x.template alloc<uint32_t>(nnzN);
    double* N_a = x.template alloc<double>(nnzN);
    double* piv = x.template alloc<double>(N);
    RowElim E{A.ia, A.ja, A.a, cnt, rmap, aux, beta, M, piv, uoff, L_ia, L_ja, L_a};
    if (M) {
        const uint32_t hmax = x.get1(P.hmax);
        if (hmax >= N) throw Error("device construction: a tail is higher than the level has nodes", 6);
        for (uint32_t h = 0; h <= hmax; h++)
            x.each(M, TailRows{E, P.queue, P.hgt, P.par, h, piv, U_ia, U_ja, U_a});
    }
    x.each(N, KeptRows{E, nkeep, N_ia, N_ja, N_a});
    double* dval = x.template alloc<double>(Md);
    x.each(Md, IsoDiag{map, aux, nkeep, dval});

    std::vector<Fetch> out;
    fetch(out, li.map, map, N);
    fetch(out, li.rmap, rmap, N);
    fetch(out, li.aux, aux, N);
    li.L = HostCSR();
    li.L.nrow = li.L.ncol = nkeep;
    fetch(out, li.L.ia, L_ia, size_t(nkeep) + 1);
    fetch(out, li.L.ja, L_ja, nnzL);
    fetch(out, li.L.a, L_a, nnzL);
    li.U = HostCSR();
    li.U.nrow = M;
    li.U.ncol = nkeep;
    fetch(out, li.U.ia, U_ia, size_t(M) + 1);
    fetch(out, li.U.ja, U_ja, nnzU);
    fetch(out, li.U.a, U_a, nnzU);
    fetch(out, li.dval, dval, Md);
    LevelMatrix nx;
    nx.n = n;
    nx.nnz = nnzN;
    nx.ia = N_ia;
    nx.ja = N_ja;
    nx.a = N_a;
    if (next_host) {
        *next_host = HostCSR();
        next_host->nrow = next_host->ncol = n;
        if (n) {
            fetch(out, next_host->ia, N_ia, size_t(n) + 1);
            fetch(out, next_host->ja, N_ja, nnzN);
            fetch(out, next_host->a, N_a, nnzN);
        }
    }
    x.get_many(out);
    return nx;
}

// Every level (cheb_constr.cc:38-161 recursion): the caller's H has its levels'
// alpha/beta set; H.levels[l + 1].A receives each next level's matrix.
template <class X, class O>
void build_levels_steps(X& x, const LevelMatrix& A0, bool tails, uint32_t max_levels, Hierarchy& H) {
    LevelMatrix A = A0;
    for (uint32_t l = 0;; l++) {
        LevelMatrix nx = level_steps<X, O>(x, A, H.levels[l].beta, tails, H.levels[l], &H.levels[l + 1].A);
        if (nx.n == 0) {
            H.nlevels = l + 1;
            return;
        }
        if (l + 1 >= max_levels) {
            std::ostringstream os;
            os << "Too many levels: " << l + 1;
            throw Error(os.str());
        }
        A = nx;
    }
}

}  // namespace steps
}  // namespace mlp

"""Pins of the oracle's multi_split construction and apply, and of the
tail order of both preconditioners, by facts computed here independently of
the oracle's code.

* The split matrix of a level (the matrix the multi_split elimination
  factors) is rebuilt densely from its definition: one-sided removal of the
  weakest off-diagonal links against the budget s = q/(1-q) c_i, taken in
  |a| order and lowered by each removal (multi_split_constr.cc:70-113), the
  diagonal c_i/(1-q) minus the kept links (sparse_lu.cc:50-73 with
  beta = 1).  The oracle's factors L, U and the next level must multiply
  back to it (L U = split matrix, permuted).
* The apply is evaluated with dense block algebra instead of triangular
  sweeps: a level's step is the inverse of the split matrix whose Schur block
  is replaced by the inverse of the next level's operator, and the level
  operator is that step iterated as Richardson on the level matrix
  (multi_split_solve.cc:59-157), over hierarchies of every depth.
* The orderings are checked against their partition rules, and small graphs
  whose permutations were worked out by hand pin the tail walks
  (cheb_mark.cc:24-91, multi_split_mark.cc:27-190).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from conftest import rng, system
from ml_prec_amd import native as N
from ml_prec_amd import spe_synth as S


def dense(ia, ja, a):
    n = len(ia) - 1
    return sp.csr_matrix((np.asarray(a, float), np.asarray(ja), np.asarray(ia)), shape=(n, n)).toarray()


def split_matrix(ia, ja, a, q):
    """(split matrix, kept-link mask, kept out-degree, kept in-degree, c/(1-q)) from the definition."""
    n = len(ia) - 1
    Ah = np.zeros((n, n))
    kept = np.zeros((n, n), dtype=bool)
    dout = np.zeros(n, dtype=int)
    din = np.zeros(n, dtype=int)
    aux = np.zeros(n)
    for i in range(n):
        b, e = int(ia[i]), int(ia[i + 1])
        cols = [int(j) for j in ja[b + 1:e]]
        vals = [float(v) for v in a[b + 1:e]]
        c = float(a[b])
        for v in vals:
            c += v
        budget = q / (1 - q) * c
        drop = set()
        for t in sorted(range(len(vals)), key=lambda t: abs(vals[t])):     # stable: ties keep storage order
            if -vals[t] <= budget:
                drop.add(t)
                budget -= -vals[t]
            else:
                break
        aux[i] = c * (1.0 / (1 - q))
        d = aux[i]
        for t, (j, v) in enumerate(zip(cols, vals)):
            if t in drop:
                continue
            kept[i, j] = True
            Ah[i, j] = v
            d -= v
            dout[i] += 1
            din[j] += 1
        Ah[i, i] = d
    return Ah, kept, dout, din, aux


def level_data(P, s):
    """Per level: the level matrix (the system, then each A_next), the oracle's arrays and info."""
    out = []
    ia, ja, a = s.ia, s.ja, s.a
    for l in range(P.nlevels):
        info, arr = P.level_arrays(l)
        out.append(dict(ia=ia, ja=ja, a=a, info=info, arr=arr))
        ia, ja, a = arr["A_ia"], arr["A_ja"], arr["A_a"]
    return out


def factor_product(info, arr):
    """L~ [U; 0 A_next] in permuted numbering over the first N-Md rows."""
    N, M, Md = info["N"], info["M"], info["Md"]
    K = N - Md
    Lt = np.eye(K)
    for i in range(K):
        for p in range(arr["L_ia"][i], arr["L_ia"][i + 1]):
            Lt[i, arr["L_ja"][p]] = arr["L_a"][p]
    Up = np.zeros((K, K))
    for i in range(M):
        for p in range(arr["U_ia"][i], arr["U_ia"][i + 1]):
            Up[i, arr["U_ja"][p]] = arr["U_a"][p]
    if info["n_next"]:
        Up[M:, M:] = dense(arr["A_ia"], arr["A_ja"], arr["A_a"])
    return Lt @ Up


def check_ordering(order, info, arr, dout, din, degree_max):
    """Partition rules of the elimination orders (multi_split_mark.cc)."""
    N, M, Md = info["N"], info["M"], info["Md"]
    perm = [int(v) for v in arr["map"]]
    assert sorted(perm) == list(range(N))
    iso = [i for i in range(N) if dout[i] == 0 and din[i] == 0]
    elim, kept, tail = perm[:M], perm[M:N - Md], perm[N - Md:]
    if order == "none":                                                       # :8-25
        assert M == 0
        assert kept == [i for i in range(N) if i not in set(iso)]
        assert tail == sorted(iso, reverse=True)
    elif order == "simple_1":                                                 # :156-190
        def bucket(i):
            return min(int(dout[i]), 3)
        first = [i for i in range(N) if dout[i] == 0 and din[i] != 0]
        rest = [i for d in (1, 2, 3) for i in range(N) if dout[i] > 0 and bucket(i) == d]
        ordered = first + rest
        assert perm[:N - Md] == ordered
        assert M == sum(1 for i in ordered if bucket(i) <= degree_max)
        assert tail == iso                                                    # ascending here
    elif order == "original":                                                 # :27-154
        # degrees only fall during the tail walk, so a node with at most two links
        # either way is eliminated, and a kept node had more than two both ways
        s_elim = set(elim)
        for i in range(N):
            if i in set(tail):
                continue
            if din[i] <= 2 or dout[i] <= 2:
                assert i in s_elim
        for i in kept:
            assert din[i] > 2 and dout[i] > 2
        assert kept == sorted(kept)
        assert tail == sorted(tail, reverse=True)
        assert all(dout[i] == 0 and din[i] == 0 for i in tail)


def ms_step_matrix(lv, l, Bnext):
    """Dense step operator of level l: inverse of the split matrix with its Schur block replaced by Bnext^-1."""
    d = lv[l]
    info, arr = d["info"], d["arr"]
    N, M, Md = info["N"], info["M"], info["Md"]
    n = N - M - Md
    Ah = d["Ah"]
    perm = arr["map"].astype(int)
    E, Kp, D = perm[:M], perm[M:M + n], perm[N - Md:]
    keep = perm[:N - Md]
    Ak = Ah[np.ix_(keep, keep)]
    Sm = np.zeros((N, N))
    if n:
        Mod = Ak.copy()
        corr = Ak[M:, :M] @ np.linalg.solve(Ak[:M, :M], Ak[:M, M:]) if M else 0.0
        Mod[M:, M:] = np.linalg.inv(Bnext) + corr
        Sm[np.ix_(keep, keep)] = np.linalg.inv(Mod)
    elif len(keep):
        Sm[np.ix_(keep, keep)] = np.linalg.inv(Ak)
    Sm[D, D] = 1.0 / np.diag(Ah)[D]
    return Sm


def ms_dense_operator(lv, l):
    """B_l as a matrix for the fixed variant: niter Richardson steps x += S (f - A x) from 0."""
    d = lv[l]
    info = d["info"]
    n = info["n_next"]
    Bn = ms_dense_operator(lv, l + 1) if n else None
    Sm = ms_step_matrix(lv, l, Bn)
    A = dense(d["ia"], d["ja"], d["a"])
    steps = max(1, info["nit"])
    B = np.zeros_like(A)
    for _ in range(steps):
        B = B + Sm @ (np.eye(len(A)) - A @ B)
    return B


def ms_dynamic_apply(lv, l, f):
    """The dynamic variant as a function (its operator depends on f): steps until ||f - A x|| / ||f|| < eps."""
    d = lv[l]
    info, arr = d["info"], d["arr"]
    N, M, Md = info["N"], info["M"], info["Md"]
    n = N - M - Md
    Ah = d["Ah"]
    perm = arr["map"].astype(int)
    keep, Kp, D = perm[:N - Md], perm[M:M + n], perm[N - Md:]
    Ak = Ah[np.ix_(keep, keep)]
    A = dense(d["ia"], d["ja"], d["a"])

    def step(g):
        x = np.zeros(N)
        gk = g[keep]
        if n:
            tE = np.linalg.solve(Ak[:M, :M], gk[:M]) if M else np.zeros(0)
            F = gk[M:] - (Ak[M:, :M] @ tE if M else 0.0)
            xK = ms_dynamic_apply(lv, l + 1, F)
            xE = np.linalg.solve(Ak[:M, :M], gk[:M] - Ak[:M, M:] @ xK) if M else np.zeros(0)
            x[keep] = np.concatenate([xE, xK])
        elif len(keep):
            x[keep] = np.linalg.solve(Ak, gk)
        x[D] = g[D] / np.diag(Ah)[D]
        return x

    x = step(f)
    eps, nit = info["eps"], info["nit"]
    init = np.linalg.norm(f)
    i = 2
    while eps or i <= nit:
        r = f - A @ x
        if eps and np.linalg.norm(r) / init < eps:
            break
        x = x + step(r)
        i += 1
    return x


def prepare(P, s, q):
    lv = level_data(P, s)
    for d in lv:
        d["Ah"], d["kept"], d["dout"], d["din"], d["aux"] = split_matrix(d["ia"], d["ja"], d["a"], q)
    return lv


CASES = [(order, dm, q) for order, dm in (("none", 3), ("original", 3), ("simple_1", 2)) for q in (0.6, 0.8)]


@pytest.mark.parametrize("order,dm,q", CASES)
def test_multi_split_construction_is_the_split_matrix_factored(order, dm, q):
    s = system("tiny_tpfa_c1")
    P = O.msplit_prec(s.ia, s.ja, s.a, q=(q,), niters=(3,), order=order, degree_max=dm)
    assert P.nlevels >= 3
    lv = prepare(P, s, q)
    for l, d in enumerate(lv):
        info, arr, Ah = d["info"], d["arr"], d["Ah"]
        N, M, Md = info["N"], info["M"], info["Md"]
        assert M + Md + info["n_next"] == N
        assert np.allclose(arr["aux"], d["aux"], rtol=1e-14, atol=0)
        if l + 1 == len(lv) and info["n_next"] == 0 and M == 0 and Md == 0:
            continue
        check_ordering(order, info, arr, d["dout"], d["din"], dm)
        perm = arr["map"].astype(int)
        keep = perm[:N - Md]
        # L U reproduces the split matrix on the kept-and-eliminated rows
        if len(keep):
            LU = factor_product(info, arr)
            ref = Ah[np.ix_(keep, keep)]
            assert np.allclose(LU, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
        # isolated rows: no kept link in or out, dval = 1 / diagonal
        iso = perm[N - Md:]
        assert not d["kept"][iso].any() and not d["kept"][:, iso].any()
        assert np.array_equal(arr["dval"], 1.0 / d["aux"][iso])


def test_split_matrix_drop_rule_against_a_mistake():
    """A budget without the 1/(1-q) scaling, or one that is not lowered by each removal, changes the kept set."""
    s = system("tiny_tpfa_c1")
    _, kept, _, _, _ = split_matrix(s.ia, s.ja, s.a, 0.6)
    A = dense(s.ia, s.ja, s.a)
    off = (A != 0) & ~np.eye(s.n, dtype=bool)
    n_drop = int((off & ~kept).sum())
    assert 0 < n_drop < int(off.sum())
    c = A.sum(1)
    # keep-all-below-budget (no lowering) drops more
    more = sum(int(np.sum((-A[i][off[i]]) <= 0.6 / 0.4 * c[i])) for i in range(s.n))
    assert more > n_drop


@pytest.mark.parametrize("order,dm,q", CASES)
def test_multi_split_apply_equals_dense_richardson_recursion(order, dm, q):
    s = system("tiny_tpfa_c1")
    P = O.msplit_prec(s.ia, s.ja, s.a, q=(q,), niters=(3,), order=order, degree_max=dm)
    assert P.nlevels >= 3
    lv = prepare(P, s, q)
    Bd = ms_dense_operator(lv, 0)
    f = rng(5).standard_normal(s.n)
    scale = np.abs(Bd).max()
    for k in range(4):
        g = f if k == 0 else rng(10 + k).standard_normal(s.n)
        assert np.allclose(P.apply(g), Bd @ g, rtol=1e-9, atol=1e-9 * scale * np.abs(g).max())
    # the whole operator on unit vectors
    B = np.stack([P.apply(np.eye(s.n)[j]) for j in range(s.n)], axis=1)
    assert np.allclose(B, Bd, rtol=1e-9, atol=1e-9 * scale)


@pytest.mark.parametrize("order,dm", [("original", 3), ("simple_1", 2)])
def test_multi_split_dynamic_apply_equals_dense_iteration(order, dm):
    s = system("tiny_tpfa_c1")
    P = O.msplit_prec(s.ia, s.ja, s.a, q=(0.6,), inner="dynamic", epss=(1e-3, 1e-3, 1e-2), order=order,
                      degree_max=dm)
    assert P.nlevels >= 3
    assert any(P.level_info(l)["eps"] > 0 for l in range(1, P.nlevels))
    lv = prepare(P, s, 0.6)
    for k in range(3):
        f = rng(20 + k).standard_normal(s.n)
        xd = ms_dynamic_apply(lv, 0, f)
        assert np.allclose(P.apply(f), xd, rtol=1e-9, atol=1e-9 * np.abs(xd).max())
    # a stopping test that never stops early matches the fixed variant with one step
    Pd = O.msplit_prec(s.ia, s.ja, s.a, q=(0.6,), inner="dynamic", epss=(1e3,), order=order, degree_max=dm)
    Pf = O.msplit_prec(s.ia, s.ja, s.a, q=(0.6,), niters=(1,), order=order, degree_max=dm)
    f = rng(30).standard_normal(s.n)
    assert np.array_equal(Pd.apply(f), Pf.apply(f))


# --------------------------------------------------------------------------
# hand-worked graphs
# --------------------------------------------------------------------------
# 13 nodes; links of strength 1 (a_ij = -1); reaction c on the diagonal (row sum = c).
# Nodes 4, 9, 11 carry c = 100 (cheap to drop), every other node c = 0.5; 10 and 12 have no links.
EDGES = [(0, 1), (0, 2), (0, 4), (2, 6), (6, 8), (4, 9), (9, 11), (4, 11), (3, 5), (5, 7)]
STRONG_C = {4: 100.0, 9: 100.0, 11: 100.0}


def hand_graph():
    n = 13
    nb = {i: [] for i in range(n)}
    for i, j in EDGES:
        nb[i].append(j)
        nb[j].append(i)
    ia, ja, a = [0], [], []
    for i in range(n):
        c = STRONG_C.get(i, 0.5)
        cols = sorted(nb[i])
        ja += [i] + cols
        a += [c + len(cols)] + [-1.0] * len(cols)
        ia.append(len(ja))
    return np.array(ia, np.uint32), np.array(ja, np.uint32), np.array(a)


# uh_cheb, sigma 3: t = -2v / (c (beta - 1)) = 1/c per link, so only the links among 4, 9, 11 are marked
# from both ends and dropped.  Tails (cheb_mark.cc:24-64), rows ascending:
#   1 -> 0 stops (0 keeps two links); 3 -> 5 -> 7 and the end node 7;
#   4 -> 0 -> 2 -> 6 -> 8 (0 is down to one link after 1 left) and the end node 8.
# Unmarked nodes without links (9, 10, 11, 12) fill the map from the back.
UH_CHEB_MAP = [1, 3, 5, 7, 4, 0, 2, 6, 8, 12, 11, 10, 9]
# multi_split original, q = 0.5: budget c per row, so 4, 9 and 11 drop all their out-links and 0 keeps 0 -> 4.
# Tails need one out-link and at most one in-link: 1 (back link from 0); 3 -> 5 -> 7 then end node 7;
# 8 -> 6 -> 2 -> 0 -> 4 (0 enters with no in-link left, its out-link to 4 has no back link) then end node 4.
# Group 2 (<= 2 links in or out) takes 9, 10, 11, 12 in row order; nothing is kept or isolated.
MSPLIT_ORIGINAL_MAP = [1, 3, 5, 7, 8, 6, 2, 0, 4, 9, 10, 11, 12]
# multi_split simple_1, degree_max 2: out-degree buckets 0 (with in-links): [4]; 1: [1, 3, 7, 8]; 2: [2, 5, 6];
# then the rest: [0]; then nodes without links [9, 10, 11, 12].  M = 8, kept {0}, Md = 4.
MSPLIT_SIMPLE1_MAP = [4, 1, 3, 7, 8, 2, 5, 6, 0, 9, 10, 11, 12]


def test_hand_worked_uh_cheb_tail_order():
    ia, ja, a = hand_graph()
    P = O.cheb_prec(ia, ja, a, sigmas=(3.0,), niters=(2,))
    info, arr = P.level_arrays(0)
    assert P.nlevels == 1
    assert arr["map"].tolist() == UH_CHEB_MAP
    assert (info["M"], info["Md"], info["n_next"]) == (9, 4, 0)
    H = N.build_uh_cheb(ia, ja, a)
    hinfo, harr = H.level_arrays(0)
    assert harr["map"].tolist() == UH_CHEB_MAP and H.nlevels == 1


def test_hand_worked_multi_split_orders():
    ia, ja, a = hand_graph()
    P = O.msplit_prec(ia, ja, a, q=(0.5,), niters=(2,), order="original")
    info, arr = P.level_arrays(0)
    assert arr["map"].tolist() == MSPLIT_ORIGINAL_MAP
    assert (info["M"], info["Md"], info["n_next"]) == (13, 0, 0)
    P = O.msplit_prec(ia, ja, a, q=(0.5,), niters=(2,), order="simple_1", degree_max=2)
    info, arr = P.level_arrays(0)
    assert arr["map"].tolist() == MSPLIT_SIMPLE1_MAP
    assert (info["M"], info["Md"], info["n_next"]) == (8, 4, 1)
    for order, dm, want in (("original", 3, MSPLIT_ORIGINAL_MAP), ("simple_1", 2, MSPLIT_SIMPLE1_MAP)):
        H = N.build_multi_split(ia, ja, a, q=(0.5,), niters=(2,), order=order, degree_max=dm)
        assert H.level_arrays(0)[1]["map"].tolist() == want


def test_block_order_interiors_at_60x220x85():
    """order_block (multi_split_mark.cc:193-252): the non-isolated nodes inside the 5x5x5 blocks of the
    60x220x85 grid, block by block (x fastest), come first.  Without a direct coarse solver the reference
    cannot order level 1 (its size is not 1,122,000): construction stops there, in the oracle and here."""
    s = system("spe10_z85")
    nx, ny, nz = 60, 220, 85
    with pytest.raises(O.OracleError, match="Wrong mesh size"):
        O.msplit_prec(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="block")
    with pytest.raises(N.MlpError, match="Wrong mesh size"):
        N.build_multi_split(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="block")
    # level 0 alone: a coarse size below level 0's and above level 1's makes level 1 the last level
    P = O.msplit_prec(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="block", coarse_n=s.n - 1)
    H = N.build_multi_split(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="block", coarse_n=s.n - 1)
    assert P.nlevels == 2 and H.nlevels == 2
    info, arr = P.level_arrays(0)
    hinfo, harr = H.level_arrays(0)
    for k in ("map", "L_ia", "L_ja", "L_a", "U_ia", "U_ja", "U_a", "dval", "A_ia", "A_ja", "A_a"):
        assert np.array_equal(arr[k], harr[k]), k
    # partition: interiors of the blocks (3x3x3 each where the block is whole), links either way present
    interior = []
    for K in range(0, nz, 5):
        for J in range(0, ny, 5):
            for I in range(0, nx, 5):
                for k in range(K + 1, min(K + 5, nz) - 1):
                    for j in range(J + 1, min(J + 5, ny) - 1):
                        for i in range(I + 1, min(I + 5, nx) - 1):
                            interior.append((k * ny + j) * nx + i)
    M, Md = info["M"], info["Md"]
    iso = set(arr["map"][s.n - Md:].tolist())
    want = [i for i in interior if i not in iso]
    assert arr["map"][:M].tolist() == want
    kept = arr["map"][M:s.n - Md]
    assert np.all(np.diff(kept.astype(np.int64)) > 0)
    assert np.all(np.diff(arr["map"][s.n - Md:].astype(np.int64)) < 0)
    assert 0 < M < s.n - Md

"""uh_cheb construction on the device (mlp_hier_build_uh_cheb_dev): every
level's link dropping, tail ordering (peeling rounds, re-rooted free trees,
walks), permutation, sparse LU and next-level matrix run as kernels, each
level's matrix staying in device memory for the next.  Bar: the hierarchy
equals the host construction's exactly (every level's sizes, permutation, L,
U, next matrix, diagonal reciprocals and row budgets, bit for bit) and the
oracle's, on the workload systems up to 60x220x85, on random matrices whose
links tie in magnitude (the order among equal magnitudes decides which links
drop), on graphs whose dropped links leave paths, stars, cycles with tails and
free trees of every shape, a single long path (the whole level one free tree),
and the same errors (a structurally asymmetric matrix, too many levels).
"""
import numpy as np
import pytest

import oracle as O
from conftest import rng, system
from ml_prec_amd import native as N
from test_construction import same_hierarchy

pytestmark = pytest.mark.gpu


def assert_same_hierarchy(Hh, Hd):
    assert Hh.nlevels == Hd.nlevels
    for l in range(Hh.nlevels):
        ih, ah = Hh.level_arrays(l)
        idv, ad = Hd.level_arrays(l)
        assert ih == idv, (l, ih, idv)
        for k in ah:
            assert np.array_equal(ah[k], ad[k]), (l, k)


@pytest.mark.parametrize("name", ["tiny", "tiny_tpfa_c1", "small_spe10", "ln5_spe", "ln5_tpfa_c1", "spe10_z12",
                                  "spe10_z85"])
def test_device_construction_same_hierarchy(ctx, name):
    s = system(name)
    assert_same_hierarchy(N.build_uh_cheb(s.ia, s.ja, s.a), N.build_uh_cheb(s.ia, s.ja, s.a, ctx=ctx))


@pytest.mark.parametrize("name", ["tiny_tpfa_c1", "ln5_tpfa_c1", "spe10_z85"])
def test_device_construction_equals_oracle(ctx, name):
    s = system(name)
    assert same_hierarchy(O.cheb_prec(s.ia, s.ja, s.a), N.build_uh_cheb(s.ia, s.ja, s.a, ctx=ctx))


def test_device_construction_parameters(ctx):
    s = system("ln5_spe")
    kw = dict(sigmas=(2.0, 4.0), niters=(3, 2), use_tails=False)
    assert_same_hierarchy(N.build_uh_cheb(s.ia, s.ja, s.a, **kw), N.build_uh_cheb(s.ia, s.ja, s.a, ctx=ctx, **kw))


def _tied_symmetric(n, k, seed):
    """Symmetric M-matrix with up to ~2k random links a row whose values come from {-0.25, -0.5, -1}, so that
    many links of a row tie in magnitude; diagonal = sum of |links| + 1 (row sum 1)."""
    g = rng(seed)
    links = {}
    for i in range(1, n):
        for j in g.choice(i, size=min(i, k), replace=False):
            links[(int(j), i)] = float(g.choice([-0.25, -0.5, -1.0]))
    rows = [dict() for _ in range(n)]
    for (j, i), v in links.items():
        rows[i][j] = v
        rows[j][i] = v
    ia, ja, a = [0], [], []
    for i in range(n):
        cols = sorted(rows[i])
        ja += [i] + cols
        a += [1.0 - sum(rows[i].values())] + [rows[i][c] for c in cols]
        ia.append(len(ja))
    return np.array(ia, np.uint32), np.array(ja, np.uint32), np.array(a)


@pytest.mark.parametrize("k", [2, 5, 9])
def test_device_construction_ties(ctx, k):
    ia, ja, a = _tied_symmetric(4000, k, 90 + k)
    assert_same_hierarchy(N.build_uh_cheb(ia, ja, a), N.build_uh_cheb(ia, ja, a, ctx=ctx))


def test_device_construction_asymmetric_structure(ctx):
    # row 0 links to 1, row 1 does not link back
    ia = np.array([0, 2, 3, 5], np.uint32)
    ja = np.array([0, 1, 1, 2, 1], np.uint32)
    a = np.array([2.0, -1.0, 2.0, 2.0, -1.0])
    msgs = []
    for ctx_arg in (None, ctx):
        with pytest.raises(N.MlpError) as e:
            N.build_uh_cheb(ia, ja, a, ctx=ctx_arg)
        msgs.append((e.value.code, str(e.value)))
    assert msgs[0] == msgs[1] and "structurally symmetric" in msgs[0][1]


def _from_links(n, links, c):
    """Skyline rows of the symmetric matrix with a_ij = -t for each link (i, j, t) and a_ii = c_i + sum t."""
    rows = [dict() for _ in range(n)]
    diag = np.array(c, dtype=np.float64)
    for i, j, t in links:
        rows[i][j] = -t
        rows[j][i] = -t
        diag[i] += t
        diag[j] += t
    ia, ja, a = [0], [], []
    for i in range(n):
        cols = sorted(rows[i])
        ja += [i] + cols
        a += [diag[i]] + [rows[i][k] for k in cols]
        ia.append(len(ja))
    return np.array(ia, np.uint32), np.array(ja, np.uint32), np.array(a)


def _forest_graph(n, seed):
    """Paths, stars and cycles with tails over a shuffled node order, joined by a few random links; link
    weights log-normal with a wide spread and 30% of the rows without a reaction term, so that link dropping
    leaves free trees, trees hanging off cycles and long chains."""
    g = rng(seed)
    perm = g.permutation(n)
    links, seen = [], set()

    def add(i, j):
        i, j = int(min(i, j)), int(max(i, j))
        if i != j and (i, j) not in seen:
            seen.add((i, j))
            links.append((i, j, float(np.exp(g.normal(0.0, 2.0)))))

    k = 0
    while k < n:
        u, ln = g.random(), int(min(n - k, 2 + int(g.random() * 12)))
        if u < 0.4:
            for t in range(1, ln):
                add(perm[k + t - 1], perm[k + t])
        elif u < 0.7:
            for t in range(1, ln):
                add(perm[k], perm[k + t])
        elif u < 0.85:
            for t in range(1, ln):
                add(perm[k + t - 1], perm[k + t])
            if ln > 2:
                add(perm[k], perm[k + ln - 1])
        k += ln
    for _ in range(n // 8):
        add(g.integers(n), g.integers(n))
    c = np.where(g.random(n) < 0.3, 0.0, np.exp(g.normal(0.0, 2.0, n)))
    return _from_links(n, links, c)


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("sigma", [1.5, 3.0, 8.0])
def test_device_construction_forests(ctx, seed, sigma):
    ia, ja, a = _forest_graph(20000, 500 + seed)
    kw = dict(sigmas=(sigma,))
    try:
        Hh = N.build_uh_cheb(ia, ja, a, **kw)
    except N.MlpError as e:                       # the same error from the device
        with pytest.raises(N.MlpError) as d:
            N.build_uh_cheb(ia, ja, a, ctx=ctx, **kw)
        assert (d.value.code, str(d.value)) == (e.code, str(e))
        return
    assert_same_hierarchy(Hh, N.build_uh_cheb(ia, ja, a, ctx=ctx, **kw))


def test_device_construction_long_path(ctx):
    """A 5000-node path: every node is a tail of one free tree rooted at its last node (the reference's
    'end node in fully tridiagonal matrix', cheb_mark.cc:56-62); 2500 peeling rounds."""
    n = 5000
    ia, ja, a = _from_links(n, [(i, i + 1, 1.0 + (i % 7) * 0.25) for i in range(n - 1)], np.full(n, 0.5))
    Hd = N.build_uh_cheb(ia, ja, a, ctx=ctx)
    assert Hd.nlevels == 1 and Hd.level_info(0)["M"] == n
    assert_same_hierarchy(N.build_uh_cheb(ia, ja, a), Hd)


def test_device_construction_too_many_levels(ctx):
    s = system("ln5_tpfa")
    with pytest.raises(N.MlpError, match="Too many levels: 30"):
        N.build_uh_cheb(s.ia, s.ja, s.a, ctx=ctx)

"""uh_cheb construction with each level's link phase on the device
(mlp_hier_build_uh_cheb_dev): row sums, the magnitude order of the links,
the budget marks, the two-sided counters and the degrees come from the GPU,
the tail walk and the sparse LU from the host.  Bar: the hierarchy equals the
host construction's exactly (every level's sizes, permutation, L, U, next
matrix, diagonal reciprocals and row budgets, bit for bit), on the workload
systems up to 60x220x85, on random matrices whose links tie in magnitude
(the order among equal magnitudes decides which links drop), and the same
error for a structurally asymmetric matrix.
"""
import numpy as np
import pytest

from conftest import rng, system
from ml_prec_amd import native as N

pytestmark = pytest.mark.gpu


def assert_same_hierarchy(Hh, Hd):
    assert Hh.nlevels == Hd.nlevels
    for l in range(Hh.nlevels):
        ih, ah = Hh.level_arrays(l)
        idv, ad = Hd.level_arrays(l)
        assert ih == idv, (l, ih, idv)
        for k in ah:
            assert np.array_equal(ah[k], ad[k]), (l, k)


@pytest.mark.parametrize("name", ["tiny", "small_spe10", "ln5_spe", "ln5_tpfa_c1", "spe10_z12", "spe10_z85"])
def test_device_links_same_hierarchy(ctx, name):
    s = system(name)
    assert_same_hierarchy(N.build_uh_cheb(s.ia, s.ja, s.a), N.build_uh_cheb(s.ia, s.ja, s.a, ctx=ctx))


def test_device_links_parameters(ctx):
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
def test_device_links_ties(ctx, k):
    ia, ja, a = _tied_symmetric(4000, k, 90 + k)
    assert_same_hierarchy(N.build_uh_cheb(ia, ja, a), N.build_uh_cheb(ia, ja, a, ctx=ctx))


def test_device_links_asymmetric_structure(ctx):
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

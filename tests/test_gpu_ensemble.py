"""Ensembles in one launch per step: systems with one 60x220 footprint and
different depths stacked into one block-diagonal matrix on the device
(mlp_spe_ensemble_matrix_dev), one uh_cheb hierarchy built on it
(mlp_hier_build_uh_cheb_mat), solved by mlp_pcg_ensemble.

A uh_cheb hierarchy of a block-diagonal matrix is the union of its blocks'
own hierarchies (link dropping is per row, tails and eliminations never cross
a block, the next level is block diagonal); with blocks of equal depth the
Chebyshev bounds, and so the whole preconditioner, are each block's own.
Bar: the stacked matrix's blocks equal the systems' own matrices bit for bit;
for systems grouped by hierarchy depth, every system's solution and iteration
count equal the single mlp_pcg solve's bit for bit (FP64, plain and flexible
with restarts, graphs on and off); the FP32 hierarchy in flexible PCG meets
the FP64 true residual with at most two iterations more than FP64 and a
solution within 1e3 eps of the FP64 one; a group of unequal depths (a valid
preconditioner, not each system's own) still converges.
"""
import numpy as np
import pytest

from ml_prec_amd import native as N
from ml_prec_amd import spe_synth as S

pytestmark = pytest.mark.gpu

SPEC = [(5, 201), (10, 202), (5, 203), (20, 204), (10, 205), (12, 206), (8, 207), (5, 208), (10, 209), (20, 210),
        (40, 211), (6, 212)]
H = (S.HX, S.HY, S.HZ)


class Member:
    def __init__(self, ctx, nz, seed, precision):
        self.nz = nz
        self.kx = S.spe10_field(60, 220, nz, S.BASE_SEED + seed).ravel()
        self.n = self.kx.size
        self.b = S.rhs(self.n, S.BASE_SEED + seed + 7919)
        self.A = N.DevMatrix.spe_on_device(ctx, ctx.to_device(self.kx), (nz, 220, 60), kind="spe", c=1.0,
                                           kz_ratio=0.1, h=H)
        self.H = N.build_uh_cheb_on_device(ctx, self.A)
        self.depth = self.H.nlevels
        self.B = N.DevPrec(ctx, self.A, self.H, precision=precision)

    def close(self):
        self.B.close()
        self.A.close()


def _members(ctx, precision=N.MLP_FP64, spec=SPEC):
    return [Member(ctx, nz, seed, precision) for nz, seed in spec]


def _stack(ctx, ms, precision=N.MLP_FP64):
    kx = np.concatenate([m.kx for m in ms])
    A = N.DevMatrix.spe_ensemble_on_device(ctx, ctx.to_device(kx), 60, 220, [m.nz for m in ms], kind="spe", c=1.0,
                                           kz_ratio=0.1, h=H)
    Hg = N.build_uh_cheb_on_device(ctx, A)
    B = N.DevPrec(ctx, A, Hg, precision=precision)
    off = np.concatenate([[0], np.cumsum([m.n for m in ms])]).astype(np.uint32)
    return A, Hg, B, off


def _groups(ms):
    g = {}
    for m in ms:
        g.setdefault(m.depth, []).append(m)
    return g


def test_stacked_matrix_blocks(ctx):
    ms = _members(ctx, spec=SPEC[:4])
    A, Hg, B, off = _stack(ctx, ms)
    ia, ja, a = A.csr()
    for m, o in zip(ms, off[:-1]):
        mia, mja, ma = m.A.csr()
        base = ia[o]
        blk = slice(int(base), int(ia[o + m.n]))
        assert np.array_equal(ia[o:o + m.n + 1] - base, mia)
        assert np.array_equal(ja[blk] - o, mja)
        assert np.array_equal(a[blk].view(np.uint64), ma.view(np.uint64))
    # the stacked hierarchy keeps every level of the deepest member
    assert Hg.nlevels == max(m.depth for m in ms)
    B.close()
    A.close()
    for m in ms:
        m.close()


def _single(ctx, m, **kw):
    b = ctx.to_device(m.b)
    x = ctx.array(m.n)
    st = N.pcg(ctx, m.A, m.B, b, x, **kw)
    return st, x.get()


@pytest.mark.parametrize("flexible,eps,use_graph", [(False, 1e-8, True), (False, 1e-8, False), (True, 1e-12, True)])
def test_ensemble_equals_single_solves(ctx, flexible, eps, use_graph):
    ms = _members(ctx)
    kw = dict(eps=eps, maxit=5000, x0_mode=N.X0_REFERENCE, check_every=4, use_graph=use_graph)
    if flexible:
        kw.update(flexible=True, true_res_every=5)
    groups = _groups(ms)
    assert len(groups) >= 2 and max(len(g) for g in groups.values()) >= 3
    for depth, g in groups.items():
        A, Hg, B, off = _stack(ctx, g)
        assert Hg.nlevels == depth
        b = ctx.to_device(np.concatenate([m.b for m in g]))
        x = ctx.array(int(off[-1]))
        x.upload(np.full(x.n, 3.0))
        rc, stats = N.pcg_ensemble(ctx, A, B, off, b, x, **kw)
        assert rc == 0, N.lib().mlp_last_error().decode()
        xs = x.get()
        for m, st, o in zip(g, stats, off[:-1]):
            st1, x1 = _single(ctx, m, **kw)
            assert st["niter"] == st1["niter"], (depth, m.nz, st["niter"], st1["niter"])
            assert np.array_equal(xs[o:o + m.n].view(np.uint64), x1.view(np.uint64)), (depth, m.nz)
            assert st["true_rel_res"] == st1["true_rel_res"] and st["final_norm"] == st1["final_norm"]
            assert st["true_rel_res"] <= eps * (1 + 1e-3) or not flexible
        B.close()
        A.close()
    for m in ms:
        m.close()


def test_ensemble_mixed_precision(ctx):
    ms64 = _members(ctx)
    ms32 = [Member(ctx, nz, seed, N.MLP_FP32) for nz, seed in SPEC]
    eps = 1e-8
    kw = dict(eps=eps, maxit=5000, x0_mode=N.X0_ZERO, check_every=4, use_graph=True)
    for depth, g in _groups(ms32).items():
        A, Hg, B, off = _stack(ctx, g, precision=N.MLP_FP32)
        b = ctx.to_device(np.concatenate([m.b for m in g]))
        x = ctx.array(int(off[-1]))
        rc, stats = N.pcg_ensemble(ctx, A, B, off, b, x, flexible=True, true_res_every=20, **kw)
        assert rc == 0, N.lib().mlp_last_error().decode()
        xs = x.get()
        for m, st, o in zip(g, stats, off[:-1]):
            m64 = next(q for q in ms64 if q.kx is not None and q.nz == m.nz and np.array_equal(q.b, m.b))
            st64, x64 = _single(ctx, m64, **kw)
            assert st["true_rel_res"] <= eps * (1 + 1e-3)
            assert st["niter"] <= st64["niter"] + 2
            rel = np.linalg.norm(xs[o:o + m.n] - x64) / np.linalg.norm(x64)
            assert rel <= 1e3 * eps, (depth, m.nz, rel)
        B.close()
        A.close()
    for m in ms64 + ms32:
        m.close()


def test_ensemble_unequal_depths_converges(ctx):
    ms = _members(ctx)
    assert len({m.depth for m in ms}) >= 2
    A, Hg, B, off = _stack(ctx, ms)
    b = ctx.to_device(np.concatenate([m.b for m in ms]))
    x = ctx.array(int(off[-1]))
    rc, stats = N.pcg_ensemble(ctx, A, B, off, b, x, eps=1e-8, maxit=5000, x0_mode=N.X0_ZERO, check_every=4)
    assert rc == 0
    for st in stats:
        assert st["true_rel_res"] <= 1e-8 * (1 + 1e-3)
    B.close()
    A.close()
    for m in ms:
        m.close()


def test_ensemble_errors(ctx):
    ms = _members(ctx, spec=SPEC[:2])
    A, Hg, B, off = _stack(ctx, ms)
    b = ctx.to_device(np.ones(int(off[-1])))
    x = ctx.array(int(off[-1]))
    for bad in ([0, 5, 5, int(off[-1])], [0, int(off[-1]) - 1], [1, int(off[-1])]):
        rc, _ = N.pcg_ensemble(ctx, A, B, bad, b, x)
        assert rc == 1          # MLP_ERR_INVALID
    B.close()
    A.close()
    for m in ms:
        m.close()


def test_ensemble_batch_of_groups_equals_single_solves():
    """Every depth group on its own context, all driven from one thread (mlp_pcg_ensemble_batch): each
    system's solution and iteration count are its single solve's, bit for bit."""
    ctx0 = N.Context(0)
    ms = _members(ctx0)
    groups = list(_groups(ms).values())
    kw = dict(eps=1e-8, maxit=5000, x0_mode=N.X0_ZERO, check_every=4, use_graph=True)
    jobs, keep = [], []
    for g in groups:
        c = N.Context(0)
        A, Hg, B, off = _stack(c, g)
        b = c.to_device(np.concatenate([m.b for m in g]))
        x = c.array(int(off[-1]))
        jobs.append((c, A, B, off, b, x))
        keep.append(g)
    rc, res = N.pcg_ensemble_batch(jobs, **kw)
    assert rc == 0
    for (c, A, B, off, b, x), g, (status, stats) in zip(jobs, keep, res):
        assert status == 0
        xs = x.get()
        for m, st, o in zip(g, stats, off[:-1]):
            st1, x1 = _single(ctx0, m, **kw)
            assert st["niter"] == st1["niter"]
            assert np.array_equal(xs[o:o + m.n].view(np.uint64), x1.view(np.uint64))
    for c, A, B, off, b, x in jobs:
        B.close()
        A.close()
    del b, x
    for c, *_ in jobs:
        c.close()
    for m in ms:
        m.close()
    ctx0.close()

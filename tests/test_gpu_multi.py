"""Many right-hand sides of one system in one batched solve (mlp_pcg_multi).

Every launch covers all right-hand sides (grid.y), each with its own vectors,
state and partial sums, so each solve must equal the single-right-hand-side
mlp_pcg on the same vector bit for bit: same iteration count, same x.  Also
the per-right-hand-side outcomes (one right-hand side that needs no iteration,
an iteration cap), the mixed-precision mode (FP32 hierarchy, flexible PCG with
true-residual replacement and restarts) and the argument checks.
"""
import numpy as np
import pytest

from conftest import rng, system
from ml_prec_amd import native as N

pytestmark = pytest.mark.gpu

# return codes of include/mlp.h
ERR_INVALID, ERR_NO_ITER, ERR_MAXIT = 1, 4, 5


def _rhs(s, R, seed):
    g = rng(seed)
    return np.stack([s.b] + [g.uniform(-1.0, 1.0, s.n) for _ in range(R - 1)])


def _single(ctx, A, B, b, **kw):
    x = ctx.array(b.size)
    st = N.pcg(ctx, A, B, ctx.to_device(b), x, **kw)
    return st, x.get()


@pytest.mark.parametrize("name,R,graph", [("spe10_z12", 5, True), ("spe10_z12", 3, False), ("ln5_tpfa_c1", 8, True)])
def test_multi_equals_single(ctx, name, R, graph):
    s = system(name)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, N.build_uh_cheb(s.ia, s.ja, s.a))
    bs = _rhs(s, R, 21)
    X = ctx.array(R * s.n)
    rc, stats = N.pcg_multi(ctx, A, B, R, ctx.to_device(bs.ravel()), X, eps=1e-8, maxit=2000, use_graph=graph)
    assert rc == 0, stats
    xs = X.get().reshape(R, s.n)
    for k in range(R):
        st, x1 = _single(ctx, A, B, bs[k], eps=1e-8, maxit=2000, x0_mode=N.X0_ZERO, use_graph=graph)
        assert stats[k]["niter"] == st["niter"], (k, stats[k]["niter"], st["niter"])
        assert np.array_equal(xs[k], x1), k
        assert stats[k]["final_norm"] == st["final_norm"]
        assert stats[k]["true_rel_res"] <= 1e-8 * (1 + 1e-3)


def test_multi_mixed_precision_equals_single(ctx):
    """FP32 hierarchy, flexible PCG, true residual every 7 iterations, then the restarts."""
    s = system("spe10_z12")
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, N.build_uh_cheb(s.ia, s.ja, s.a), precision=N.MLP_FP32)
    R = 4
    bs = _rhs(s, R, 24)
    X = ctx.array(R * s.n)
    kw = dict(eps=1e-10, maxit=2000, flexible=True, true_res_every=7)
    rc, stats = N.pcg_multi(ctx, A, B, R, ctx.to_device(bs.ravel()), X, **kw)
    assert rc == 0, stats
    xs = X.get().reshape(R, s.n)
    for k in range(R):
        st, x1 = _single(ctx, A, B, bs[k], x0_mode=N.X0_ZERO, use_graph=True, **kw)
        assert stats[k]["niter"] == st["niter"], (k, stats[k]["niter"], st["niter"])
        assert np.array_equal(xs[k], x1), k
        assert stats[k]["true_rel_res"] <= 1e-9


def test_multi_reference_start_and_repeat(ctx):
    s = system("spe10_z12")
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, N.build_uh_cheb(s.ia, s.ja, s.a))
    R = 3
    bs = _rhs(s, R, 22)
    X = ctx.array(R * s.n)
    for _ in range(2):                      # the second call replays the cached graph
        rc, stats = N.pcg_multi(ctx, A, B, R, ctx.to_device(bs.ravel()), X, eps=1e-8, maxit=2000,
                                x0_mode=N.X0_REFERENCE)
        assert rc == 0
    xs = X.get().reshape(R, s.n)
    for k in range(R):
        st, x1 = _single(ctx, A, B, bs[k], eps=1e-8, maxit=2000, x0_mode=N.X0_REFERENCE, use_graph=True)
        assert stats[k]["niter"] == st["niter"]
        assert np.array_equal(xs[k], x1)


def test_multi_per_rhs_outcomes(ctx):
    s = system("spe10_z12")
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, N.build_uh_cheb(s.ia, s.ja, s.a))
    bs = _rhs(s, 3, 23)
    bs[1] = 0.0                             # "No need to iterate" for this one only
    X = ctx.array(3 * s.n)
    rc, stats = N.pcg_multi(ctx, A, B, 3, ctx.to_device(bs.ravel()), X, eps=1e-8, maxit=2000)
    assert rc == ERR_NO_ITER
    assert [st["status"] for st in stats] == [0, ERR_NO_ITER, 0]
    assert stats[0]["true_rel_res"] <= 1e-8 * (1 + 1e-3) and stats[2]["true_rel_res"] <= 1e-8 * (1 + 1e-3)
    rc, stats = N.pcg_multi(ctx, A, B, 2, ctx.to_device(bs[[0, 2]].ravel()), X, eps=1e-12, maxit=3)
    assert rc == ERR_MAXIT
    assert all(st["status"] == ERR_MAXIT and st["niter"] == 3 for st in stats)


def test_multi_argument_checks(ctx):
    s = system("spe10_z12")
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, N.build_uh_cheb(s.ia, s.ja, s.a))
    b = ctx.to_device(np.tile(s.b, 2))
    X = ctx.array(2 * s.n)
    for R in (0, 65):
        rc, _ = N.pcg_multi(ctx, A, B, R, b, X)
        assert rc == ERR_INVALID
    o_rc, _ = N.pcg_multi(ctx, A, N.DevPrec(ctx, A, kind=N.DIAG), 2, b, X)
    assert o_rc == ERR_INVALID

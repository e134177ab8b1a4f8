"""Concurrent PCG solves (mlp_pcg_batch) on the device.

Bar: every job's solution and iteration count are bitwise those of a single
mlp_pcg call on the same system (same kernels, same reduction order), and the
FP64 jobs agree with the CPU oracle within the PCG bars of test_gpu_parity.
"""
import numpy as np
import pytest

import oracle as O
from ml_prec_amd import native as N
from ml_prec_amd import spe_synth as S

pytestmark = pytest.mark.gpu

SPEC = [(5, 101), (8, 102), (12, 103), (5, 104), (20, 105)]


def _systems():
    return [S.ensemble_system(nz, S.BASE_SEED + seed) for nz, seed in SPEC]


def _setup(systems, precision=N.MLP_FP64):
    out = []
    for s in systems:
        ctx = N.Context(0)
        A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
        B = N.DevPrec(ctx, A, N.build_uh_cheb(s.ia, s.ja, s.a), precision=precision)
        out.append((ctx, A, B, ctx.to_device(s.b), ctx.array(s.n)))
    return out


def _close(jobs):
    for ctx, A, B, b, x in jobs:
        B.close()
        A.close()
    for ctx, *_ in jobs:
        ctx.close()


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("mixed", [False, True])
def test_batch_equals_single_solves(use_graph, mixed):
    systems = _systems()
    prec = N.MLP_FP32 if mixed else N.MLP_FP64
    kw = dict(eps=1e-8, maxit=5000, x0_mode=N.X0_REFERENCE, check_every=5, use_graph=use_graph)
    if mixed:
        kw.update(flexible=True, true_res_every=20)
    jobs = _setup(systems, prec)
    single = []
    for ctx, A, B, b, x in jobs:
        st = N.pcg(ctx, A, B, b, x, **kw)
        single.append((st, x.get()))
    for _, _, _, _, x in jobs:
        x.upload(np.full(x.n, 7.0))
    rc, stats = N.pcg_batch(jobs, **kw)
    assert rc == 0
    for (st1, x1), st2, (_, _, _, _, x) in zip(single, stats, jobs):
        assert st2["rc"] == 0
        assert st2["niter"] == st1["niter"]
        assert np.array_equal(x.get(), x1)
        assert st2["true_rel_res"] == st1["true_rel_res"]
    _close(jobs)


def test_batch_vs_oracle():
    systems = _systems()[:3]
    jobs = _setup(systems)
    rc, stats = N.pcg_batch(jobs, eps=1e-8, maxit=5000, x0_mode=N.X0_ZERO, check_every=8, use_graph=True)
    assert rc == 0
    for s, st, (_, _, _, _, x) in zip(systems, stats, jobs):
        P = O.cheb_prec(s.ia, s.ja, s.a)
        xo, so = O.pcg(s.ia, s.ja, s.a, s.b, P, eps=1e-8, x0=np.zeros(s.n), maxit=5000)
        assert abs(int(st["niter"]) - so["niter"]) <= 1
        xg = x.get()
        tol = 1e-10 if st["niter"] == so["niter"] else 1e-6
        assert np.linalg.norm(xg - xo) <= tol * np.linalg.norm(xo)
    _close(jobs)


def test_batch_job_errors_are_per_job():
    systems = _systems()[:3]
    jobs = _setup(systems)
    ctx, A, B, b, x = jobs[1]
    b.upload(np.zeros(x.n))                          # b = 0, x0 = 0 -> "No need to iterate"
    rc, stats = N.pcg_batch(jobs, eps=1e-8, maxit=5000, x0_mode=N.X0_ZERO)
    assert rc == 4
    assert [st["rc"] for st in stats] == [0, 4, 0]
    assert stats[0]["true_rel_res"] <= 1e-8 and stats[2]["true_rel_res"] <= 1e-8
    _close(jobs)


def test_batch_rejects_shared_context():
    s = _systems()[0]
    ctx = N.Context(0)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, kind=N.DIAG)
    b, x1, x2 = ctx.to_device(s.b), ctx.array(s.n), ctx.array(s.n)
    rc, _ = N.pcg_batch([(ctx, A, B, b, x1), (ctx, A, B, b, x2)])
    assert rc == 1
    B.close()
    A.close()
    ctx.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_shared_hierarchy_many_rhs(mixed):
    """One matrix and one hierarchy, several right-hand sides solved concurrently: the jobs share
    the matrix and use clones of one preconditioner (mlp_prec_clone); each result is bitwise the
    single solve with the source preconditioner.  The source is closed before the batch runs, so
    the clones keep the shared device data alive."""
    s = S.ensemble_system(12, S.BASE_SEED + 110)
    prec = N.MLP_FP32 if mixed else N.MLP_FP64
    kw = dict(eps=1e-8, maxit=5000, x0_mode=N.X0_ZERO, check_every=6, use_graph=True)
    if mixed:
        kw.update(flexible=True, true_res_every=20)
    ctx0 = N.Context(0)
    A = N.DevMatrix(ctx0, s.ia, s.ja, s.a)
    B0 = N.DevPrec(ctx0, A, N.build_uh_cheb(s.ia, s.ja, s.a), precision=prec)
    rhs = [np.random.default_rng(120 + k).standard_normal(s.n) for k in range(4)]
    single = []
    for f in rhs:
        x = ctx0.array(s.n)
        st = N.pcg(ctx0, A, B0, ctx0.to_device(f), x, **kw)
        single.append((st["niter"], x.get()))
    ctxs = [N.Context(0) for _ in rhs]
    clones = [B0.clone(c) for c in ctxs]
    B0.close()
    jobs = [(c, A, Bc, c.to_device(f), c.array(s.n)) for c, Bc, f in zip(ctxs, clones, rhs)]
    rc, stats = N.pcg_batch(jobs, **kw)
    assert rc == 0
    for (niter, x1), st, job in zip(single, stats, jobs):
        assert st["niter"] == niter
        assert np.array_equal(job[4].get(), x1)
    for Bc in clones:
        Bc.close()
    A.close()
    for c in ctxs:
        c.close()
    ctx0.close()

"""The level chain in two-wavefront blocks (components of 65..128 nodes).

Most SPE10-shaped 60x220x85 fields have a level-3 component of more than 64
nodes (the adjoint's gradient field: 76), which no one-wavefront chain block
holds: by default such a hierarchy runs level kernels and the fused last pair,
and the chain in 128-thread blocks is built when asked for (MLP_CHAIN_BS=128,
or the autotuner's candidate).  Bar: the schedules are the ones described; the
wide chain's FP64 apply is bitwise the oracle's; PCG with it is bitwise the
solve with the default schedule, iterations equal; the FP32 hierarchy's apply
stays within its error bar; the tuner times the wide chain and keeps the
fastest candidate; forcing the wide chain where the narrow one fits changes no
bit; batched right-hand sides (mlp_pcg_multi) equal the single solves.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rng, system
from ml_prec_amd import native as N
from ml_prec_amd import spe_synth as S

pytestmark = pytest.mark.gpu

_GRAD = {}


def gradient_field_system():
    """The adjoint benchmark's second field (60x220x85), its matrix and right-hand side."""
    if not _GRAD:
        nz = 85
        kx = S.spe10_field(60, 220, nz, S.BASE_SEED + 40)
        ia, ja, a = N.spe_assemble(kx.reshape((nz, 220, 60)), kind="spe", c=1.0, kz_ratio=0.1,
                                   h=(S.HX, S.HY, S.HZ))
        _GRAD.update(ia=ia, ja=ja, a=a, n=len(ia) - 1, b=S.rhs(len(ia) - 1, S.BASE_SEED + 1))
    return _GRAD


def test_wide_chain_on_the_gradient_field(ctx, monkeypatch):
    s = gradient_field_system()
    A = N.DevMatrix(ctx, s["ia"], s["ja"], s["a"])
    H = N.build_uh_cheb(s["ia"], s["ja"], s["a"])
    BL = N.DevPrec(ctx, A, H)                                   # the default: no narrow chain fits
    assert (BL.schedule() & 0xff) == N.SCHED_PAIR, BL.schedule()
    monkeypatch.setenv("MLP_CHAIN_BS", "128")
    B = N.DevPrec(ctx, A, H)
    B32 = N.DevPrec(ctx, A, H, precision=N.MLP_FP32)
    monkeypatch.delenv("MLP_CHAIN_BS")
    sched = B.schedule()
    assert sched & N.SCHED_CHAIN and sched & N.SCHED_WIDE_CHAIN, sched
    f = rng(11).standard_normal(s["n"])
    x = ctx.array(s["n"])
    B.apply(ctx.to_device(f), x)
    P = O.cheb_prec(s["ia"], s["ja"], s["a"])
    assert np.array_equal(x.get().view(np.uint64), P.apply(f).view(np.uint64))
    assert B32.schedule() & N.SCHED_WIDE_CHAIN
    x32 = ctx.array(s["n"])
    B32.apply(ctx.to_device(f), x32)
    ref = x.get()
    assert np.linalg.norm(x32.get() - ref) / np.linalg.norm(ref) < 1e-6
    # PCG with the wide chain against the default schedule
    out = []
    for Bk in (B, BL):
        xs = ctx.array(s["n"])
        st = N.pcg(ctx, A, Bk, ctx.to_device(s["b"]), xs, eps=1e-10, maxit=2000, x0_mode=N.X0_ZERO,
                   use_graph=True, check_every=8)
        out.append((st["niter"], xs.get()))
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1].view(np.uint64), out[1][1].view(np.uint64))
    # the tuner times the wide chain among its candidates and keeps the fastest
    T = N.DevPrec.tuned(ctx, A, H, precision=N.MLP_FP32, reps=5)
    cands = T.tune["candidates"]
    assert any(c & N.SCHED_WIDE_CHAIN for c, _ in cands), cands
    assert cands[T.tune["chosen"]][1] == min(t for _, t in cands)
    for o in (B, B32, BL, T, A):
        o.close()


@pytest.mark.parametrize("name", ["spe10_z12", "spe10_z40"])
def test_forced_wide_chain_changes_no_bit(ctx, monkeypatch, name):
    s = system(name)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    H = N.build_uh_cheb(s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, H)
    monkeypatch.setenv("MLP_CHAIN_BS", "128")
    W = N.DevPrec(ctx, A, H)
    monkeypatch.delenv("MLP_CHAIN_BS")
    assert B.schedule() & N.SCHED_CHAIN and not B.schedule() & N.SCHED_WIDE_CHAIN
    assert W.schedule() & N.SCHED_CHAIN and W.schedule() & N.SCHED_WIDE_CHAIN
    f = rng(12).standard_normal(s.n)
    xb, xw = ctx.array(s.n), ctx.array(s.n)
    B.apply(ctx.to_device(f), xb)
    W.apply(ctx.to_device(f), xw)
    P = O.cheb_prec(s.ia, s.ja, s.a)
    ref = P.apply(f).view(np.uint64)
    assert np.array_equal(xb.get().view(np.uint64), ref)
    assert np.array_equal(xw.get().view(np.uint64), ref)
    # three right-hand sides in one batched solve equal the single solves
    R = 3
    bs = np.concatenate([s.b * (1.0 + k) for k in range(R)])
    X = ctx.array(R * s.n)
    rc, stats = N.pcg_multi(ctx, A, W, R, ctx.to_device(bs), X, eps=1e-8, maxit=1000, x0_mode=N.X0_ZERO)
    assert rc == 0
    Xh = X.get().reshape(R, s.n)
    for k in range(R):
        xk = ctx.array(s.n)
        st = N.pcg(ctx, A, W, ctx.to_device(bs[k * s.n:(k + 1) * s.n]), xk, eps=1e-8, maxit=1000,
                   x0_mode=N.X0_ZERO, use_graph=True)
        assert st["niter"] == stats[k]["niter"]
        assert np.array_equal(xk.get().view(np.uint64), Xh[k].view(np.uint64))
    for o in (B, W, A):
        o.close()

"""The launch-schedule autotuner (mlp_prec_from_hier_tuned): candidate
schedules of one hierarchy timed on the device, the fastest kept.

Bar: every distinct candidate is reported with a positive time and the kept
one is the fastest; the tuned FP64 apply is bitwise the oracle's (every
schedule gives the reference's arithmetic); the tuned FP32 apply stays within
the FP32 hierarchy's error bar; a PCG solve with the tuned preconditioner is
bitwise the solve with the default one (FP64).
"""
import numpy as np
import pytest

import oracle as O
from conftest import rng, system
from ml_prec_amd import native as N

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["spe10_z12", "spe10_z85"])
def test_tuned_fp64_apply_is_the_oracles(ctx, name):
    s = system(name)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    H = N.build_uh_cheb(s.ia, s.ja, s.a)
    B = N.DevPrec.tuned(ctx, A, H, precision=N.MLP_FP64, reps=5)
    cands = B.tune["candidates"]
    assert len(cands) >= 3 and len({c for c, _ in cands}) == len(cands)
    times = [t for _, t in cands]
    assert min(times) > 0 and times[B.tune["chosen"]] == min(times)
    f = rng(5).standard_normal(s.n)
    x = ctx.array(s.n)
    B.apply(ctx.to_device(f), x)
    P = O.cheb_prec(s.ia, s.ja, s.a)
    assert np.array_equal(x.get().view(np.uint64), P.apply(f).view(np.uint64))
    B.close()
    A.close()


def test_tuned_fp32_apply_and_solve(ctx):
    s = system("spe10_z12")
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    H = N.build_uh_cheb(s.ia, s.ja, s.a)
    T32 = N.DevPrec.tuned(ctx, A, H, precision=N.MLP_FP32, reps=5)
    D64 = N.DevPrec(ctx, A, H, precision=N.MLP_FP64)
    T64 = N.DevPrec.tuned(ctx, A, H, precision=N.MLP_FP64, reps=5)
    f = rng(6).standard_normal(s.n)
    x32, x64 = ctx.array(s.n), ctx.array(s.n)
    T32.apply(ctx.to_device(f), x32)
    D64.apply(ctx.to_device(f), x64)
    ref = x64.get()
    assert np.linalg.norm(x32.get() - ref) / np.linalg.norm(ref) < 1e-6
    kw = dict(eps=1e-8, maxit=1000, x0_mode=N.X0_ZERO, use_graph=True)
    b = ctx.to_device(s.b)
    xa, xb = ctx.array(s.n), ctx.array(s.n)
    sa = N.pcg(ctx, A, D64, b, xa, **kw)
    sb = N.pcg(ctx, A, T64, b, xb, **kw)
    assert sa["niter"] == sb["niter"]
    assert np.array_equal(xa.get().view(np.uint64), xb.get().view(np.uint64))
    for B in (T32, D64, T64):
        B.close()
    A.close()

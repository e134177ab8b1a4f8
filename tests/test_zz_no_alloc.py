"""mlp_pcg_noalloc against mlp_pcg through the C ABI: two calls in a row (the second on a smaller system, so
the shared static workspace holds the first call's larger contents), each bitwise equal to mlp_pcg on the same
operands; the FP32-hierarchy flexible variant likewise; the size limit and the refused modes are errors."""
import numpy as np
import pytest

from conftest import rng, system
from ml_prec_amd import native as N
from ml_prec_amd import spe_synth as S

pytestmark = pytest.mark.gpu


def _pair(ctx, s, precision=N.MLP_FP64, **kw):
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, N.build_uh_cheb(s.ia, s.ja, s.a), precision=precision)
    b = ctx.to_device(s.b)
    x1, x2 = ctx.array(s.n), ctx.array(s.n)
    ref = N.pcg(ctx, A, B, b, x1, use_graph=True, **kw)
    got = N.pcg_noalloc(ctx, A, B, b, x2, **kw)
    assert got["niter"] == ref["niter"]
    assert np.array_equal(x2.get(), x1.get())
    for k in ("init_norm", "final_norm", "eps_used", "true_rel_res"):
        assert got[k] == ref[k], k
    B.close()
    A.close()
    return got


def test_noalloc_equals_pcg_two_calls_second_smaller(ctx):
    big = system("spe10_z12")                     # 158,400 rows
    small = system("ln5_spe")                     # 66,000 rows
    st = _pair(ctx, big, eps=1e-8, maxit=5000, x0_mode=N.X0_ZERO, check_every=8)
    assert st["true_rel_res"] <= 1e-8 * (1 + 1e-3)
    st = _pair(ctx, small, eps=1e-8, maxit=5000, x0_mode=N.X0_ZERO, check_every=8)
    assert st["true_rel_res"] <= 1e-8 * (1 + 1e-3)


def test_noalloc_flexible_fp32_equals_pcg(ctx):
    _pair(ctx, system("spe10_z12"), precision=N.MLP_FP32, eps=1e-10, maxit=5000, x0_mode=N.X0_ZERO, flexible=True,
          true_res_every=20, check_every=8)
    _pair(ctx, system("tiny"), precision=N.MLP_FP32, eps=1e-10, maxit=5000, x0_mode=N.X0_ZERO, flexible=True,
          true_res_every=20, check_every=8)


def test_noalloc_given_x0_and_errors(ctx):
    s = system("spe10_z12")
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, N.build_uh_cheb(s.ia, s.ja, s.a))
    x0 = rng(7).standard_normal(s.n)
    b = ctx.to_device(s.b)
    x1, x2 = ctx.to_device(x0), ctx.to_device(x0)
    r1 = N.pcg(ctx, A, B, b, x1, eps=1e-8, maxit=5000, x0_mode=N.X0_GIVEN)
    r2 = N.pcg_noalloc(ctx, A, B, b, x2, eps=1e-8, maxit=5000, x0_mode=N.X0_GIVEN)
    assert r1["niter"] == r2["niter"] and np.array_equal(x1.get(), x2.get())
    # "No need to iterate" as mlp_pcg reports it
    z = ctx.to_device(np.zeros(s.n))
    st = N.pcg_noalloc(ctx, A, B, z, x2, x0_mode=N.X0_ZERO, raise_on_error=False)
    assert st["rc"] == 4
    # refused modes
    st = N.pcg_noalloc(ctx, A, B, b, x2, x0_mode=N.X0_REFERENCE, raise_on_error=False)
    assert st["rc"] == 1
    B.close()
    A.close()
    # more rows than the static workspace holds: rejected before any launch
    big = S.make_system(60, 220, 90, 5, "lognormal", "spe", 1.0)
    assert big.n > 1179648
    A = N.DevMatrix(ctx, big.ia, big.ja, big.a)
    B = N.DevPrec(ctx, A, kind=N.DIAG)
    bb, xx = ctx.to_device(big.b), ctx.array(big.n)
    st = N.pcg_noalloc(ctx, A, B, bb, xx, raise_on_error=False)
    assert st["rc"] == 1 and "static workspace" in N.lib().mlp_last_error().decode()
    B.close()
    A.close()

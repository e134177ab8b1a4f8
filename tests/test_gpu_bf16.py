"""The BF16 variant of the FP32 hierarchy (precision MLP_FP32_BF16): level 0 keeps the vectors of its
Chebyshev iteration on level 1 (F, u1, a1 u1 and x2; cheb_solve.cc:45-101) in BF16, every kernel computing
in FP32.  Its error is stated against the FP64 oracle:

* one apply at 60x220x85 against the oracle's FP64 apply: the relative error is bounded, and it is larger
  than the plain FP32 hierarchy's (the variant does store in BF16);
* FP64 flexible PCG with the variant at 60x220x85 to 1e-8 and 1e-10: the FP64 true residual meets the
  tolerance, within a few iterations of the FP64 oracle's count;
* several right-hand sides in one batched solve are each bitwise the single solve;
* multi_split hierarchies and uh_cheb hierarchies with ncheb != 2 on level 0 are refused.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rng, system
from ml_prec_amd import native as N

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def z85():
    s = system("spe10_z85")
    return s, O.cheb_prec(s.ia, s.ja, s.a), N.build_uh_cheb(s.ia, s.ja, s.a)


def _apply(ctx, B, f):
    x = ctx.array(len(f))
    B.apply(ctx.to_device(f), x)
    return x.get()


def test_bf16_apply_error_vs_fp64_oracle(ctx, z85):
    s, P, H = z85
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B32 = N.DevPrec(ctx, A, H, precision=N.MLP_FP32)
    Bbf = N.DevPrec(ctx, A, H, precision=N.MLP_FP32_BF16)
    e32, ebf = [], []
    for seed in (1, 2, 3):
        f = rng(seed).standard_normal(s.n)
        ref = P.apply(f)
        nr = np.linalg.norm(ref)
        e32.append(np.linalg.norm(_apply(ctx, B32, f) - ref) / nr)
        ebf.append(np.linalg.norm(_apply(ctx, Bbf, f) - ref) / nr)
    print(f"apply rel err vs FP64 oracle: FP32 {max(e32):.3e}, FP32+BF16 {max(ebf):.3e}; launches "
          f"{Bbf.launches_per_apply()} (FP32 {B32.launches_per_apply()}), algorithmic bytes "
          f"{Bbf.bytes_per_apply():.4g} (FP32 {B32.bytes_per_apply():.4g})")
    assert max(e32) <= 1e-6
    # BF16 keeps 8 significant bits (unit roundoff 2^-9 ~ 2e-3) and only the level-1 share of the result
    # passes through it (2.9e-4 measured at this size)
    assert max(ebf) <= 2e-3
    assert min(ebf) > 10 * max(e32)
    # the variant runs level 0 on its own two kernels (its BF16 vectors), the levels below as in FP32
    assert Bbf.launches_per_apply() <= B32.launches_per_apply() + 2
    assert Bbf.bytes_per_apply() < B32.bytes_per_apply()
    Bbf.close()
    B32.close()
    A.close()


@pytest.mark.parametrize("eps", [1e-8, 1e-10])
def test_bf16_flexible_pcg_full_size(ctx, z85, eps):
    s, P, H = z85
    xo, so = O.pcg(s.ia, s.ja, s.a, s.b, P, eps=eps, x0=np.zeros(s.n))
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, H, precision=N.MLP_FP32_BF16)
    x = ctx.array(s.n)
    st = N.pcg(ctx, A, B, ctx.to_device(s.b), x, eps=eps, maxit=5000, flexible=True, true_res_every=20,
               use_graph=True)
    xg = x.get()
    rr = np.linalg.norm(s.b - O.spmv(s.ia, s.ja, s.a, xg)) / st["init_norm"]
    err = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    print(f"FP32+BF16 60x220x85 eps {eps:g}: {st['niter']} iterations (oracle FP64 {so['niter']}), true rel "
          f"res {rr:.3e}, error vs FP64 oracle {err:.3e}")
    assert rr <= eps * (1 + 1e-3)
    assert st["niter"] <= so["niter"] + 6
    assert err <= 1e3 * eps, err
    B.close()
    A.close()


def test_bf16_batched_rhs_equal_single_solves(ctx):
    s = system("spe10_z12")
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, N.build_uh_cheb(s.ia, s.ja, s.a), precision=N.MLP_FP32_BF16)
    R = 3
    bs = np.stack([s.b] + [rng(40 + k).standard_normal(s.n) for k in range(R - 1)])
    X = ctx.array(R * s.n)
    rc, sts = N.pcg_multi(ctx, A, B, R, ctx.to_device(bs.ravel()), X, eps=1e-8, maxit=2000, flexible=True,
                          true_res_every=20)
    assert rc == 0
    Xg = X.get().reshape(R, s.n)
    for k in range(R):
        x = ctx.array(s.n)
        st = N.pcg(ctx, A, B, ctx.to_device(bs[k]), x, eps=1e-8, maxit=2000, flexible=True, true_res_every=20)
        assert sts[k]["niter"] == st["niter"]
        assert np.array_equal(Xg[k], x.get())
    B.close()
    A.close()


def test_bf16_refuses_other_hierarchies(ctx):
    s = system("spe10_z12")
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    for H in (N.build_multi_split(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="simple_1", degree_max=2),
              N.build_uh_cheb(s.ia, s.ja, s.a, sigmas=(2.0, 3.0), niters=(1, 3))):
        with pytest.raises(N.MlpError, match="BF16"):
            N.DevPrec(ctx, A, H, precision=N.MLP_FP32_BF16)
    with pytest.raises(N.MlpError, match="precision"):
        N.DevPrec(ctx, A, N.build_uh_cheb(s.ia, s.ja, s.a), precision=3)
    A.close()

"""The integration path of the drop-in boundary: a hierarchy built elsewhere (here the oracle's, the arrays
that Prec::Level / MultiSplitPrec::Level hold: cheb_prec.h:17-40, multi_split_prec.h:19-44) is handed to
mlp_prec_upload level by level, and its device apply must equal the oracle's apply bit for bit (FP64).

Also the dynamic multi_split variant (inner_iter = dynamic, level_eps; multi_split_solve.cc:97-99,119-134),
built by the library and applied on the device, against the oracle.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rng, system
from ml_prec_amd import native as N

pytestmark = pytest.mark.gpu


def oracle_levels(P):
    """The oracle hierarchy as mlp_level_desc dicts: level l's matrix is level l-1's A_next."""
    out = []
    prev = None
    for l in range(P.nlevels):
        info, arr = P.level_arrays(l)
        d = dict(N=info["N"], M=info["M"], Md=info["Md"], nit=info["nit"], lmin=info["lmin"], lmax=info["lmax"],
                 eps=info["eps"], map=arr["map"], L_ia=arr["L_ia"], L_ja=arr["L_ja"], L_a=arr["L_a"],
                 U_ia=arr["U_ia"], U_ja=arr["U_ja"], U_a=arr["U_a"], dval=arr["dval"])
        if prev is not None:
            d.update(A_ia=prev["A_ia"], A_ja=prev["A_ja"], A_a=prev["A_a"])
        out.append(d)
        prev = arr
    return out


def _apply(ctx, B, f):
    x = ctx.array(len(f))
    B.apply(ctx.to_device(f), x)
    return x.get()


CASES = [
    ("uh_cheb", "spe10_z12", dict(sigmas=(3.0,), niters=(2,))),
    ("uh_cheb", "spe10_z12", dict(sigmas=(2.0, 3.0), niters=(1, 3))),
    ("multi_split", "spe10_z12", dict(q=(0.8,), niters=(5,), order="simple_1", degree_max=2)),
    ("multi_split", "ln5_tpfa_c1", dict(q=(0.6,), niters=(3,), order="original", degree_max=3)),
    ("multi_split", "tiny_tpfa_c1", dict(q=(0.6,), inner="dynamic", epss=(1e-3, 1e-3, 1e-2), order="original",
                                         degree_max=3)),
    ("multi_split", "spe10_z12", dict(q=(0.8,), inner="dynamic", epss=(1e-2,), order="simple_1", degree_max=2)),
]


@pytest.mark.parametrize("kind,name,kw", CASES)
def test_upload_of_oracle_levels_is_bitexact(ctx, kind, name, kw):
    s = system(name)
    P = O.cheb_prec(s.ia, s.ja, s.a, **kw) if kind == "uh_cheb" else O.msplit_prec(s.ia, s.ja, s.a, **kw)
    assert P.nlevels >= 2
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec.upload(ctx, A, N.UH_CHEB if kind == "uh_cheb" else N.MULTI_SPLIT, oracle_levels(P))
    for seed in (1, 2):
        f = rng(seed).standard_normal(s.n)
        got, ref = _apply(ctx, B, f), P.apply(f)
        assert np.array_equal(got, ref), (kind, name, np.abs(got - ref).max())
    B.close()
    A.close()


def test_upload_fp32_close_to_fp64(ctx):
    """The same oracle levels uploaded in FP32 apply within FP32 rounding of the FP64 apply."""
    s = system("spe10_z12")
    P = O.cheb_prec(s.ia, s.ja, s.a)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec.upload(ctx, A, N.UH_CHEB, oracle_levels(P), precision=N.MLP_FP32)
    f = rng(3).standard_normal(s.n)
    got, ref = _apply(ctx, B, f), P.apply(f)
    assert np.linalg.norm(got - ref) <= 1e-4 * np.linalg.norm(ref)
    B.close()
    A.close()


def test_upload_rejects_inconsistent_levels(ctx):
    s = system("spe10_z12")
    P = O.cheb_prec(s.ia, s.ja, s.a)
    lv = oracle_levels(P)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    bad = [dict(L) for L in lv]
    bad[1]["N"] = bad[1]["N"] + 1                      # does not chain with level 0's kept rows
    with pytest.raises(N.MlpError) as e:
        N.DevPrec.upload(ctx, A, N.UH_CHEB, bad)
    assert e.value.code == 1                           # MLP_ERR_INVALID
    bad = [dict(L) for L in lv]
    bad[0]["N"] = s.n - 1                              # level 0 must match A
    with pytest.raises(N.MlpError):
        N.DevPrec.upload(ctx, A, N.UH_CHEB, bad)
    A.close()


@pytest.mark.parametrize("name,kw", [
    ("spe10_z12", dict(q=(0.8,), inner="dynamic", eps=(1e-2,), order="simple_1", degree_max=2)),
    ("tiny_tpfa_c1", dict(q=(0.6,), inner="dynamic", eps=(1e-3, 1e-3, 1e-2), order="original", degree_max=3)),
])
def test_dynamic_multi_split_apply_and_pcg(ctx, name, kw):
    s = system(name)
    okw = dict(kw)
    okw["epss"] = okw.pop("eps")
    P = O.msplit_prec(s.ia, s.ja, s.a, **okw)
    H = N.build_multi_split(s.ia, s.ja, s.a, **kw)
    assert H.nlevels == P.nlevels and any(H.level_info(l)["eps"] > 0 for l in range(H.nlevels))
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, H)
    f = rng(4).standard_normal(s.n)
    assert np.array_equal(_apply(ctx, B, f), P.apply(f))
    xo, so = O.pcg(s.ia, s.ja, s.a, s.b, P, eps=1e-8, x0=np.zeros(s.n), maxit=5000)
    x = ctx.array(s.n)
    st = N.pcg(ctx, A, B, ctx.to_device(s.b), x, eps=1e-8, maxit=5000, use_graph=True)
    assert abs(int(st["niter"]) - so["niter"]) <= 1
    tol = 1e-10 if st["niter"] == so["niter"] else 1e-6
    assert np.linalg.norm(x.get() - xo) <= tol * np.linalg.norm(xo)
    B.close()
    A.close()

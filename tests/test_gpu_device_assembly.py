"""The device front end of the adjoint: the seven-point matrix assembled on the
device (mlp_spe_matrix_dev, spe.cc:69-131) and the uh_cheb hierarchy built on
the device from that matrix handle (mlp_hier_build_uh_cheb_mat).  Bar: the
matrix equals the host assembly's bit for bit (both kinds, grids with a
dimension of 1 that take the sliced layout, the full 60x220x85 grid), its SpMV
equals the uploaded matrix's, the hierarchy equals the host construction's and
the oracle's, and the adjoint gradient through the device front end equals the
host front end's bit for bit.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rng
from ml_prec_amd import native as N
from ml_prec_amd import spe_synth as S
from test_construction import same_hierarchy
from test_gpu_device_construction import assert_same_hierarchy

pytestmark = pytest.mark.gpu


def _field(shape, seed):
    nz, ny, nx = shape
    return S.spe10_field(nx, ny, nz, seed) if ny >= 20 else np.exp(rng(seed).normal(0.0, 2.0, shape))


@pytest.mark.parametrize("shape,kind,c", [((3, 4, 5), "spe", 1.0), ((2, 2, 2), "tpfa", 0.0), ((1, 4, 5), "spe", 1.0),
                                          ((4, 1, 6), "tpfa", 1.0), ((5, 3, 1), "spe", 0.5),
                                          ((12, 220, 60), "tpfa", 0.0), ((85, 220, 60), "spe", 1.0)])
def test_device_assembly_equals_host(ctx, shape, kind, c):
    kx = np.ascontiguousarray(_field(shape, S.BASE_SEED + 60), dtype=np.float64).reshape(shape)
    ia, ja, a = N.spe_assemble(kx, kind=kind, c=c, kz_ratio=0.1, h=(S.HX, S.HY, S.HZ))
    A = N.DevMatrix.spe_on_device(ctx, ctx.to_device(kx.ravel()), shape, kind=kind, c=c, kz_ratio=0.1,
                                  h=(S.HX, S.HY, S.HZ))
    # the solve layout first (no host copy yet), then the CSR
    x = rng(3).standard_normal(ia.shape[0] - 1)
    y1, y2 = ctx.array(x.size), ctx.array(x.size)
    A.spmv(ctx.to_device(x), y1)
    U = N.DevMatrix(ctx, ia, ja, a)
    U.spmv(ctx.to_device(x), y2)
    assert np.array_equal(y1.get().view(np.uint64), y2.get().view(np.uint64))
    assert A.info()[2] == U.info()[2]                          # the same layout, so the same stored bytes
    dia, dja, da = A.csr()
    assert np.array_equal(dia, ia) and np.array_equal(dja, ja)
    assert np.array_equal(da.view(np.uint64), a.view(np.uint64))
    U.close()
    A.close()


@pytest.mark.parametrize("nz", [12, 85])
def test_hierarchy_from_device_matrix(ctx, nz):
    shape = (nz, 220, 60)
    kx = S.spe10_field(60, 220, nz, S.BASE_SEED + 61).ravel()
    A = N.DevMatrix.spe_on_device(ctx, ctx.to_device(kx), shape, kind="spe", c=1.0, kz_ratio=0.1,
                                  h=(S.HX, S.HY, S.HZ))
    Hd = N.build_uh_cheb_on_device(ctx, A)
    ia, ja, a = N.spe_assemble(kx.reshape(shape), kind="spe", c=1.0, kz_ratio=0.1, h=(S.HX, S.HY, S.HZ))
    assert_same_hierarchy(N.build_uh_cheb(ia, ja, a), Hd)
    if nz == 12:
        assert same_hierarchy(O.cheb_prec(ia, ja, a), Hd)
    # an uploaded matrix handle works too (its host copy goes up)
    U = N.DevMatrix(ctx, ia, ja, a)
    assert_same_hierarchy(Hd, N.build_uh_cheb_on_device(ctx, U))
    U.close()
    A.close()


def test_adjoint_device_front_end_equals_host(ctx):
    from ml_prec_amd.adjoint import SpeAdjointModel, misfit
    shape = (12, 220, 60)
    n = int(np.prod(shape))
    g = rng(45)
    theta = np.log(S.spe10_field(60, 220, 12, S.BASE_SEED + 62).ravel())
    b = S.rhs(n, S.BASE_SEED + 63)
    obs = g.choice(n, 200, replace=False)
    d = g.standard_normal(obs.size)
    out = []
    for device in (True, False):
        m = SpeAdjointModel(ctx, shape, b, eps=1e-10, device=device)
        p = m.forward(theta)
        _, gp = misfit(p, obs, d)
        out.append((p, m.vjp(gp), m.stats[-2][1]["niter"], m.stats[-1][1]["niter"]))
        m.close()
    (p1, g1, f1, a1), (p2, g2, f2, a2) = out
    assert (f1, a1) == (f2, a2)
    assert np.array_equal(p1.view(np.uint64), p2.view(np.uint64))
    assert np.array_equal(g1.view(np.uint64), g2.view(np.uint64))

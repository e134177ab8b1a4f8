"""The seven-point matrix layout against the sliced ELL one (GPU).

A symmetric matrix of a full seven-point grid is stored as its diagonal and
three lower-neighbour values per row; every other matrix keeps the sliced ELL
layout (MLP_NO_STENCIL=1 forces it).  Both compute each row in the reference's
storage order, so SpMV, residual and whole PCG solves must agree bit for bit,
and both must equal the oracle's SpMV.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rng, system
from ml_prec_amd import native as N
from ml_prec_amd import spe_synth as S

pytestmark = pytest.mark.gpu


def dev(ctx, x):
    return ctx.to_device(np.ascontiguousarray(x, dtype=np.float64))


def _pair(ctx, monkeypatch, s):
    A1 = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    monkeypatch.setenv("MLP_NO_STENCIL", "1")
    A0 = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    monkeypatch.delenv("MLP_NO_STENCIL")
    return A1, A0


@pytest.mark.parametrize("name", ["ln5_spe", "ln5_tpfa", "spe10_z12", "small_spe10"])
def test_stencil_spmv_and_residual(ctx, monkeypatch, name):
    s = system(name)
    A1, A0 = _pair(ctx, monkeypatch, s)
    n = s.n
    assert A1.info()[2] == n * 33          # stencil layout: 4 values + 1 presence byte per row
    assert A0.info()[2] > n * 33
    for seed in (1, 2):
        x = rng(seed).standard_normal(n)
        y1, y0 = ctx.array(n), ctx.array(n)
        A1.spmv(dev(ctx, x), y1)
        A0.spmv(dev(ctx, x), y0)
        ref = O.spmv(s.ia, s.ja, s.a, x)
        assert np.array_equal(y1.get(), ref) and np.array_equal(y0.get(), ref)
        r1, r0 = ctx.array(n), ctx.array(n)
        A1.residual(dev(ctx, s.b), dev(ctx, x), r1)
        A0.residual(dev(ctx, s.b), dev(ctx, x), r0)
        assert np.array_equal(r1.get(), r0.get())


def test_stencil_pcg_equals_sliced(ctx, monkeypatch):
    s = system("spe10_z12")
    A1, A0 = _pair(ctx, monkeypatch, s)
    H = N.build_uh_cheb(s.ia, s.ja, s.a)
    out = []
    for A in (A1, A0):
        B = N.DevPrec(ctx, A, H)
        x = ctx.array(s.n)
        st = N.pcg(ctx, A, B, dev(ctx, s.b), x, eps=1e-8, maxit=2000, x0_mode=N.X0_ZERO, use_graph=True)
        out.append((st["niter"], x.get()))
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("dims", [(5, 7, 3), (7, 4, 2), (6, 5, 3), (25, 37, 9), (30, 21, 8)])
def test_stencil_row_pairs_any_shape(ctx, monkeypatch, dims):
    """The seven-point kernels take two rows per thread: one 16-byte load per
    pair of values when n, nx and nx*ny are even ((6,5,3), (30,21,8)), the rows
    one by one otherwise (odd n or odd nx).  Both equal the oracle's SpMV, and
    a Jacobi PCG on either layout gives the same iterates bit for bit (the fused
    SpMV + (Ap, p) kernel adds its products in the same order for both)."""
    nx, ny, nz = dims
    s = S.make_system(nx, ny, nz, 31, "lognormal", "spe", 1.0, sigma=1.5)
    A1, A0 = _pair(ctx, monkeypatch, s)
    n = s.n
    assert A1.info()[2] == n * 33 and A0.info()[2] > n * 33
    x = rng(5).standard_normal(n)
    y1, y0, r1 = ctx.array(n), ctx.array(n), ctx.array(n)
    A1.spmv(dev(ctx, x), y1)
    A0.spmv(dev(ctx, x), y0)
    ref = O.spmv(s.ia, s.ja, s.a, x)
    assert np.array_equal(y1.get(), ref) and np.array_equal(y0.get(), ref)
    A1.residual(dev(ctx, s.b), dev(ctx, x), r1)
    assert np.array_equal(r1.get(), s.b - ref)
    out = []
    for A in (A1, A0):
        B = N.DevPrec(ctx, A, kind=N.DIAG)
        xs = ctx.array(n)
        st = N.pcg(ctx, A, B, dev(ctx, s.b), xs, eps=1e-10, maxit=5000, x0_mode=N.X0_ZERO, use_graph=False)
        out.append((st["niter"], xs.get()))
        B.close()
    assert out[0][0] == out[1][0] and out[0][0] > 5
    assert np.array_equal(out[0][1], out[1][1])
    A1.close()
    A0.close()


def test_non_stencil_matrices_keep_slices(ctx):
    s = system("spe10_z12")
    n = s.n
    # one value off symmetry
    a = s.a.copy()
    a[s.ia[5] + 1] *= 1.0 + 2.0 ** -40
    A = N.DevMatrix(ctx, s.ia, s.ja, a)
    assert A.info()[2] > n * 33
    x = rng(3).standard_normal(n)
    y = ctx.array(n)
    A.spmv(dev(ctx, x), y)
    assert np.array_equal(y.get(), O.spmv(s.ia, s.ja, a, x))
    # a flat grid (nz = 1) has only five-point rows
    k = np.exp(rng(4).standard_normal((1, 20, 30)))
    ia, ja, av = S.assemble(k, kind="spe", c=1.0)
    B = N.DevMatrix(ctx, ia, ja, av)
    assert B.info()[2] > 600 * 33
    x = rng(5).standard_normal(600)
    y = ctx.array(600)
    B.spmv(dev(ctx, x), y)
    assert np.array_equal(y.get(), O.spmv(ia, ja, av, x))


@pytest.mark.parametrize("precision", [N.MLP_FP64, N.MLP_FP32])
def test_pageable_staging_same_results(ctx, monkeypatch, precision):
    """Uploads staged through pageable memory (MLP_STAGE_MAX=0, the path of blocks too large for the pinned
    buffer) give the same device data: SpMV and preconditioner apply bitwise equal to the pinned path."""
    s = system("spe10_z12")
    H = N.build_uh_cheb(s.ia, s.ja, s.a)
    f = ctx.to_device(rng(31).standard_normal(s.n))
    out = []
    for cap in (None, "0"):
        if cap is None:
            monkeypatch.delenv("MLP_STAGE_MAX", raising=False)
        else:
            monkeypatch.setenv("MLP_STAGE_MAX", cap)
        A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
        B = N.DevPrec(ctx, A, H, precision=precision)
        y, z = ctx.array(s.n), ctx.array(s.n)
        A.spmv(f, y)
        B.apply(f, z)
        out.append((y.get(), z.get()))
        B.close()
        A.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])

"""Nonsymmetric matrices on the device: GSPrec, CompositePrec, SymPrec and GMRES.

Bars: GS and Composite applies bit-exact against the oracle (the sweep keeps
the reference's per-row order); SymPrec (an inner PCG to 1e-14) within 1e-10;
GMRES iteration counts within one of the oracle and solutions within 1e-9.
The matrices are the reference's nonsymmetric M-matrices (spe.cc:138-182,
tau = 0.3) of seeded SPE10-shaped fields.
"""
import functools

import numpy as np
import pytest

import oracle as O
from conftest import rng, system
from ml_prec_amd import native as N
from ml_prec_amd import spe_synth as S

pytestmark = pytest.mark.gpu


@functools.lru_cache(maxsize=None)
def nonsym(name):
    s = system(name)
    return s, S.make_nonsymmetric(s.ia, s.ja, s.a, 0.3)


def dev(ctx, x):
    return ctx.to_device(np.ascontiguousarray(x, dtype=np.float64))


MS = dict(q=(0.8,), niters=(5,), order="original", degree_max=3)


@pytest.mark.parametrize("name", ["ln5_spe", "spe10_z12"])
def test_gs_apply_bitexact(ctx, name):
    s, a = nonsym(name)
    A = N.DevMatrix(ctx, s.ia, s.ja, a)
    B = N.DevPrec(ctx, A, kind=N.GS)
    f = rng(61).standard_normal(s.n)
    x = ctx.array(s.n)
    B.apply(dev(ctx, f), x)
    assert np.array_equal(x.get(), O.gs_prec(s.ia, s.ja, a).apply(f))


def test_composite_apply_bitexact(ctx):
    s, a = nonsym("ln5_spe")
    A = N.DevMatrix(ctx, s.ia, s.ja, a)
    H = N.build_multi_split(s.ia, s.ja, a, **MS)
    B = N.DevPrec(ctx, A, H, kind=N.COMPOSITE)
    P = O.composite_prec(O.gs_prec(s.ia, s.ja, a), O.msplit_prec(s.ia, s.ja, a, **MS))
    f = rng(62).standard_normal(s.n)
    x = ctx.array(s.n)
    B.apply(dev(ctx, f), x)
    assert np.array_equal(x.get(), P.apply(f))


def test_sym_prec_apply(ctx):
    s, a = nonsym("ln5_spe")
    A = N.DevMatrix(ctx, s.ia, s.ja, a)
    B = N.DevPrec(ctx, A, kind=N.SYM)
    P = O.sym_prec(s.ia, s.ja, a)
    f = rng(63).standard_normal(s.n)
    x = ctx.array(s.n)
    B.apply(dev(ctx, f), x)
    ref = P.apply(f)
    assert np.linalg.norm(x.get() - ref) <= 1e-10 * np.linalg.norm(ref)


@pytest.mark.parametrize("prec", ["composite", "gs", "sym"])
def test_gmres_nonsym_vs_oracle(ctx, prec):
    s, a = nonsym("ln5_spe")
    A = N.DevMatrix(ctx, s.ia, s.ja, a)
    if prec == "composite":
        B = N.DevPrec(ctx, A, N.build_multi_split(s.ia, s.ja, a, **MS), kind=N.COMPOSITE)
        P = O.composite_prec(O.gs_prec(s.ia, s.ja, a), O.msplit_prec(s.ia, s.ja, a, **MS))
    elif prec == "gs":
        B = N.DevPrec(ctx, A, kind=N.GS)
        P = O.gs_prec(s.ia, s.ja, a)
    else:
        B = N.DevPrec(ctx, A, kind=N.SYM)
        P = O.sym_prec(s.ia, s.ja, a)
    xo, so = O.gmres(s.ia, s.ja, a, s.b, P, eps=1e-8, x0=None, maxit=3000)
    x = ctx.array(s.n)
    sg = N.gmres(ctx, A, B, dev(ctx, s.b), x, eps=1e-8, maxit=3000, x0_mode=N.X0_REFERENCE)
    assert abs(int(sg["niter"]) - so["niter"]) <= 1, (sg["niter"], so["niter"])
    assert np.linalg.norm(x.get() - xo) <= 1e-9 * np.linalg.norm(xo)


def _random_lower(n, k, seed):
    """Skyline CSR of an n x n M-matrix whose row i has min(i, k) lower entries (random columns) and their
    mirrors, with random nonsymmetric values and a dominant positive diagonal."""
    g = rng(seed)
    rows = [set() for _ in range(n)]
    for i in range(1, n):
        for j in g.choice(i, size=min(i, k), replace=False):
            rows[i].add(int(j))
            rows[int(j)].add(i)
    ia, ja, a = [0], [], []
    for i in range(n):
        cols = sorted(rows[i])
        off = -g.uniform(0.1, 1.0, len(cols))
        ja += [i] + cols
        a += [float(-off.sum() + 1.0)] + off.tolist()
        ia.append(len(ja))
    return np.array(ia, np.uint32), np.array(ja, np.uint32), np.array(a)


@pytest.mark.parametrize("k", [1, 3, 6, 12])
def test_gs_apply_row_widths(ctx, k):
    """Up to 4 lower entries a row use that many fixed slots, 5..8 use 8 slots, more use row pointers:
    every width gives the oracle's sweep bit for bit."""
    ia, ja, a = _random_lower(3000, k, 70 + k)
    A = N.DevMatrix(ctx, ia, ja, a)
    B = N.DevPrec(ctx, A, kind=N.GS)
    f = rng(64).standard_normal(len(ia) - 1)
    x = ctx.array(len(ia) - 1)
    B.apply(dev(ctx, f), x)
    assert np.array_equal(x.get(), O.gs_prec(ia, ja, a).apply(f))


def _grid7(nx, ny, nz, seed, drop=0.2):
    """Nonsymmetric seven-point M-matrix on an nx x ny x nz grid (x fastest) with a share `drop` of the
    couplings left out, so rows miss some of their lower neighbours."""
    g = rng(seed)
    n = nx * ny * nz
    ia, ja, a = [0], [], []
    for i in range(n):
        x, y, z = i % nx, (i // nx) % ny, i // (nx * ny)
        nb = []
        for d, ok in ((nx * ny, z > 0), (nx, y > 0), (1, x > 0)):
            if ok and g.uniform() >= drop:
                nb.append(i - d)
        for d, ok in ((1, x + 1 < nx), (nx, y + 1 < ny), (nx * ny, z + 1 < nz)):
            if ok and g.uniform() >= drop:
                nb.append(i + d)
        off = -g.uniform(0.1, 1.0, len(nb))
        ja += [i] + nb
        a += [float(-off.sum() + 0.5)] + off.tolist()
        ia.append(len(ja))
    return np.array(ia, np.uint32), np.array(ja, np.uint32), np.array(a)


@pytest.mark.parametrize("dims", [(13, 9, 7), (8, 8, 8), (5, 30, 17), (60, 4, 3)])
def test_gs_apply_grid_bricks(ctx, dims):
    """Seven-point lower triangles take the brick sweep (8x8x8 bricks, partial at the grid's far faces, rows
    with missing lower neighbours): bit for bit the oracle's sweep."""
    ia, ja, a = _grid7(*dims, seed=sum(dims))
    A = N.DevMatrix(ctx, ia, ja, a)
    B = N.DevPrec(ctx, A, kind=N.GS)
    f = rng(65).standard_normal(len(ia) - 1)
    x = ctx.array(len(ia) - 1)
    B.apply(dev(ctx, f), x)
    assert np.array_equal(x.get(), O.gs_prec(ia, ja, a).apply(f))

"""GPU parity of the three block paths of the level kernels.

A level's component blocks run one of three sweeps, chosen per block when the
level is planned: owned-node (a block no larger than the workgroup, values in
LDS), staged (a larger block whose rows and values are staged in LDS) and
global (a block too large for the LDS budget, values in global memory).  The
default launch shape (64-thread blocks of at most 64 nodes on levels whose
components all fit, else 256-thread blocks) takes the owned-node path almost
everywhere; the MLP_LEVEL_BLOCK / MLP_LEVEL_CAP / MLP_LEVEL_LDS knobs move
blocks onto the other two (on spe10_z12: "staged" puts every block but one on
the staged path, "global" nearly every block on the global path, "mixed"
gives each level a mix of the three), and the multi_split hierarchy with the simple_1
order and degree_max 3 has one elimination component of about 25k rows,
which only the global path can run.  Every case must match the oracle
bit-exactly (the paths differ only in where values live, not in the per-row
operation order).
"""
import numpy as np
import pytest

import oracle as O
from conftest import rng, system
from ml_prec_amd import native as N

pytestmark = pytest.mark.gpu

SHAPES = {
    "staged": {"MLP_LEVEL_BLOCK": "64"},
    "global": {"MLP_LEVEL_BLOCK": "256", "MLP_LEVEL_LDS": "2048"},
    "mixed": {"MLP_LEVEL_BLOCK": "64", "MLP_LEVEL_CAP": "512", "MLP_LEVEL_LDS": "24576"},
}


def dev(ctx, x):
    return ctx.to_device(np.ascontiguousarray(x, dtype=np.float64))


def _apply_check(ctx, s, P, H, seeds=(7, 8)):
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, H)
    for seed in seeds:
        f = rng(seed).standard_normal(s.n)
        x = ctx.array(s.n)
        B.apply(dev(ctx, f), x)
        got, ref = x.get(), P.apply(f)
        assert np.array_equal(got, ref), (seed, np.abs(got - ref).max())


@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_uh_cheb_block_paths(ctx, monkeypatch, shape):
    for k, v in SHAPES[shape].items():
        monkeypatch.setenv(k, v)
    s = system("spe10_z12")
    P = O.cheb_prec(s.ia, s.ja, s.a, sigmas=(3.0,), niters=(3,))
    H = N.build_uh_cheb(s.ia, s.ja, s.a, sigmas=(3.0,), niters=(3,))
    _apply_check(ctx, s, P, H)


@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_multi_split_block_paths(ctx, monkeypatch, shape):
    for k, v in SHAPES[shape].items():
        monkeypatch.setenv(k, v)
    s = system("spe10_z12")
    P = O.msplit_prec(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="original", degree_max=3)
    H = N.build_multi_split(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="original", degree_max=3)
    _apply_check(ctx, s, P, H)


def test_giant_component_apply(ctx):
    """One ~25k-row elimination component: the coarsest-level solve from global memory."""
    s = system("ln5_tpfa_c1")
    P = O.msplit_prec(s.ia, s.ja, s.a, q=(0.8,), niters=(3,), order="simple_1", degree_max=3)
    H = N.build_multi_split(s.ia, s.ja, s.a, q=(0.8,), niters=(3,), order="simple_1", degree_max=3)
    _apply_check(ctx, s, P, H, seeds=(9,))


def test_global_blocks_pcg_graph(ctx, monkeypatch):
    """PCG through a captured graph with global-memory blocks on every level."""
    monkeypatch.setenv("MLP_LEVEL_BLOCK", "256")
    monkeypatch.setenv("MLP_LEVEL_LDS", "2048")
    s = system("spe10_z12")
    b = rng(11).standard_normal(s.n)
    xo, so = O.pcg(s.ia, s.ja, s.a, b, O.cheb_prec(s.ia, s.ja, s.a), eps=1e-8, x0=np.zeros(s.n), maxit=5000)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, N.build_uh_cheb(s.ia, s.ja, s.a))
    x = ctx.array(s.n)
    st = N.pcg(ctx, A, B, dev(ctx, b), x, eps=1e-8, maxit=5000, x0_mode=N.X0_ZERO, use_graph=True)
    assert abs(int(st["niter"]) - so["niter"]) <= 1, (st["niter"], so["niter"])
    got = x.get()
    tol = 1e-10 if st["niter"] == so["niter"] else 1e-6
    assert np.linalg.norm(got - xo) <= tol * np.linalg.norm(xo)
    assert st["true_rel_res"] <= 1e-8 * (1 + 1e-3)

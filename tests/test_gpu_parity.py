"""GPU parity of the HIP path (through the C ABI) against the CPU oracle.

Bars:
* SpMV, residual, daxpy/dscal, and the uh_cheb / multi_split preconditioner
  apply in FP64: bit-exact (the kernels keep the reference's per-row order and
  are compiled without FMA contraction);
* dot / nrm2: relative 1e-13 (a parallel reduction rounds differently from the
  sequential sum);
* PCG / GMRES in FP64: iteration counts within +-1 of the oracle; solutions
  within 1e-10 relative when the counts agree and within 1e-6 when they differ
  by one (the reductions differ only by rounding); the FP64 true residual of
  the device solution within eps (1 + 1e-3) of the initial residual;
* mixed precision (FP32 hierarchy, FP64 flexible PCG): FP64 true relative
  residual <= 1e-8 and solution error vs the FP64 oracle <= 1e-6 relative.
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


@pytest.mark.parametrize("name", ["ln5_spe", "ln5_tpfa", "spe10_z12"])
def test_spmv_residual_bitexact(ctx, name):
    s = system(name)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    x = rng(1).standard_normal(s.n)
    y = ctx.array(s.n)
    A.spmv(dev(ctx, x), y)
    ref = O.spmv(s.ia, s.ja, s.a, x)
    assert np.array_equal(y.get(), ref)
    r = ctx.array(s.n)
    A.residual(dev(ctx, s.b), dev(ctx, x), r)
    assert np.array_equal(r.get(), s.b - ref)


def test_blas1(ctx):
    n = 100_003
    g = rng(2)
    x, y = g.standard_normal(n), g.standard_normal(n)
    dx, dy = dev(ctx, x), dev(ctx, y)
    d = N.ddot(ctx, dx, dy)
    assert abs(d - O.ddot(x, y)) <= 1e-13 * np.abs(x * y).sum()
    idx = np.arange(1000, dtype=np.float64)
    di = dev(ctx, idx)
    assert N.ddot(ctx, di, di) == 999 * 1000 * 1999 / 6          # exact: small integers
    assert abs(N.dnrm2(ctx, dx) - np.sqrt(O.ddot(x, x))) <= 1e-14 * np.sqrt(O.ddot(x, x))
    for alpha in (0.37, 1.0, 1.0 + 2.0 ** -52, -1.0):
        dy2 = dev(ctx, y)
        N.daxpy(ctx, alpha, dx, dy2)
        assert np.array_equal(dy2.get(), O.daxpy(alpha, x, y))
        dx2 = dev(ctx, x)
        N.dscal(ctx, alpha, dx2)
        assert np.array_equal(dx2.get(), O.dscal(alpha, x))


CHEB_CASES = [("ln5_tpfa_c1", (3.0,), (2,)), ("spe10_z12", (3.0,), (2,)), ("spe10_z12", (3.0,), (3,)),
              ("spe10_z12", (2.0, 3.0), (1, 2)), ("spe10_z40", (3.0,), (4,))]


@pytest.mark.parametrize("name,sigmas,niters", CHEB_CASES)
def test_uh_cheb_apply_bitexact(ctx, name, sigmas, niters):
    s = system(name)
    P = O.cheb_prec(s.ia, s.ja, s.a, sigmas=sigmas, niters=niters)
    H = N.build_uh_cheb(s.ia, s.ja, s.a, sigmas=sigmas, niters=niters)
    assert H.nlevels == P.nlevels
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, H)
    for seed in (3, 4):
        f = rng(seed).standard_normal(s.n)
        x = ctx.array(s.n)
        B.apply(dev(ctx, f), x)
        got, ref = x.get(), P.apply(f)
        assert np.array_equal(got, ref), (name, np.abs(got - ref).max())


@pytest.mark.parametrize("name,sigmas,niters", CHEB_CASES[:2] + [CHEB_CASES[3], ("spe10_z40", (3.0,), (2,))])
def test_uh_cheb_apply_without_chain_bitexact(ctx, name, sigmas, niters, monkeypatch):
    """The same applies on the launch schedule without the chain (MLP_NO_CHAIN=1): the level kernels and the
    fused last pair.  Both schedules give the oracle's bits; the chain needs fewer launches."""
    s = system(name)
    P = O.cheb_prec(s.ia, s.ja, s.a, sigmas=sigmas, niters=niters)
    H = N.build_uh_cheb(s.ia, s.ja, s.a, sigmas=sigmas, niters=niters)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    monkeypatch.setenv("MLP_NO_CHAIN", "1")
    B0 = N.DevPrec(ctx, A, H)
    monkeypatch.delenv("MLP_NO_CHAIN")
    B1 = N.DevPrec(ctx, A, H)
    f = rng(5).standard_normal(s.n)
    for B in (B0, B1):
        x = ctx.array(s.n)
        B.apply(dev(ctx, f), x)
        assert np.array_equal(x.get(), P.apply(f)), name
    print(f"{name} ncheb {niters}: launches per apply {B0.launches_per_apply()} without the chain, "
          f"{B1.launches_per_apply()} with it")
    assert B1.launches_per_apply() < B0.launches_per_apply()
    B1.close()
    B0.close()
    A.close()


MS_CASES = [("spe10_z12", "simple_1", 2, (5,)), ("spe10_z12", "original", 3, (5,)),
            ("spe10_z12", "none", 3, (2, 3)), ("ln5_tpfa_c1", "original", 3, (3,)),
            ("small_spe10", "simple_1", 2, (4,))]


@pytest.mark.parametrize("name,order,degree_max,niters", MS_CASES)
def test_multi_split_apply_bitexact(ctx, name, order, degree_max, niters):
    s = system(name)
    P = O.msplit_prec(s.ia, s.ja, s.a, q=(0.8,), niters=niters, order=order, degree_max=degree_max)
    H = N.build_multi_split(s.ia, s.ja, s.a, q=(0.8,), niters=niters, order=order, degree_max=degree_max)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, H)
    f = rng(5).standard_normal(s.n)
    x = ctx.array(s.n)
    B.apply(dev(ctx, f), x)
    got, ref = x.get(), P.apply(f)
    assert np.array_equal(got, ref), (name, order, np.abs(got - ref).max())


def test_msplit_config_file_apply(ctx, tmp_path):
    s = system("spe10_z12")
    cfg = tmp_path / "msplit.cfg"
    cfg.write_text("q = 0.8\ninner_iter = fixed\nlevel_niter = 5\nelim_order = simple_1\n"
                   "elim_stop = degree\ndegree_max = 2\nmax_levels = 15\n")
    H = N.build_multi_split(s.ia, s.ja, s.a, config_path=str(cfg))
    P = O.msplit_prec(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="simple_1", degree_max=2)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, H)
    f = rng(6).standard_normal(s.n)
    x = ctx.array(s.n)
    B.apply(dev(ctx, f), x)
    assert np.array_equal(x.get(), P.apply(f))


def assert_solve_parity(s, xg, sg, xo, so, eps, implicit=False):
    """Iterations within one; the solution within 1e-10 (same count) or 1e-6 (one apart) of the oracle's; the
    FP64 true residual, recomputed here, within eps (1 + 1e-3) of the initial residual.  GMRES stops on its
    implicit residual (gmres.cc:104-110), whose true residual may lie above eps in the oracle too
    (implicit=True): then the true residual must match the oracle's within 1e-3 relative."""
    d = abs(int(sg["niter"]) - int(so["niter"]))
    assert d <= 1, (sg["niter"], so["niter"])
    tol = 1e-10 if d == 0 else 1e-6
    assert np.linalg.norm(xg - xo) <= tol * np.linalg.norm(xo), np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    rg = np.linalg.norm(s.b - O.spmv(s.ia, s.ja, s.a, xg))
    if implicit:
        ro = np.linalg.norm(s.b - O.spmv(s.ia, s.ja, s.a, xo))
        assert abs(rg - ro) <= 1e-3 * ro, (rg, ro)
    else:
        assert rg <= eps * (1 + 1e-3) * sg["init_norm"], rg / sg["init_norm"]


def _pcg_pair(ctx, s, prec="cheb", x0_ref=True, eps=1e-8, use_graph=False):
    if prec == "cheb":
        P = O.cheb_prec(s.ia, s.ja, s.a)
        H = N.build_uh_cheb(s.ia, s.ja, s.a)
    else:
        P = O.msplit_prec(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="simple_1", degree_max=2)
        H = N.build_multi_split(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="simple_1", degree_max=2)
    xo, so = O.pcg(s.ia, s.ja, s.a, s.b, P, eps=eps, x0=None if x0_ref else np.zeros(s.n), maxit=5000)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, H)
    x = ctx.array(s.n)
    sg = N.pcg(ctx, A, B, dev(ctx, s.b), x, eps=eps, maxit=5000,
               x0_mode=N.X0_REFERENCE if x0_ref else N.X0_ZERO, use_graph=use_graph)
    return xo, so, x.get(), sg


@pytest.mark.parametrize("name", ["ln5_tpfa_c1", "spe10_z12"])
@pytest.mark.parametrize("x0_ref", [True, False])
def test_pcg_uh_cheb_vs_oracle(ctx, name, x0_ref):
    s = system(name)
    xo, so, xg, sg = _pcg_pair(ctx, s, "cheb", x0_ref)
    assert_solve_parity(s, xg, sg, xo, so, 1e-8)


def test_pcg_graph_equals_stream(ctx):
    s = system("spe10_z12")
    _, _, x1, s1 = _pcg_pair(ctx, s, "cheb", True, use_graph=False)
    _, _, x2, s2 = _pcg_pair(ctx, s, "cheb", True, use_graph=True)
    assert s1["niter"] == s2["niter"]
    assert np.array_equal(x1, x2)


def test_pcg_multi_split_vs_oracle(ctx):
    s = system("spe10_z12")
    xo, so, xg, sg = _pcg_pair(ctx, s, "msplit", True)
    assert_solve_parity(s, xg, sg, xo, so, 1e-8)


def test_gmres_vs_oracle(ctx):
    s = system("spe10_z12")
    P = O.msplit_prec(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="simple_1", degree_max=2)
    H = N.build_multi_split(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="simple_1", degree_max=2)
    xo, so = O.gmres(s.ia, s.ja, s.a, s.b, P, eps=1e-8, x0=None, maxit=3000)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, H)
    x = ctx.array(s.n)
    sg = N.gmres(ctx, A, B, dev(ctx, s.b), x, eps=1e-8, maxit=3000, x0_mode=N.X0_REFERENCE)
    assert_solve_parity(s, x.get(), sg, xo, so, 1e-8, implicit=True)


def test_pcg_errors(ctx):
    s = system("ln5_spe")
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, kind=N.DIAG)
    # b = A x0 with x0 = 0 -> b = 0 -> init norm 0 -> "No need to iterate"
    x = ctx.array(s.n)
    with pytest.raises(N.MlpError) as e:
        N.pcg(ctx, A, B, dev(ctx, np.zeros(s.n)), x, x0_mode=N.X0_ZERO)
    assert e.value.code == 4
    # a negated matrix is negative definite -> "<Ap, p> <= 0"
    An = N.DevMatrix(ctx, s.ia, s.ja, -s.a)
    Bi = N.DevPrec(ctx, An, kind=N.ID)
    with pytest.raises(N.MlpError) as e:
        N.pcg(ctx, An, Bi, dev(ctx, s.b), x, x0_mode=N.X0_ZERO)
    assert e.value.code == 3


def test_diag_prec_pcg_vs_oracle(ctx):
    s = system("ln5_tpfa")                      # the pure Dirichlet matrix, Jacobi-preconditioned
    P = O.diag_prec(s.ia, s.ja, s.a)
    xo, so = O.pcg(s.ia, s.ja, s.a, s.b, P, eps=1e-8, x0=np.zeros(s.n), maxit=20000)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, kind=N.DIAG)
    x = ctx.array(s.n)
    sg = N.pcg(ctx, A, B, dev(ctx, s.b), x, eps=1e-8, maxit=20000)
    assert_solve_parity(s, x.get(), sg, xo, so, 1e-8)


def test_mixed_precision_pcg(ctx):
    s = system("spe10_z40")
    P = O.cheb_prec(s.ia, s.ja, s.a)
    xo, so = O.pcg(s.ia, s.ja, s.a, s.b, P, eps=1e-10, x0=np.zeros(s.n), maxit=5000)
    H = N.build_uh_cheb(s.ia, s.ja, s.a)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, H, precision=N.MLP_FP32)
    x = ctx.array(s.n)
    sg = N.pcg(ctx, A, B, dev(ctx, s.b), x, eps=1e-8, maxit=5000, flexible=True, true_res_every=20)
    assert sg["true_rel_res"] <= 1e-8
    err = np.linalg.norm(x.get() - xo) / np.linalg.norm(xo)
    assert err <= 1e-6, err
    assert sg["niter"] <= so["niter"] + 10


def test_full_size_pcg_fp64(ctx):
    s = system("spe10_z85")
    xo, so, xg, sg = _pcg_pair(ctx, s, "cheb", False, use_graph=True)
    assert_solve_parity(s, xg, sg, xo, so, 1e-8)


def _cheb_solver_pair(ctx, s, eps=1e-8, use_graph=True, check_every=0):
    P = O.cheb_prec(s.ia, s.ja, s.a)
    H = N.build_uh_cheb(s.ia, s.ja, s.a)
    li = H.level_info(0)
    assert (li["lmin"], li["lmax"]) == (P.level_info(0)["lmin"], P.level_info(0)["lmax"])
    xo, so = O.cheb_solver(s.ia, s.ja, s.a, s.b, P, li["lmin"], li["lmax"], eps=eps, x0=None, maxit=20000)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, H)
    x = ctx.array(s.n)
    sg = N.cheb_solver(ctx, A, B, li["lmin"], li["lmax"], dev(ctx, s.b), x, eps=eps, maxit=20000,
                       x0_mode=N.X0_REFERENCE, use_graph=use_graph, check_every=check_every)
    return xo, so, x.get(), sg


@pytest.mark.parametrize("name", ["ln5_tpfa_c1", "spe10_z12"])
def test_cheb_solver_vs_oracle(ctx, name):
    """ChebSolver: the coefficients depend only on the iteration count, so the
    iterates are bit-exact; only the stop test (a reduction) may differ by one."""
    s = system(name)
    xo, so, xg, sg = _cheb_solver_pair(ctx, s)
    assert abs(int(sg["niter"]) - so["niter"]) <= 1, (sg["niter"], so["niter"])
    if sg["niter"] == so["niter"]:
        assert np.array_equal(xg, xo)
    else:
        assert np.linalg.norm(xg - xo) <= 1e-6 * np.linalg.norm(xo)
    assert sg["true_rel_res"] <= 1e-8 * (1 + 1e-3)


def test_cheb_solver_graph_equals_stream(ctx):
    s = system("spe10_z12")
    _, _, x1, s1 = _cheb_solver_pair(ctx, s, use_graph=False, check_every=5)
    _, _, x2, s2 = _cheb_solver_pair(ctx, s, use_graph=True, check_every=7)
    assert s1["niter"] == s2["niter"]
    assert np.array_equal(x1, x2)


def test_cheb_solver_maxit(ctx):
    s = system("spe10_z12")
    H = N.build_uh_cheb(s.ia, s.ja, s.a)
    li = H.level_info(0)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, H)
    x = ctx.array(s.n)
    with pytest.raises(N.MlpError) as e:
        N.cheb_solver(ctx, A, B, li["lmin"], li["lmax"], dev(ctx, s.b), x, eps=1e-12, maxit=3)
    assert e.value.code == 5
    P = O.cheb_prec(s.ia, s.ja, s.a)
    xo, so = O.cheb_solver(s.ia, s.ja, s.a, s.b, P, li["lmin"], li["lmax"], eps=1e-12, x0=np.zeros(s.n), maxit=3,
                           check=False)
    assert so["rc"] != 0
    assert np.array_equal(x.get(), xo)


def _msplit(s):
    P = O.msplit_prec(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="simple_1", degree_max=2)
    H = N.build_multi_split(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="simple_1", degree_max=2)
    return P, H


@pytest.mark.parametrize("use_graph", [False, True])
def test_simple_solver_vs_oracle(ctx, use_graph):
    """SimpleSolver with the multi_split preconditioner: iterates bit-exact,
    stop test within one iteration."""
    s = system("spe10_z12")
    P, H = _msplit(s)
    xo, so = O.simple_solver(s.ia, s.ja, s.a, s.b, P, eps=1e-8, x0=None, maxit=5000)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, H)
    x = ctx.array(s.n)
    sg = N.simple_solver(ctx, A, B, dev(ctx, s.b), x, eps=1e-8, maxit=5000, x0_mode=N.X0_REFERENCE,
                         use_graph=use_graph)
    assert abs(int(sg["niter"]) - so["niter"]) <= 1, (sg["niter"], so["niter"])
    if sg["niter"] == so["niter"]:
        assert np.array_equal(x.get(), xo)
    else:
        assert np.linalg.norm(x.get() - xo) <= 1e-6 * np.linalg.norm(xo)
    assert sg["true_rel_res"] <= 1e-8 * (1 + 1e-3)


def test_simple_solver_errors(ctx):
    s = system("ln5_spe")
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, kind=N.DIAG)
    x = ctx.array(s.n)
    # one iteration is enough -> "No need to iterate", x = B b all the same
    with pytest.raises(N.MlpError) as e:
        N.simple_solver(ctx, A, B, dev(ctx, s.b), x, eps=0.999999, x0_mode=N.X0_ZERO)
    assert e.value.code == 4
    P = O.diag_prec(s.ia, s.ja, s.a)
    xo, so = O.simple_solver(s.ia, s.ja, s.a, s.b, P, eps=0.999999, x0=np.zeros(s.n), check=False)
    assert so["niter"] == 1 and np.array_equal(x.get(), xo)
    with pytest.raises(N.MlpError) as e:
        N.simple_solver(ctx, A, B, dev(ctx, s.b), x, eps=1e-12, maxit=4, x0_mode=N.X0_ZERO)
    assert e.value.code == 5


def test_full_size_applies_bitexact(ctx):
    """At 60x220x85: the uh_cheb and multi_split (project/msplit.cfg settings) FP64 applies equal the oracle's bit
    for bit, and so does GS on the nonsymmetric variant of the matrix (the brick sweep)."""
    s = system("spe10_z85")
    f = rng(91).standard_normal(s.n)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    cases = [(N.DevPrec(ctx, A, N.build_uh_cheb(s.ia, s.ja, s.a)), O.cheb_prec(s.ia, s.ja, s.a)),
             (N.DevPrec(ctx, A, N.build_multi_split(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="simple_1",
                                                    degree_max=2)),
              O.msplit_prec(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="simple_1", degree_max=2))]
    for B, P in cases:
        x = ctx.array(s.n)
        B.apply(dev(ctx, f), x)
        assert np.array_equal(x.get(), P.apply(f)), B.kind
        B.close()
    A.close()
    a_ns = S.make_nonsymmetric(s.ia, s.ja, s.a, 0.3)
    A = N.DevMatrix(ctx, s.ia, s.ja, a_ns)
    B = N.DevPrec(ctx, A, kind=N.GS)
    x = ctx.array(s.n)
    B.apply(dev(ctx, f), x)
    assert np.array_equal(x.get(), O.gs_prec(s.ia, s.ja, a_ns).apply(f))


@pytest.mark.parametrize("precision", [N.MLP_FP64, N.MLP_FP32])
def test_flexible_pcg_multi_split(ctx, precision):
    """Flexible PCG (beta from (r, z - z_prev)) with the multi_split hierarchy, FP64 or FP32: beside the parity
    bar (the reference's PCG has no flexible beta), it must reach the tolerance on the true residual, agree with
    the oracle's solution to the accuracy both solves have, and need no more iterations than the oracle's PCG."""
    s = system("spe10_z12")
    P, H = _msplit(s)
    xo, so = O.pcg(s.ia, s.ja, s.a, s.b, P, eps=1e-8, x0=np.zeros(s.n), maxit=5000)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, H, precision=precision)
    x = ctx.array(s.n)
    sg = N.pcg(ctx, A, B, dev(ctx, s.b), x, eps=1e-8, maxit=5000, flexible=True, true_res_every=20)
    assert sg["true_rel_res"] <= 1e-8
    assert np.linalg.norm(x.get() - xo) <= 1e-5 * np.linalg.norm(xo)
    assert sg["niter"] <= so["niter"]


def test_flexible_pcg_reports_missed_tolerance(ctx):
    """Flexible PCG restarts from the FP64 true residual at most 8 times; a tolerance below what FP64 can
    reach leaves the true residual above eps, which must be reported as MLP_ERR_NO_CONV (8) with the stats
    filled, not as success.  eps = 1e-17 is raised by the reference's floor (solvers.cc:49-58) to an absolute
    residual of 1e-14, below the rounding of b - A x over 158,400 rows (about 1e-16 |b_i| each)."""
    s = system("spe10_z12")
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, N.build_uh_cheb(s.ia, s.ja, s.a), precision=N.MLP_FP32)
    x = ctx.array(s.n)
    st = N.pcg(ctx, A, B, dev(ctx, s.b), x, eps=1e-17, maxit=20000, flexible=True, true_res_every=20,
               raise_on_error=False)
    assert st["rc"] == 8 and st["status"] == 8, st
    assert st["true_rel_res"] > st["eps_used"]
    assert st["true_rel_res"] < 1e-9                     # the solve itself went as far as FP64 allows
    # the same solve at a reachable tolerance succeeds
    st = N.pcg(ctx, A, B, dev(ctx, s.b), x, eps=1e-10, maxit=20000, flexible=True, true_res_every=20)
    assert st["rc"] == 0 and st["true_rel_res"] <= 1e-10

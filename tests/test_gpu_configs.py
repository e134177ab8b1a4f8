"""The workload configurations at their stated sizes on the device, against the oracle.

* Exact two-point-flux matrix (pure Dirichlet, 60x220x85, 1,122,000 cells): SpMV bit-exact; Jacobi PCG
  capped at 40 iterations equal to the oracle's capped solve, and the converged solve against the oracle's
  result kept in tests/golden (its 2,668 iterations take the oracle about 90 s).
* FP32 hierarchy inside FP64 flexible PCG at 60x220x85 to 1e-8 and 1e-10: the FP64 true residual meets the
  tolerance and the solution is within a bound of the FP64 oracle solution.
* 64 systems with depths drawn from {5, 10, 20, 40, 85}: the concurrent solves are bitwise the single solves,
  and three of them (FP64 hierarchy) agree with the oracle; the same 64 stacked into one block-diagonal matrix per
  hierarchy depth, each group's schedule autotuned, solved one launch per step per group
  (mlp_pcg_ensemble_batch): FP64 bitwise the single solves, the FP32 hierarchy in flexible PCG within the FP64
  true-residual bar; 32 right-hand sides of the 60x220x85 system in one batched solve, each bitwise its single
  solve, one against the oracle.
* Adjoint gradient at 1,122,000 cells: 200 observations, 20 seeded directions, central differences
  (h = 1e-6) within 1e-5 relative, and the device gradient against the FP64 gradient from the oracle's solves.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import rng, system
from ml_prec_amd import native as N
from ml_prec_amd import spe_synth as S

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def true_rel(s, x, init=None):
    r = s.b - O.spmv(s.ia, s.ja, s.a, x)
    return np.linalg.norm(r) / (np.linalg.norm(s.b) if init is None else init)


@pytest.fixture(scope="module")
def tpfa85():
    return S.spe10_system(kind="tpfa", c=0.0)


def test_tpfa_full_size_spmv_and_capped_jacobi(ctx, tpfa85):
    s = tpfa85
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    x = rng(81).standard_normal(s.n)
    y = ctx.array(s.n)
    A.spmv(ctx.to_device(x), y)
    assert np.array_equal(y.get(), O.spmv(s.ia, s.ja, s.a, x))
    B = N.DevPrec(ctx, A, kind=N.DIAG)
    xo, so = O.pcg(s.ia, s.ja, s.a, s.b, O.diag_prec(s.ia, s.ja, s.a), eps=1e-8, x0=np.zeros(s.n), maxit=40,
                   check=False)
    xd = ctx.array(s.n)
    st = N.pcg(ctx, A, B, ctx.to_device(s.b), xd, eps=1e-8, maxit=40, raise_on_error=False)
    assert st["rc"] == 5 and st["niter"] == 40 == so["niter"]
    assert np.linalg.norm(xd.get() - xo) <= 1e-12 * np.linalg.norm(xo)
    B.close()
    A.close()


def test_tpfa_full_size_jacobi_pcg_vs_golden(ctx, tpfa85):
    s = tpfa85
    g = np.load(os.path.join(GOLDEN, "tpfa_jacobi_60x220x85.npz"), allow_pickle=False)
    assert float(np.sum(s.a)) == float(g["a_sum"]) and float(np.sum(s.b)) == float(g["b_sum"])
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, kind=N.DIAG)
    xd = ctx.array(s.n)
    st = N.pcg(ctx, A, B, ctx.to_device(s.b), xd, eps=1e-8, maxit=200000, use_graph=True)
    x = xd.get()
    print(f"Jacobi PCG 60x220x85 TPFA: {st['niter']} iterations (oracle {int(g['niter'])}), "
          f"true rel res {st['true_rel_res']:.3e}")
    # 2,668 iterations on a contrast-1e6 matrix: reduction rounding moves the stop by a few iterations
    assert abs(int(st["niter"]) - int(g["niter"])) <= max(1, int(g["niter"]) // 200)
    assert true_rel(s, x, st["init_norm"]) <= 1e-8 * (1 + 1e-3)
    stride = int(g["stride"])
    err = np.linalg.norm(x[::stride] - g["x_samples"]) / np.linalg.norm(g["x_samples"])
    assert err <= 1e-6, err
    B.close()
    A.close()


@pytest.mark.parametrize("eps", [1e-8, 1e-10])
def test_mixed_precision_full_size(ctx, eps):
    s = system("spe10_z85")
    P = O.cheb_prec(s.ia, s.ja, s.a)
    xo, so = O.pcg(s.ia, s.ja, s.a, s.b, P, eps=eps, x0=np.zeros(s.n))
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    H = N.build_uh_cheb(s.ia, s.ja, s.a)
    B32 = N.DevPrec(ctx, A, H, precision=N.MLP_FP32)
    x = ctx.array(s.n)
    st = N.pcg(ctx, A, B32, ctx.to_device(s.b), x, eps=eps, maxit=5000, flexible=True, true_res_every=20,
               use_graph=True)
    xg = x.get()
    rr = true_rel(s, xg, st["init_norm"])
    err = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    print(f"mixed 60x220x85 eps {eps:g}: {st['niter']} iterations (oracle FP64 {so['niter']}), true rel res "
          f"{rr:.3e}, error vs FP64 oracle {err:.3e}")
    assert rr <= eps * (1 + 1e-3)
    assert st["niter"] <= so["niter"] + 2
    # two solutions that both meet eps: their difference is bounded by the error the tolerance allows
    assert err <= 1e3 * eps, err
    B32.close()
    A.close()


def _jobs(systems, precision):
    out = []
    for s in systems:
        c = N.Context(0)
        A = N.DevMatrix(c, s.ia, s.ja, s.a)
        B = N.DevPrec(c, A, N.build_uh_cheb(s.ia, s.ja, s.a), precision=precision)
        out.append((c, A, B, c.to_device(s.b), c.array(s.n)))
    return out


def _close(jobs):
    for c, A, B, b, x in jobs:
        B.close()
        A.close()
    for c, *_ in jobs:
        c.close()


def test_ensemble_64_systems():
    spec = S.ensemble_spec(64)
    assert {nz for nz, _ in spec} == set(S.ENSEMBLE_NZ)
    systems = [S.ensemble_system(nz, seed) for nz, seed in spec]
    kw = dict(eps=1e-8, maxit=5000, x0_mode=N.X0_ZERO, check_every=8, use_graph=True, flexible=True,
              true_res_every=20)
    jobs = _jobs(systems, N.MLP_FP32)
    single = []
    for c, A, B, b, x in jobs:
        st = N.pcg(c, A, B, b, x, **kw)
        single.append((st["niter"], x.get()))
        x.upload(np.zeros(x.n))
    rc, stats = N.pcg_batch(jobs, **kw)
    assert rc == 0
    for (niter, x1), st, job, s in zip(single, stats, jobs, systems):
        assert st["rc"] == 0 and st["niter"] == niter
        assert np.array_equal(job[4].get(), x1)
        assert st["true_rel_res"] <= 1e-8
    _close(jobs)
    # three members, one of each of the three smallest depths, with the FP64 hierarchy against the oracle
    picks = []
    for nz in (5, 10, 20):
        picks.append(next(i for i, (z, _) in enumerate(spec) if z == nz))
    sub = [systems[i] for i in picks]
    jobs = _jobs(sub, N.MLP_FP64)
    rc, stats = N.pcg_batch(jobs, eps=1e-8, maxit=5000, x0_mode=N.X0_ZERO, check_every=8, use_graph=True)
    assert rc == 0
    for s, st, job in zip(sub, stats, jobs):
        xo, so = O.pcg(s.ia, s.ja, s.a, s.b, O.cheb_prec(s.ia, s.ja, s.a), eps=1e-8, x0=np.zeros(s.n))
        assert abs(int(st["niter"]) - so["niter"]) <= 1
        tol = 1e-10 if st["niter"] == so["niter"] else 1e-6
        assert np.linalg.norm(job[4].get() - xo) <= tol * np.linalg.norm(xo)
    _close(jobs)


def test_ensemble_64_systems_stacked_and_tuned():
    spec = S.ensemble_spec(64)
    hh = (S.HX, S.HY, S.HZ)
    fields = [S.spe10_field(60, 220, nz, seed).ravel() for nz, seed in spec]
    rhs = [S.rhs(k.size, seed + 7919) for k, (_, seed) in zip(fields, spec)]
    kw = dict(eps=1e-8, maxit=5000, x0_mode=N.X0_ZERO, check_every=8, use_graph=True)
    # every system alone: its own matrix and hierarchy built on the device, FP64 PCG
    ctx0 = N.Context(0)
    single, depth = [], []
    for (nz, _), kx, b in zip(spec, fields, rhs):
        A = N.DevMatrix.spe_on_device(ctx0, ctx0.to_device(kx), (nz, 220, 60), kind="spe", c=1.0, kz_ratio=0.1,
                                      h=hh)
        H = N.build_uh_cheb_on_device(ctx0, A)
        B = N.DevPrec(ctx0, A, H)
        x = ctx0.array(kx.size)
        st = N.pcg(ctx0, A, B, ctx0.to_device(b), x, **kw)
        single.append((st["niter"], x.get()))
        depth.append(H.nlevels)
        B.close()
        A.close()
    groups = {}
    for i, d in enumerate(depth):
        groups.setdefault(d, []).append(i)
    assert len(groups) >= 3
    # one stacked matrix and tuned hierarchy per depth, each group on a context of its own
    stacks = []
    for d, idx in sorted(groups.items()):
        c = N.Context(0)
        A = N.DevMatrix.spe_ensemble_on_device(c, c.to_device(np.concatenate([fields[i] for i in idx])), 60, 220,
                                               [spec[i][0] for i in idx], kind="spe", c=1.0, kz_ratio=0.1, h=hh)
        H = N.build_uh_cheb_on_device(c, A)
        assert H.nlevels == d
        off = np.concatenate([[0], np.cumsum([fields[i].size for i in idx])]).astype(np.uint32)
        B64 = N.DevPrec.tuned(c, A, H, precision=N.MLP_FP64, reps=3)
        B32 = N.DevPrec.tuned(c, A, H, precision=N.MLP_FP32, reps=3)
        stacks.append((c, A, B64, B32, idx, off, c.to_device(np.concatenate([rhs[i] for i in idx])),
                       c.array(int(off[-1]))))
    for mixed in (False, True):
        jobs = [(c, A, B32 if mixed else B64, off, b, x) for c, A, B64, B32, idx, off, b, x in stacks]
        rc, res = N.pcg_ensemble_batch(jobs, **(dict(kw, flexible=True, true_res_every=20) if mixed else kw))
        assert rc == 0
        for (c, A, B64, B32, idx, off, b, x), (status, stats) in zip(stacks, res):
            assert status == 0
            xs = x.get()
            for i, st, o in zip(idx, stats, off[:-1]):
                niter, x1 = single[i]
                xi = xs[o:o + x1.size]
                if mixed:
                    assert st["true_rel_res"] <= 1e-8 and st["niter"] <= niter + 2
                    assert np.linalg.norm(xi - x1) <= 1e-5 * np.linalg.norm(x1)
                else:
                    assert st["niter"] == niter
                    assert np.array_equal(xi.view(np.uint64), x1.view(np.uint64))
    ctxs = [st[0] for st in stacks]
    for c, A, B64, B32, idx, off, b, x in stacks:
        B64.close()
        B32.close()
        A.close()
    del jobs, res, b, x
    stacks.clear()                                  # the group vectors are freed before their contexts close
    for c in ctxs:
        c.close()
    ctx0.close()


def test_32_rhs_full_size(ctx):
    s = system("spe10_z85")
    R = 32
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    H = N.build_uh_cheb(s.ia, s.ja, s.a)
    B = N.DevPrec(ctx, A, H)
    F = np.concatenate([s.b] + [rng(200 + k).uniform(-1, 1, s.n) for k in range(R - 1)])
    b = ctx.to_device(F)
    x = ctx.array(R * s.n)
    rc, stats = N.pcg_multi(ctx, A, B, R, b, x, eps=1e-8, maxit=5000, x0_mode=N.X0_ZERO, check_every=8)
    assert rc == 0
    X = x.get().reshape(R, s.n)
    for k in range(R):
        xs = ctx.array(s.n)
        st = N.pcg(ctx, A, B, ctx.to_device(F[k * s.n:(k + 1) * s.n]), xs, eps=1e-8, maxit=5000, use_graph=True,
                   check_every=8)
        assert st["niter"] == stats[k]["niter"]
        assert np.array_equal(xs.get(), X[k]), k
    xo, so = O.pcg(s.ia, s.ja, s.a, s.b, O.cheb_prec(s.ia, s.ja, s.a), eps=1e-8, x0=np.zeros(s.n))
    assert abs(int(stats[0]["niter"]) - so["niter"]) <= 1
    tol = 1e-10 if stats[0]["niter"] == so["niter"] else 1e-6
    assert np.linalg.norm(X[0] - xo) <= tol * np.linalg.norm(xo)
    B.close()
    A.close()


def grad_formula(shape, kx, p, lam, kzr=0.1):
    """-lam^T dA/d(ln kx_i) p face by face (the formula pinned against assembly differences on the CPU)."""
    nz, ny, nx = shape
    k3, p3, l3 = kx.reshape(shape), p.reshape(shape), lam.reshape(shape)
    g = np.zeros(shape)
    for axis, fac, kr in ((2, 1 / S.HX**2, 1.0), (1, 1 / S.HY**2, 1.0), (0, 1 / S.HZ**2, kzr)):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[axis] = slice(None, -1)
        hi[axis] = slice(1, None)
        ka, kb = kr * k3[tuple(lo)], kr * k3[tuple(hi)]
        T = 2.0 / (1.0 / ka + 1.0 / kb) * fac
        w = (l3[tuple(lo)] - l3[tuple(hi)]) * (p3[tuple(lo)] - p3[tuple(hi)])
        g[tuple(lo)] -= T * T / (2 * fac * ka) * w
        g[tuple(hi)] -= T * T / (2 * fac * kb) * w
    return g.ravel()


def test_adjoint_gradient_full_size(ctx):
    from ml_prec_amd.adjoint import SpeAdjointModel, misfit
    nx, ny, nz = 60, 220, 85
    shape = (nz, ny, nx)
    n = nx * ny * nz
    g = rng(401)
    theta_true = np.log(S.spe10_field(nx, ny, nz, S.BASE_SEED + 4).ravel())
    theta = np.log(S.spe10_field(nx, ny, nz, S.BASE_SEED + 1).ravel())
    b = S.rhs(n, S.BASE_SEED + 6)
    obs = g.choice(n, 200, replace=False)
    model = SpeAdjointModel(ctx, shape, b, eps=1e-10)
    p_true = model.forward(theta_true)
    d = p_true[obs] + g.normal(0, 0.01, obs.size) * np.abs(p_true[obs])
    p = model.forward(theta)
    L, gp = misfit(p, obs, d)
    grad = model.vjp(gp)
    # FP64 CPU gradient from the oracle's forward and adjoint solves
    kx = np.exp(theta)
    ia, ja, a = N.spe_assemble(kx.reshape(shape), kind="spe", c=1.0)
    P = O.cheb_prec(ia, ja, a)
    po, _ = O.pcg(ia, ja, a, b, P, eps=1e-10, x0=np.zeros(n))
    _, gpo = misfit(po, obs, d)
    lo, _ = O.pcg(ia, ja, a, gpo, P, eps=1e-10, x0=np.zeros(n))
    gref = grad_formula(shape, kx, po, lo)
    gerr = np.linalg.norm(grad - gref) / np.linalg.norm(gref)
    # central differences with forward solves to 1e-13, so that the difference quotient resolves 1e-5
    fdm = SpeAdjointModel(ctx, shape, b, eps=1e-13)
    h = 1e-6
    worst = 0.0
    for k in range(20):
        v = rng(500 + k).standard_normal(n)
        Lp, _ = misfit(fdm.forward(theta + h * v), obs, d)
        Lm, _ = misfit(fdm.forward(theta - h * v), obs, d)
        fd = (Lp - Lm) / (2 * h)
        worst = max(worst, abs(fd - grad @ v) / abs(fd))
    print(f"adjoint 60x220x85: gradient vs oracle-solve FP64 gradient {gerr:.3e}, worst central difference "
          f"{worst:.3e} over 20 directions")
    assert gerr <= 1e-8
    assert worst <= 1e-5
    fdm.close()
    model.close()

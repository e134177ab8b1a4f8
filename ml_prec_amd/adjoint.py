"""Adjoint gradients through the device solve.

The model: A(theta) p = b with A the reference's two-point-flux matrix
(modules/mesh/spe.cc:69-131, reaction coefficient c, ky = kx, kz = 0.1 kx) and
theta = ln kx per cell.  A misfit L(p) is differentiated through the solve:

    dL/dtheta_i = -lam^T (dA/dtheta_i) p,    A lam = dL/dp   (A is symmetric),

so one gradient costs the forward solve plus one adjoint solve with the same
hierarchy, and one pass of k_spe_grad.  The forward and adjoint solves are
PCG with the uh_cheb preconditioner built for A(theta).  By default the
matrix is assembled on the device from kx = exp(theta) and the hierarchy is
built there from it (device=False assembles and builds on the host; both give
the same matrix and hierarchy bit for bit).

`SpeAdjointModel` is the plain interface; `spe_solve_function()` wraps it as a
torch.autograd.Function (CPU float64 tensors in and out; all solver work runs
on the device through the C ABI).
"""
from __future__ import annotations

import time

import numpy as np

from . import native as N
from . import spe_synth as S

KZ_RATIO = 0.1


class SpeAdjointModel:
    """Solves A(theta) p = b on the device and returns adjoint gradients.

    shape = (nz, ny, nx); b is the right-hand side (float64[N]); eps the PCG
    tolerance of both solves.
    """

    def __init__(self, ctx: N.Context, shape, b, c=1.0, eps=1e-10, maxit=20000, device=True):
        self.ctx = ctx
        self.device = bool(device)
        self.shape = tuple(int(v) for v in shape)
        self.n = int(np.prod(self.shape))
        self.b = np.ascontiguousarray(b, dtype=np.float64)
        if self.b.shape != (self.n,):
            raise ValueError("b has the wrong size")
        self.c = float(c)
        self.eps = float(eps)
        self.maxit = int(maxit)
        self.faces = (1.0 / (S.HX * S.HX), 1.0 / (S.HY * S.HY), 1.0 / (S.HZ * S.HZ))
        self.db = ctx.to_device(self.b)
        self._state = None          # (theta, A, B, kx, p) of the last forward solve
        self.stats = []
        self.phases = {}            # host wall time (ms) of the last forward's phases

    def _assemble(self, theta):
        kx = np.exp(np.asarray(theta, dtype=np.float64))
        ia, ja, a = N.spe_assemble(kx.reshape(self.shape), kind="spe", c=self.c, kz_ratio=KZ_RATIO,
                                   h=(S.HX, S.HY, S.HZ))
        return kx, ia, ja, a

    def close(self):
        if self._state:
            _, A, B, _, _ = self._state
            B.close()
            A.close()
            self._state = None

    def forward(self, theta):
        """p(theta) as a host array; keeps the matrix and hierarchy for the adjoint."""
        self.close()
        t0 = time.perf_counter()
        if self.device:
            kx = np.exp(np.asarray(theta, dtype=np.float64))
            dkx = self.ctx.to_device(kx)
            A = N.DevMatrix.spe_on_device(self.ctx, dkx, self.shape, kind="spe", c=self.c, kz_ratio=KZ_RATIO,
                                          h=(S.HX, S.HY, S.HZ))
            t1 = time.perf_counter()
            H = N.build_uh_cheb_on_device(self.ctx, A)
        else:
            kx, ia, ja, a = self._assemble(theta)
            A = N.DevMatrix(self.ctx, ia, ja, a)
            dkx = self.ctx.to_device(kx)
            t1 = time.perf_counter()
            H = N.build_uh_cheb(ia, ja, a)
        t2 = time.perf_counter()
        B = N.DevPrec(self.ctx, A, H)
        t3 = time.perf_counter()
        p = self.ctx.array(self.n)
        st = N.pcg(self.ctx, A, B, self.db, p, eps=self.eps, maxit=self.maxit, x0_mode=N.X0_ZERO, use_graph=True,
                   check_every=8)
        self.phases = {"matrix": 1e3 * (t1 - t0), "construction": 1e3 * (t2 - t1), "plan_upload": 1e3 * (t3 - t2)}
        self.stats.append(("forward", st))
        self._state = (np.array(theta, dtype=np.float64), A, B, dkx, p)
        return p.get()

    def vjp(self, dL_dp):
        """dL/dtheta for the last forward solve, given dL/dp (host array)."""
        if not self._state:
            raise RuntimeError("forward() must run before vjp()")
        _, A, B, dkx, p = self._state
        g = np.ascontiguousarray(dL_dp, dtype=np.float64)
        lam = self.ctx.array(self.n)
        if np.any(g):
            st = N.pcg(self.ctx, A, B, self.ctx.to_device(g), lam, eps=self.eps, maxit=self.maxit,
                       x0_mode=N.X0_ZERO, use_graph=True, check_every=8)
            self.stats.append(("adjoint", st))
        else:
            lam.upload(np.zeros(self.n))
        grad = self.ctx.array(self.n)
        N.spe_adjoint_grad(self.ctx, self.shape, self.faces, KZ_RATIO, dkx, p, lam, grad)
        return grad.get()


def misfit(p, obs, d):
    """L = sum_k (p[obs_k] - d_k)^2 and dL/dp."""
    r = p[obs] - d
    g = np.zeros_like(p)
    np.add.at(g, obs, 2.0 * r)
    return float(r @ r), g


def spe_solve_function():
    """torch.autograd.Function: p = SpeSolve.apply(theta, model); backward runs the adjoint solve."""
    import torch

    class SpeSolve(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, model):
            ctx.model = model
            p = model.forward(theta.detach().cpu().numpy())
            return torch.from_numpy(p)

        @staticmethod
        def backward(ctx, grad_p):
            g = ctx.model.vjp(grad_p.detach().cpu().numpy())
            return torch.from_numpy(g), None

    return SpeSolve

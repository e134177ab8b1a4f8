"""Writes tpfa_jacobi_60x220x85.npz: the oracle's Jacobi-preconditioned PCG on the pure-Dirichlet
two-point-flux matrix of the 60x220x85 SPE10-shaped field (seed BASE_SEED + 1, x0 = 0, eps 1e-8).

The oracle needs about 90 s for its 2,668 iterations, too long for a test, so the result is kept as a
fixture: the iteration count, the final recurrence norm, the initial norm, ||x|| and every 997th entry of x.
Run from the repository root:  python tests/golden/make_tpfa_jacobi_60x220x85.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O                                   # noqa: E402
from ml_prec_amd import spe_synth as S               # noqa: E402

STRIDE = 997


def main():
    s = S.spe10_system(kind="tpfa", c=0.0)
    P = O.diag_prec(s.ia, s.ja, s.a)
    x, st = O.pcg(s.ia, s.ja, s.a, s.b, P, eps=1e-8, x0=np.zeros(s.n), maxit=200000)
    np.savez(os.path.join(HERE, "tpfa_jacobi_60x220x85.npz"), niter=np.array(st["niter"]),
             init_norm=np.array(st["init_norm"]), final_norm=np.array(st["final_norm"]),
             x_norm=np.array(np.linalg.norm(x)), stride=np.array(STRIDE), x_samples=x[::STRIDE].copy(),
             a_sum=np.array(float(np.sum(s.a))), b_sum=np.array(float(np.sum(s.b))))
    print("niter", st["niter"], "||x||", np.linalg.norm(x))


if __name__ == "__main__":
    main()

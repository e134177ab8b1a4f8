"""The C++ drop-in API (ml_prec_amd/cpp/ml_prec.h) end to end on the GPU: a
driver that follows the reference's main.cc sequence loads a CSR BINARY matrix
(diagonal not first, so SkylineMatrix::load must transform it), builds the
preconditioner from a Config / prec config file and runs PCGSolver or
GMRESSolver from the reference's generate_x0 start."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import system
from ml_prec_amd import build as B
from ml_prec_amd import spe_synth as S

pytestmark = pytest.mark.gpu


def plain_csr(s):
    """Ascending columns with the diagonal in place (the layout CSRMatrix::load reads)."""
    ia, ja, a = [0], [], []
    for i in range(s.n):
        cols = s.ja[s.ia[i]:s.ia[i + 1]]
        vals = s.a[s.ia[i]:s.ia[i + 1]]
        o = np.argsort(cols, kind="stable")
        ja.extend(cols[o])
        a.extend(vals[o])
        ia.append(len(ja))
    return np.array(ia, np.uint32), np.array(ja, np.uint32), np.array(a)


def run_demo(tmp_path, s, prec, solver, eps, conf="-", prec_arg=""):
    ia, ja, a = plain_csr(s)
    assert not np.array_equal(ja, s.ja)
    S.write_csr_binary(tmp_path / "A.bin", ia, ja, a)
    with open(tmp_path / "b.txt", "w") as f:
        f.write(f"# Vector size\n{s.n}\n# Vector values\n")
        f.write("\n".join(f"{v:.17e}" for v in s.b) + "\n")
    out = tmp_path / "x.bin"
    cmd = [B.DEMO, str(tmp_path / "A.bin"), str(tmp_path / "b.txt"), str(out), prec, solver, str(eps), conf]
    if prec_arg:
        cmd.append(prec_arg)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    toks = r.stdout.split()
    data = np.fromfile(out, dtype=np.float64)
    return int(toks[1]), float(toks[3]), data[:s.n], data[s.n:]


def test_dropin_pcg_uh_cheb(tmp_path):
    s = system("spe10_z12")
    b_read = np.array([float(f"{v:.17e}") for v in s.b])
    assert np.array_equal(b_read, s.b)
    niter, rel, x, z = run_demo(tmp_path, s, "uh_cheb", "pcg", 1e-8)
    P = O.cheb_prec(s.ia, s.ja, s.a)
    xo, so = O.pcg(s.ia, s.ja, s.a, s.b, P, eps=1e-8, x0=None, maxit=5000)
    assert abs(niter - so["niter"]) <= 1
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)
    assert np.array_equal(z, P.apply(s.b))           # PrecBase::solve through the device


def test_dropin_gmres_multi_split_config_file(tmp_path):
    s = system("spe10_z12")
    conf = tmp_path / "msplit.cfg"
    conf.write_text("q = 0.8\ninner_iter = fixed\nlevel_niter = 5\nelim_order = simple_1\n"
                    "elim_stop = degree\ndegree_max = 2\ncoarse_n = 0\nmax_levels = 15\n")
    niter, rel, x, z = run_demo(tmp_path, s, "multi_split", "gmres", 1e-8, str(conf))
    P = O.msplit_prec(s.ia, s.ja, s.a, q=(0.8,), niters=(5,), order="simple_1", degree_max=2)
    xo, so = O.gmres(s.ia, s.ja, s.a, s.b, P, eps=1e-8, x0=None, maxit=3000)
    assert abs(niter - so["niter"]) <= 1
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)
    assert np.array_equal(z, P.apply(s.b))


def test_dropin_mixed_precision(tmp_path):
    s = system("spe10_z12")
    niter, rel, x, z = run_demo(tmp_path, s, "uh_cheb", "pcg", 1e-8, "-", "fp32")
    P = O.cheb_prec(s.ia, s.ja, s.a)
    zo = P.apply(s.b)
    assert np.linalg.norm(z - zo) <= 1e-5 * np.linalg.norm(zo)


MSPLIT_CFG = ("q = 0.8\ninner_iter = fixed\nlevel_niter = 5\nelim_order = simple_1\n"
              "elim_stop = degree\ndegree_max = 2\ncoarse_n = 0\nmax_levels = 15\n")
MSPLIT_KW = dict(q=(0.8,), niters=(5,), order="simple_1", degree_max=2)


def test_dropin_cheb_and_simple_solvers(tmp_path):
    s = system("spe10_z12")
    _, _, x, _ = run_demo(tmp_path, s, "uh_cheb", "cheb", 1e-8)
    P = O.cheb_prec(s.ia, s.ja, s.a)
    li = P.level_info(0)
    xo, so = O.cheb_solver(s.ia, s.ja, s.a, s.b, P, li["lmin"], li["lmax"], eps=1e-8, x0=None, maxit=20000)
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)
    conf = tmp_path / "msplit.cfg"
    conf.write_text(MSPLIT_CFG)
    niter, _, x, _ = run_demo(tmp_path, s, "multi_split", "simple", 1e-8, str(conf))
    P = O.msplit_prec(s.ia, s.ja, s.a, **MSPLIT_KW)
    xo, so = O.simple_solver(s.ia, s.ja, s.a, s.b, P, eps=1e-8, x0=None, maxit=5000)
    assert abs(niter - so["niter"]) <= 1
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)


@pytest.mark.parametrize("prec", ["comp", "gs", "sym_split"])
def test_dropin_gmres_nonsymmetric(tmp_path, prec):
    s0 = system("ln5_spe")
    a = S.make_nonsymmetric(s0.ia, s0.ja, s0.a, 0.3)
    s = S.System(**{**s0.__dict__, "a": a})
    conf = tmp_path / "msplit.cfg"
    conf.write_text(MSPLIT_CFG)
    niter, _, x, z = run_demo(tmp_path, s, prec, "gmres", 1e-8, str(conf))
    if prec == "comp":
        P = O.composite_prec(O.gs_prec(s.ia, s.ja, a), O.msplit_prec(s.ia, s.ja, a, **MSPLIT_KW))
    elif prec == "gs":
        P = O.gs_prec(s.ia, s.ja, a)
    else:
        P = O.sym_prec(s.ia, s.ja, a)
    xo, so = O.gmres(s.ia, s.ja, a, s.b, P, eps=1e-8, x0=None, maxit=3000)
    assert abs(niter - so["niter"]) <= 1
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)
    zo = P.apply(s.b)
    if prec == "sym_split":
        assert np.linalg.norm(z - zo) <= 1e-10 * np.linalg.norm(zo)
    else:
        assert np.array_equal(z, zo)

"""ctypes wrapper around the C oracle (``liboracle.so``) -- TEST INFRASTRUCTURE ONLY.

Only the test suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module.  The HIP product path is tested against
it and never calls it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

UH_CHEB, MULTI_SPLIT, DIAG, ID, GS, COMPOSITE, SYM = 1, 2, 3, 4, 5, 6, 7
ORDERS = {"none": 0, "original": 1, "block": 2, "simple_1": 3}

_lib = None
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, no FMA contraction)."""
    if force or not os.path.exists(LIB_PATH):
        cmd = ["make", "-s", "-C", HERE] + (["-B"] if force else [])
        subprocess.run(cmd, check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        L.or_cheb_build.restype = C.c_void_p
        L.or_cheb_build.argtypes = [C.c_uint32, _u32p, _u32p, _f64p, _f64p, C.c_uint32, _u32p, C.c_uint32,
                                    C.c_int, C.c_int, C.c_char_p, C.c_size_t]
        L.or_msplit_build.restype = C.c_void_p
        L.or_msplit_build.argtypes = [C.c_uint32, _u32p, _u32p, _f64p, _f64p, C.c_uint32, _u32p, C.c_uint32,
                                      _f64p, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_uint32,
                                      C.c_uint32, C.c_char_p, C.c_size_t]
        L.or_diag_build.restype = C.c_void_p
        L.or_diag_build.argtypes = [C.c_uint32, _u32p, _u32p, _f64p]
        L.or_id_build.restype = C.c_void_p
        L.or_id_build.argtypes = [C.c_uint32]
        L.or_gs_build.restype = C.c_void_p
        L.or_gs_build.argtypes = [C.c_uint32, _u32p, _u32p, _f64p, C.c_char_p, C.c_size_t]
        L.or_composite_make.restype = C.c_void_p
        L.or_composite_make.argtypes = [C.c_void_p, C.c_void_p]
        L.or_sym_build.restype = C.c_void_p
        L.or_sym_build.argtypes = [C.c_uint32, _u32p, _u32p, _f64p, _f64p, C.c_uint32, _u32p, C.c_uint32,
                                   C.c_int, C.c_int, C.c_char_p, C.c_size_t]
        L.or_sym_matrix.argtypes = [C.c_void_p, _f64p]
        L.or_prec_free.argtypes = [C.c_void_p]
        L.or_prec_apply.argtypes = [C.c_void_p, _f64p, _f64p]
        L.or_prec_nlevels.restype = C.c_uint32
        L.or_prec_nlevels.argtypes = [C.c_void_p]
        L.or_prec_level_info.argtypes = [C.c_void_p, C.c_uint32, _u32p, _f64p]
        L.or_prec_level_array.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]
        for nm in ("or_pcg_csr", "or_gmres_csr"):
            f = getattr(L, nm)
            f.argtypes = [C.c_uint32, _u32p, _u32p, _f64p, _f64p, C.c_void_p, _f64p, _f64p,
                          C.c_double, C.c_int, C.c_uint32, C.c_char_p, C.c_size_t]
        L.or_cheb_solver_csr.argtypes = [C.c_uint32, _u32p, _u32p, _f64p, C.c_double, C.c_double, _f64p,
                                         C.c_void_p, _f64p, _f64p, C.c_double, C.c_int, C.c_uint32,
                                         C.c_char_p, C.c_size_t]
        L.or_simple_solver_csr.argtypes = [C.c_uint32, _u32p, _u32p, _f64p, _f64p, C.c_void_p, _f64p, _f64p,
                                           C.c_double, C.c_int, C.c_uint32, C.c_char_p, C.c_size_t]
        for nm in ("or_spmv_csr", "or_sym_spmv_csr"):
            f = getattr(L, nm)
            f.argtypes = [C.c_uint32, _u32p, _u32p, _f64p, _f64p, _f64p]
        L.or_ddot.restype = C.c_double
        L.or_ddot.argtypes = [_f64p, _f64p, C.c_uint32]
        L.or_dnrm2.restype = C.c_double
        L.or_dnrm2.argtypes = [_f64p, C.c_uint32]
        L.or_daxpy.argtypes = [C.c_double, _f64p, _f64p, C.c_uint32]
        L.or_dscal.argtypes = [C.c_double, _f64p, C.c_uint32]
        L.or_generate_x0.argtypes = [_f64p, C.c_uint32]
        L.or_cheb.restype = C.c_double
        L.or_cheb.argtypes = [C.c_double, C.c_uint32]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _u32(x):
    return np.ascontiguousarray(x, dtype=np.uint32)


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


class Prec:
    """An oracle preconditioner hierarchy (uh_cheb, multi_split, diag or id)."""

    def __init__(self, handle, kind, n):
        self.h = handle
        self.kind = kind
        self.n = n

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_prec_free(self.h)
            self.h = None

    def apply(self, f):
        x = np.zeros(self.n)
        if lib().or_prec_apply(self.h, _f64(f), x):
            raise OracleError("preconditioner apply failed")
        return x

    @property
    def nlevels(self):
        return int(lib().or_prec_nlevels(self.h))

    def level_info(self, l):
        ints = np.zeros(9, dtype=np.uint32)
        dbls = np.zeros(5)
        if lib().or_prec_level_info(self.h, l, ints, dbls):
            raise IndexError(l)
        keys = ["N", "nnz", "M", "Md", "nit", "nnzL", "nnzU", "n_next", "nnz_next"]
        d = {k: int(v) for k, v in zip(keys, ints)}
        d.update(alpha=dbls[0], beta=dbls[1], lmin=dbls[2], lmax=dbls[3], eps=dbls[4])
        return d

    def level_arrays(self, l):
        info = self.level_info(l)
        N, M, Md = info["N"], info["M"], info["Md"]
        nL = N - Md
        out = {}

        def get(which, dtype, size):
            arr = np.zeros(max(size, 1), dtype=dtype)
            lib().or_prec_level_array(self.h, l, which, arr.ctypes.data_as(C.c_void_p))
            return arr[:size]

        out["map"] = get(0, np.uint32, N)
        out["rmap"] = get(11, np.uint32, N)
        out["aux"] = get(12, np.float64, N)
        out["L_ia"] = get(1, np.uint32, nL + 1)
        out["L_ja"] = get(2, np.uint32, info["nnzL"])
        out["L_a"] = get(3, np.float64, info["nnzL"])
        out["U_ia"] = get(4, np.uint32, M + 1)
        out["U_ja"] = get(5, np.uint32, info["nnzU"])
        out["U_a"] = get(6, np.float64, info["nnzU"])
        out["dval"] = get(7, np.float64, Md)
        n = info["n_next"]
        out["A_ia"] = get(8, np.uint32, n + 1 if n else 0)
        out["A_ja"] = get(9, np.uint32, info["nnz_next"])
        out["A_a"] = get(10, np.float64, info["nnz_next"])
        return info, out


def cheb_prec(ia, ja, a, sigmas=(3.0,), niters=(2,), use_tails=True, optimize_storage=True):
    """Prec(A, cfg) -- the default "uh_cheb" preconditioner (defaults of project/main_helper.cc:62-79)."""
    n = len(ia) - 1
    err = C.create_string_buffer(512)
    h = lib().or_cheb_build(n, _u32(ia), _u32(ja), _f64(a), _f64(sigmas), len(sigmas), _u32(niters),
                            len(niters), int(use_tails), int(optimize_storage), err, 512)
    if not h:
        raise OracleError(err.value.decode())
    return Prec(h, UH_CHEB, n)


def msplit_prec(ia, ja, a, q=(0.8,), niters=(5,), epss=(1e-2,), inner="fixed", order="original",
                use_tails=True, degree_max=3, max_levels=15, coarse_n=0):
    """MultiSplitPrec(A, cfg) with config-file values given directly
    (defaults of modules/prec/multi_split/multi_split_constr.cc:185-228)."""
    n = len(ia) - 1
    err = C.create_string_buffer(512)
    h = lib().or_msplit_build(n, _u32(ia), _u32(ja), _f64(a), _f64(q), len(q), _u32(niters), len(niters),
                              _f64(epss), len(epss), int(inner == "dynamic"), ORDERS[order], int(use_tails),
                              degree_max, max_levels, coarse_n, err, 512)
    if not h:
        raise OracleError(err.value.decode())
    return Prec(h, MULTI_SPLIT, n)


def gs_prec(ia, ja, a):
    """GSPrec(A): one forward Gauss-Seidel sweep."""
    n = len(ia) - 1
    err = C.create_string_buffer(512)
    h = lib().or_gs_build(n, _u32(ia), _u32(ja), _f64(a), err, 512)
    if not h:
        raise OracleError(err.value.decode())
    return Prec(h, GS, n)


def composite_prec(first, second):
    """CompositePrec: x = second(first(f)); the new object owns both stages."""
    h = lib().or_composite_make(first.h, second.h)
    P = Prec(h, COMPOSITE, first.n)
    first.h = second.h = None              # freed with the composite
    return P


def sym_prec(ia, ja, a, sigmas=(3.0,), niters=(2,), use_tails=True, optimize_storage=True):
    """SymPrec(A, cfg): uh_cheb on max(A, A^T), applied as PCG to 1e-14."""
    n = len(ia) - 1
    err = C.create_string_buffer(512)
    h = lib().or_sym_build(n, _u32(ia), _u32(ja), _f64(a), _f64(sigmas), len(sigmas), _u32(niters), len(niters),
                           int(use_tails), int(optimize_storage), err, 512)
    if not h:
        raise OracleError(err.value.decode())
    P = Prec(h, SYM, n)
    return P


def sym_matrix(P, nnz):
    out = np.zeros(nnz)
    if lib().or_sym_matrix(P.h, out):
        raise OracleError("not a SymPrec")
    return out


def diag_prec(ia, ja, a):
    n = len(ia) - 1
    return Prec(lib().or_diag_build(n, _u32(ia), _u32(ja), _f64(a)), DIAG, n)


def id_prec(n):
    return Prec(lib().or_id_build(n), ID, n)


def _solve(fn, ia, ja, a, b, prec, eps, x0=None, maxit=0):
    n = len(ia) - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    st = np.zeros(6)
    err = C.create_string_buffer(512)
    rc = fn(n, _u32(ia), _u32(ja), _f64(a), _f64(b), prec.h, x, st, eps, 0 if x0 is None else 1, maxit, err, 512)
    stats = dict(niter=int(st[0]), t_sol=st[1], init_norm=st[2], final_norm=st[3], eps=st[4], t_prec=st[5], rc=rc,
                 err=err.value.decode())
    return x, stats


def pcg(ia, ja, a, b, prec, eps=1e-8, x0=None, maxit=0, check=True):
    """PCGSolver.  x0=None reproduces the reference's generate_x0 start."""
    x, st = _solve(lib().or_pcg_csr, ia, ja, a, b, prec, eps, x0, maxit)
    if check and st["rc"]:
        raise OracleError(st["err"])
    return x, st


def gmres(ia, ja, a, b, prec, eps=1e-8, x0=None, maxit=0, check=True):
    x, st = _solve(lib().or_gmres_csr, ia, ja, a, b, prec, eps, x0, maxit)
    if check and st["rc"]:
        raise OracleError(st["err"])
    return x, st


def cheb_solver(ia, ja, a, b, prec, lmin, lmax, eps=1e-8, x0=None, maxit=0, check=True):
    """ChebSolver (needs the uh_cheb bounds Prec::lmin()/lmax())."""
    n = len(ia) - 1
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    st = np.zeros(6)
    err = C.create_string_buffer(512)
    rc = lib().or_cheb_solver_csr(n, _u32(ia), _u32(ja), _f64(a), lmin, lmax, _f64(b), prec.h, x, st, eps,
                                  0 if x0 is None else 1, maxit, err, 512)
    stats = dict(niter=int(st[0]), t_sol=st[1], init_norm=st[2], final_norm=st[3], eps=st[4], rc=rc,
                 err=err.value.decode())
    if check and rc:
        raise OracleError(stats["err"])
    return x, stats


def simple_solver(ia, ja, a, b, prec, eps=1e-8, x0=None, maxit=0, check=True):
    """SimpleSolver (preconditioned Richardson)."""
    x, st = _solve(lib().or_simple_solver_csr, ia, ja, a, b, prec, eps, x0, maxit)
    if check and st["rc"]:
        raise OracleError(st["err"])
    return x, st


def spmv(ia, ja, a, x):
    n = len(ia) - 1
    y = np.zeros(n)
    lib().or_spmv_csr(n, _u32(ia), _u32(ja), _f64(a), _f64(x), y)
    return y


def sym_spmv(ia, ja, a, x):
    n = len(ia) - 1
    y = np.zeros(n)
    lib().or_sym_spmv_csr(n, _u32(ia), _u32(ja), _f64(a), _f64(x), y)
    return y


def generate_x0(n):
    x = np.zeros(n)
    lib().or_generate_x0(x, n)
    return x


def ddot(x, y):
    return lib().or_ddot(_f64(x), _f64(y), len(x))


def dnrm2(x):
    return lib().or_dnrm2(_f64(x), len(x))


def daxpy(alpha, x, y):
    y = _f64(y).copy()
    lib().or_daxpy(alpha, _f64(x), y, len(x))
    return y


def dscal(alpha, x):
    x = _f64(x).copy()
    lib().or_dscal(alpha, x, len(x))
    return x


def cheb(x, k):
    return lib().or_cheb(x, k)

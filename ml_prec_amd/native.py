"""ctypes binding of ``libmlp.so`` (the C ABI in include/mlp.h).

The library is the only implementation of the solve path: when it is missing
every entry point raises, there is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmlp.so")

MLP_FP64, MLP_FP32, MLP_FP32_BF16 = 0, 1, 2
UH_CHEB, MULTI_SPLIT, DIAG, ID, GS, COMPOSITE, SYM = 1, 2, 3, 4, 5, 6, 7
SCHED_CHAIN, SCHED_PAIR, SCHED_WIDE_CHAIN = 1, 2, 4        # mlp_prec_schedule bits
X0_ZERO, X0_GIVEN, X0_REFERENCE = 0, 1, 2
ORDERS = {"none": 0, "original": 1, "block": 2, "simple_1": 3}
STATUS = {0: "OK", 1: "INVALID", 2: "HIP", 3: "NOT_SPD", 4: "NO_ITER", 5: "MAXIT", 6: "CONSTRUCT", 7: "CONFIG",
          8: "NO_CONV"}

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = C.c_void_p


class MlpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{STATUS.get(code, code)}] {msg}")
        self.code = code
        self.msg = msg


class LevelInfo(C.Structure):
    _fields_ = [("N", C.c_uint32), ("nnz", C.c_uint32), ("M", C.c_uint32), ("Md", C.c_uint32),
                ("nit", C.c_uint32), ("nnz_L", C.c_uint32), ("nnz_U", C.c_uint32),
                ("n_next", C.c_uint32), ("nnz_next", C.c_uint32),
                ("alpha", C.c_double), ("beta_or_q", C.c_double), ("lmin", C.c_double),
                ("lmax", C.c_double), ("eps", C.c_double)]


class MsplitParams(C.Structure):
    _fields_ = [("q", C.POINTER(C.c_double)), ("nq", C.c_uint32),
                ("niters", C.POINTER(C.c_uint32)), ("nniters", C.c_uint32),
                ("eps", C.POINTER(C.c_double)), ("neps", C.c_uint32),
                ("inner_dynamic", C.c_int), ("order", C.c_int), ("use_tails", C.c_int),
                ("degree_max", C.c_uint32), ("max_levels", C.c_uint32), ("coarse_n", C.c_uint32)]


class SolveOpts(C.Structure):
    _fields_ = [("eps", C.c_double), ("maxit", C.c_uint32), ("x0_mode", C.c_int), ("flexible", C.c_int),
                ("true_res_every", C.c_uint32), ("check_every", C.c_uint32), ("use_graph", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("niter", C.c_uint32), ("init_norm", C.c_double), ("final_norm", C.c_double),
                ("eps_used", C.c_double), ("true_rel_res", C.c_double), ("t_solve_ms", C.c_double),
                ("status", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LevelDesc(C.Structure):
    """mlp_level_desc: one level of a hierarchy as plain arrays (Prec::Level / MultiSplitPrec::Level)."""
    _fields_ = [("N", C.c_uint32), ("M", C.c_uint32), ("Md", C.c_uint32), ("nit", C.c_uint32),
                ("lmin", C.c_double), ("lmax", C.c_double), ("eps", C.c_double),
                ("map", _vp), ("L_ia", _vp), ("L_ja", _vp), ("L_a", _vp),
                ("U_ia", _vp), ("U_ja", _vp), ("U_a", _vp), ("dval", _vp),
                ("A_ia", _vp), ("A_ja", _vp), ("A_a", _vp)]


class EnsJob(C.Structure):
    _fields_ = [("ctx", _vp), ("A", _vp), ("B", _vp), ("nseg", C.c_uint32), ("seg_off", _vp), ("d_b", _vp),
                ("d_x", _vp), ("stats", C.POINTER(SolveStats)), ("status", C.c_int)]


class TuneReport(C.Structure):
    _fields_ = [("ncand", C.c_uint32), ("chosen", C.c_uint32), ("sched", C.c_uint32 * 8), ("us", C.c_double * 8)]


class PcgJob(C.Structure):
    _fields_ = [("ctx", _vp), ("A", _vp), ("B", _vp), ("d_b", _vp), ("d_x", _vp), ("stats", SolveStats),
                ("status", C.c_int)]


EXPORTS = {
    "mlp_last_error": (C.c_char_p, []),
    "mlp_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "mlp_ctx_destroy": (C.c_int, [_vp]),
    "mlp_ctx_sync": (C.c_int, [_vp]),
    "mlp_ctx_stream": (C.c_int, [_vp, C.POINTER(_vp)]),
    "mlp_dev_alloc": (C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "mlp_dev_free": (C.c_int, [_vp, _vp]),
    "mlp_copy_h2d": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "mlp_copy_d2h": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "mlp_csr_upload": (C.c_int, [_vp, C.c_uint32, _u32p, _u32p, _f64p, C.POINTER(_vp)]),
    "mlp_mat_destroy": (C.c_int, [_vp]),
    "mlp_mat_info": (C.c_int, [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]),
    "mlp_spmv": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mlp_residual": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "mlp_ddot": (C.c_int, [_vp, C.c_uint32, _vp, _vp, C.POINTER(C.c_double)]),
    "mlp_dnrm2": (C.c_int, [_vp, C.c_uint32, _vp, C.POINTER(C.c_double)]),
    "mlp_daxpy": (C.c_int, [_vp, C.c_uint32, C.c_double, _vp, _vp]),
    "mlp_dscal": (C.c_int, [_vp, C.c_uint32, C.c_double, _vp]),
    "mlp_hier_build_uh_cheb": (C.c_int, [C.c_uint32, _u32p, _u32p, _f64p, _f64p, C.c_uint32, _u32p, C.c_uint32,
                                         C.c_int, C.POINTER(_vp)]),
    "mlp_hier_build_uh_cheb_dev": (C.c_int, [_vp, C.c_uint32, _u32p, _u32p, _f64p, _f64p, C.c_uint32, _u32p,
                                            C.c_uint32, C.c_int, C.POINTER(_vp)]),
    "mlp_hier_build_multi_split": (C.c_int, [C.c_uint32, _u32p, _u32p, _f64p, C.c_char_p, C.POINTER(MsplitParams),
                                             C.POINTER(_vp)]),
    "mlp_hier_destroy": (C.c_int, [_vp]),
    "mlp_hier_nlevels": (C.c_int, [_vp, C.POINTER(C.c_uint32)]),
    "mlp_hier_level_info": (C.c_int, [_vp, C.c_uint32, C.POINTER(LevelInfo)]),
    "mlp_hier_level_array": (C.c_int, [_vp, C.c_uint32, C.c_int, _vp]),
    "mlp_prec_upload": (C.c_int, [_vp, C.c_int, C.c_uint32, _vp, _vp, C.c_int, C.POINTER(_vp)]),
    "mlp_prec_from_hier": (C.c_int, [_vp, _vp, _vp, C.c_int, C.POINTER(_vp)]),
    "mlp_prec_simple": (C.c_int, [_vp, C.c_int, _vp, C.POINTER(_vp)]),
    "mlp_prec_gs": (C.c_int, [_vp, _vp, C.POINTER(_vp)]),
    "mlp_prec_composite": (C.c_int, [_vp, _vp, _vp, C.POINTER(_vp)]),
    "mlp_prec_sym_set_eps": (C.c_int, [_vp, C.c_double]),
    "mlp_prec_clone": (C.c_int, [_vp, _vp, C.POINTER(_vp)]),
    "mlp_prec_sym": (C.c_int, [_vp, _vp, _f64p, C.c_uint32, _u32p, C.c_uint32, C.c_int, C.POINTER(_vp)]),
    "mlp_prec_destroy": (C.c_int, [_vp]),
    "mlp_prec_apply": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mlp_pcg": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "mlp_gmres": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "mlp_pcg_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "mlp_spe_adjoint_grad": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                       C.c_double, _vp, _vp, _vp, _vp]),
    "mlp_pcg_batch": (C.c_int, [C.c_uint32, C.POINTER(PcgJob), C.POINTER(SolveOpts)]),
    "mlp_pcg_multi": (C.c_int, [_vp, _vp, _vp, C.c_uint32, _vp, _vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "mlp_cheb_solver": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, _vp, _vp, C.POINTER(SolveOpts),
                                  C.POINTER(SolveStats)]),
    "mlp_simple_solver": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "mlp_prec_bytes": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_uint32)]),
    "mlp_prec_schedule": (C.c_int, [_vp, C.POINTER(C.c_uint32)]),
    "mlp_time_spmv": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint32, C.c_size_t, C.POINTER(C.c_double)]),
    "mlp_time_prec_apply": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint32, C.POINTER(C.c_double)]),
    "mlp_spe_nnz": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_uint32]),
    "mlp_spe_matrix_dev": (C.c_int, [_vp, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, _vp, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_double, C.POINTER(_vp)]),
    "mlp_mat_get_csr": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mlp_spe_ensemble_matrix_dev": (C.c_int, [_vp, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, _u32p, _vp,
                                              C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                              C.POINTER(_vp)]),
    "mlp_pcg_ensemble": (C.c_int, [_vp, _vp, _vp, C.c_uint32, _u32p, _vp, _vp, C.POINTER(SolveOpts),
                                   C.POINTER(SolveStats)]),
    "mlp_pcg_ensemble_batch": (C.c_int, [C.c_uint32, _vp, C.POINTER(SolveOpts)]),
    "mlp_prec_from_hier_tuned": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_uint32, C.POINTER(TuneReport),
                                           C.POINTER(_vp)]),
    "mlp_hier_build_uh_cheb_mat": (C.c_int, [_vp, _vp, _f64p, C.c_uint32, _u32p, C.c_uint32, C.c_int,
                                            C.POINTER(_vp)]),
    "mlp_spe_assemble": (C.c_int, [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, _vp, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_double, _vp, _vp, _vp]),
}

_lib = None


def lib():
    """Load libmlp.so; raise if it is not built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -c 'import __graft_entry__ as g; g.build()'` "
                              "or `python ml_prec_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise MlpError(rc, lib().mlp_last_error().decode())


def _u32(x):
    return np.ascontiguousarray(x, dtype=np.uint32)


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


# ---------------------------------------------------------------------------
# host hierarchy (construction runs on the host, no GPU needed)
# ---------------------------------------------------------------------------
class Hierarchy:
    ARRAYS = {"map": (0, np.uint32), "L_ia": (1, np.uint32), "L_ja": (2, np.uint32), "L_a": (3, np.float64),
              "U_ia": (4, np.uint32), "U_ja": (5, np.uint32), "U_a": (6, np.float64), "dval": (7, np.float64),
              "A_ia": (8, np.uint32), "A_ja": (9, np.uint32), "A_a": (10, np.float64), "rmap": (11, np.uint32),
              "aux": (12, np.float64)}

    def __init__(self, handle, kind, n):
        self.h = handle
        self.kind = kind
        self.n = n

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.mlp_hier_destroy(self.h)
            self.h = None

    @property
    def nlevels(self):
        v = C.c_uint32()
        check(lib().mlp_hier_nlevels(self.h, C.byref(v)))
        return v.value

    def level_info(self, l):
        li = LevelInfo()
        check(lib().mlp_hier_level_info(self.h, l, C.byref(li)))
        return {k: getattr(li, k) for k, _ in LevelInfo._fields_}

    def level_arrays(self, l):
        info = self.level_info(l)
        N, M, Md = info["N"], info["M"], info["Md"]
        sizes = {"map": N, "rmap": N, "aux": N, "L_ia": N - Md + 1, "L_ja": info["nnz_L"], "L_a": info["nnz_L"],
                 "U_ia": M + 1, "U_ja": info["nnz_U"], "U_a": info["nnz_U"], "dval": Md,
                 "A_ia": info["n_next"] + 1 if info["n_next"] else 0, "A_ja": info["nnz_next"],
                 "A_a": info["nnz_next"]}
        out = {}
        for name, (which, dt) in self.ARRAYS.items():
            arr = np.zeros(max(sizes[name], 1), dtype=dt)
            check(lib().mlp_hier_level_array(self.h, l, which, arr.ctypes.data_as(_vp)))
            out[name] = arr[:sizes[name]]
        return info, out


def build_uh_cheb(ia, ja, a, sigmas=(3.0,), niters=(2,), use_tails=True, ctx=None):
    """Prec(A, cfg) construction (defaults: sigma 3, ncheb 2, tails on).  With a Context, every level is built
    on its device (mlp_hier_build_uh_cheb_dev: link dropping, tail order, elimination); the hierarchy is the
    same bit for bit."""
    h = _vp()
    n = len(ia) - 1
    args = (n, _u32(ia), _u32(ja), _f64(a), _f64(sigmas), len(sigmas), _u32(niters), len(niters), int(use_tails),
            C.byref(h))
    if ctx is None:
        check(lib().mlp_hier_build_uh_cheb(*args))
    else:
        check(lib().mlp_hier_build_uh_cheb_dev(ctx.h, *args))
    return Hierarchy(h, UH_CHEB, n)


def build_uh_cheb_on_device(ctx, A: "DevMatrix", sigmas=(3.0,), niters=(2,), use_tails=True):
    """Prec(A, cfg) built on the device from the matrix handle (mlp_hier_build_uh_cheb_mat): level 0 read in
    device memory when A was assembled there."""
    h = _vp()
    check(lib().mlp_hier_build_uh_cheb_mat(ctx.h, A.h, _f64(sigmas), len(sigmas), _u32(niters), len(niters),
                                           int(use_tails), C.byref(h)))
    return Hierarchy(h, UH_CHEB, A.n)


def build_multi_split(ia, ja, a, config_path=None, q=(0.8,), niters=(5,), eps=(1e-2,), inner="fixed",
                      order="original", use_tails=True, degree_max=3, max_levels=15, coarse_n=0):
    """MultiSplitPrec(A, cfg) construction; config_path reads a prec config file."""
    qa, na, ea = _f64(q), _u32(niters), _f64(eps)
    p = MsplitParams(qa.ctypes.data_as(C.POINTER(C.c_double)), len(qa),
                     na.ctypes.data_as(C.POINTER(C.c_uint32)), len(na),
                     ea.ctypes.data_as(C.POINTER(C.c_double)), len(ea),
                     int(inner == "dynamic"), ORDERS[order], int(use_tails), degree_max, max_levels, coarse_n)
    h = _vp()
    n = len(ia) - 1
    path = config_path.encode() if config_path else None
    check(lib().mlp_hier_build_multi_split(n, _u32(ia), _u32(ja), _f64(a), path, C.byref(p), C.byref(h)))
    return Hierarchy(h, MULTI_SPLIT, n)


# ---------------------------------------------------------------------------
# device objects
# ---------------------------------------------------------------------------
class Context:
    def __init__(self, device=0):
        self.h = _vp()
        check(lib().mlp_ctx_create(device, C.byref(self.h)))

    def close(self):
        """Destroy the context; objects created on it must be closed first (their
        device memory is released with the context otherwise)."""
        if getattr(self, "h", None):
            lib().mlp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        if _lib is not None:
            self.close()

    def sync(self):
        check(lib().mlp_ctx_sync(self.h))

    def array(self, n, dtype=np.float64):
        return DevArray(self, n, dtype)

    def to_device(self, x):
        x = np.ascontiguousarray(x)
        d = DevArray(self, x.shape[0], x.dtype)
        d.upload(x)
        return d


class DevArray:
    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        self.ptr = _vp()
        check(lib().mlp_dev_alloc(ctx.h, max(self.nbytes, 8), C.byref(self.ptr)))

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None and self.ctx.h:
            _lib.mlp_dev_free(self.ctx.h, self.ptr)
            self.ptr = None

    def upload(self, x):
        x = np.ascontiguousarray(x, dtype=self.dtype)
        assert x.shape[0] == self.n
        check(lib().mlp_copy_h2d(self.ctx.h, self.ptr, x.ctypes.data_as(_vp), self.nbytes))

    def get(self):
        out = np.empty(self.n, dtype=self.dtype)
        check(lib().mlp_copy_d2h(self.ctx.h, out.ctypes.data_as(_vp), self.ptr, self.nbytes))
        return out


class DevMatrix:
    def __init__(self, ctx, ia, ja, a):
        self.ctx = ctx
        self.n = len(ia) - 1
        self.h = _vp()
        check(lib().mlp_csr_upload(ctx.h, self.n, _u32(ia), _u32(ja), _f64(a), C.byref(self.h)))

    @classmethod
    def spe_on_device(cls, ctx, kx: "DevArray", shape, kind="spe", c=0.0, kz_ratio=0.1, h=(20.0, 10.0, 2.0)):
        """The seven-point matrix of a (nz, ny, nx) field assembled on the device from device-resident kx
        (mlp_spe_matrix_dev); the same matrix as spe_assemble's."""
        nz, ny, nx = (int(v) for v in shape)
        if kx.n != nx * ny * nz:
            raise ValueError("kx has the wrong size")
        k = {"spe": 0, "tpfa": 1}.get(kind)
        if k is None:
            raise ValueError(f"unknown kind {kind!r}")
        m = cls.__new__(cls)
        m.ctx = ctx
        m.n = nx * ny * nz
        m.h = _vp()
        check(lib().mlp_spe_matrix_dev(ctx.h, k, nx, ny, nz, kx.ptr, kz_ratio, c, h[0], h[1], h[2], C.byref(m.h)))
        return m

    @classmethod
    def spe_ensemble_on_device(cls, ctx, kx: "DevArray", nx, ny, nzs, kind="spe", c=0.0, kz_ratio=0.1,
                               h=(20.0, 10.0, 2.0)):
        """Systems of nx*ny*nz cells (nz from nzs) stacked along z into one block-diagonal matrix assembled on
        the device (mlp_spe_ensemble_matrix_dev); kx holds their fields one after another."""
        nzs = np.ascontiguousarray(nzs, dtype=np.uint32)
        if kx.n != nx * ny * int(nzs.sum()):
            raise ValueError("kx has the wrong size")
        k = {"spe": 0, "tpfa": 1}.get(kind)
        if k is None:
            raise ValueError(f"unknown kind {kind!r}")
        m = cls.__new__(cls)
        m.ctx = ctx
        m.n = kx.n
        m.h = _vp()
        check(lib().mlp_spe_ensemble_matrix_dev(ctx.h, k, nx, ny, len(nzs), nzs, kx.ptr, kz_ratio, c, h[0], h[1],
                                                h[2], C.byref(m.h)))
        return m

    def csr(self):
        """The matrix's Skyline CSR as host arrays (ia, ja, a)."""
        n, nnz, _ = self.info()
        ia = np.empty(n + 1, dtype=np.uint32)
        ja = np.empty(max(nnz, 1), dtype=np.uint32)
        a = np.empty(max(nnz, 1), dtype=np.float64)
        check(lib().mlp_mat_get_csr(self.h, ia.ctypes.data_as(_vp), ja.ctypes.data_as(_vp), a.ctypes.data_as(_vp)))
        return ia, ja[:nnz], a[:nnz]

    def close(self):
        # a matrix outlives nothing: destroy it only while its context is alive
        if getattr(self, "h", None) and _lib is not None and self.ctx.h:
            _lib.mlp_mat_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def spmv(self, x: DevArray, y: DevArray):
        check(lib().mlp_spmv(self.ctx.h, self.h, x.ptr, y.ptr))

    def residual(self, b: DevArray, x: DevArray, r: DevArray):
        check(lib().mlp_residual(self.ctx.h, self.h, b.ptr, x.ptr, r.ptr))

    def time_spmv(self, x: DevArray, y: DevArray, reps=20, flush_bytes=0):
        ms = C.c_double()
        check(lib().mlp_time_spmv(self.ctx.h, self.h, x.ptr, y.ptr, reps, flush_bytes, C.byref(ms)))
        return ms.value

    def info(self):
        n, nnz, by = C.c_uint32(), C.c_uint32(), C.c_uint64()
        check(lib().mlp_mat_info(self.h, C.byref(n), C.byref(nnz), C.byref(by)))
        return n.value, nnz.value, by.value


class DevPrec:
    def __init__(self, ctx, A: DevMatrix, hier: Hierarchy | None = None, kind=None, precision=MLP_FP64):
        self.ctx = ctx
        self.A = A
        self.h = _vp()
        self.hier = hier
        if kind == COMPOSITE:                      # GSPrec(A) then the multi_split hierarchy `hier` of A
            check(lib().mlp_prec_composite(ctx.h, A.h, hier.h, C.byref(self.h)))
            self.kind = kind
        elif hier is not None:
            check(lib().mlp_prec_from_hier(ctx.h, hier.h, A.h, precision, C.byref(self.h)))
            self.kind = hier.kind
        elif kind == GS:
            check(lib().mlp_prec_gs(ctx.h, A.h, C.byref(self.h)))
            self.kind = kind
        elif kind == SYM:                          # SymPrec with the default uh_cheb parameters
            check(lib().mlp_prec_sym(ctx.h, A.h, _f64(np.array([3.0])), 1, _u32(np.array([2])), 1, 1,
                                     C.byref(self.h)))
            self.kind = kind
        else:
            check(lib().mlp_prec_simple(ctx.h, kind, A.h, C.byref(self.h)))
            self.kind = kind
        self.precision = precision

    @classmethod
    def tuned(cls, ctx, A: DevMatrix, hier: Hierarchy, precision=MLP_FP64, reps=10):
        """The hierarchy with the fastest of the candidate launch schedules (mlp_prec_from_hier_tuned); the
        report is kept in .tune: [(schedule bits, us per apply)] and the index chosen."""
        p = cls.__new__(cls)
        p.ctx, p.A, p.hier, p.kind, p.precision = ctx, A, hier, hier.kind, precision
        p.h = _vp()
        rep = TuneReport()
        check(lib().mlp_prec_from_hier_tuned(ctx.h, hier.h, A.h, precision, reps, C.byref(rep), C.byref(p.h)))
        p.tune = {"candidates": [(int(rep.sched[k]), float(rep.us[k])) for k in range(min(rep.ncand, 8))],
                  "chosen": int(rep.chosen)}
        return p

    def close(self):
        if getattr(self, "h", None) and _lib is not None and self.ctx.h:
            _lib.mlp_prec_destroy(self.h)
        self.h = None

    @classmethod
    def upload(cls, ctx, A: DevMatrix, kind, levels, precision=MLP_FP64):
        """mlp_prec_upload: a uh_cheb or multi_split hierarchy given as per-level dicts with the keys of
        Hierarchy.level_arrays (map, L_ia, L_ja, L_a, U_ia, U_ja, U_a, dval; A_ia, A_ja, A_a = this level's
        matrix, unused for level 0) and the scalars N, M, Md, nit, lmin, lmax, eps."""
        keep = []
        descs = (LevelDesc * len(levels))()
        for d, L in zip(descs, levels):
            d.N, d.M, d.Md, d.nit = int(L["N"]), int(L["M"]), int(L["Md"]), int(L["nit"])
            d.lmin, d.lmax, d.eps = float(L.get("lmin", 0.0)), float(L.get("lmax", 0.0)), float(L.get("eps", 0.0))
            for name, dt in (("map", np.uint32), ("L_ia", np.uint32), ("L_ja", np.uint32), ("L_a", np.float64),
                             ("U_ia", np.uint32), ("U_ja", np.uint32), ("U_a", np.float64), ("dval", np.float64),
                             ("A_ia", np.uint32), ("A_ja", np.uint32), ("A_a", np.float64)):
                v = L.get(name)
                if v is None:
                    continue
                arr = np.ascontiguousarray(v, dtype=dt)
                if arr.size == 0:
                    arr = np.zeros(1, dtype=dt)
                keep.append(arr)
                setattr(d, name, arr.ctypes.data)
        p = cls.__new__(cls)
        p.ctx, p.A, p.hier, p.kind, p.precision, p.h = ctx, A, None, kind, precision, _vp()
        check(lib().mlp_prec_upload(ctx.h, kind, len(levels), descs, A.h, precision, C.byref(p.h)))
        return p

    def clone(self, ctx):
        """A copy on `ctx` sharing this hierarchy's device data, with its own level vectors."""
        c = DevPrec.__new__(DevPrec)
        c.ctx, c.A, c.hier, c.kind, c.precision, c.h = ctx, self.A, self.hier, self.kind, self.precision, _vp()
        check(lib().mlp_prec_clone(ctx.h, self.h, C.byref(c.h)))
        return c

    def __del__(self):
        self.close()

    def apply(self, f: DevArray, x: DevArray):
        check(lib().mlp_prec_apply(self.ctx.h, self.h, f.ptr, x.ptr))

    def time_apply(self, f: DevArray, x: DevArray, reps=10):
        ms = C.c_double()
        check(lib().mlp_time_prec_apply(self.ctx.h, self.h, f.ptr, x.ptr, reps, C.byref(ms)))
        return ms.value

    def bytes_per_apply(self):
        b = C.c_double()
        n = C.c_uint32()
        check(lib().mlp_prec_bytes(self.h, C.byref(b), C.byref(n)))
        return b.value

    def launches_per_apply(self):
        """Kernels the last apply launched (0 before the first apply)."""
        b = C.c_double()
        n = C.c_uint32()
        check(lib().mlp_prec_bytes(self.h, C.byref(b), C.byref(n)))
        return n.value

    def schedule(self):
        """The launch schedule built: SCHED_* bits | forced level block size << 8."""
        v = C.c_uint32()
        check(lib().mlp_prec_schedule(self.h, C.byref(v)))
        return v.value


def _opts(eps, maxit, x0_mode, flexible, true_res_every, check_every, use_graph):
    return SolveOpts(eps, maxit, x0_mode, int(flexible), true_res_every, check_every, int(use_graph))


def pcg(ctx, A, B, b: DevArray, x: DevArray, eps=1e-8, maxit=0, x0_mode=X0_ZERO, flexible=False,
        true_res_every=0, check_every=16, use_graph=False, raise_on_error=True):
    """PCGSolver on the device; returns the stats dict."""
    o = _opts(eps, maxit, x0_mode, flexible, true_res_every, check_every, use_graph)
    st = SolveStats()
    rc = lib().mlp_pcg(ctx.h, A.h, B.h, b.ptr, x.ptr, C.byref(o), C.byref(st))
    d = st.as_dict()
    d["rc"] = rc
    if rc and raise_on_error:
        raise MlpError(rc, lib().mlp_last_error().decode())
    return d


def pcg_noalloc(ctx, A, B, b: DevArray, x: DevArray, eps=1e-8, maxit=0, x0_mode=X0_ZERO, flexible=False,
                true_res_every=0, check_every=16, raise_on_error=True):
    """mlp_pcg_noalloc: PCGSolver with its workspace in the library's static device array."""
    o = _opts(eps, maxit, x0_mode, flexible, true_res_every, check_every, False)
    st = SolveStats()
    rc = lib().mlp_pcg_noalloc(ctx.h, A.h, B.h, b.ptr, x.ptr, C.byref(o), C.byref(st))
    d = st.as_dict()
    d["rc"] = rc
    if rc and raise_on_error:
        raise MlpError(rc, lib().mlp_last_error().decode())
    return d


def pcg_batch(jobs, eps=1e-8, maxit=0, x0_mode=X0_ZERO, flexible=False, true_res_every=0, check_every=16,
              use_graph=False):
    """Concurrent PCG solves.  jobs: sequence of (ctx, A, B, b, x), one context per job.
    Returns (rc, [stats dict with 'status' per job])."""
    o = _opts(eps, maxit, x0_mode, flexible, true_res_every, check_every, use_graph)
    arr = (PcgJob * len(jobs))()
    for i, (ctx, A, B, b, x) in enumerate(jobs):
        arr[i].ctx, arr[i].A, arr[i].B, arr[i].d_b, arr[i].d_x = ctx.h, A.h, B.h, b.ptr, x.ptr
    rc = lib().mlp_pcg_batch(len(jobs), arr, C.byref(o))
    out = []
    for j in arr:
        d = j.stats.as_dict()
        d["rc"] = j.status
        out.append(d)
    return rc, out


def pcg_multi(ctx, A, B, nrhs, b: DevArray, x: DevArray, eps=1e-8, maxit=0, x0_mode=X0_ZERO, flexible=False,
              true_res_every=0, check_every=16, use_graph=True):
    """nrhs PCG solves of one system at once (mlp_pcg_multi); b and x hold nrhs vectors of length n one after
    another.  Returns (rc, [stats dict per right-hand side])."""
    o = _opts(eps, maxit, x0_mode, flexible, true_res_every, check_every, use_graph)
    arr = (SolveStats * nrhs)()
    rc = lib().mlp_pcg_multi(ctx.h, A.h, B.h, nrhs, b.ptr, x.ptr, C.byref(o), arr)
    return rc, [st.as_dict() for st in arr]


def pcg_ensemble(ctx, A, B, seg_off, b: DevArray, x: DevArray, eps=1e-8, maxit=0, x0_mode=X0_ZERO, flexible=False,
                 true_res_every=0, check_every=16, use_graph=True):
    """mlp_pcg_ensemble: every segment [seg_off[y], seg_off[y+1]) of the block-diagonal system solved as its own
    PCG, all in the same launches.  Returns (rc, [stats per segment])."""
    off = np.ascontiguousarray(seg_off, dtype=np.uint32)
    S = len(off) - 1
    st = (SolveStats * S)()
    o = _opts(eps, maxit, x0_mode, flexible, true_res_every, check_every, use_graph)
    rc = lib().mlp_pcg_ensemble(ctx.h, A.h, B.h, S, off, b.ptr, x.ptr, C.byref(o), st)
    return rc, [s.as_dict() for s in st]


def pcg_ensemble_batch(jobs, eps=1e-8, maxit=0, x0_mode=X0_ZERO, flexible=False, true_res_every=0, check_every=16,
                       use_graph=True):
    """mlp_pcg_ensemble_batch over jobs = [(ctx, A, B, seg_off, b, x)], each context used once.  Returns
    (rc, [(status, [stats per segment]) per job])."""
    n = len(jobs)
    arr = (EnsJob * n)()
    keep = []
    for k, (ctx, A, B, off, b, x) in enumerate(jobs):
        off = np.ascontiguousarray(off, dtype=np.uint32)
        st = (SolveStats * (len(off) - 1))()
        keep.append((off, st))
        arr[k] = EnsJob(ctx.h, A.h, B.h, len(off) - 1, off.ctypes.data_as(_vp), b.ptr, x.ptr,
                        C.cast(st, C.POINTER(SolveStats)), 0)
    o = _opts(eps, maxit, x0_mode, flexible, true_res_every, check_every, use_graph)
    rc = lib().mlp_pcg_ensemble_batch(n, C.cast(arr, _vp), C.byref(o))
    return rc, [(arr[k].status, [s.as_dict() for s in keep[k][1]]) for k in range(n)]


def gmres(ctx, A, B, b: DevArray, x: DevArray, eps=1e-8, maxit=0, x0_mode=X0_ZERO, raise_on_error=True):
    """GMRESSolver (restart 30) on the device; returns the stats dict."""
    o = _opts(eps, maxit, x0_mode, False, 0, 0, False)
    st = SolveStats()
    rc = lib().mlp_gmres(ctx.h, A.h, B.h, b.ptr, x.ptr, C.byref(o), C.byref(st))
    d = st.as_dict()
    d["rc"] = rc
    if rc and raise_on_error:
        raise MlpError(rc, lib().mlp_last_error().decode())
    return d


def cheb_solver(ctx, A, B, lmin, lmax, b: DevArray, x: DevArray, eps=1e-8, maxit=0, x0_mode=X0_ZERO,
                check_every=0, use_graph=True, raise_on_error=True):
    """ChebSolver on the device over [lmin, lmax]; returns the stats dict."""
    o = _opts(eps, maxit, x0_mode, False, 0, check_every, use_graph)
    st = SolveStats()
    rc = lib().mlp_cheb_solver(ctx.h, A.h, B.h, float(lmin), float(lmax), b.ptr, x.ptr, C.byref(o), C.byref(st))
    d = st.as_dict()
    d["rc"] = rc
    if rc and raise_on_error:
        raise MlpError(rc, lib().mlp_last_error().decode())
    return d


def simple_solver(ctx, A, B, b: DevArray, x: DevArray, eps=1e-8, maxit=0, x0_mode=X0_ZERO, check_every=0,
                  use_graph=True, raise_on_error=True):
    """SimpleSolver (x += B(b - A x)) on the device; returns the stats dict."""
    o = _opts(eps, maxit, x0_mode, False, 0, check_every, use_graph)
    st = SolveStats()
    rc = lib().mlp_simple_solver(ctx.h, A.h, B.h, b.ptr, x.ptr, C.byref(o), C.byref(st))
    d = st.as_dict()
    d["rc"] = rc
    if rc and raise_on_error:
        raise MlpError(rc, lib().mlp_last_error().decode())
    return d


def spe_adjoint_grad(ctx, shape, face_factors, kz_ratio, kx: DevArray, p: DevArray, lam: DevArray,
                     grad: DevArray):
    """grad = -lam^T dA/d(ln kx) p for the two-point-flux assembly; shape = (nz, ny, nx)."""
    nz, ny, nx = shape
    fx, fy, fz = face_factors
    check(lib().mlp_spe_adjoint_grad(ctx.h, nx, ny, nz, fx, fy, fz, kz_ratio, kx.ptr, p.ptr, lam.ptr, grad.ptr))


def ddot(ctx, x: DevArray, y: DevArray):
    v = C.c_double()
    check(lib().mlp_ddot(ctx.h, x.n, x.ptr, y.ptr, C.byref(v)))
    return v.value


def dnrm2(ctx, x: DevArray):
    v = C.c_double()
    check(lib().mlp_dnrm2(ctx.h, x.n, x.ptr, C.byref(v)))
    return v.value


def daxpy(ctx, alpha, x: DevArray, y: DevArray):
    check(lib().mlp_daxpy(ctx.h, x.n, alpha, x.ptr, y.ptr))


def dscal(ctx, alpha, x: DevArray):
    check(lib().mlp_dscal(ctx.h, x.n, alpha, x.ptr))


def spe_assemble(kx3, kind="spe", c=0.0, kz_ratio=0.1, h=(20.0, 10.0, 2.0)):
    """Seven-point matrix of a (nz, ny, nx) kx field on the host (mlp_spe_assemble): returns ia, ja, a."""
    kx3 = np.ascontiguousarray(kx3, dtype=np.float64)
    if kx3.ndim != 3:
        raise ValueError("kx must be a (nz, ny, nx) array")
    nz, ny, nx = kx3.shape
    k = {"spe": 0, "tpfa": 1}.get(kind)
    if k is None:
        raise ValueError(f"unknown kind {kind!r}")
    nnz = int(lib().mlp_spe_nnz(nx, ny, nz))
    ia = np.empty(nx * ny * nz + 1, dtype=np.uint32)
    ja = np.empty(nnz, dtype=np.uint32)
    a = np.empty(nnz, dtype=np.float64)
    check(lib().mlp_spe_assemble(k, nx, ny, nz, kx3.ctypes.data, kz_ratio, c, h[0], h[1], h[2],
                                 ia.ctypes.data, ja.ctypes.data, a.ctypes.data))
    return ia, ja, a

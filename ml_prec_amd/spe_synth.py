"""Seeded synthetic SPE10-shaped pressure systems.

The reference builds its 7-point matrix from the SPE10 permeability file
(``modules/mesh/spe.cc:11-67``), which is a network download and not available
here.  This module generates permeability fields of the same shape from a seed
and assembles the two-point-flux matrix in the reference's storage convention:

* cells are numbered x fastest, then y, then z (``modules/mesh/mesh.h:253-263``,
  the ``XYZ`` layout);
* each row stores the diagonal first and then the off-diagonals in ascending
  column order z-, y-, x-, x+, y+, z+ (``modules/mesh/spe.cc:101-125``), i.e.
  the Skyline layout ``SkylineMatrix`` expects (``modules/matrix/matrix.h:269``);
* the diagonal is accumulated neighbour by neighbour in that same order, as the
  ``ADD`` macro does (``modules/mesh/spe.cc:81-87``).

Two transmissibility conventions are provided:

``kind="tpfa"``
    Face transmissibility = harmonic mean of the two cells' permeability times
    face area over cell spacing, with Dirichlet p=0 on the x=0 and x=max faces
    through a half-cell transmissibility added to the diagonal.  An optional
    accumulation term ``c`` is added to every diagonal.
``kind="spe"``
    The reference's own formula ``a_ij = -2/(1/k_i+1/k_j)/h^2`` with the
    reaction coefficient ``c`` on every diagonal (``modules/mesh/spe.cc:69-131``,
    default ``c = 1``, ``project/main_helper.cc:65``).

Every function is deterministic for a given seed (numpy PCG64 streams).
"""
from __future__ import annotations

import dataclasses
import hashlib

import numpy as np
import scipy.ndimage as ndi

HX, HY, HZ = 20.0, 10.0, 2.0  # cell sizes in ft (modules/mesh/spe.cc:14)
BASE_SEED = 20261203


@dataclasses.dataclass
class System:
    """A Skyline-CSR matrix plus right-hand side and the fields behind it."""

    nx: int
    ny: int
    nz: int
    ia: np.ndarray  # uint32[N+1]
    ja: np.ndarray  # uint32[nnz]
    a: np.ndarray   # float64[nnz]
    b: np.ndarray   # float64[N]
    kx: np.ndarray  # float64[N]  (ky = kx, kz = 0.1 kx)
    kind: str
    c: float

    @property
    def n(self) -> int:
        return int(self.ia.shape[0] - 1)

    @property
    def nnz(self) -> int:
        return int(self.ja.shape[0])

    def digest(self) -> str:
        h = hashlib.sha256()
        for arr in (self.ia, self.ja, self.a, self.b):
            h.update(np.ascontiguousarray(arr).tobytes())
        return h.hexdigest()


def _gaussian_field(rng: np.random.Generator, shape_zyx, corr_zyx) -> np.ndarray:
    """White noise convolved with a separable Gaussian, renormalised to N(0,1)."""
    g = rng.standard_normal(shape_zyx)
    f = ndi.gaussian_filter(g, sigma=corr_zyx, mode="wrap")
    f -= f.mean()
    f /= f.std()
    return f


def lognormal_field(nx, ny, nz, seed, sigma=0.5):
    """ln k = sigma * g with g correlated over 8x8x2 cells (x, y, z)."""
    rng = np.random.default_rng(seed)
    g = _gaussian_field(rng, (nz, ny, nx), (2.0, 8.0, 8.0))
    return np.exp(sigma * g)


def spe10_field(nx, ny, nz, seed, n_top=35):
    """Top ``n_top`` layers log-normal (sigma=2), the rest channelised.

    Channel indicator: a Gaussian field correlated over 20x2x1 cells (x, y, z)
    thresholded at its 70th percentile; channel ln k ~ N(ln 1000, 0.5),
    background ln k ~ N(ln 0.01, 1.0).
    """
    rng = np.random.default_rng(seed)
    k = np.empty((nz, ny, nx))
    ntop = min(n_top, nz)
    if ntop:
        g = _gaussian_field(rng, (ntop, ny, nx), (2.0, 8.0, 8.0))
        k[:ntop] = np.exp(2.0 * g)
    nbot = nz - ntop
    if nbot:
        ind = _gaussian_field(rng, (nbot, ny, nx), (1.0, 2.0, 20.0))
        chan = ind > np.percentile(ind, 70.0)
        lk = np.where(
            chan,
            np.log(1000.0) + 0.5 * rng.standard_normal(ind.shape),
            np.log(0.01) + 1.0 * rng.standard_normal(ind.shape),
        )
        k[ntop:] = np.exp(lk)
    return k


def _hm(k1, k2):
    # written as in the reference's ADD macro so that T(i,j) == T(j,i) bitwise
    return 2.0 / (1.0 / k1 + 1.0 / k2)


def assemble(kx3, kind="tpfa", c=0.0):
    """Assemble the Skyline-CSR 7-point matrix from a (nz,ny,nx) kx field."""
    nz, ny, nx = kx3.shape
    n = nx * ny * nz
    kx3 = np.asarray(kx3, dtype=np.float64)
    kz3 = 0.1 * kx3
    if kind == "tpfa":
        fx, fy, fz = HY * HZ / HX, HX * HZ / HY, HX * HY / HZ

        def scale(t, f, h):
            return t * f
    elif kind == "spe":
        # spe.cc:80-86 (the ADD macro): the harmonic mean divided by h*h
        fx, fy, fz = None, None, None

        def scale(t, f, h):
            return t / (h * h)
    else:
        raise ValueError(f"unknown kind {kind!r}")

    # face transmissibilities, indexed by the lower cell of the face
    tx = np.zeros((nz, ny, nx))
    ty = np.zeros((nz, ny, nx))
    tz = np.zeros((nz, ny, nx))
    tx[:, :, :-1] = scale(_hm(kx3[:, :, :-1], kx3[:, :, 1:]), fx, HX)
    ty[:, :-1, :] = scale(_hm(kx3[:, :-1, :], kx3[:, 1:, :]), fy, HY)
    tz[:-1, :, :] = scale(_hm(kz3[:-1], kz3[1:]), fz, HZ)

    idx = np.arange(n, dtype=np.int64).reshape(nz, ny, nx)
    # neighbour (column, value, present) in the reference's ascending order
    zero = np.zeros((nz, ny, nx))
    cols, vals, pres = [], [], []

    def nb(shift_axis, delta, t):
        col = np.full((nz, ny, nx), -1, dtype=np.int64)
        val = zero.copy()
        p = np.zeros((nz, ny, nx), dtype=bool)
        sl_dst = [slice(None)] * 3
        sl_src = [slice(None)] * 3
        if delta < 0:
            sl_dst[shift_axis] = slice(1, None)
            sl_src[shift_axis] = slice(None, -1)
            tv = t[tuple(sl_src)]
        else:
            sl_dst[shift_axis] = slice(None, -1)
            sl_src[shift_axis] = slice(1, None)
            tv = t[tuple(sl_dst)]
        col[tuple(sl_dst)] = idx[tuple(sl_src)]
        val[tuple(sl_dst)] = tv
        p[tuple(sl_dst)] = True
        cols.append(col.ravel())
        vals.append(val.ravel())
        pres.append(p.ravel())

    nb(0, -1, tz)  # z-
    nb(1, -1, ty)  # y-
    nb(2, -1, tx)  # x-
    nb(2, +1, tx)  # x+
    nb(1, +1, ty)  # y+
    nb(0, +1, tz)  # z+

    diag = np.full(n, float(c))
    if kind == "tpfa":
        # Dirichlet p=0 on x=0 and x=max: half-cell transmissibility k*(dy*dz)/(dx/2)
        tb = np.zeros((nz, ny, nx))
        tb[:, :, 0] += kx3[:, :, 0] * (HY * HZ) / (HX / 2)
        if nx > 1:
            tb[:, :, -1] += kx3[:, :, -1] * (HY * HZ) / (HX / 2)
        else:
            tb[:, :, 0] *= 2.0
        diag = diag + tb.ravel()
    for v, p in zip(vals, pres):
        diag = diag + np.where(p, v, 0.0)

    C = np.stack([np.arange(n, dtype=np.int64)] + cols, axis=1)       # (n,7)
    V = np.stack([diag] + [-v for v in vals], axis=1)
    P = np.stack([np.ones(n, dtype=bool)] + pres, axis=1)
    counts = P.sum(axis=1)
    ia = np.zeros(n + 1, dtype=np.uint32)
    np.cumsum(counts, out=ia[1:])
    ja = C[P].astype(np.uint32)
    a = V[P].astype(np.float64)
    return ia, ja, a


def make_nonsymmetric(ia, ja, a, tau, seed=1):
    """The reference's nonsymmetric M-matrix (modules/mesh/spe.cc:138-182): row by row, every off-diagonal
    a_ij is changed by d = random(-tau, tau) * a_ij (a_ij -= d, a_ii += d; include/tools.h:41-44 draws
    -tau + 2 tau random()/RAND_MAX), so row sums are kept.  The draws come from glibc random() seeded with
    `seed` (1 is glibc's state at program start, the reference's case).  Returns a new value array."""
    import ctypes
    import ctypes.util
    if not tau:
        raise ValueError("Nonsymmetricity parameter is 0")
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.random.restype = ctypes.c_long
    libc.srandom.argtypes = [ctypes.c_uint]
    rand_max = 2147483647
    libc.srandom(seed)
    ia = np.asarray(ia, dtype=np.int64)
    a = np.array(a, dtype=np.float64)
    n = ia.shape[0] - 1
    for i in range(n):
        dind = ia[i]
        for j in range(ia[i] + 1, ia[i + 1]):
            r = -tau + (tau - (-tau)) * (float(libc.random()) / rand_max)
            d = r * a[j]
            a[j] -= d
            a[dind] += d
    return a


def rhs(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, n)


def make_system(nx, ny, nz, seed=BASE_SEED, field="lognormal", kind="tpfa", c=0.0, sigma=0.5):
    """Generate permeability, assemble A and draw b ~ U[-1,1] from ``seed``."""
    if field == "lognormal":
        k = lognormal_field(nx, ny, nz, seed, sigma=sigma)
    elif field == "spe10":
        k = spe10_field(nx, ny, nz, seed)
    else:
        raise ValueError(field)
    ia, ja, a = assemble(k, kind=kind, c=c)
    b = rhs(nx * ny * nz, seed + 7919)
    return System(nx, ny, nz, ia, ja, a, b, k.ravel(), kind, c)


def lognormal_system(seed: int | None = None, kind="tpfa", c=0.0, nz: int = 5):
    """The small log-normal workload: 60x220x5 cells (nz overridable), ln k with sigma 0.5, seed BASE_SEED."""
    return make_system(60, 220, nz, BASE_SEED if seed is None else seed, "lognormal", kind, c, 0.5)


def spe10_system(seed: int | None = None, kind="tpfa", c=0.0, nz: int | None = None):
    """The SPE10-shaped workload: 60x220x85 cells (nz overridable), seed BASE_SEED + 1."""
    return make_system(60, 220, 85 if nz is None else nz, BASE_SEED + 1 if seed is None else seed, "spe10", kind, c)


ENSEMBLE_NZ = (5, 10, 20, 40, 85)


def ensemble_spec(count=64, base_seed=BASE_SEED + 3):
    """The ensemble: `count` systems, system i with its own field seed base_seed + 1000 + i and a layer count
    drawn uniformly from ENSEMBLE_NZ by a generator seeded with base_seed.  Returns [(nz, seed)]."""
    g = np.random.default_rng(base_seed)
    nzs = g.choice(np.array(ENSEMBLE_NZ), size=count)
    return [(int(nz), base_seed + 1000 + i) for i, nz in enumerate(nzs)]


def ensemble_system(nz, seed, kind="spe", c=1.0):
    """One member of the config-3 ensemble: the SPE10-shaped 60x220xnz field of `seed`."""
    return make_system(60, 220, nz, seed, "spe10", kind, c)


def write_csr_binary(path, ia, ja, a, ncol=None):
    """CSR BINARY layout: uint nrow, uint ncol, ia[n+1], ja[nnz], double a[nnz]
    (``modules/matrix/csrmatrix.cc:314-320``)."""
    n = ia.shape[0] - 1
    with open(path, "wb") as f:
        np.array([n, n if ncol is None else ncol], dtype=np.uint32).tofile(f)
        np.asarray(ia, dtype=np.uint32).tofile(f)
        np.asarray(ja, dtype=np.uint32).tofile(f)
        np.asarray(a, dtype=np.float64).tofile(f)


def read_csr_binary(path):
    """Inverse of :func:`write_csr_binary` (``modules/matrix/csrmatrix.cc:174-190``)."""
    with open(path, "rb") as f:
        nrow, ncol = np.fromfile(f, dtype=np.uint32, count=2)
        ia = np.fromfile(f, dtype=np.uint32, count=int(nrow) + 1)
        nnz = int(ia[-1])
        ja = np.fromfile(f, dtype=np.uint32, count=nnz)
        a = np.fromfile(f, dtype=np.float64, count=nnz)
    if ia.shape[0] != nrow + 1 or ja.shape[0] != nnz or a.shape[0] != nnz:
        raise IOError(f"truncated CSR BINARY file {path}")
    return int(nrow), int(ncol), ia, ja, a

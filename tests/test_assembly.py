"""Native seven-point assembly (mlp_spe_assemble) against the numpy generator and
a loop-by-loop restatement of SPEMesh::construct_matrix (modules/mesh/spe.cc:69-131).

Bit-exact: the same rows, columns and values as the numpy generator for both
kinds, on grids with every boundary case (single cells, lines, slabs) and the
60x220x85 SPE10 grid; on a small grid, the same values as a direct transcription
of the reference's ADD macro; the matrix is symmetric and its row sums are c up
to rounding (spe kind).
"""
import numpy as np
import pytest

from ml_prec_amd import native as N
from ml_prec_amd import spe_synth as S


def _field(shape, seed):
    return np.exp(np.random.default_rng(seed).normal(0.0, 2.0, shape))


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 1, 7), (1, 5, 1), (4, 1, 1), (3, 4, 5), (2, 7, 3)])
@pytest.mark.parametrize("kind,c", [("spe", 1.0), ("spe", 0.0), ("tpfa", 0.0), ("tpfa", 1.0)])
def test_matches_numpy_generator(shape, kind, c):
    k = _field(shape, 5)
    ia0, ja0, a0 = S.assemble(k, kind=kind, c=c)
    ia1, ja1, a1 = N.spe_assemble(k, kind=kind, c=c)
    assert np.array_equal(ia0, ia1) and np.array_equal(ja0, ja1)
    assert np.array_equal(a0.view(np.uint64), a1.view(np.uint64))


def test_full_spe10_grid_matches_numpy():
    k = S.spe10_field(60, 220, 85, S.BASE_SEED + 1)
    ia0, ja0, a0 = S.assemble(k, kind="spe", c=1.0)
    ia1, ja1, a1 = N.spe_assemble(k, kind="spe", c=1.0)
    assert np.array_equal(ia0, ia1) and np.array_equal(ja0, ja1)
    assert np.array_equal(a0.view(np.uint64), a1.view(np.uint64))


def _reference_loops(k3, c, h=(20.0, 10.0, 2.0)):
    """spe.cc:69-131 in the XYZ order, one push_back at a time (kx = ky, kz = 0.1 kx)."""
    nz, ny, nx = k3.shape
    kx = k3.ravel()
    kz = 0.1 * kx
    idx = lambda i, j, kk: i + nx * (j + ny * kk)
    ia, ja, a = [0], [], []
    for kk in range(nz):
        for j in range(ny):
            for i in range(nx):
                i0 = idx(i, j, kk)
                ja.append(i0)
                a.append(c)
                dind = len(a) - 1

                def add(di, dj, dk, kv, hh):
                    i1 = idx(i + di, j + dj, kk + dk)
                    ja.append(i1)
                    v = 2 / (1 / kv[i0] + 1 / kv[i1]) / (hh * hh)
                    a.append(-v)
                    a[dind] += v

                if kk: add(0, 0, -1, kz, h[2])
                if j: add(0, -1, 0, kx, h[1])
                if i: add(-1, 0, 0, kx, h[0])
                if i < nx - 1: add(1, 0, 0, kx, h[0])
                if j < ny - 1: add(0, 1, 0, kx, h[1])
                if kk < nz - 1: add(0, 0, 1, kz, h[2])
                ia.append(len(ja))
    return np.array(ia), np.array(ja), np.array(a)


def test_matches_reference_macro():
    k = _field((3, 4, 5), 9)
    ia0, ja0, a0 = _reference_loops(k, 1.0)
    ia1, ja1, a1 = N.spe_assemble(k, kind="spe", c=1.0)
    assert np.array_equal(ia0, ia1) and np.array_equal(ja0, ja1)
    assert np.array_equal(a0.view(np.uint64), a1.view(np.uint64))


def test_symmetric_and_row_sums():
    import scipy.sparse as sp
    k = _field((4, 6, 5), 13)
    ia, ja, a = N.spe_assemble(k, kind="spe", c=1.0)
    n = ia.size - 1
    A = sp.csr_matrix((a, ja, ia), shape=(n, n))
    assert (A - A.T).count_nonzero() == 0
    assert np.allclose(A @ np.ones(n), 1.0, rtol=0, atol=1e-12 * np.abs(a).max())


def test_bad_arguments():
    with pytest.raises(ValueError):
        N.spe_assemble(np.ones((2, 2)))
    with pytest.raises(ValueError):
        N.spe_assemble(np.ones((2, 2, 2)), kind="nope")
    with pytest.raises(N.MlpError):
        N.spe_assemble(np.ones((0, 2, 2)))

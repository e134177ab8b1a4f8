"""Synthetic SPE10-shaped systems and the CSR BINARY format (csrmatrix.cc:172-190, 302-320)."""
import numpy as np
import scipy.sparse as sp

from ml_prec_amd import spe_synth as S


def test_config_sizes_and_symmetry():
    s0 = S.lognormal_system()
    assert (s0.n, s0.nnz) == (66_000, 432_800)
    s1 = S.spe10_system()
    assert (s1.n, s1.nnz) == (1_122_000, 7_780_000)
    assert s1.kx.max() / s1.kx.min() >= 1e6                       # channel contrast
    for s in (s0, S.lognormal_system(kind="spe", c=1.0)):
        A = sp.csr_matrix((s.a, s.ja, s.ia), shape=(s.n, s.n))
        assert (A != A.T).nnz == 0                               # bitwise symmetric
        assert np.all(s.ja[s.ia[:-1]] == np.arange(s.n))         # Skyline: diagonal first
        for i in range(0, s.n, 997):
            row = s.ja[s.ia[i] + 1:s.ia[i + 1]]
            assert np.all(np.diff(row.astype(np.int64)) > 0)     # then ascending columns


def test_reference_formula_row_sums_equal_c():
    s = S.lognormal_system(kind="spe", c=1.0)
    rs = np.add.reduceat(s.a, s.ia[:-1].astype(np.int64))
    assert np.allclose(rs, 1.0, atol=1e-12)


def test_determinism():
    assert S.lognormal_system().digest() == S.lognormal_system().digest()
    assert S.lognormal_system().digest() != S.lognormal_system(seed=5).digest()


def test_csr_binary_roundtrip(tmp_path):
    s = S.make_system(6, 10, 3, 7, "spe10", "spe", 1.0)
    p = tmp_path / "A.bin"
    S.write_csr_binary(p, s.ia, s.ja, s.a)
    raw = p.read_bytes()
    assert len(raw) == 8 + 4 * (s.n + 1) + 12 * s.nnz
    assert np.frombuffer(raw[:8], np.uint32).tolist() == [s.n, s.n]
    n, m, ia, ja, a = S.read_csr_binary(p)
    assert n == m == s.n
    assert np.array_equal(ia, s.ia) and np.array_equal(ja, s.ja) and np.array_equal(a, s.a)

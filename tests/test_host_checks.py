"""Host-side checks of the construction code that need a C++ driver (no GPU)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def tail_walk_exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("cpp") / "tail_walk_check")
    src = os.path.join(HERE, "cpp", "tail_walk_check.cc")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", "-o", out, src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_tail_walk_rejects_inconsistent_link_counters(tail_walk_exe):
    r = subprocess.run([tail_walk_exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Row 0: could not find remaining link" in r.stdout
    assert r.stdout.startswith("OK 6 ")                    # MLP_ERR_CONSTRUCT


def _fmt_e(v):
    return "%.15e" % v


@pytest.fixture(scope="module")
def io_exe(tmp_path_factory):
    import ml_prec_amd.native as N
    out = str(tmp_path_factory.mktemp("io") / "io_check")
    src = os.path.join(HERE, "cpp", "io_check.cc")
    libdir = os.path.dirname(N.LIB_PATH)
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-o", out, src, "-L" + libdir, "-lmlp", "-Wl,-rpath," + libdir],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_matrix_file_formats(io_exe, tmp_path):
    """BINARY, ASCII, MatrixMarket and HYPRE as csrmatrix.cc:172-355 writes and reads them: the dumps' text,
    exactly, and the round trips through the readers (the MatrixMarket reader skips a comment line after the
    header, which the reference's own dump does not write, so the test inserts one)."""
    import numpy as np
    from conftest import system
    from ml_prec_amd import spe_synth as S
    s = system("tiny")
    a = s.a.copy()
    a[::7] *= 1.0 + 2.0 ** -40                       # values that need all 16 printed digits
    binp = str(tmp_path / "A.bin")
    S.write_csr_binary(binp, s.ia, s.ja, a)
    prefix = str(tmp_path / "A")
    # the MatrixMarket file with the comment line, entries in reverse order and one duplicate (later wins)
    n = s.n
    ent = [(i + 1, int(s.ja[p]) + 1, float(a[p])) for i in range(n) for p in range(s.ia[i], s.ia[i + 1])]
    first = ent[0]
    lines = ["%%MatrixMarket matrix coordinate real general", "% written by the test", f"{n} {n} {len(ent) + 1}"]
    lines += [f"{i} {j} {v!r}" for i, j, v in [(first[0], first[1], 99.0)] + ent[::-1]]
    open(prefix + ".mmc", "w").write("\n".join(lines) + "\n")
    r = subprocess.run([io_exe, binp, prefix], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    res = dict(line.split() for line in r.stdout.splitlines())
    assert res == {"ascii_roundtrip": "1", "binary_roundtrip": "1", "hypre_csr_refused": "1",
                   "hypre_load_refused": "1", "mm_roundtrip": "1"}, res
    nnz = int(s.ia[-1])
    # ASCII: three commented sections, values in scientific notation with 15 decimals (csrmatrix.cc:322-334)
    want = ["# rows cols nonzeros", f"{n} {n} {nnz}", "# row_ptr"] + [str(v) for v in s.ia] + ["# col_ind"]
    want += [str(v) for v in s.ja] + ["# value"] + [_fmt_e(v) for v in a]
    assert open(prefix + ".ascii").read().splitlines() == want
    # MatrixMarket: 1-based, fixed notation with 15 decimals (csrmatrix.cc:346-353)
    want = ["%%MatrixMarket matrix coordinate real general", f"{n} {n} {nnz}"]
    want += ["%d %d %.15f" % (i, j, v) for i, j, v in ent]
    assert open(prefix + ".mm").read().splitlines() == want
    # HYPRE: index ranges, then 0-based "i j a" in storage order, scientific (csrmatrix.cc:335-345)
    want = [f"0 {n - 1} 0 {n - 1}"] + ["%d %d %s" % (i - 1, j - 1, _fmt_e(v)) for i, j, v in ent]
    assert open(prefix + ".hypre").read().splitlines() == want
    assert np.isfinite(a).all()


@pytest.fixture(scope="module")
def chain_exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("chain") / "chain_check")
    src = os.path.join(HERE, "cpp", "chain_check.cc")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-pthread", "-o", out, src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def _write_levels(path, H):
    """The hierarchy's level arrays (with each level's A_next) in the layout chain_check.cc reads."""
    import numpy as np
    with open(path, "wb") as fh:
        np.array([H.nlevels], np.uint32).tofile(fh)
        for l in range(H.nlevels):
            info, arr = H.level_arrays(l)
            np.array([info["N"], info["M"], info["Md"], info["nnz_L"], info["nnz_U"], info["n_next"],
                      info["nnz_next"]], np.uint32).tofile(fh)
            for k in ("map", "L_ia", "L_ja", "L_a", "U_ia", "U_ja", "U_a", "dval", "A_ia", "A_ja", "A_a"):
                arr[k].tofile(fh)


@pytest.mark.parametrize("name", ["spe10_z12", "ln5_tpfa_c1", "spe10_z40"])
def test_chain_launches_equal_recursion(chain_exe, tmp_path, name):
    """The chain program (plan.cc: plan_chain) and the descent/ascent launch sequence, executed on the host slot
    by slot as the device kernels do (node data from the packed words), give Prec::solve's result
    (cheb_solve.cc:13-123, plain recursion over the level arrays) bit for bit, for chains starting at every level,
    in one- and two-wavefront blocks; every packed word agrees with the full arrays."""
    from conftest import system
    from ml_prec_amd import native as N
    s = system(name)
    path = str(tmp_path / "levels.bin")
    _write_levels(path, N.build_uh_cheb(s.ia, s.ja, s.a))
    r = subprocess.run([chain_exe, path], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.rstrip().endswith("OK"), r.stdout + r.stderr
    assert " 0 of " in r.stdout


@pytest.fixture(scope="module")
def steps_exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("steps") / "steps_check")
    src = os.path.join(HERE, "cpp", "steps_check.cc")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-pthread", "-include", "set", "-include",
                        "array", "-o", out, src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_device_construction_steps_on_host(steps_exe, tmp_path):
    """The device construction's steps (cheb_steps.h) run as host loops, every step over its indices in a
    shuffled order: the hierarchy equals the host builder's bit for bit on the workload systems up to
    60x220x85 (with and without tails), on 40 seeded random grids and forest-shaped graphs, and on a
    1000-node path; the same errors where the host builder throws."""
    from conftest import system
    from ml_prec_amd import spe_synth as S
    files = []
    for name in ["tiny", "tiny_tpfa_c1", "small_spe10", "ln5_spe", "ln5_tpfa_c1", "spe10_z12", "spe10_z85"]:
        s = system(name)
        f = str(tmp_path / (name + ".bin"))
        S.write_csr_binary(f, s.ia, s.ja, s.a)
        files.append(f)
    r = subprocess.run([steps_exe, "--random", "40"] + files, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.rstrip().endswith("ALL OK 0")
    assert r.stdout.count("\nOK ") + r.stdout.startswith("OK ") == 2 * len(files) + 40 + 1

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

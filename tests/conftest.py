import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


@pytest.fixture(scope="session")
def ctx():
    from ml_prec_amd import native as N
    c = N.Context(0)
    yield c
    c.close()


_SYSTEMS = {}


def system(name):
    """Small seeded systems shared by the tests (cached per session)."""
    from ml_prec_amd import spe_synth as S
    if name not in _SYSTEMS:
        if name == "ln5_spe":
            s = S.lognormal_system(kind="spe", c=1.0)
        elif name == "ln5_tpfa_c1":
            s = S.lognormal_system(kind="tpfa", c=1.0)
        elif name == "ln5_tpfa":
            s = S.lognormal_system(kind="tpfa", c=0.0)
        elif name == "spe10_z12":
            s = S.make_system(60, 220, 12, S.BASE_SEED + 11, "spe10", "spe", 1.0)
        elif name == "spe10_z40":
            s = S.make_system(60, 220, 40, S.BASE_SEED + 12, "spe10", "spe", 1.0)
        elif name == "spe10_z85":
            s = S.spe10_system(kind="spe", c=1.0)
        elif name == "tiny":
            s = S.make_system(6, 10, 3, 7, "spe10", "spe", 1.0)
        elif name == "tiny_tpfa_c1":
            s = S.make_system(6, 10, 3, 8, "lognormal", "tpfa", 1.0, sigma=1.5)
        elif name == "small_spe10":
            s = S.make_system(12, 20, 5, 9, "spe10", "spe", 1.0)
        else:
            raise KeyError(name)
        _SYSTEMS[name] = s
    return _SYSTEMS[name]


def rng(seed=0):
    return np.random.default_rng(seed)

"""Shared test helpers: the gpu marker, the engine's Python mirror, the oracle and
the golden fixtures written by the compiled reference."""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_ccnm():
    spec = importlib.util.spec_from_file_location(
        "ccnm", os.path.join(ROOT, "coupled-cluster-nuclearmatter_amd", "ccnm.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ccnm"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def ccnm():
    return load_ccnm()


def golden(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def fixture_lattice(g):
    """Oracle lattice from a fixture's stored orbital table (reference numbering)."""
    import ccd_oracle as co
    oi, od = g["orb_int"], g["orb_dbl"]
    return co.Lattice(oi[0], oi[1], oi[2], oi[3], oi[4], od[0], od[1], od[2], od[3],
                      int(g["params"][0]), float(g["box"]), float(g["volume"]), kf=float(g["kf"]),
                      rho=float(g["rho"]), nmax=int(g["params"][1]))


def fixture_interaction(g, lat):
    import ccd_oracle as co
    idx = g["v_idx"].astype(np.int64)
    recs = [(idx[:, 1] - 1, idx[:, 2] - 1, idx[:, 3] - 1, idx[:, 4] - 1, g["v_val"])]
    return co.TableInteraction(lat.nsp, recs)

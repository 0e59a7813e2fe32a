import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "nested-parallelism_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def load_pkg():
    name = "nested_parallelism_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def npa():
    return load_pkg()


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def ctx(npa):
    c = npa.Context(0)
    yield c
    c.close()


def random_csr(rng, n, m, wmin=1, wmax=100, self_loops=True):
    """Random directed multigraph in CSR (test inputs)."""
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    if not self_loops and n > 1:
        bad = src == dst
        dst[bad] = (dst[bad] + 1) % n
    w = rng.integers(wmin, wmax + 1, m)
    order = np.argsort(src, kind="stable")
    src, dst, w = src[order], dst[order], w[order]
    row = np.zeros(n + 1, np.int64)
    np.add.at(row, src + 1, 1)
    row = np.cumsum(row)
    return row.astype(np.int32), dst.astype(np.int32), w.astype(np.int32)


def hub_csr(rng, n, m, hub_sizes, inward=False):
    """Random multigraph plus rows (inward=False) or in-edge lists (inward=True) of exactly
    hub_sizes[i] entries at vertex i, on top of m random edges (hub vertices get no random
    entries of that kind): rows longer than one wavefront's share, split into chunks."""
    src = rng.integers(len(hub_sizes), n, m)
    dst = rng.integers(0, n, m)
    if inward:
        dst = rng.integers(len(hub_sizes), n, m)
        src = rng.integers(0, n, m)
    hs, hd = [src], [dst]
    for i, k in enumerate(hub_sizes):
        other = rng.integers(0, n, k)
        hs.append(other if inward else np.full(k, i))
        hd.append(np.full(k, i) if inward else other)
    src, dst = np.concatenate(hs), np.concatenate(hd)
    order = np.argsort(src, kind="stable")
    src, dst = src[order], dst[order]
    row = np.zeros(n + 1, np.int64)
    np.add.at(row, src + 1, 1)
    return np.cumsum(row).astype(np.int32), dst.astype(np.int32)


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))

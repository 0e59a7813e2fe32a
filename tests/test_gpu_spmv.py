"""np_spmv (SpMV_GPU) against the oracle: exact with the reference's inputs (quarter-integer
products, exact sums in any order), within 1e-12 relative for random float64 values."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG_DIR, hub_csr, load_golden, random_csr

pytestmark = pytest.mark.gpu


# rows of 4095 / 4096 / 4097 / 9000 / 20000 entries: the last medium row and chunked rows
HUBS = [4095, 4096, 4097, 9000, 20000]


@pytest.mark.parametrize("kind", ["rmat", "grid", "multigraph", "hubs"])
def test_spmv_matches_oracle(npa, ctx, orc, kind):
    if kind == "hubs":
        row, col = hub_csr(np.random.default_rng(85), 6000, 30000, HUBS)
        G = npa.Graph.from_csr(ctx, row, col, np.ones(len(col), np.int32))
    elif kind == "rmat":
        G = npa.Graph.rmat(ctx, 14, 16, wmin=1, wmax=1, seed=81)
    elif kind == "grid":
        G = npa.Graph.grid(ctx, 64, 80, None, 1, 1, seed=82)
    else:
        rng = np.random.default_rng(83)
        row, col, _ = random_csr(rng, 4000, 30000, 1, 1)
        G = npa.Graph.from_csr(ctx, row, col, np.ones(len(col), np.int32))
    row, col, _ = G.download()
    data, x = orc.spmv_reference_values(G.n, G.m)
    y, _ = ctx.spmv(G, data, x)
    assert np.array_equal(y, orc.spmv(row, col, data, x))
    rng = np.random.default_rng(84)
    d2, x2 = rng.standard_normal(G.m), rng.standard_normal(G.n)
    y2, _ = ctx.spmv(G, d2, x2)
    ref = orc.spmv(row, col, d2, x2)
    assert np.allclose(y2, ref, rtol=1e-12, atol=1e-12 * float(np.abs(ref).max()))
    G.free()


def test_spmv_without_entries_and_with_empty_rows(npa, ctx, orc):
    """No entries at all (y = 0), then a matrix whose rows are empty but for one."""
    G = npa.Graph.from_csr(ctx, np.zeros(8, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32))
    y, _ = ctx.spmv(G, np.zeros(0), np.arange(7, dtype=np.float64))
    assert np.array_equal(y, np.zeros(7))
    G.free()
    row = np.array([0, 0, 0, 5, 5, 5], np.int32)
    col = np.array([4, 0, 4, 1, 2], np.int32)
    G = npa.Graph.from_csr(ctx, row, col, np.ones(5, np.int32))
    data, x = orc.spmv_reference_values(5, 5)
    y, _ = ctx.spmv(G, data, x)
    assert np.array_equal(y, orc.spmv(row, col, data, x))
    assert y[0] == y[1] == y[3] == y[4] == 0.0
    G.free()


def test_gpu_spmv_cli_matches_oracle(tmp_path, orc):
    g = load_golden("d9_rand.npz")
    out = tmp_path / "y.txt"
    r = subprocess.run([os.path.join(PKG_DIR, "gpu-spmv"), "-f", "0", "-i", os.path.join(GOLDEN, str(g["path"])),
                        "-s", "6", "-o", str(out)], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    n = len(g["row"]) - 1
    data, x = orc.spmv_reference_values(n, len(g["col"]))
    ref = orc.spmv(g["row"], g["col"], data, x)
    got = np.array([float(l.split("=")[1]) for l in out.read_text().splitlines()])
    assert np.array_equal(got, ref)

"""np_color (GRAPH_COLOR_GPU) against the oracle: both methods bit-exact (int32 colors)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG_DIR, load_golden, random_csr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["rmat", "rmat_plain", "er", "grid", "multigraph", "tiny"])
def test_color_both_methods_match_oracle(npa, ctx, orc, kind):
    if kind == "rmat":
        G = npa.Graph.rmat(ctx, 13, 16, wmin=1, wmax=1, seed=61)
    elif kind == "rmat_plain":                       # ids not permuted: hubs at small ids
        G = npa.Graph.rmat(ctx, 12, 16, wmin=1, wmax=1, seed=62, permute=False)
    elif kind == "er":
        G = npa.Graph.er(ctx, 3000, 30000, 1, 1, seed=63)
    elif kind == "grid":
        G = npa.Graph.grid(ctx, 40, 50, None, 1, 1, seed=64)
    elif kind == "multigraph":
        rng = np.random.default_rng(65)
        row, col, _ = random_csr(rng, 2000, 12000, 1, 1)
        col[::5] = col[::5] % 7                       # parallel edges into small ids
        G = npa.Graph.from_csr(ctx, row, col, np.ones(len(col), np.int32))
    else:
        G = npa.Graph.from_csr(ctx, np.array([0, 1], np.int32), np.array([0], np.int32), np.ones(1, np.int32))
    row, col, _ = G.download()
    ref = orc.color(row, col)
    for method in (0, 1):
        got, st = ctx.color(G, method)
        assert np.array_equal(got, ref), (kind, method)
        assert st.iterations >= int(ref.max())
    G.free()


def test_color_rejects_bad_method(npa, ctx):
    G = npa.Graph.rmat(ctx, 8, 4, wmin=1, wmax=1, seed=66)
    with pytest.raises(npa.NpError):
        ctx.color(G, 2)
    G.free()


@pytest.mark.parametrize("solution", ["7", "12"])
def test_gpu_graph_color_cli_matches_oracle(tmp_path, orc, solution):
    g = load_golden("d9_rand.npz")
    out = tmp_path / "colors.txt"
    r = subprocess.run([os.path.join(PKG_DIR, "gpu-graph-color"), "-f", "0", "-i", os.path.join(GOLDEN, str(g["path"])),
                        "-s", solution, "-o", str(out)], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Execution time:" in r.stderr
    ref = orc.color(g["row"], g["col"])
    assert out.read_text().splitlines() == ["Node %d: %d" % (i + 1, c) for i, c in enumerate(ref)]

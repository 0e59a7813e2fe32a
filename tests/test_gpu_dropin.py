"""The reference-named entry points (libnpdropin.so) through their command-line drivers:
gpu-sssp / gpu-bfs-rec read the reference's file formats, call SSSP_GPU / BFS_REC_GPU N=10
times and print what the reference's drivers print; gpu-bc calls BC_GPU once and prints the
values with outputBC."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG_DIR, load_golden

pytestmark = pytest.mark.gpu


def _run(exe, args, cwd, env=None):
    e = dict(os.environ)
    if env:
        e.update(env)
    r = subprocess.run([os.path.join(PKG_DIR, exe)] + args, cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120, env=e)
    return r


@pytest.mark.parametrize("level", ["0", "1", "2"])
def test_gpu_sssp_cli_output_cost_matches_reference(tmp_path, level):
    out = tmp_path / "costs.txt"
    r = _run("gpu-sssp", ["-f", "0", "-i", os.path.join(GOLDEN, "graphs", "d9_small.gr"), "-s", "6",
                          "-o", str(out)], tmp_path, {"NP_CONSOLIDATE_LEVEL": level})
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("iteration:") == 10          # printed once per SSSP_GPU call
    assert out.read_text() == open(os.path.join(GOLDEN, "d9_small.costs.txt")).read()


@pytest.mark.parametrize("name", ["d9_rand", "d10_small", "slndc_small"])
@pytest.mark.parametrize("solution", ["0", "3", "6"])
def test_gpu_sssp_cli_formats(tmp_path, name, solution):
    g = load_golden(name + ".npz")
    out = tmp_path / "costs.txt"
    r = _run("gpu-sssp", ["-f", str(int(g["fmt"])), "-i", os.path.join(GOLDEN, str(g["path"])),
                          "-s", solution, "-o", str(out), "-v"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "CUDA kernel execution" in r.stdout
    lines = out.read_text().splitlines()
    src = int(g["source"])
    k = list(g["sources"]).index(src)
    expect = ["From Node %d -> %d is %d" % (src + 1, i + 1, d) for i, d in enumerate(g["dist"][k])]
    assert lines == expect


@pytest.mark.parametrize("name", ["d9_rand", "d10_small"])
def test_gpu_bc_cli_matches_oracle(tmp_path, orc, name):
    """BC_GPU over every vertex (solution 7), values printed with outputBC's "%f"."""
    g = load_golden(name + ".npz")
    out = tmp_path / "bc.txt"
    r = _run("gpu-bc", ["-f", str(int(g["fmt"])), "-i", os.path.join(GOLDEN, str(g["path"])), "-s", "7",
                        "-o", str(out)], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "Execution time:" in r.stdout
    lines = out.read_text().splitlines()
    assert [l.split(":")[0] for l in lines] == ["Node %d" % (i + 1) for i in range(len(g["row"]) - 1)]
    got = np.array([float(l.split(":")[1]) for l in lines])
    ref = orc.bc(g["row"], g["col"])
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-6 + 5e-7), float(np.max(np.abs(got - ref)))


@pytest.mark.parametrize("solution", ["0", "1", "2", "3"])
def test_gpu_bfs_rec_cli_validates(tmp_path, solution):
    g = load_golden("d9_rand.npz")
    out = tmp_path / "levels.txt"
    r = _run("gpu-bfs-rec", ["-f", "0", "-i", os.path.join(GOLDEN, str(g["path"])), "-s", solution,
                             "-o", str(out)], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "PASS: GPU bfs rec !" in r.stdout
    k = list(g["sources"]).index(int(g["source"]))
    lv = [int(l.split()[-1]) for l in out.read_text().splitlines()]
    assert np.array_equal(np.array(lv, np.int64), g["level"][k].astype(np.int64))


def test_reader_failure_exits_like_reference(tmp_path):
    bad = tmp_path / "bad.gr"
    bad.write_text("p sp 3 2\na 1 2 1\n")
    r = _run("gpu-sssp", ["-f", "0", "-i", str(bad), "-s", "6"], tmp_path)
    assert r.returncode == 0            # EXIT macro of common/global.h exits with status 0
    assert "info:" in r.stderr and "header says" in r.stderr


def test_cli_beside_a_process_holding_the_cooperative_queue(tmp_path, npa, ctx, orc):
    """This process runs persistent (cooperative) solves, then the CLI process solves on the
    same device: it gets exact results whichever drive it is given."""
    g = load_golden("d9_rand.npz")
    G = npa.Graph.from_csr(ctx, g["row"], g["col"], g["w"])
    src = int(g["source"])
    d, st = ctx.sssp(G, src, npa.Sched(level=1))
    k = list(g["sources"]).index(src)
    assert np.array_equal(d, g["dist"][k]) and st.drive == npa.DRIVE_PERSISTENT
    out = tmp_path / "costs.txt"
    r = _run("gpu-sssp", ["-f", "0", "-i", os.path.join(GOLDEN, str(g["path"])), "-s", "6",
                          "-o", str(out)], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = out.read_text().splitlines()
    assert lines == ["From Node %d -> %d is %d" % (src + 1, i + 1, x) for i, x in enumerate(g["dist"][k])]
    G.free()


REF_BIN = os.path.join(os.path.dirname(PKG_DIR), "oracle", "_ref")


def _ref_driver(name):
    p = os.path.join(REF_BIN, "ref-gpu-" + name)
    if not os.access(p, os.X_OK):
        pytest.skip("relinked reference driver %s not built (oracle/Makefile needs the reference sources)" % name)
    return p


@pytest.mark.parametrize("name", ["d9_rand", "d10_small", "slndc_small"])
@pytest.mark.parametrize("solution", ["0", "1", "2", "3"])
def test_reference_bfs_rec_main_relinked(tmp_path, name, solution):
    """The reference's own bfs_rec.cpp main(), unchanged, linked against libnpdropin.so: it runs
    BFS_REC_GPU N=10 times and checks the levels with its own bfs_cpu + validateArrays."""
    exe = _ref_driver("bfs_rec")
    g = load_golden(name + ".npz")
    r = subprocess.run([exe, "-f", str(int(g["fmt"])), "-i", os.path.join(GOLDEN, str(g["path"])), "-s", solution],
                       cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "PASS: GPU bfs rec !" in r.stdout, r.stdout


@pytest.mark.parametrize("solution", [str(s) for s in range(7)])
def test_reference_sssp_main_relinked(tmp_path, solution):
    """The reference's own sssp.cpp main(), unchanged, linked against libnpdropin.so: reads the
    DIMACS9 file, builds the CSR and calls SSSP_GPU N=10 times with the given solution."""
    exe = _ref_driver("sssp")
    r = subprocess.run([exe, "-f", "0", "-i", os.path.join(GOLDEN, "graphs", "d9_rand.gr"), "-s", solution, "-v"],
                       cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "CUDA kernel execution" in r.stdout
    if solution == "6":
        assert r.stdout.count("iteration:") == 10

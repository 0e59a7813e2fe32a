"""The drop-in boundary without a source change: the reference's own GPU drivers
(main() of apps/<app>/src/gpu/<app>.cpp, unchanged, with the reference's headers
common/global.h and apps/<app>/src/gpu/<app>.h) compile and link against libnpdropin.so
instead of the reference's CUDA wrappers, with no undefined symbol: the documented link recipe,
run as written.  Skipped where the reference sources are absent."""
import os
import subprocess

import pytest

from conftest import PKG_DIR

REF = "/root/reference"
APPS = [("sssp", "sssp"), ("bfs-rec", "bfs_rec"), ("bc", "bc"), ("graph-color", "graph_color"),
        ("pagerank", "pagerank"), ("SpMV", "SpMV")]

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(REF, "common", "global.h")),
                                reason="reference sources not present")


@pytest.mark.parametrize("app,name", APPS)
def test_reference_driver_links_against_dropin(tmp_path, app, name):
    src = os.path.join(REF, "apps", app, "src", "gpu", name + ".cpp")
    exe = str(tmp_path / ("gpu-" + name))
    cmd = ["g++", "-O1", "-w", "-I" + os.path.join(REF, "common"), "-I" + os.path.dirname(src), src,
           "-L" + PKG_DIR, "-lnpdropin", "-lnpsssp", "-Wl,-rpath," + PKG_DIR, "-Wl,--no-undefined", "-o", exe]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    # every symbol the driver needs resolves in the drop-in libraries (loaded, not run)
    r = subprocess.run(["ldd", "-r", exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    bad = [l for l in r.stdout.splitlines() if "undefined symbol" in l or "not found" in l]
    assert not bad, bad
    assert "libnpdropin.so" in r.stdout


def test_relinked_drivers_built_for_the_gpu_tests():
    """oracle/Makefile builds the same relinked drivers into oracle/_ref/ (git-ignored; they
    travel with the tree) for the GPU test that runs them."""
    here = os.path.join(os.path.dirname(PKG_DIR), "oracle", "_ref")
    for _, name in APPS:
        assert os.access(os.path.join(here, "ref-gpu-" + name), os.X_OK), name

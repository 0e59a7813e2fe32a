"""The C-ABI libraries load and export every entry point their headers declare (no GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import PKG_DIR, ROOT


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(np_[a-z0-9_]+)\s*\(", text)))


def _exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if l.strip()}


def test_c_abi_exports_every_declared_symbol(npa):
    npa.load()
    syms = _exported(os.path.join(PKG_DIR, "libnpsssp.so"))
    declared = _declared("np_sssp.h")
    assert len(declared) > 40
    missing = [d for d in declared if d not in syms]
    assert not missing, missing
    # the Python mirror binds exactly the declared set
    assert sorted(npa.EXPORTS) == declared


def test_bounds_checked_build_exports_the_same_abi():
    """libnpsssp_bounds.so (every frontier-queue and delayed-buffer write checked against its
    segment) is a drop-in for libnpsssp.so: the same exports, and it says which build it is."""
    blib = os.path.join(PKG_DIR, "libnpsssp_bounds.so")
    syms = _exported(blib)
    missing = [d for d in _declared("np_sssp.h") if d not in syms]
    assert not missing, missing
    for path, tag in ((blib, True), (os.path.join(PKG_DIR, "libnpsssp.so"), False)):
        L = ctypes.CDLL(path)
        L.np_version.restype = ctypes.c_char_p
        assert (b"bounds-checked" in L.np_version()) == tag, path


def test_dropin_exports_reference_names():
    syms = _exported(os.path.join(PKG_DIR, "libnpdropin.so"))
    # C++ names of common/global.h and apps/sssp/src/gpu/sssp.h / bfs_rec.h
    need = ["_Z8SSSP_GPUv", "_Z11BFS_REC_GPUv", "_Z16readInputDIMACS9v", "_Z17readInputDIMACS10v",
            "_Z14readInputSLNDCv", "_Z9readInputv", "_Z10convertCSRv", "_Z10outputCostP8_IO_FILE",
            "_Z5clearv", "_Z7gettimev", "_Z10gettime_msv", "graph", "noNodeTotal", "noEdgeTotal", "source",
            "fp", "VERBOSE", "DEBUG", "config", "ker_exe_time", "h2d_memcpy_time", "d2h_memcpy_time",
            "_Z6BC_GPUv", "_Z8outputBCP8_IO_FILE", "bc", "_Z15GRAPH_COLOR_GPUv", "_Z11outputColorP8_IO_FILE",
            "_Z12PAGERANK_GPUv", "_Z10outputRankP8_IO_FILE", "_Z11invertGraphv", "_Z8SpMV_GPUv"]
    missing = [n for n in need if n not in syms]
    assert not missing, missing


def test_cli_binaries_built():
    for exe in ("gpu-sssp", "gpu-bfs-rec", "gpu-bc", "gpu-graph-color", "gpu-pagerank", "gpu-spmv"):
        p = os.path.join(PKG_DIR, exe)
        assert os.path.exists(p) and os.access(p, os.X_OK)


def test_no_device_is_a_clean_error(npa):
    import torch  # noqa: F401  (device count only; does not initialise HIP)
    if torch.cuda.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(npa.NpError) as e:
        npa.Context(0)
    assert e.value.code == -7


def test_sched_defaults(npa):
    s = npa.default_sched(1)
    assert (s.level, s.log2_t_lo, s.log2_t_hi, s.drive, s.frontier) == (1, 6, 6, 0, 0)
    s = npa.default_sched(9)
    assert s.level == 1


def test_struct_layouts_match_header(npa, tmp_path):
    """np_sched, np_stats and np_trace_rec as the C compiler lays them out from the header
    equal the ctypes mirror (size and every field offset)."""
    structs = {"np_sched": npa.Sched, "np_stats": npa.Stats, "np_trace_rec": npa.TraceRec}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "np_sssp.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True).stdout.split("\n")
    got = {}
    for l in out:
        if l.strip():
            parts = l.split()
            got[(parts[0], parts[1])] = int(parts[2])
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)

"""Builds the gfx950 library of the SSSP nested-parallelism path, in-tree.

    python nested-parallelism_amd/build.py        (or __graft_entry__.build())

Outputs next to this file:
    libnpsssp.so     HIP kernels + C ABI (include/np_sssp.h)
    libnpsssp_bounds.so   the same with every frontier-queue and delayed-buffer write bounds-checked
                          (NP_BOUNDS_CHECK=1; a violation is skipped and fails the solve with NP_ERR_BOUNDS)
    libnpdropin.so   the reference's C++ entry points (include/np_dropin.h) over the C ABI
    gpu-sssp, gpu-bfs-rec, gpu-bc, gpu-graph-color, gpu-pagerank, gpu-spmv   command-line drivers with the
                            reference's flags
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INC = os.path.join(os.path.dirname(HERE), "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

ARCH = ["--offload-arch=gfx950"]
# packed-fp32 instructions are kept out of every HIP object
HIP_FLAGS = ARCH + ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I" + INC]
HOST_FLAGS = HIP_FLAGS  # hipcc compiles .cpp as HIP as well: same target and flags

HIP_SRC = ["np_kernels.hip", "np_gen.hip", "np_bc.hip", "np_color.hip", "np_pagerank.hip"]
HOST_SRC = ["np_api.cpp", "np_graph_io.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s" % (" ".join(cmd), out))
    return out


def _obj(src):
    return os.path.join(BUILD, os.path.splitext(src)[0] + ".o")


BOUNDS_SRC = ["np_kernels.hip", "np_api.cpp"]  # compiled a second time with NP_BOUNDS_CHECK=1


def _compile(src, bounds=False):
    path = os.path.join(CSRC, src)
    obj = _obj(src)
    extra = []
    if bounds:
        obj = obj[:-2] + "_bounds.o"
        extra = ["-DNP_BOUNDS_CHECK=1"]
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + extra + ["-c", path, "-o", obj]
    else:
        cmd = [HIPCC] + HOST_FLAGS + extra + ["-c", path, "-o", obj]
    return _run(cmd)


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = [(x, False) for x in HIP_SRC + HOST_SRC + ["np_dropin.cpp"]] + [(x, True) for x in BOUNDS_SRC]
    with ThreadPoolExecutor(max_workers=4) as ex:
        outs = list(ex.map(lambda j: _compile(*j), jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    lib = os.path.join(HERE, "libnpsssp.so")
    tmp = lib + ".tmp"
    _run([HIPCC] + ARCH + ["-shared", "-fPIC", "-o", tmp] + [_obj(s) for s in HIP_SRC + HOST_SRC])
    os.replace(tmp, lib)
    blib = os.path.join(HERE, "libnpsssp_bounds.so")
    bobjs = [_obj(s)[:-2] + "_bounds.o" if s in BOUNDS_SRC else _obj(s) for s in HIP_SRC + HOST_SRC]
    _run([HIPCC] + ARCH + ["-shared", "-fPIC", "-o", blib + ".tmp"] + bobjs)
    os.replace(blib + ".tmp", blib)
    dropin = os.path.join(HERE, "libnpdropin.so")
    _run([HIPCC, "-shared", "-fPIC", "-o", dropin + ".tmp", _obj("np_dropin.cpp"),
          "-L" + HERE, "-lnpsssp", "-Wl,-rpath,$ORIGIN"])
    os.replace(dropin + ".tmp", dropin)
    for exe, flag in (("gpu-sssp", "-DNP_CLI_SSSP"), ("gpu-bfs-rec", "-DNP_CLI_BFS"), ("gpu-bc", "-DNP_CLI_BC"),
                      ("gpu-graph-color", "-DNP_CLI_COLOR"), ("gpu-pagerank", "-DNP_CLI_PR"),
                      ("gpu-spmv", "-DNP_CLI_SPMV")):
        o = os.path.join(BUILD, exe + ".o")
        _run([HIPCC] + HOST_FLAGS + [flag, "-c", os.path.join(CSRC, "np_cli.cpp"), "-o", o])
        _run([HIPCC, "-o", os.path.join(HERE, exe), o, "-L" + HERE, "-lnpdropin", "-lnpsssp",
              "-Wl,-rpath,$ORIGIN"])
    return lib


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", os.path.join(HERE, "libnpsssp.so"))

"""Host logic of bench.py's live traffic measurement (no GPU): the counter-file reader that
cuts one solve out of a profiled child run, and the --sched form of a schedule."""
import importlib.util
import os

import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("np_bench_host", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


HEAD = "Correlation_Id,Dispatch_Id,Agent_Id,Queue_Id,Kernel_Name,Counter_Name,Counter_Value,Start_Timestamp,End_Timestamp\n"


def _write(path, dispatches):
    """dispatches: (name, [values of its counter instances]) in launch order."""
    with open(path, "w") as f:
        f.write(HEAD)
        for i, (name, vals) in enumerate(dispatches):
            for v in vals:
                f.write('%d,%d,1,1,"%s",FETCH_SIZE,%s,%d,%d\n' % (i, i + 1, name, v, 1000 + 10 * i, 1005 + 10 * i))


def test_solve_counter_cuts_the_solve_between_markers(bench, tmp_path):
    copy = "__amd_rocclr_copyBuffer.kd"
    d = [("np::k_gen_rmat(...)", [7.0]), (copy, [3.0]),                 # workload build
         ("np::k_pk_compact(...)", [11.0, 1.0]),                          # warm-up solve
         ("np::k_reached_work(int const*, int const*, int, int, unsigned long long*)", [2.0]),
         (copy, [0.5]),                                                   # its result copy: not the solve's
         ("np::k_fill_i32(...)", [100.0, 100.0]), ("np::k_pk_compact(...)", [5.0, 6.0]),
         ("np::k_decide(...)", [1.0]), ("np::k_pk_expand<false>(...)", [400.0, 350.0]),
         ("np::k_reached_work(int const*, int const*, int, int, unsigned long long*)", [2.0]),
         (copy, [0.5]), (copy, [524288.0]), (copy, [524311.0]), (copy, [524300.0])]
    f = tmp_path / "run_counter_collection.csv"
    _write(f, d)
    total, copy_kib, n = bench.solve_counter([str(f)])
    assert total == 100 + 100 + 5 + 6 + 1 + 400 + 350
    assert copy_kib == 524300.0                                           # median of the last three copies
    assert n == 4


def test_solve_counter_reports_what_is_missing(bench, tmp_path):
    f = tmp_path / "run_counter_collection.csv"
    _write(f, [("np::k_pk_expand<false>(...)", [1.0]), ("__amd_rocclr_copyBuffer.kd", [1.0])])
    assert isinstance(bench.solve_counter([str(f)]), str)
    _write(f, [("np::k_reached_work(...)", [1.0]), ("np::k_reached_work(...)", [1.0])] +
           [("__amd_rocclr_copyBuffer.kd", [1.0])] * 3)
    assert bench.solve_counter([str(f)]) == "no dispatch between the markers"


def test_sched_arg_round_trips(bench, npa):
    s = npa.Sched(level=npa.LEVEL_LEARNED, log2_t_lo=4, log2_t_hi=10, wgs_per_cu=2, drive=npa.DRIVE_GRAPH,
                  delta=4004, pack_hold=11)
    v = [int(x) for x in bench.sched_arg(s).split(",")]
    t = npa.Sched(level=v[0], log2_t_lo=v[1], log2_t_hi=v[2], wgs_per_cu=v[3], drive=v[4], delta=v[5],
                  pack_hold=v[6])
    assert repr(t) == repr(s)

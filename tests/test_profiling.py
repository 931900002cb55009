"""The rocprofv3 summarizer (astrodrift_amd/profiling.py) on synthetic CSVs in
the layout rocprofv3 writes: the per-kernel table, and the long-dispatch
filter that separates full launches from short trial runs of one kernel."""

from __future__ import annotations

import csv

from astrodrift_amd import profiling

TRACE_FIELDS = ["Kind", "Agent_Id", "Queue_Id", "Stream_Id", "Thread_Id", "Dispatch_Id",
                "Kernel_Id", "Kernel_Name", "Correlation_Id", "Start_Timestamp", "End_Timestamp"]


def _write_trace(d, dispatches):
    d.mkdir()
    with open(d / "run_kernel_trace.csv", "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=TRACE_FIELDS)
        w.writeheader()
        t = 1_000_000
        for k, (name, ms) in enumerate(dispatches):
            w.writerow({"Kind": "KERNEL_DISPATCH", "Agent_Id": "Agent 1", "Queue_Id": 1,
                        "Stream_Id": 0, "Thread_Id": 1, "Dispatch_Id": k, "Kernel_Id": 1,
                        "Kernel_Name": name, "Correlation_Id": k, "Start_Timestamp": t,
                        "End_Timestamp": t + int(ms * 1e6)})
            t += int(ms * 1e6) + 1000


def test_long_dispatches_filters_and_orders(tmp_path, capsys):
    integ = "void drift::dop853_persistent_kernel<3, float>(drift::PlainArgs)"
    unstage = "drift::unstage_kernel(double const*, double*, long, int)"
    # two short trial launches, then two full launches with their transposes
    _write_trace(tmp_path / "trace", [(integ, 5.0), (unstage, 0.8), (integ, 5.1),
                                      (integ, 40.0), (unstage, 19.0), (integ, 41.0),
                                      (unstage, 19.5)])
    assert profiling.summarize([str(tmp_path / "trace")], min_ms=10.0) == 0
    out = capsys.readouterr().out
    rows = [line for line in out.splitlines() if line.startswith("| `")]
    assert rows[0].startswith(f"| `{integ}` | 2 | 40.500 | 81.000 |")
    assert rows[1].startswith(f"| `{unstage}` | 2 | 19.250 | 38.500 |")
    order = [line for line in out.splitlines() if line.startswith("| 0 |") or
             line.startswith("| 1 |") or line.startswith("| 2 |") or line.startswith("| 3 |")]
    assert order == ["| 0 | `dop853_persistent_kernel<3, float>` | 40.000 |",
                     "| 1 | `unstage_kernel` | 19.000 |",
                     "| 2 | `dop853_persistent_kernel<3, float>` | 41.000 |",
                     "| 3 | `unstage_kernel` | 19.500 |"]


def test_kernel_stats_table(tmp_path, capsys):
    d = tmp_path / "trace"
    d.mkdir()
    with open(d / "run_kernel_stats.csv", "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=["Name", "Calls", "TotalDurationNs", "AverageNs",
                                           "Percentage"])
        w.writeheader()
        w.writerow({"Name": "small", "Calls": 4, "TotalDurationNs": 4e6, "AverageNs": 1e6,
                    "Percentage": 3.0})
        w.writerow({"Name": "big", "Calls": 2, "TotalDurationNs": 240e6, "AverageNs": 120e6,
                    "Percentage": 97.0})
    assert profiling.summarize([str(d)]) == 0
    rows = [line for line in capsys.readouterr().out.splitlines() if line.startswith("| `")]
    assert rows == ["| `big` | 2 | 120.000 | 240.000 | 97.00% |",
                    "| `small` | 4 | 1.000 | 4.000 | 3.00% |"]

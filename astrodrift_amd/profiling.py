"""Profiling driver and summarizer for rocprofv3 runs of the benchmark kernel.

Run (one pass each; counters in runs of their own):

    rocprofv3 --kernel-trace --stats --output-format csv -d D/trace -o run -- \\
        python3 -m astrodrift_amd.profiling
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d D/fetch -o run -- \\
        python3 -m astrodrift_amd.profiling
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d D/write -o run -- \\
        python3 -m astrodrift_amd.profiling

The run launches, on one stream: a calibration copy of a known size (torch's
copy of a 2 GiB float64 tensor: 2 GiB read, 2 GiB written), then one
integration of the 1,048,576-orbit Milky-Way DOP853 workload (rtol = atol =
1e-12, 1,001 outputs over t = 0..50) with the persistent schedule and the
default (cost) orbit ordering and block size (512 threads: packed dense
output; --block 256 runs the inline form).  The copy's counter values give the factor that
turns FETCH_SIZE / WRITE_SIZE into bytes for the integrator's dispatch.

Summarize (reads the CSV files rocprofv3 wrote, prints markdown):

    python3 -m astrodrift_amd.profiling --summarize D/trace [D/fetch D/write [D/sq ...]]
"""

from __future__ import annotations

import argparse
import csv
import math
import sys
from pathlib import Path

# the persistent integrator: dop853_persistent_packed_kernel (the default
# launch) or, with --block 64..256, dop853_persistent_kernel
INTEGRATOR = "dop853_persistent"
COPY = "__amd_rocclr_copyBuffer"
ALGORITHMIC_BYTES_PER_ORBIT_SAMPLE = 48  # one (x, y, z, vx, vy, vz) record


def algorithmic_bytes(n: int, nt: int) -> int:
    """Bytes one launch must move: state0 in (48 n), the time grid (8 nt), the
    trajectories out (48 n nt) and the statistics out (16 n)."""
    return 48 * n + 8 * nt + ALGORITHMIC_BYTES_PER_ORBIT_SAMPLE * n * nt + 16 * n


def run(args) -> int:
    import torch

    from . import _native, workloads

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    n_copy = int(args.copy_gib * (1 << 30) // 8)
    if n_copy > 0:
        a = torch.ones(n_copy, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(f"[profiling] calibration copy: {n_copy * 8} bytes read, {n_copy * 8} bytes written",
              flush=True)
        del a, b
        torch.cuda.empty_cache()

    if args.ragged_jobs > 0:
        return run_ragged(args.ragged_jobs)
    if args.dopr54_uniform > 0:
        return run_dopr54_uniform(args.dopr54_uniform)
    rtol = args.rtol if args.rtol else None
    w = workloads.mw_disk(n=args.n) if rtol is None else workloads.mw_disk(n=args.n, rtol=rtol,
                                                                          atol=rtol)
    s0 = torch.from_numpy(w.state0).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    out = torch.empty((w.nt, w.n, 6), dtype=torch.float64, device=dev)
    st = torch.zeros((w.n, 4), dtype=torch.int32, device=dev)
    lt = math.log(w.rtol)
    _native.launch_device(_native.METHOD_DOP853, s0.data_ptr(), tt.data_ptr(), out.data_ptr(),
                          w.n, w.nt, lt, lt, w.potential, st.data_ptr(), schedule=args.schedule,
                          block_size=args.block,
                          precision=_native.PRECISION_MIXED if args.mixed else _native.PRECISION_FP64,
                          staging=_native.STAGING_STAGED if args.staged else _native.STAGING_AUTO)
    torch.cuda.synchronize()
    acc = int(st[:, 0].sum().item())
    print(f"[profiling] integrator: n={w.n} nt={w.nt} accepted={acc} "
          f"algorithmic bytes={algorithmic_bytes(w.n, w.nt)}", flush=True)
    return 0


def run_ragged(njobs: int) -> int:
    """One drift_ragged_device call on the first njobs jobs of the ragged mix
    (256-thread blocks, no tuner; one launch per method), inputs and outputs
    resident in HBM.  Prints each method's algorithmic FLOPs."""
    import numpy as np
    import torch

    from . import _native, flops, workloads

    dev = torch.device("cuda:0")
    jobs = workloads.ragged_mix(njobs)
    pots = list(workloads.RAGGED_POTENTIALS)
    keep, descs = [], []
    for j in jobs:
        s0 = torch.from_numpy(j.state0).to(dev)
        tt = torch.from_numpy(j.times).to(dev)
        out = torch.empty((j.nt, j.n, 6), dtype=torch.float64, device=dev)
        st = torch.zeros((j.n, 4), dtype=torch.int32, device=dev)
        keep.append((s0, tt, out, st))
        lt = math.log(j.rtol)
        method = _native.METHOD_DOP853 if j.method == "DOP853" else _native.METHOD_DOPR54
        descs.append(_native.DeviceJob(method, j.potential, j.n, j.nt, lt, lt, s0.data_ptr(),
                                       tt.data_ptr(), out.data_ptr(), st.data_ptr(),
                                       span_hint=float(j.times[-1] - j.times[0])))
    torch.cuda.synchronize()
    _native.ragged_device(descs, pots, block_size=256)
    torch.cuda.synchronize()
    flop = {"DOP853": 0.0, "DOPR54": 0.0}
    for j, (_, _, _, st) in zip(jobs, keep):
        s = st.cpu().numpy()
        spec = pots[j.potential]
        flop[j.method] += (flops.dop853_flops(s, j.nt, spec) if j.method == "DOP853"
                           else flops.dopr54_flops(s, spec))
    print(f"[profiling] ragged mix: {njobs} jobs, {sum(j.n for j in jobs)} orbits, algorithmic "
          f"FLOPs DOP853 {flop['DOP853']:.4e} DOPR54 {flop['DOPR54']:.4e}", flush=True)
    return 0


def run_dopr54_uniform(n: int) -> int:
    """DOPR54 on n disk orbits under each potential of the ragged mix (Kepler,
    Plummer, Milky Way; rtol = atol = 1e-10, 101 outputs over t = 0..50), one
    launch each with the default (plain, cost-ordered) schedule.  Prints each
    launch's accepted steps and algorithmic FLOPs."""
    import torch

    from . import _native, flops, workloads

    dev = torch.device("cuda:0")
    times = workloads.grid_times(0.0, 50.0, 101)
    tt = torch.from_numpy(times).to(dev)
    lt = math.log(1e-10)
    for k, spec in enumerate(workloads.RAGGED_POTENTIALS):
        y0 = workloads.disk_orbits(spec, n, seed=7 + k)
        s0 = torch.from_numpy(y0).to(dev)
        out = torch.empty((len(times), n, 6), dtype=torch.float64, device=dev)
        st = torch.zeros((n, 4), dtype=torch.int32, device=dev)
        _native.launch_device(_native.METHOD_DOPR54, s0.data_ptr(), tt.data_ptr(), out.data_ptr(),
                              n, len(times), lt, lt, spec, st.data_ptr())
        torch.cuda.synchronize()
        s = st.cpu().numpy()
        print(f"[profiling] DOPR54 {('Kepler', 'Plummer', 'Milky Way')[k]}: n={n} "
              f"accepted={int(s[:, 0].astype('int64').sum())} "
              f"algorithmic FLOPs={flops.dopr54_flops(s, spec):.4e}", flush=True)
        del out, st, s0
    return 0


def derived(c: dict[str, float], waves: int = 2) -> list[tuple[str, str, str]]:
    """Ratios of an SQ counter set (whichever inputs are present); `waves` is
    the kernel's resident waves per SIMD."""
    rows = []
    if "SQ_INSTS_SALU" in c and "SQ_INSTS_VALU" in c:
        rows.append(("SALU per VALU instruction", f"{c['SQ_INSTS_SALU'] / c['SQ_INSTS_VALU']:.3f}",
                     "INSTS_SALU / INSTS_VALU"))
    if "SQ_ACTIVE_INST_VALU" in c and "SQ_WAVE_CYCLES" in c:
        rows.append(("VALU busy",
                     f"{100 * c['SQ_ACTIVE_INST_VALU'] / (c['SQ_WAVE_CYCLES'] / waves):.1f}%",
                     f"ACTIVE_INST_VALU / (WAVE_CYCLES / {waves} waves per SIMD)"))
    if "SQ_THREAD_CYCLES_VALU" in c and "SQ_ACTIVE_INST_VALU" in c:
        rows.append(("active lanes per VALU instruction",
                     f"{100 * c['SQ_THREAD_CYCLES_VALU'] / (64 * c['SQ_ACTIVE_INST_VALU']):.1f}%",
                     "THREAD_CYCLES_VALU / (64 ACTIVE_INST_VALU)"))
    f64 = [k for k in ("SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_ADD_F64",
                       "SQ_INSTS_VALU_TRANS_F64") if k in c]
    if "SQ_WAIT_INST_LDS" in c and "SQ_WAVE_CYCLES" in c:
        rows.append(("waiting to issue an LDS instruction",
                     f"{100 * c['SQ_WAIT_INST_LDS'] / c['SQ_WAVE_CYCLES']:.2f}% of wave cycles",
                     "WAIT_INST_LDS / WAVE_CYCLES"))
    if len(f64) == 4 and "SQ_INSTS_VALU" in c:
        tot = sum(c[k] for k in f64)
        rows.append(("FP64 share of VALU instructions", f"{100 * tot / c['SQ_INSTS_VALU']:.1f}%",
                     "(FMA + MUL + ADD + TRANS F64) / INSTS_VALU"))
    return rows


def _csv_rows(d: Path, suffix: str) -> list[dict]:
    files = sorted(Path(d).rglob(f"*{suffix}"))
    if not files:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    rows = []
    for f in files:
        with open(f, newline="") as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def _counters(d: Path) -> dict[str, dict[str, float]]:
    """'kernel [dispatch id]' -> counter -> value summed over the dispatch's
    instances (one per XCD / shader engine)."""
    per_dispatch: dict[tuple[str, str], dict[str, float]] = {}
    for r in _csv_rows(d, "counter_collection.csv"):
        key = (r["Dispatch_Id"], r["Kernel_Name"])
        c = per_dispatch.setdefault(key, {})
        c[r["Counter_Name"]] = c.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return {f"{name} [dispatch {did}]": c for (did, name), c in per_dispatch.items()}


def _pick(counters: dict, needle: str, counter: str) -> float:
    vals = [c[counter] for k, c in counters.items() if needle in k and counter in c]
    if not vals:
        raise KeyError(f"{counter} of {needle} not found")
    return max(vals)  # the large dispatch of that name


def _short(name: str) -> str:
    return name if len(name) < 90 else name[:87] + "..."


def long_dispatches(trace: Path, min_ms: float) -> None:
    """Per-kernel table over the dispatches of at least min_ms, then those
    dispatches in launch order (separates full launches from short trial runs
    of the same kernel)."""
    rows = []
    for r in _csv_rows(trace, "kernel_trace.csv"):
        ms = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
        if ms >= min_ms:
            rows.append((int(r["Start_Timestamp"]), r["Kernel_Name"], ms))
    rows.sort()
    total = sum(ms for _, _, ms in rows) or 1.0
    groups: dict[str, list[float]] = {}
    for _, name, ms in rows:
        groups.setdefault(name, []).append(ms)
    print(f"Dispatches of at least {min_ms:g} ms:")
    print()
    print("| kernel | calls | avg duration (ms) | total (ms) | share |")
    print("|---|---|---|---|---|")
    for name, v in sorted(groups.items(), key=lambda kv: -sum(kv[1])):
        print(f"| `{_short(name)}` | {len(v)} | {sum(v) / len(v):.3f} | {sum(v):.3f} | "
              f"{100.0 * sum(v) / total:.2f}% |")
    print()
    print("| # | kernel | duration (ms) |")
    print("|---|---|---|")
    for k, (_, name, ms) in enumerate(rows):
        short = name.split("(")[0].replace("void ", "").replace("drift::", "")
        print(f"| {k} | `{short}` | {ms:.3f} |")


def kernel_counters(dirs: list[str], needle: str, waves: int = 2) -> int:
    """SQ counters of every dispatch whose name contains `needle`, per counter
    folder, merged per dispatch order, with derived ratios."""
    merged: dict[str, dict[str, float]] = {}
    for d in dirs:
        for k, c in _counters(Path(d)).items():
            if needle in k:
                merged.setdefault(k.split(" [dispatch")[0], {}).update(c)
    if not merged:
        raise KeyError(f"no dispatch matching {needle}")
    for name, c in merged.items():
        print(f"Counters of `{_short(name)}`:")
        print()
        print("| counter | value |")
        print("|---|---|")
        for k, v in sorted(c.items()):
            print(f"| {k} | {v:.6e} |")
        print()
        print("| quantity | value | how |")
        print("|---|---|---|")
        for q, v, how in derived(c, waves):
            print(f"| {q} | {v} | {how} |")
        print()
    return 0


def summarize(dirs: list[str], min_ms: float = 0.0) -> int:
    trace = Path(dirs[0])
    if min_ms > 0.0:
        long_dispatches(trace, min_ms)
        return 0
    print("| kernel | calls | avg duration (ms) | total (ms) | share |")
    print("|---|---|---|---|---|")
    stats = _csv_rows(trace, "kernel_stats.csv")
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:8]:
        print(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e6:.3f} | "
              f"{float(r['TotalDurationNs']) / 1e6:.3f} | {float(r['Percentage']):.2f}% |")
    if len(dirs) == 1:
        return 0
    fetch, write, *extra = [Path(d) for d in dirs[1:]]
    fc, wc = _counters(fetch), _counters(write)
    copy_bytes = 2 * (1 << 30)
    f_copy = _pick(fc, COPY, "FETCH_SIZE")
    w_copy = _pick(wc, COPY, "WRITE_SIZE")
    f_int = _pick(fc, INTEGRATOR, "FETCH_SIZE")
    w_int = _pick(wc, INTEGRATOR, "WRITE_SIZE")
    f_factor = copy_bytes / f_copy
    w_factor = copy_bytes / w_copy
    fetch_b = f_int * f_factor
    write_b = w_int * w_factor
    alg = algorithmic_bytes(1 << 20, 1001)
    print()
    print("| quantity | value |")
    print("|---|---|")
    print(f"| calibration copy FETCH_SIZE (2 GiB read) | {f_copy:.1f} -> {f_factor:.4f} bytes/unit |")
    print(f"| calibration copy WRITE_SIZE (2 GiB written) | {w_copy:.1f} -> {w_factor:.4f} bytes/unit |")
    print(f"| integrator FETCH_SIZE | {f_int:.1f} -> {fetch_b / 1e9:.3f} GB |")
    print(f"| integrator WRITE_SIZE | {w_int:.1f} -> {write_b / 1e9:.3f} GB |")
    print(f"| traffic per launch (fetch + write) | {(fetch_b + write_b) / 1e9:.3f} GB |")
    print(f"| algorithmic bytes per launch | {alg / 1e9:.3f} GB |")
    print(f"| traffic / algorithmic | {(fetch_b + write_b) / alg:.3f} |")
    print()
    print(f"traffic_bytes_per_launch: {int(round(fetch_b + write_b))}")
    for d in extra:
        cc = _counters(d)
        print()
        print(f"Counters from `{d.name}` (integrator dispatch):")
        print()
        print("| counter | value |")
        print("|---|---|")
        for k, c in cc.items():
            if INTEGRATOR in k:
                for name, v in sorted(c.items()):
                    print(f"| {name} | {v:.6e} |")
                break
    return 0


def main() -> int:
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--n", type=int, default=1 << 20)
    p.add_argument("--schedule", type=int, default=1, help="0 plain, 1 persistent")
    p.add_argument("--block", type=int, default=0,
                   help="threads per block (0: the library default, 512 for this launch)")
    p.add_argument("--rtol", type=float, default=0.0,
                   help="rtol = atol of the 1M-orbit launch (0: its 1e-12)")
    p.add_argument("--mixed", action="store_true", help="mixed precision (FP32 stage forces)")
    p.add_argument("--staged", action="store_true",
                   help="staged output (orbit-blocked samples, then the transpose)")
    p.add_argument("--copy-gib", type=float, default=2.0, help="0 skips the calibration copy")
    p.add_argument("--summarize", nargs="+", metavar="DIR",
                   help="trace folder, then optionally the fetch and write folders and more "
                        "counter folders")
    p.add_argument("--ragged-jobs", type=int, default=0,
                   help="run the first N jobs of the ragged mix (one drift_ragged_device call) "
                        "instead of the 1M-orbit launch")
    p.add_argument("--dopr54-uniform", type=int, default=0, metavar="N",
                   help="run DOPR54 on N disk orbits under each ragged-mix potential instead "
                        "of the 1M-orbit launch")
    p.add_argument("--kernel", metavar="NAME",
                   help="with --summarize: the SQ counters of the dispatches whose name contains "
                        "NAME, from every folder given, with derived ratios")
    p.add_argument("--waves", type=int, default=2,
                   help="with --kernel: the kernel's resident waves per SIMD (for VALU busy)")
    p.add_argument("--min-ms", type=float, default=0.0,
                   help="with --summarize and a trace folder alone: only dispatches of at "
                        "least this many ms, grouped and in launch order")
    args = p.parse_args()
    if args.summarize and args.kernel:
        return kernel_counters(args.summarize, args.kernel, args.waves)
    if args.summarize:
        if len(args.summarize) == 2:
            p.error("--summarize takes the trace folder, then the fetch and write folders")
        if args.min_ms > 0.0 and len(args.summarize) != 1:
            p.error("--min-ms takes the trace folder alone")
        return summarize(args.summarize, args.min_ms)
    return run(args)


if __name__ == "__main__":
    sys.exit(main())

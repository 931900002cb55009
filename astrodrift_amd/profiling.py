"""One profiling pass of the benchmark kernel, for rocprofv3.

    rocprofv3 --kernel-trace --stats -- python3 -m astrodrift_amd.profiling
    rocprofv3 --pmc FETCH_SIZE -- python3 -m astrodrift_amd.profiling
    rocprofv3 --pmc WRITE_SIZE -- python3 -m astrodrift_amd.profiling

Launches, in this order on one stream: a calibration copy of a known size
(torch's copy of a 2 GiB float64 tensor: 2 GiB read, 2 GiB written), then
one integration of the 1,048,576-orbit Milky-Way DOP853 workload (rtol =
atol = 1e-12, 1,001 outputs over t = 0..50) with the persistent schedule.
The copy's counter values give the factor that turns FETCH_SIZE / WRITE_SIZE
into bytes for the integrator's dispatch.
"""

from __future__ import annotations

import argparse
import math
import sys


def main() -> int:
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--n", type=int, default=1 << 20)
    p.add_argument("--schedule", type=int, default=1, help="0 plain, 1 persistent")
    p.add_argument("--copy-gib", type=float, default=2.0, help="0 skips the calibration copy")
    args = p.parse_args()

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

    w = workloads.mw_disk(n=args.n)
    s0 = torch.from_numpy(w.state0).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    out = torch.empty((w.nt, w.n, 6), dtype=torch.float64, device=dev)
    st = torch.zeros((w.n, 4), dtype=torch.int32, device=dev)
    lt = math.log(w.rtol)
    _native.launch_device(_native.METHOD_DOP853, s0.data_ptr(), tt.data_ptr(), out.data_ptr(),
                          w.n, w.nt, lt, lt, w.potential, st.data_ptr(), schedule=args.schedule)
    torch.cuda.synchronize()
    acc = int(st[:, 0].sum().item())
    algorithmic_bytes = 48 * w.n + 8 * w.nt + 48 * w.n * w.nt + 16 * w.n
    print(f"[profiling] integrator: n={w.n} nt={w.nt} accepted={acc} "
          f"algorithmic bytes={algorithmic_bytes}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

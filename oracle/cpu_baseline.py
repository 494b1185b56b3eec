"""CPU baseline of the SpMV (test infrastructure, run by bench.py's
cpu_baseline leg as a child process): the oracle's OpenMP restatement of
monolish's OSS CRS row loop (crs_matvec.hpp:101-108) timed on this host with
pinned threads, so that the number does not move with the scheduler.

    python -m oracle.cpu_baseline --matrix 7pt192 --seconds 10

The parent sets nothing; this script derives the thread count (physical
cores visible to the process, capped by OMP_NUM_THREADS when set), exports
OMP_NUM_THREADS / OMP_PROC_BIND=close / OMP_PLACES=cores before libgomp is
loaded, builds the matrix with the library's host generators, and prints one
JSON line: GB/s of algorithmic bytes (median of >= 5 SpMVs after 1
warm-up), threads, CPU model and the sample.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import subprocess
import sys
import time


def cpu_topology():
    """(physical cores among the CPUs this process may run on, model name)."""
    allowed = os.sched_getaffinity(0)
    cores = set()
    model = "unknown"
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True,
                             text=True).stdout
        for line in out.splitlines():
            if line.startswith("#"):
                continue
            cpu, core, sock = (int(v) for v in line.split(",")[:3])
            if cpu in allowed:
                cores.add((sock, core))
        for line in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, ValueError):
        pass
    return (len(cores) or len(allowed)), model


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--matrix", default="7pt192", choices=["7pt192"])
    ap.add_argument("--seconds", type=float, default=10.0)
    args = ap.parse_args()
    phys, model = cpu_topology()
    cap = os.environ.get("OMP_NUM_THREADS")
    threads = min(phys, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else phys
    # before the oracle (and with it libgomp) is loaded
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ["OMP_PROC_BIND"] = "close"
    os.environ["OMP_PLACES"] = "cores"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import numpy as np  # noqa: F401

    import monolish_amd as m  # host generators only: no device call here
    import oracle

    rp, ci, val = m.stencil3d(192, 192, 192, 7)
    x = m.random_vector(len(rp) - 1, -1.0, 1.0, 2001)
    rows, nnz = len(rp) - 1, len(val)
    nbytes = nnz * 12 + (rows + 1) * 4 + 2 * rows * 8
    oracle.spmv(rp, ci, val, x)  # warm-up (first touch, thread start)
    ts = []
    t_end = time.perf_counter() + args.seconds
    while len(ts) < 5 or (time.perf_counter() < t_end and len(ts) < 200):
        t0 = time.perf_counter()
        oracle.spmv(rp, ci, val, x)
        ts.append(time.perf_counter() - t0)
    med = statistics.median(ts)
    print(json.dumps({
        "value": round(nbytes / med / 1e9, 2), "unit": "GB/s", "cores": oracle.num_threads(),
        "kind": "port",
        "sample": (f"oracle (CPU restatement of the reference OpenMP CRS row loop) on the full "
                   f"192^3 7-point FP64 matrix, median of {len(ts)} SpMVs after 1 warm-up "
                   f"({med * 1e3:.1f} ms each); {oracle.num_threads()} threads, "
                   f"OMP_PROC_BIND=close OMP_PLACES=cores, {phys} physical cores visible, "
                   f"CPU: {model}"),
        "cpu_model": model, "physical_cores_visible": phys,
        "spmv_ms_median": round(med * 1e3, 3), "spmv_ms_min": round(min(ts) * 1e3, 3),
        "spmv_ms_max": round(max(ts) * 1e3, 3)}))


if __name__ == "__main__":
    main()

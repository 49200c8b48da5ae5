"""Benchmark of the MOCASSIN Monte Carlo iteration on one MI355X.

A step is one transport pass of the workload's packets (energyPacketDriver)
over the grid with inputs resident in HBM.  The timed region brackets K
steps with device synchronisation; `value` is packets/s of the mixed-precision
path, the FP64 path is timed beside it, and the oracle (plain C, FP64) is
timed on a bounded sample of the same workload on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload nebula32]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def build_model(name: str):
    from mocassin_xray_amd import nebula
    if name == "nebula32":
        return nebula.frozen_nebula(n=32, nbins=1000, n_packets=10_000_000)
    if name == "hsphere16":
        return nebula.frozen_hsphere(n=16, nbins=400, n_packets=1_000_000)
    raise SystemExit(f"unknown workload {name}")


def transport_bytes(st: dict, nbins: int, real_bytes: int, n: int) -> float:
    """Algorithmic HBM bytes of n packets: per segment kappa + cell id + J RMW;
    per emission totalLines + CDF bisection; per escape the tally RMW."""
    seg = real_bytes + 4 + 16
    emit = 4 + real_bytes * math.ceil(math.log2(nbins))
    return st["segments"] * seg + st["emissions"] * emit + st["escaped"] * 16.0


def profiled_traffic(workload: str):
    """HBM bytes per launch of k_transport<float> measured by the FETCH_SIZE and
    WRITE_SIZE counter runs recorded in profiles/ (None if not profiled)."""
    f = ROOT / "profiles" / f"transport_{workload}.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def time_path(m, prec: str, steps: int, warmup: int, n: int):
    import torch
    from mocassin_xray_amd import MocGPU
    g = MocGPU(m.tables, prec)
    g.grid_attach(m.grid)
    for k in ("opacity", "Ne", "recPDF", "totalLines"):
        g.upload(k, m.fields[k])
    for w in range(warmup):
        g.reset_tallies()
        g.energyPacketDriver(1, n, seed=1000 + w, ctr_base=w)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for s in range(steps):
        g.reset_tallies()
        stats.append(g.energyPacketDriver(1, n, seed=2000 + s, ctr_base=s))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    g.close()
    return dt, stats


def copy_calibration(nbytes: int = 1 << 30, reps: int = 3):
    """Plain device copy of known size: FETCH_SIZE/WRITE_SIZE of this kernel
    calibrate the counters (bytes per KiB reported)."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copy_calibration: {reps} copies of {nbytes} bytes", file=sys.stderr, flush=True)


def cpu_baseline(m, target_s: float = 12.0):
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    threads = min(os.cpu_count() or 1, 16)
    n = 20000
    t0 = time.perf_counter()
    pyoracle.transport(m.tables, m.grid, m.fields, n, 5, nthreads=threads)
    dt = time.perf_counter() - t0
    n2 = int(min(max(n * target_s / max(dt, 1e-3), n), 5_000_000))
    t0 = time.perf_counter()
    pyoracle.transport(m.tables, m.grid, m.fields, n2, 6, nthreads=threads)
    dt2 = time.perf_counter() - t0
    return {"value": n2 / dt2, "unit": "packets/s", "cores": threads, "kind": "port",
            "sample": f"{n2} packets of the same workload, oracle transport (C, FP64, OpenMP)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="nebula32")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["mixed", "fp64"], default=None,
                    help="time one path only (for profiler runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain 1 GiB device copy (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the transport shards by packets; "
              "see the sharding design); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.init()
    if args.copy_calibration:
        copy_calibration()
        return
    m = build_model(args.workload)
    n = m.n_packets
    print(f"bench: workload {m.name} cells={m.grid.nCells} bins={m.tables.nbins} packets={n}",
          file=sys.stderr, flush=True)
    if args.only != "fp64":
        dt_mix, st_mix = time_path(m, "mixed", args.steps, args.warmup, n)
        print(f"bench: mixed {args.steps * n / dt_mix:.3e} packets/s", file=sys.stderr,
              flush=True)
    if args.only != "mixed":
        dt_64, st_64 = time_path(m, "fp64", args.steps, args.warmup, n)
        print(f"bench: fp64  {args.steps * n / dt_64:.3e} packets/s", file=sys.stderr,
              flush=True)
    if args.only is not None:
        return
    value = args.steps * n / dt_mix
    kms = sum(s["kernel_ms"] for s in st_mix) / len(st_mix)
    bytes_mix = sum(transport_bytes(s, m.tables.nbins, 4, n) for s in st_mix) / len(st_mix)
    achieved = bytes_mix / (kms * 1e-3) / 1e9
    seg_per_pkt = sum(s["segments"] for s in st_mix) / (len(st_mix) * n)
    emit_per_pkt = sum(s["emissions"] for s in st_mix) / (len(st_mix) * n)
    out = {
        "metric": "packets/s (transport, mixed precision)",
        "value": value,
        "unit": "packets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_mix / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64acc",
        "data": "synthetic",
        "config": {"workload": m.name, "cells": m.grid.nCells, "nbins": m.tables.nbins,
                   "packets": n, "fields": "frozen Stromgren-sphere opacity/emission",
                   "parallelism": "single GPU"},
        "fp64_packets_per_s": args.steps * n / dt_64,
        "mixed_over_fp64": dt_64 / dt_mix,
        "segments_per_packet": seg_per_pkt,
        "emissions_per_packet": emit_per_pkt,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic(args.workload),
                     "kernel": "k_transport<float>", "kernel_ms": kms},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(m)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

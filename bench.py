"""Benchmark of the MOCASSIN Monte Carlo iteration on one MI355X.

A step is one Monte Carlo iteration of the workload (iterateMC): opacity and
diffuse emission of every cell, the workload's energy packets through the
grid, normalisation of J and the ionisation/thermal balance of every cell,
with all state resident in HBM.  W untimed iterations run first from the
reference's starting guess, then exactly K iterations are timed between
device synchronisations.  With --graph, iterations from the second on replay
one captured HIP graph of the iteration (moc_set_graph), and one more eager
iteration after the timed ones gives the per-phase kernel times.  `value` is packets/s of the whole iteration on the
mixed-precision path (FP32 transport, FP64 balance); the all-FP64 path runs
the same iterations beside it, and the difference of the two states after the
same iterations is reported.  The oracle (plain C, FP64, OpenMP) is timed on
a bounded sample of the same workload on the host cores.

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
SEED = 20261203


def build_model(name: str):
    from mocassin_xray_amd import nebula
    if name == "nebula32":
        return nebula.gas_nebula(n=32, nbins=1000, n_packets=10_000_000)
    if name == "hsphere16":
        return nebula.gas_hsphere(n=16, nbins=400, n_packets=1_000_000)
    if name == "dustneb64":
        return nebula.gas_dust_nebula(n=64, nbins=700, n_packets=100_000_000)
    if name == "xray128":
        return nebula.xray_nebula(n=128, nbins=2000, n_packets=1_000_000_000)
    if name == "xray32":
        return nebula.xray_nebula(n=32, nbins=2000, n_packets=10_000_000)
    if name == "batch256":
        return nebula.batch_nebulae(n_grids=256, nbins=400)
    raise SystemExit(f"unknown workload {name}")


def transport_bytes(segments: float, emissions: float, escaped: float, nbins: int,
                    real_bytes: int, dense: bool) -> float:
    """Algorithmic HBM bytes of a transport launch: per segment kappa + J
    read-modify-write (+ the cell id from the active map unless every cell
    is active and the id is computed); per emission totalLines + the CDF
    bisection; per escape the tally read-modify-write."""
    seg = real_bytes + (0 if dense else 4) + 16
    emit = 4 + real_bytes * math.ceil(math.log2(nbins))
    return segments * seg + emissions * emit + escaped * 16.0


def dense_grid(g) -> bool:
    """Every cell active and numbered 1 + iz + nz*(iy + ny*ix): the transport
    then computes cell ids instead of reading the active map."""
    import numpy as np
    nx, ny, nz = g.active.shape
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    return bool(np.array_equal(g.active, 1 + iz + nz * (iy + ny * ix)))


def profiled(workload: str, key: str):
    """A figure of the k_transport<float> counter runs recorded in profiles/
    (traffic_bytes_per_launch: HBM bytes per launch from FETCH_SIZE and
    WRITE_SIZE; l2_hit_rate: TCC hits / requests), None if not profiled."""
    f = ROOT / "profiles" / f"transport_{workload}.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def run_path(m, prec: str, steps: int, warmup: int, seed: int = SEED, outputs: bool = True,
             graph: bool = False, tune: bool = True):
    """W warmup and K timed iterations; with graph, iterations from the second
    on replay one captured graph (captured in the warmup when W >= 2).  With
    tune, the transport's launch parameters come from the package's
    autotuner, run on the model after the last warmup iteration (the first
    with graph; outside the timed region; the parameters change no packet).  After the state is read,
    one more iteration runs eagerly for the phase times."""
    import torch
    from mocassin_xray_amd import MocGPU
    with MocGPU(m.tables, prec) as g:
        m.load(g)
        g.set_graph(graph)
        tuning = g.get_tuning()
        tuned = None
        for w in range(1, warmup + 1):
            g.iterateMC(w, seed)
            print(f"bench: {prec} warmup iteration {w} done", file=sys.stderr, flush=True)
            # after the last warmup iteration (the state the timed ones start
            # from); with graph replay after the first, so that the capture in
            # the second keeps the tuned parameters
            if tune and w == (1 if graph else warmup):
                from mocassin_xray_amd.autotune import tune_transport
                npk = int(min(float(m.tables.nPhotons[0]), 2e7))
                r = tune_transport(g, packets=npk, repeats=2)
                tuning = g.get_tuning()
                tuned = {"default_ms": r["default_ms"], "tuned_ms": r["ms"], "packets": npk,
                         "seconds": r["seconds"]}
                print(f"bench: {prec} transport tuned {r['best']} ({r['default_ms']:.2f} -> "
                      f"{r['ms']:.2f} ms per {npk} packets)", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        stats = []
        t0 = time.perf_counter()
        for s in range(warmup + 1, warmup + steps + 1):
            stats.append(g.iterateMC(s, seed))
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        keys = ["Te", "Ne", "ionDen", "Jste", "opacity", "recPDF", "totalLines"]
        if getattr(m, "dust", None) is not None:
            keys += ["Tdust", "scaOpac"]
        if big_model(m):               # (cell, bin) planes of tens of GB stay on the device
            keys = ["Te", "Ne", "ionDen"]
        state = {k: g.download(k) for k in keys}
        state["tuning"] = tuning
        state["autotune"] = tuned
        if graph:  # the phases of one more iteration, eager (a graph replay times only the whole)
            g.set_graph(False)
            state["eager_stats"] = g.iterateMC(warmup + steps + 1, seed)
        if outputs and not hasattr(m, "grids"):  # the output reductions take one nebula per context
            out_ms = {}
            outs = [("volume_averages", g.volume_averages), ("line_luminosities", g.line_luminosities),
                    ("sed", g.sed)]
            if big_model(m):  # the line luminosities sweep every cell once per ion group
                outs = [outs[0], outs[2]]
            for name, fn in outs:
                t1 = time.perf_counter()
                res = fn()
                out_ms[name] = (time.perf_counter() - t1) * 1e3
                if name == "volume_averages":
                    state["volavg"] = res
            state["output_ms"] = out_ms
        elif not hasattr(m, "grids"):
            state["volavg"] = g.volume_averages()
    if graph:
        # phase times and the transport counters of the same eager launch
        stats = [dict(st, **{k: state["eager_stats"][k] for k in (
            "opacity_ms", "emission_ms", "transport_ms", "update_ms", "segments", "emissions",
            "escaped")}) for st in stats]
    return dt, stats, state


def n_cells(m) -> int:
    return sum(g.nCells for g in m.grids) if hasattr(m, "grids") else m.grid.nCells


def big_model(m) -> bool:
    """More than 5e8 (cell, bin) entries: its planes are not copied to the host."""
    return (n_cells(m) + 1) * m.tables.nbins > 500_000_000


def state_error(m, a: dict, b: dict) -> dict:
    """Difference of two model states: volume-averaged T_e and ion fractions
    (weights n_H dV), and a chi2-like J distance over cells."""
    import numpy as np
    from mocassin_xray_amd import tables as T
    grids = m.grids if hasattr(m, "grids") else [m.grid]
    dV = np.concatenate([T.cell_volumes(g.xAxis, g.yAxis, g.zAxis, g.active, True)[1:]
                         for g in grids])
    w = dV * m.state["Hden"][1:]
    Ta, Tb = a["Te"][1:], b["Te"][1:]
    te = abs(np.sum(w * Ta) - np.sum(w * Tb)) / np.sum(w * Tb)
    Xa, Xb = a["ionDen"][1:], b["ionDen"][1:]
    xa = np.einsum("c,ces->es", w, Xa) / w.sum()
    xb = np.einsum("c,ces->es", w, Xb) / w.sum()
    big = xb > 1e-3
    xrel = float(np.max(np.abs(xa[big] - xb[big]) / xb[big]))
    out = {"Te_volavg_rel": float(te), "ion_frac_volavg_max_rel": xrel, "J_cell_rel_rms": None}
    if "volavg" in a and "volavg" in b:
        # outputGas definitions: <Te> per ion and <X> (Harrington et al. 1982)
        (Ta_i, Xa_i), (Tb_i, Xb_i) = a["volavg"], b["volavg"]
        on = Xb_i > 1e-3
        out["ionic_Te_max_rel"] = float(np.max(np.abs(Ta_i[on] - Tb_i[on]) / Tb_i[on]))
        out["ionic_frac_max_rel"] = float(np.max(np.abs(Xa_i[on] - Xb_i[on]) / Xb_i[on]))
    if "Jste" in a and "Jste" in b:
        Ja, Jb = a["Jste"][1:].sum(axis=1), b["Jste"][1:].sum(axis=1)
        hit = (Ja + Jb) > 0
        chi = float(np.mean((Ja[hit] - Jb[hit]) ** 2 / (0.5 * (Ja[hit] + Jb[hit])) ** 2))
        out["J_cell_rel_rms"] = math.sqrt(chi)
    return out


def atomic_calibration(nbytes: int = 1 << 28) -> float | None:
    """The device's rate of scattered no-return FP64 atomic additions
    (moc_calibrate_atomics) into a buffer the size of a J plane."""
    import ctypes as C
    from mocassin_xray_amd import _lib
    rate = C.c_double(0.0)
    if _lib.lib().moc_calibrate_atomics(nbytes, C.byref(rate)) != 0:
        return None
    return rate.value


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


def cpu_baseline_batch(m, n_members: int = 8):
    """The oracle on a bounded sample of a batch: the first n_members grids,
    each through one whole iteration on its own from its starting state
    (opacity, emission, its packets, J normalisation, balance), scaled by
    packets to the whole batch."""
    import numpy as np
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as O
    threads = min(os.cpu_count() or 1, 16)
    E, Z, NS = m.elems, m.abund, m.nstages
    t_tot, pk = 0.0, 0
    for k in range(n_members):
        mk = m.member(k)
        s0 = mk.state
        t0 = time.perf_counter()
        op = O.opacity(m.xsec, m.tables.KNsigmaT, True, E, Z, NS, s0["ionDen"], s0["Ne"], s0["Te"],
                       s0["Hden"])
        em = O.emission(m.et, m.bt, E, Z, NS, s0["Te"], s0["Ne"], s0["ionDen"], nthreads=threads)
        fields = {"opacity": op, "Ne": s0["Ne"], "recPDF": em["recPDF"],
                  "totalLines": em["totalLines"]}
        J = O.transport(mk.tables, mk.grid, fields, mk.n_packets, 5, nthreads=threads)[0]
        J *= 1.e-9 / 8.0 if mk.grid.lgSymmetricXYZ else 1.e-9
        O.update(m.bt, E, Z, NS, True, J, s0["Te"], s0["Ne"], s0["Hden"], s0["ionDen"],
                 nthreads=threads)
        t_tot += time.perf_counter() - t0
        pk += mk.n_packets
    return {"value": pk / t_tot, "unit": "packets/s", "cores": threads, "kind": "port",
            "sample": (f"the first iteration of the first {n_members} of {len(m.grids)} grids of "
                       f"{m.name}, each run whole by the oracle (C, FP64, OpenMP): {pk} packets, "
                       f"{sum(g.nCells for g in m.grids[:n_members])} cells")}


def cpu_baseline(m, gs: dict, target_s: float = 20.0):
    """The oracle's iteration on the host cores, timed on a bounded sample:
    opacity, emission and balance on a subset of the cells and the transport
    of a subset of the packets over the full grid (with the device's last
    fields), each scaled to the whole iteration; reported as packets/s of the
    whole iteration."""
    import numpy as np
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as O
    threads = min(os.cpu_count() or 1, 16)
    C1 = m.grid.nCells + 1
    E, Z, NS = m.elems, m.abund, m.nstages
    k = max(C1 // 1024, 1)
    sub = np.r_[0, np.arange(1, C1, k)]
    scale = (C1 - 1) / (sub.size - 1)
    part = lambda a: np.ascontiguousarray(a[sub])
    H = m.state["Hden"]
    dust = getattr(m, "dust", None)
    nu, wf = m.tables.nuArray, m.tables.widFlx
    t0 = time.perf_counter()
    gas = O.opacity(m.xsec, m.tables.KNsigmaT, True, E, Z, NS, part(gs["ionDen"]),
                    part(gs["Ne"]), part(gs["Te"]), part(H))
    if dust is not None:
        O.dust_opacity(dust, nu, wf, part(m.Ndust), part(gs["Tdust"]), gasOpac=gas)
    t_op = (time.perf_counter() - t0) * scale
    t0 = time.perf_counter()
    O.emission(m.et, m.bt, E, Z, NS, part(gs["Te"]), part(gs["Ne"]), part(gs["ionDen"]),
               nthreads=threads,
               dust=None if dust is None else (dust, part(gs["Tdust"]), part(m.Ndust), part(H)))
    t_em = (time.perf_counter() - t0) * scale
    t0 = time.perf_counter()
    grains = None
    if dust is not None and getattr(m, "photoelectric", False):
        grains = (dust, part(gs["Tdust"]), part(m.Ndust))
    up = O.update(m.bt, E, Z, NS, True, part(gs["Jste"]), part(gs["Te"]), part(gs["Ne"]),
                  part(H), part(gs["ionDen"]), nthreads=threads, grains=grains)
    if dust is not None:
        O.dust_temperature(dust, nu, wf, part(m.Ndust), part(gs["Tdust"]), part(gs["Jste"]),
                           dustOnly=False, collD=up.get("collD"), peD=up.get("peD"))
    t_up = (time.perf_counter() - t0) * scale
    fields = {"opacity": gs["opacity"], "Ne": gs["Ne"], "recPDF": gs["recPDF"],
              "totalLines": gs["totalLines"]}
    if dust is not None:
        fields.update(scaOpac=gs["scaOpac"], dustOnly=False)
    n = 20000
    t0 = time.perf_counter()
    O.transport(m.tables, m.grid, fields, n, 5, nthreads=threads)
    dt = time.perf_counter() - t0
    n2 = int(min(max(n * target_s / max(dt, 1e-3), n), 5_000_000, m.n_packets))
    t0 = time.perf_counter()
    O.transport(m.tables, m.grid, fields, n2, 6, nthreads=threads)
    t_tr = (time.perf_counter() - t0) * m.n_packets / n2
    t_iter = t_op + t_em + t_tr + t_up
    return {"value": m.n_packets / t_iter, "unit": "packets/s", "cores": threads, "kind": "port",
            "sample": (f"one iteration of {m.name} extrapolated from the oracle (C, FP64, OpenMP) "
                       f"on {sub.size - 1} of {C1 - 1} cells (opacity, emission, balance) and "
                       f"{n2} of {m.n_packets} packets (transport)"),
            "phase_s": {"opacity": t_op, "emission": t_em, "transport": t_tr, "balance": t_up},
            "balance_cells_per_s": (C1 - 1) / t_up}


def cpu_baseline_big(m, gs: dict, st_mix: list, target_s: float = 20.0):
    """The oracle's iteration for a model too large to copy its (cell, bin)
    planes to the host: opacity, emission and balance are timed on 1024
    sampled cells of the device state and scaled to every cell; the transport
    is timed per segment on the same model at 32^3 (its fields from one device
    iteration) and scaled by the segments this model's packets walk (device
    counters).  Extrapolated, as labelled in `sample`."""
    import numpy as np
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as O
    from mocassin_xray_amd import MocGPU, nebula
    threads = min(os.cpu_count() or 1, 16)
    C1 = m.grid.nCells + 1
    E, Z, NS = m.elems, m.abund, m.nstages
    k = max(C1 // 1024, 1)
    sub = np.r_[0, np.arange(1, C1, k)]
    scale = (C1 - 1) / (sub.size - 1)
    part = lambda a: np.ascontiguousarray(a[sub])
    H = m.state["Hden"]
    t0 = time.perf_counter()
    O.opacity(m.xsec, m.tables.KNsigmaT, True, E, Z, NS, part(gs["ionDen"]), part(gs["Ne"]),
              part(gs["Te"]), part(H))
    t_op = (time.perf_counter() - t0) * scale
    t0 = time.perf_counter()
    O.emission(m.et, m.bt, E, Z, NS, part(gs["Te"]), part(gs["Ne"]), part(gs["ionDen"]), nthreads=threads)
    t_em = (time.perf_counter() - t0) * scale
    # the balance needs J of the sampled cells: from the small model below
    small = nebula.xray_nebula(n=32, nbins=m.tables.nbins, n_packets=1_000_000)
    with MocGPU(small.tables, "fp64") as g:
        small.load(g)
        g.iterateMC(1, SEED)
        fs = {k2: g.download(k2) for k2 in ("opacity", "recPDF", "totalLines", "Jste", "Ne", "Te",
                                              "ionDen")}
    Cs = small.grid.nCells + 1
    ks = max(Cs // 1024, 1)
    subs = np.r_[0, np.arange(1, Cs, ks)]
    t0 = time.perf_counter()
    O.update(small.bt, E, Z, NS, True, np.ascontiguousarray(fs["Jste"][subs]), fs["Te"][subs],
             fs["Ne"][subs], small.state["Hden"][subs], np.ascontiguousarray(fs["ionDen"][subs]),
             nthreads=threads)
    t_up = (time.perf_counter() - t0) * (C1 - 1) / (subs.size - 1)
    fields = {k2: fs[k2] for k2 in ("opacity", "Ne", "recPDF", "totalLines")}
    n = 20000
    t0 = time.perf_counter()
    O.transport(small.tables, small.grid, fields, n, 5, nthreads=threads)
    dt = time.perf_counter() - t0
    n2 = int(min(max(n * target_s / max(dt, 1e-3), n), 2_000_000))
    t0 = time.perf_counter()
    _, _, cnt, _ = O.transport(small.tables, small.grid, fields, n2, 6, nthreads=threads)
    per_seg = (time.perf_counter() - t0) / cnt["segments"]
    segs = mean(st_mix, "segments")
    t_tr = per_seg * segs
    t_iter = t_op + t_em + t_tr + t_up
    return {"value": m.n_packets / t_iter, "unit": "packets/s", "cores": threads, "kind": "port",
            "sample": (f"one iteration of {m.name} extrapolated from the oracle (C, FP64, OpenMP): "
                       f"opacity and emission on {sub.size - 1} of {C1 - 1} cells of the device state, "
                       f"balance on {subs.size - 1} cells of the same model at 32^3, transport "
                       f"per segment ({n2} packets over the 32^3 model) times the "
                       f"{segs:.3e} segments of this model's iteration"),
            "phase_s": {"opacity": t_op, "emission": t_em, "transport": t_tr, "balance": t_up},
            "balance_cells_per_s": (C1 - 1) / t_up}


def mean(stats, key):
    return sum(s[key] for s in stats) / len(stats)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="nebula32")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-band", action="store_true",
                    help="skip the second FP64 run that measures the Monte Carlo noise band")
    ap.add_argument("--only", choices=["mixed", "fp64"], default=None,
                    help="time one path only (for profiler runs)")
    ap.add_argument("--graph", action="store_true",
                    help="replay iterations 2.. as one captured HIP graph (moc_set_graph)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default transport launch parameters (no autotuner run)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain 1 GiB device copy (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this benchmark measures one GPU only (the transport shards by packets "
              "and the balance by cells, but no multi-GPU path is built); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.init()
    if args.copy_calibration:
        copy_calibration()
        return
    m = build_model(args.workload)
    n = m.n_packets
    C = n_cells(m)
    print(f"bench: workload {m.name} cells={C} bins={m.tables.nbins} packets={n}",
          file=sys.stderr, flush=True)
    res = {}
    for prec in ("mixed", "fp64"):
        if args.only not in (None, prec):
            continue
        dt, st, state = run_path(m, prec, args.steps, args.warmup, graph=args.graph,
                                 tune=not args.no_tune)
        res[prec] = (dt, st, state)
        print(f"bench: {prec} {args.steps * n / dt:.3e} packets/s, "
              f"{dt / args.steps * 1e3:.1f} ms/iteration", file=sys.stderr, flush=True)
    if args.only is not None:
        return
    dt_mix, st_mix, s_mix = res["mixed"]
    dt_64, st_64, s_64 = res["fp64"]
    band = None
    if not args.no_band:
        # Monte Carlo noise band: the FP64 path again with another packet
        # stream (seed), the same iterations from the same start
        _, _, s_64b = run_path(m, "fp64", args.steps, args.warmup, seed=SEED + 1, outputs=False,
                               graph=args.graph, tune=not args.no_tune)
        print("bench: fp64 second-seed run done (noise band)", file=sys.stderr, flush=True)
        band = state_error(m, s_64b, s_64)
    kms = mean(st_mix, "transport_ms")
    dense = all(dense_grid(g) for g in m.grids) if hasattr(m, "grids") else dense_grid(m.grid)
    bytes_mix = transport_bytes(mean(st_mix, "segments"), mean(st_mix, "emissions"),
                                mean(st_mix, "escaped"), m.tables.nbins, 4, dense)
    achieved = bytes_mix / (kms * 1e-3) / 1e9
    # the transport's other bound: one J tally per path segment, each a
    # scattered FP64 atomic addition unless the LDS hot block takes it; the
    # device's rate of such additions is measured here, live
    atomic_rate = atomic_calibration()
    tallies_per_s = mean(st_mix, "segments") / (kms * 1e-3)
    out = {
        "metric": "packets/s of the Monte Carlo iteration (mixed precision) + balance cells/s",
        "value": args.steps * n / dt_mix,
        "unit": "packets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_mix / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 transport + f64 balance",
        "data": "synthetic",
        "config": {"workload": m.name, "cells": C, "nbins": m.tables.nbins, "packets": n,
                   "elements": m.elems, "nstages": m.nstages,
                   "step": "one MC iteration: opacity, emission, transport, balance",
                   "parallelism": "single GPU"},
        "balance_cells_per_s": C / (mean(st_mix, "update_ms") * 1e-3),
        "transport_packets_per_s": n / (kms * 1e-3),
        "phase_ms": {k: mean(st_mix, k + "_ms") for k in ("opacity", "emission", "transport",
                                                         "update")},
        "fp64": {"packets_per_s": args.steps * n / dt_64,
                 "transport_packets_per_s": n / (mean(st_64, "transport_ms") * 1e-3),
                 "balance_cells_per_s": C / (mean(st_64, "update_ms") * 1e-3),
                 "phase_ms": {k: mean(st_64, k + "_ms") for k in ("opacity", "emission",
                                                                  "transport", "update")}},
        "output_ms": s_mix.get("output_ms"),
        "graph_replay": args.graph,
        "transport_launch": {k: s_mix["tuning"][k] for k in ("emitBatch", "segsPerTrip",
                                                             "transportWaves", "hotJ")},
        "transport_autotune": s_mix.get("autotune"),
        "mixed_over_fp64": dt_64 / dt_mix,
        "mixed_vs_fp64_error": state_error(m, s_mix, s_64),
        "mc_noise_band": band,
        "converged_percent": st_mix[-1]["convPercent"],
        "segments_per_packet": mean(st_mix, "segments") / n,
        "emissions_per_packet": mean(st_mix, "emissions") / n,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled(args.workload, "traffic_bytes_per_launch"),
                     "l2_hit_rate": profiled(args.workload, "l2_hit_rate"),
                     "kernel": "k_transport<float>", "kernel_ms": kms,
                     "tally_bound": {"tallies_per_s": tallies_per_s,
                                     "device_atomic_adds_per_s": atomic_rate,
                                     "frac": tallies_per_s / atomic_rate if atomic_rate else None}},
    }
    if band is not None:
        err = out["mixed_vs_fp64_error"]
        out["mixed_within_band"] = {
            k: bool(err[k] <= max(0.01, 2.0 * band[k])) for k in ("Te_volavg_rel", "ion_frac_volavg_max_rel")}
    if not args.no_cpu and hasattr(m, "grids"):
        out["cpu_baseline"] = cpu_baseline_batch(m)
    elif not args.no_cpu and not big_model(m):
        out["cpu_baseline"] = cpu_baseline(m, s_mix)
    elif not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_big(m, s_mix, st_mix)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

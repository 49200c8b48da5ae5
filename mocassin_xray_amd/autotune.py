"""Launch-parameter tuning of the packet transport on a live model.

The transport kernel has four launch parameters (moc_set_tuning): how many
diffuse packets a wave waits for before it emits them together, how many
path segments a flying lane steps per trip through the wave's loop, how
many waves one launch keeps resident, and (mixed form) whether the J
tallies of the cells around the star are summed in LDS first.  None of them
changes which packets are drawn (every packet keeps its own Philox stream),
so the choice is pure speed and depends on the model: optically thick,
line-rich grids re-emit often and favour larger emission batches; thin grids
favour longer trips; the LDS sums pay where many tallies land near the star.

tune_transport times candidates by coordinate descent (the LDS sums, waves,
emission batch, then segments per trip, in two passes: the parameters
interact -- on the 32^3 nebula the LDS sums pay only with short trips --
so the second pass revisits each one at the others' new values), each with
the same packets of star 1, and leaves the fastest set on the context.  The transport adds to the J and
escape tallies, so the tuner zeroes them when it is done: run it before an
iteration (moc_iterate starts by zeroing them too), not between an iteration
and the reading of its J.
"""
from __future__ import annotations

import json
import time
from pathlib import Path

WAVES = (1024, 2048, 4096, 8192, 16384)
EMIT_BATCH = (8, 16, 24, 32, 48, 64)
SEGS_PER_TRIP = (4, 8, 16, 24, 32, 48, 64)
HOT_J = (0, 1)  # the mixed form's LDS sums of the hot J block off / on


def _time(g, n: int, seed: int, repeats: int) -> float:
    best = float("inf")
    for r in range(repeats):
        st = g.energyPacketDriver(1, n, seed, ctr_base=0xfff00 + r)
        best = min(best, st["kernel_ms"])
    return best


def tune_transport(g, packets: int | None = None, seed: int = 77, repeats: int = 2,
                   store: str | Path | None = None, passes: int = 2) -> dict:
    """Tune the transport launch of MocGPU context g on its attached model.

    packets: packets of star 1 per timed launch (default: its nPhotons, at
    most 2 million); repeats: launches per candidate (the fastest counts);
    store: a JSON file that receives the choice and every timing.
    Returns {"best": {...}, "ms": best time, "default_ms": time of the
    starting parameters, "trials": [...]}."""
    n = int(packets if packets is not None else min(int(g.tables.nPhotons[0]), 2_000_000))
    if n < 1:
        raise ValueError("tune_transport: no packets to time")
    start = g.get_tuning()
    cur = {k: start[k] for k in ("emitBatch", "segsPerTrip", "transportWaves", "hotJ")}
    trials = []
    t0 = time.perf_counter()

    def trial(params: dict) -> float:
        g.set_tuning(**params)
        ms = _time(g, n, seed, repeats)
        trials.append({**params, "ms": ms})
        return ms

    default_ms = trial(cur)
    best_ms = default_ms
    seen = {tuple(sorted(cur.items())): default_ms}
    for _ in range(max(1, passes)):
        for key, values in (("hotJ", HOT_J), ("transportWaves", WAVES), ("emitBatch", EMIT_BATCH),
                            ("segsPerTrip", SEGS_PER_TRIP)):
            for v in values:
                if v == cur[key]:
                    continue
                cand = {**cur, key: v}
                sig = tuple(sorted(cand.items()))
                ms = seen[sig] if sig in seen else trial(cand)
                seen[sig] = ms
                if ms < best_ms:
                    best_ms, cur = ms, cand
    g.set_tuning(**cur)
    g.reset_tallies()
    out = {"best": cur, "ms": best_ms, "default_ms": default_ms, "packets": n,
           "seconds": time.perf_counter() - t0, "trials": trials}
    if store is not None:
        Path(store).write_text(json.dumps(out, indent=1))
    return out

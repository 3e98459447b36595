"""Launch-geometry autotuner for the IS-SPA passes.

The field pass splits each window's atoms into `splits` pieces and the force
pass splits each window's MC points into `chunks` pieces; both trade
parallelism (enough waves for 256 CUs) against the FP64 partial traffic and
the reduction work.  The kernels' work items are equal, so a grid runs in
whole rounds of the 8192 resident-wave slots (256 CUs x 4 SIMDs x 8 waves):
the wave targets tried are multiples of a round.  The library's automatic
choice aims at a target wave count shared among windows in proportion to
their work; this tuner measures
candidates on the device with isspa_model_profile (HIP events around
back-to-back launches of each stage) and keeps the fastest.

Two candidate families:
  * target wave counts (the automatic per-window choice at several targets),
    the only sensible knob for batches of uneven windows;
  * explicit (splits, chunks) pairs around the automatic choice, for a single
    system.
The choice is deterministic given the timings; ties keep the earlier
candidate.  Changing the tiling rebuilds the launch tables, so tune before
capturing a graph.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence, Tuple

PASS_STAGES = ("field", "epilogue", "force", "reduce")


@dataclasses.dataclass
class TuneResult:
    kind: str                 # "waves" or "tiling"
    value: Tuple[int, ...]    # (waves,) or (splits, chunks)
    ms: float                 # field + epilogue + force + reduce, per pass
    table: List[Tuple[str, Tuple[int, ...], float]]


def _measure(g, reps: int) -> float:
    prof = g.profile(reps)
    return sum(prof[k] for k in PASS_STAGES)


def tune(g, reps: int = 3, waves: Sequence[int] = (8192, 16384, 24576, 32768, 49152, 65536,
                                                  98304, 131072),
         tilings: Optional[Sequence[Tuple[int, int]]] = None, log=None) -> TuneResult:
    """Pick the fastest launch geometry for GpuSystem `g`'s IS-SPA model and
    leave it applied.  `tilings` defaults to a grid around the automatic
    choice for a single system and to none for a batch."""
    table = []
    for wv in waves:
        g.set_tiling(0, 0)
        g.set_target_waves(wv)
        ms = _measure(g, reps)
        table.append(("waves", (wv,), ms))
        if log:
            log(f"autotune waves={wv}: {ms:.4f} ms  tiling {g.tiling()}")
    if tilings is None and g.windows is None:
        g.set_tiling(0, 0)
        g.set_target_waves(32768)
        s0, c0 = g.tiling()
        tilings = sorted({(max(1, int(s0 * a)), max(1, int(c0 * b)))
                          for a in (0.5, 1.0, 2.0) for b in (0.5, 1.0, 2.0)})
    for sc in tilings or ():
        g.set_tiling(*sc)
        ms = _measure(g, reps)
        table.append(("tiling", tuple(sc), ms))
        if log:
            log(f"autotune tiling={sc}: {ms:.4f} ms")
    best = min(range(len(table)), key=lambda k: (table[k][2], k))
    kind, value, ms = table[best]
    if kind == "waves":
        g.set_tiling(0, 0)
        g.set_target_waves(value[0])
    else:
        g.set_target_waves(32768)
        g.set_tiling(*value)
    return TuneResult(kind, value, ms, table)

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
    which share the parallelism among uneven windows by their work;
  * explicit (splits, chunks) pairs: a x0.5/x1/x2 grid around the tiling of
    the best wave target, plus the pair of the best split count and the best
    chunk count found separately.  The field pass and its epilogue depend only
    on the split count and the force pass and its reduction only on the chunk
    count, so each is searched on its own stages: for a single system over
    counts whose grid fills at least 85% of its last round of wave slots, for
    a batch over x0.25..x2 of the largest window's counts (an explicit tiling
    applies to every window, clamped to its size).
The three fastest candidates are timed again with three times the
repetitions (with `confirm`, the five fastest with eight times) and the
fastest of that second timing is kept; the choice is
deterministic given the timings (ties keep the earlier candidate).  With
`confirm` (engine_step_ms, engine_forces_ms) that second timing covers the
five fastest and measures what the caller will run, such as the MD engine's
graph-replayed step, whose all-pairs kernel overlaps the IS-SPA passes on a
second stream and so favours some tilings over others with the same IS-SPA
time.  Changing the tiling rebuilds the launch tables, so capture the graph
that will be run after tuning (engine_step_ms captures its own, and a later
capture replaces it).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence, Tuple

PASS_STAGES = ("field", "epilogue", "force", "reduce")
FIELD_STAGES = ("field", "epilogue")
FORCE_STAGES = ("force", "reduce")
WAVE_SLOTS = 8192        # 256 CUs x 4 SIMDs x 8 resident waves
WAVES_PER_BLOCK = 4      # 256-thread blocks


@dataclasses.dataclass
class TuneResult:
    kind: str                 # "waves" or "tiling"
    value: Tuple[int, ...]    # (waves,) or (splits, chunks)
    ms: float                 # the confirmation timing of the choice, ms
    table: List[Tuple[str, Tuple[int, ...], float]]
    measure: str = "IS-SPA pass"  # what `ms` times: "IS-SPA pass" or the confirmation's measure


def _measure(g, reps: int, stages=PASS_STAGES) -> float:
    prof = g.profile(reps)
    return sum(prof[k] for k in stages)


def engine_forces_ms(g, reps: int) -> float:
    """Mean ms of the MD engine's whole force evaluation (bonded, IS-SPA,
    all-pairs on its side stream, assembly, restraint), timed with HIP events
    around `reps` back-to-back calls on the current stream."""
    import torch
    g.forces()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        g.forces()
    b.record()
    b.synchronize()
    return a.elapsed_time(b) / reps


engine_forces_ms.measure = "engine forces"


def engine_step_ms(g, reps: int) -> float:
    """Mean ms of one graph-replayed MD step: the engine's whole step
    (forces, integration, step advance) captured with md_engine_capture and
    replayed `reps` times, timed with HIP events on the replay stream.  This is
    the work a dynamics run repeats; it ranks tilings differently from the
    eager force evaluation, whose launches and side-stream joins are host
    driven.  Positions, velocities and the step counter are restored after,
    so tuning leaves the system as it found it."""
    import torch
    x0, v0 = g.x.clone(), g.v.clone()
    g.set_step(0)
    g.capture(1)
    g.replay(1)
    side = g._side()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(side)
    g.replay(reps)
    b.record(side)
    b.synchronize()
    ms = a.elapsed_time(b) / reps
    g.x.copy_(x0)
    g.v.copy_(v0)
    g.set_step(0)
    torch.cuda.synchronize()
    return ms


engine_step_ms.measure = "engine step"


def effective_pieces(n: int, pieces: int) -> int:
    """The piece count the library uses for `pieces` requested pieces of `n`
    items: pieces of ceil(n / pieces) items (isspa_model_set_tiling)."""
    pieces = max(1, min(pieces, n))
    per = -(-n // pieces)
    return -(-n // per)


def round_fill(blocks: int, pieces: int) -> float:
    """Fraction of the last round of wave slots that a grid of blocks x pieces
    256-thread blocks fills (1.0 = whole rounds)."""
    rounds = blocks * pieces * WAVES_PER_BLOCK / WAVE_SLOTS
    last = rounds - math.floor(rounds)
    return 1.0 if last == 0.0 else last


def round_aligned(n: int, blocks: int, base: int, lo: float = 0.5, hi: float = 2.0,
                  min_fill: float = 0.85, limit: int = 6) -> List[int]:
    """Piece counts of `n` items between lo*base and hi*base whose grid
    (`blocks` blocks per piece) fills at least `min_fill` of its last round,
    the `limit` nearest to `base`, plus `base` itself."""
    found = set()
    for p in range(max(1, int(base * lo)), max(2, int(base * hi)) + 1):
        pe = effective_pieces(n, p)
        if round_fill(blocks, pe) >= min_fill:
            found.add(pe)
    picked = sorted(found, key=lambda p: (abs(p - base), p))[:limit]
    return sorted(set(picked) | {effective_pieces(n, base)})


def tune(g, reps: int = 3, waves: Sequence[int] = (8192, 16384, 24576, 32768, 49152, 65536,
                                                  98304, 131072),
         tilings: Optional[Sequence[Tuple[int, int]]] = None, log=None,
         confirm=None) -> TuneResult:
    """Pick the fastest launch geometry for GpuSystem `g`'s IS-SPA model and
    leave it applied.  `tilings` defaults to the explicit candidates described
    above; pass () for wave targets only."""
    table = []
    for wv in waves:
        g.set_tiling(0, 0)
        g.set_target_waves(wv)
        ms = _measure(g, reps)
        table.append(("waves", (wv,), ms))
        if log:
            log(f"autotune waves={wv}: {ms:.4f} ms  tiling {g.tiling()}")
    if tilings is None:
        # centre the explicit candidates on the best wave target's tiling (for a
        # batch: the largest window's; an explicit tiling applies to every
        # window, clamped to its size)
        best_wv = min(table, key=lambda t: t[2])[1][0] if table else 32768
        g.set_tiling(0, 0)
        g.set_target_waves(best_wv)
        s0, c0 = g.tiling()
        grid = {(max(1, int(s0 * a)), max(1, int(c0 * b)))
                for a in (0.5, 1.0, 2.0) for b in (0.5, 1.0, 2.0)}
        if g.windows is None:
            n, P = g.system.n_atoms, g.system.n_points
            split_c = round_aligned(n, -(-P // 256), s0)
            chunk_c = round_aligned(P, -(-n // 256), c0)
        else:
            split_c = sorted({max(1, s0 // 4), max(1, s0 // 2), s0, 2 * s0})
            chunk_c = sorted({max(1, c0 // 4), max(1, c0 // 2), c0, 2 * c0})
        best = []
        for kind, cands, stages, pick in (
                ("field-splits", split_c, FIELD_STAGES, lambda v: (v, c0)),
                ("force-chunks", chunk_c, FORCE_STAGES, lambda v: (s0, v))):
            stage_ms = []
            for v in cands:
                g.set_tiling(*pick(v))
                ms = _measure(g, reps, stages)
                stage_ms.append((ms, v))
                table.append((kind, (v,), ms))
                if log:
                    log(f"autotune {kind}={v}: {ms:.4f} ms ({' + '.join(stages)})")
            best.append(min(stage_ms)[1])
        grid.add(tuple(best))
        tilings = sorted(grid)
    for sc in tilings or ():
        g.set_tiling(*sc)
        ms = _measure(g, reps)
        table.append(("tiling", tuple(sc), ms))
        if log:
            log(f"autotune tiling={sc}: {ms:.4f} ms")
    whole = [k for k in range(len(table)) if table[k][0] in ("waves", "tiling")]
    # the leading candidates differ by about the timing noise: time the best
    # three again with three times the repetitions and decide on that
    final = []
    # a confirmation that measures more than the IS-SPA pass re-ranks more
    # candidates: the pass ranking alone is a noisier guide to it
    for k in sorted(whole, key=lambda k: (table[k][2], k))[:5 if confirm else 3]:
        kind, value, _ = table[k]
        _apply(g, kind, value)
        # the whole force evaluation varies more from call to call than the
        # IS-SPA stages: it is timed over 8x the repetitions (16 calls, ~0.14 s
        # at 19,998 atoms)
        ms = confirm(g, 8 * reps) if confirm else _measure(g, 3 * reps)
        final.append((ms, k))
        table.append((kind + "-confirm", value, ms))
        if log:
            log(f"autotune confirm {kind}={value}: {ms:.4f} ms")
    ms, best = min(final)
    kind, value, _ = table[best]
    _apply(g, kind, value)
    return TuneResult(kind, value, ms, table,
                      getattr(confirm, "measure", "confirmation") if confirm else "IS-SPA pass")


def _apply(g, kind: str, value) -> None:
    if kind == "waves":
        g.set_tiling(0, 0)
        g.set_target_waves(value[0])
    else:
        g.set_target_waves(32768)
        g.set_tiling(*value)

"""IS-SPA MD throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 2] [--cpu-threads T] [--no-cpu]

A step is one full MD step of the benchmark workload -- bonded terms, IS-SPA
MC points + field + force passes, all-pairs nonbond, leap-frog + Andersen --
replayed from a captured hipGraph (10 steps per graph when 10 divides the
warm-up and timed step counts, else 1) with every input resident in HBM.
Prints ONE JSON line on stdout (diagnostics go to stderr).

Roofline: the dominant kernel is the IS-SPA force pass.  Its algorithmic
FLOPs per launch are 70 per (atom, point) pair (the counting convention of
the workload definition: add/mul 1, FMA 2, rsqrt 1, compares and selects 0,
per-point and per-atom constants hoisted), i.e. 70 N P; the average launch
time comes from HIP events around back-to-back launches on the stream the
kernel runs on (isspa_model_profile).  Peak: 157.3 TFLOP/s FP32 vector.

Batched umbrella windows (reported beside the headline, "batched"): 64
windows of uneven size (32 two-atom LJ pairs, 24 chains of 22 atoms, 8
systems of 220 atoms; each with a k = 10 distance restraint, centres 3-15 A),
nMC=20, run as ONE batched model whose steps are graph replays, against the
plain per-window loop (64 single-window engines, replayed back to back on one
stream), both with the same number of steps per graph (--steps-per-graph).  Both report aggregate window-ns/day
(sum over windows of ns/day at dt = 2 fs); the launch geometry of the batched
model is chosen by isspa_amd.autotune.

The reference's own published run ("reference_published_run"): ADI2, 44
atoms, nMC=100, one distance restraint, the whole MD step on the engine at
50 steps per graph replay, in ns/day beside the GTX 1080 figure of
timing_tests/ADI2_US/window_6.5/ADI2.run0.log (also alone: --adi2-only).

CPU baseline: the reference's own src/isspa_force.omp.c (built in place into
oracle/_ref by oracle/Makefile), cpu_ispa_force on the same 19,998-atom
geometry with nMC=20 and the workload's per-type parabola and LJ-solvent
parameters, timed on this host's cores.  It computes a cheaper LJ-only,
single-pass model; it is a reported baseline, not the target.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "isspa-force-cuda_amd"))

METRIC = "IS-SPA MD steps/s at 20k atoms nMC=20; batched 64-window ns/day vs plain"
FP32_PEAK = 157.3e12
FLOP_FORCE_PAIR = 70      # force pass, per (atom, point), outside rmax
FLOP_FIELD_PAIR = 24      # field pass, per (point, atom), outside rmax
FLOP_NB_PAIR = 36         # nonbond, per unordered pair


TRAFFIC_JSON = "k_force_19998atoms_nmc20_traffic.json"
COUNTERS_JSON = "isspa_passes_19998atoms_nmc20_valu_counters.json"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_at(chunks, path):
    """HBM-side bytes of one k_force launch at `chunks` force chunks, from the
    committed FETCH_SIZE / WRITE_SIZE passes (copy-calibrated) at several
    chunk counts.  k_force reads the point records and writes 6 FP64 partials
    per (chunk, atom): its WRITE_SIZE is linear in the chunk count and its
    FETCH_SIZE nearly constant, so both are interpolated linearly between the
    two profiled counts around `chunks`.  Returns (bytes or None outside the
    profiled range, note, summary file), or None without the file."""
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        tr = json.load(fh)
    pts = sorted(tr["points"], key=lambda q: q["chunks"])
    traffic, note = None, (f"outside the profiled chunk counts "
                           f"{pts[0]['chunks']}-{pts[-1]['chunks']}; this run used {chunks}")
    for lo, hi in zip(pts, pts[1:]):
        if lo["chunks"] <= chunks <= hi["chunks"]:
            t = (chunks - lo["chunks"]) / max(1, hi["chunks"] - lo["chunks"])
            kib = lambda k: lo[k] + t * (hi[k] - lo[k])
            traffic = int(round((kib("fetch_kib") * tr["fetch_scale"] +
                                 kib("write_kib") * tr["write_scale"]) * 1024))
            note = (f"FETCH_SIZE/WRITE_SIZE interpolated between the profiled {lo['chunks']} and "
                    f"{hi['chunks']} force chunks (this run {chunks})")
            break
    return traffic, note, tr["summary"]


def counter_check(path):
    """Per wave-pair VALU instructions and executed-to-algorithmic FLOP of
    the force and field passes, from the committed counter pass (or None)."""
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        vc = json.load(fh)
    out = {k: {"valu_per_wave_pair": vc[k]["valu_per_wave_pair"],
               "executed_flop_per_pair": vc[k]["executed_flop_per_pair"],
               "algorithmic_flop_per_pair": vc[k]["algorithmic_flop_per_pair"],
               "executed_over_algorithmic": vc[k]["executed_over_algorithmic"]}
           for k in ("k_force", "k_field")}
    out["source"] = vc["summary"]
    return out


def cpu_model():
    """the host CPU's model name (/proc/cpuinfo), or the platform string"""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def host_threads():
    """threads the CPU baseline uses: the cores this process may run on, at
    most 16 (the host share of one GPU on the measurement machines)"""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(system, threads, trials=3, budget_s=30.0):
    """The reference OpenMP cpu_ispa_force on the workload geometry: one call
    is one step's IS-SPA solvent force over N x P point-atom pairs.  Returns
    (median seconds per call, kind, trials run) or None if neither library is
    built."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    ref = po.ref_lib()
    kind = "reference"
    fn = ref.cpu_ispa_force if ref is not None else None
    if fn is None:
        fn, kind = po.lib().orc_cpu_ispa_force, "port"
    par = system.type_params
    T = len(par)
    w = np.array([par[t]["w"] for t in range(T)], np.float32)
    x0 = np.array([par[t]["x0"] for t in range(T)], np.float32)
    g0 = np.array([par[t]["g0"] for t in range(T)], np.float32)
    al = np.array([par[t]["alpha"] for t in range(T)], np.float32)
    gr2 = np.array([[par[t]["r_lo"] ** 2, par[t]["r_hi"] ** 2] for t in range(T)], np.float32).ravel()
    A = np.array([par[t]["lj_A"] for t in range(T)], np.float32)
    B = np.array([par[t]["lj_B"] for t in range(T)], np.float32)
    xyz = np.ascontiguousarray(system.xyzq[:, :3]).ravel()
    times = []
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        po.cpu_ispa_force(fn, xyz, w, x0, g0, gr2, al, A, B, system.ityp, system.n_mc, threads,
                          seed=1)
        times.append(time.perf_counter() - t0)
        log(f"cpu baseline trial {len(times)}: {times[-1]:.2f} s ({threads} threads, {kind})")
        if len(times) >= trials or time.perf_counter() - t_start + times[-1] > budget_s:
            break
    times.sort()
    return times[len(times) // 2], kind, len(times)


def _time_replays(launch, steps):
    """wall seconds for `steps` calls of launch(), bracketed by synchronisations"""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        launch()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def batched_bench(steps, plain_rounds, warmup, tune_reps, per_graph=1):
    """Batched 64-window step vs the plain per-window loop on the same GPU.
    Both capture `per_graph` consecutive steps in each graph."""
    import numpy as np
    import torch
    from isspa_amd import autotune, batch, capi
    from isspa_amd.engine import GpuSystem

    windows = batch.umbrella_windows()
    merged = batch.merge_windows(windows)
    W = len(windows)
    g = GpuSystem(merged, "cuda:0")
    tr = autotune.tune(g, reps=tune_reps, waves=(2048, 4096, 8192, 16384, 32768, 65536, 131072),
                       log=log, confirm=autotune.engine_step_ms)
    log(f"batched autotune: {tr.kind} {tr.value} -> {tr.ms:.4f} ms per {tr.measure}")
    K = max(1, int(per_graph))
    steps = max(K, steps - steps % K)
    run_plain_loop = plain_rounds > 0
    plain_rounds = max(K, plain_rounds - plain_rounds % K)
    g.set_step(0)
    g.capture(K)
    st = torch.cuda.Stream()
    eng = g.make_engine()

    def run_batch():
        capi.check(capi.lib.md_engine_replay(eng, 1, st.cuda_stream))

    with torch.cuda.stream(st):
        _time_replays(run_batch, max(1, warmup // K))
        t_b = _time_replays(run_batch, steps // K)
    if not np.all(np.isfinite(g.x.cpu().numpy())):
        raise RuntimeError("batched run produced non-finite positions")
    if not run_plain_loop:
        ms_b = 1e3 * t_b / steps
        log(f"batched: {ms_b:.4f} ms per step of all {W} windows "
            f"({W * 0.1728 * steps / t_b:.1f} window-ns/day); plain loop not run")
        g.close()
        return {"windows": W, "atoms": merged.n_atoms, "steps": steps, "steps_per_graph": K,
                "ms_per_step": round(ms_b, 5),
                "window_ns_per_day": round(W * 0.1728 * steps / t_b, 2),
                "autotune": {"kind": tr.kind, "value": list(tr.value), "ms": round(tr.ms, 5),
                             "measure": tr.measure}}
    plain = []
    for w, s in enumerate(windows):
        p = GpuSystem(s, "cuda:0", system_id=w)
        p.set_step(0)
        p.capture(K)
        plain.append((p, p.make_engine()))

    def run_plain():
        for _, e in plain:
            capi.check(capi.lib.md_engine_replay(e, 1, st.cuda_stream))

    with torch.cuda.stream(st):
        _time_replays(run_plain, max(1, warmup // (10 * K)))
        t_p = _time_replays(run_plain, plain_rounds // K)
    ms_b = 1e3 * t_b / steps
    ms_p = 1e3 * t_p / plain_rounds
    nsd_b = W * 0.1728 * steps / t_b
    nsd_p = W * 0.1728 * plain_rounds / t_p
    log(f"batched: {ms_b:.4f} ms per step of all {W} windows ({nsd_b:.1f} window-ns/day); "
        f"plain loop: {ms_p:.4f} ms ({nsd_p:.1f} window-ns/day); ratio {nsd_b / nsd_p:.2f}")
    kinds = [w.name.split("_")[1] for w in windows]
    out = {
        "workload": f"{W} umbrella windows: {kinds.count('ljpair')} two-atom LJ pairs, "
                    f"{kinds.count('chain')} x 22-atom chain, {kinds.count('big')} x 220-atom "
                    f"(10 chains); k=10 distance restraint, centres 3-15 A; nMC={merged.n_mc}",
        "windows": W, "atoms": merged.n_atoms, "steps": steps, "plain_rounds": plain_rounds,
        "steps_per_graph": K,
        "ms_per_step": round(ms_b, 5),
        "window_ns_per_day": round(nsd_b, 2),
        "plain_loop_ms_per_step": round(ms_p, 5),
        "plain_loop_window_ns_per_day": round(nsd_p, 2),
        "speedup_vs_plain_loop": round(nsd_b / nsd_p, 3),
        "autotune": {"kind": tr.kind, "value": list(tr.value), "ms": round(tr.ms, 5),
                     "measure": tr.measure, "candidates": len(tr.table)},
    }
    for p, _ in plain:
        p.close()
    g.close()
    return out


ADI2_GTX1080_NS_PER_DAY = 1270.51  # timing_tests/ADI2_US/window_6.5/ADI2.run0.log:66


def adi2_bench(steps, warmup, per_graph):
    """The reference's own published timing run: ADI2 (44 atoms, two 22-atom
    chains), nMC = 100, box 200 A, one centre-of-mass distance restraint
    (atoms 1-22 against 23-44, k = 20, x0 = 6.5 A), 298 K, pnu = 0.01,
    dt = 2 fs (timing_tests/ADI2_US/window_6.5/ADI2.run0.log).  The whole MD
    step runs on the engine, replayed from a graph of per_graph steps."""
    import torch
    from isspa_amd import capi, io
    from isspa_amd.engine import GpuSystem
    from isspa_amd.system import UmbrellaRestraint
    d = os.path.join(ROOT, "tests", "golden", "adi2")
    s = io.load_system(os.path.join(d, "ADI2.prmtop"), os.path.join(d, "ADI2.isspatop"),
                       os.path.join(d, "ADI2.rst7"), lbox=200.0, n_mc=100, pnu=0.01,
                       seed=34669, name="ADI2")
    s.us = UmbrellaRestraint.from_groups(range(0, 22), range(22, 44), s.true_mass, 20.0, 6.5)
    g = GpuSystem(s, "cuda:0")
    g.set_step(0)
    g.capture(per_graph)
    eng = g.make_engine()
    st = torch.cuda.Stream()
    n = max(1, steps // per_graph)
    with torch.cuda.stream(st):
        # the library's own replay loop: one graph launch per per_graph steps
        _time_replays(lambda: capi.check(capi.lib.md_engine_replay(
            eng, max(1, warmup // per_graph), st.cuda_stream)), 1)
        wall = _time_replays(lambda: capi.check(capi.lib.md_engine_replay(
            eng, n, st.cuda_stream)), 1)
    g.close()
    sec = wall / (n * per_graph)
    log(f"ADI2: {sec * 1e6:.2f} us per step ({per_graph} steps per graph), "
        f"{0.1728 / sec:.0f} ns/day")
    return {"workload": "ADI2 44 atoms, nMC=100, box 200 A, one distance restraint (k=20, x0=6.5 A); "
                        "whole MD step on the engine, graph-replayed",
            "steps": n * per_graph, "steps_per_graph": per_graph,
            "us_per_step": round(sec * 1e6, 3), "ns_per_day": round(0.1728 / sec, 1),
            "reference_gtx1080_ns_per_day": ADI2_GTX1080_NS_PER_DAY,
            "ratio_vs_reference": round(0.1728 / sec / ADI2_GTX1080_NS_PER_DAY, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", type=int, default=2,
                    help="plain workload index: 0 one chain, 1 1,100 atoms, 2 19,998 atoms")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=5)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched comparison and print its JSON object")
    ap.add_argument("--batch-steps", type=int, default=2000)
    ap.add_argument("--plain-rounds", type=int, default=100,
                    help="rounds of the per-window loop in the batched comparison (0: skip it)")
    ap.add_argument("--steps-per-graph", type=int, default=10,
                    help="steps captured per graph: the headline step (when it divides --steps "
                         "and --warmup, else 1), the batched comparison (both sides) and "
                         "--adi2-only")
    ap.add_argument("--adi2-only", action="store_true",
                    help="run only the reference's published ADI2 configuration on the engine "
                         "and print its JSON object")
    ap.add_argument("--adi2-steps", type=int, default=20000,
                    help="steps of the ADI2 configuration (also timed beside the headline)")
    ap.add_argument("--tiling", default="", help="explicit field splits,force chunks (no tuning)")
    ap.add_argument("--target-waves", type=int, default=0, help="automatic tiling at this target")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run an untimed copy of this many MiB (counter calibration)")
    ap.add_argument("--traffic-sweep", default="",
                    help="after the timed steps, one untimed launch of each IS-SPA stage at each "
                         "of these force chunk counts (comma-separated; counter passes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only; --gpus {args.gpus} is not supported "
              "(the per-GPU shard design is documented, no multi-GPU path runs here)",
              file=sys.stderr)
        sys.exit(2)

    if args.adi2_only:
        print(json.dumps({"adi2": adi2_bench(args.adi2_steps, 200, args.steps_per_graph)}))
        return

    if args.batch_only:
        print(json.dumps({"batched": batched_bench(args.batch_steps, args.plain_rounds,
                                                   args.warmup, tune_reps=3,
                                                   per_graph=args.steps_per_graph)}))
        return

    import numpy as np
    import torch
    from isspa_amd import autotune, synth
    from isspa_amd.engine import GpuSystem

    t0 = time.perf_counter()
    s = synth.workload(args.workload)
    log(f"workload {s.name}: N={s.n_atoms} nMC={s.n_mc} P={s.n_points} "
        f"box={s.lbox} built in {time.perf_counter() - t0:.1f}s")
    g = GpuSystem(s, "cuda:0")
    tune = None
    if args.tiling:
        g.set_tiling(*(int(v) for v in args.tiling.split(",")))
    elif args.target_waves:
        g.set_target_waves(args.target_waves)
    elif not args.no_tune:
        # candidates ranked by the IS-SPA pass, the leaders confirmed on the
        # graph-replayed MD step that is timed below (its all-pairs kernel
        # overlaps the IS-SPA passes)
        tune = autotune.tune(g, reps=3, log=log, confirm=autotune.engine_step_ms)
        log(f"autotune: {tune.kind} {tune.value} -> {tune.ms:.4f} ms per {tune.measure}")
    splits, chunks = g.tiling()
    if args.copy_mib:
        # untimed plain copy of known size: the calibration kernel for the
        # FETCH_SIZE / WRITE_SIZE counters
        nbytes = args.copy_mib << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {nbytes} bytes read, {nbytes} bytes written")
        del src, dst
    # K steps per graph replay when K divides both the warm-up and the timed
    # steps, so exactly --steps steps are timed
    K = args.steps_per_graph
    if K < 1 or args.steps % K or args.warmup % K:
        K = 1
    g.set_step(0)
    g.capture(K)
    g.replay(args.warmup // K)
    torch.cuda.synchronize()
    side = g._side()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record(side)
    g.replay(args.steps // K)
    ev1.record(side)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ev_ms = ev0.elapsed_time(ev1)
    x = g.x.cpu().numpy()
    if not np.all(np.isfinite(x)):
        log("non-finite positions after the timed steps")
        sys.exit(1)
    steps_per_s = args.steps / wall
    ms_step = 1e3 * wall / args.steps
    log(f"timed {args.steps} steps: wall {wall:.3f}s  events {ev_ms / 1e3:.3f}s  "
        f"{steps_per_s:.2f} steps/s  {ms_step:.3f} ms/step")

    prof = g.profile(args.profile_reps)
    log("stage ms: " + ", ".join(f"{k} {v:.4f}" for k, v in prof.items()))
    if args.traffic_sweep:
        # k_force launches at other chunk counts, told apart in a counter pass
        # by their grid (chunks x receiver blocks x 256 work-items)
        rblocks = (s.n_atoms + 255) // 256
        for c in (int(v) for v in args.traffic_sweep.split(",")):
            g.set_tiling(splits, c)
            _, c_eff = g.tiling()
            g.profile(1)
            log(f"traffic sweep: {c_eff} force chunks, k_force grid {c_eff * rblocks * 256}")
        g.set_tiling(splits, chunks)
    N, P = s.n_atoms, s.n_points
    force_flop = FLOP_FORCE_PAIR * N * P
    achieved = force_flop / (prof["force"] * 1e-3)
    field_tflops = FLOP_FIELD_PAIR * N * P / (prof["field"] * 1e-3) / 1e12
    step_flop = (FLOP_FIELD_PAIR + FLOP_FORCE_PAIR) * N * P + FLOP_NB_PAIR * N * (N - 1) / 2
    out = {
        "metric": METRIC,
        "value": round(steps_per_s, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{s.name}: {N // 22} seeded copies of the 22-atom chain "
                        f"({N} atoms), nMC={s.n_mc}, box {s.lbox:g} A; bonded + IS-SPA + "
                        f"all-pairs nonbond + leap-frog/Andersen, hipGraph replays of {K} step(s)",
            "n_atoms": N, "n_mc": s.n_mc, "mc_points": P, "dt_fs": 2.0,
            "field_splits": splits, "force_chunks": chunks, "steps_per_graph": K,
            "seeds": synth.workload_seeds(args.workload),
        },
        "ns_per_day": round(0.1728 * steps_per_s, 4),
        "step_tflops": round(step_flop * steps_per_s / 1e12, 3),
        "roofline": {
            "kernel": "k_force (IS-SPA force pass)",
            "bound": "valu",
            "achieved": round(achieved / 1e12, 3),
            "peak": FP32_PEAK / 1e12,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK, 4),
            "traffic": None,
            "flop_per_launch": force_flop,
            "launch_ms": round(prof["force"], 4),
        },
        "stage_ms": {k: round(v, 4) for k, v in prof.items()},
        "field_pass_tflops": round(field_tflops, 3),
    }
    if tune is not None:
        out["config"]["autotune"] = {"kind": tune.kind, "value": list(tune.value),
                                     "ms": round(tune.ms, 4), "measure": tune.measure,
                                     "candidates": len(tune.table)}
    # HBM-side bytes per k_force launch and the VALU-instruction counter pass
    # of the same kernels, from the committed profiles
    if args.workload == 2:
        tr = traffic_at(chunks, os.path.join(ROOT, "profiles", TRAFFIC_JSON))
        if tr is not None:
            out["roofline"]["traffic"], out["roofline"]["traffic_note"], \
                out["roofline"]["traffic_source"] = tr
        cc = counter_check(os.path.join(ROOT, "profiles", COUNTERS_JSON))
        if cc is not None:
            out["roofline"]["counter_check"] = cc
    g.close()
    if not args.no_batch:
        out["batched"] = batched_bench(args.batch_steps, args.plain_rounds, args.warmup,
                                       tune_reps=3, per_graph=args.steps_per_graph)
        # the one configuration the reference publishes timings for (ns/day on
        # a GTX 1080), beside the headline: a different metric, so it does not
        # set vs_baseline
        out["reference_published_run"] = adi2_bench(args.adi2_steps, 200, 50)
    if not args.no_cpu:
        # the reference's own OpenMP file times 1..10 threads
        # (src/isspa_force.omp.c:172-199): here 1 thread (one call) and the
        # host's core share (median of 3 calls)
        threads = args.cpu_threads or host_threads()
        pairs = N * P
        res = cpu_baseline(s, threads)
        res1 = cpu_baseline(s, 1, trials=1) if threads > 1 else res
        if res is not None:
            t_cpu, kind, trials = res
            t_one = res1[0]
            out["cpu_baseline"] = {
                "value": round(1.0 / t_cpu, 5), "unit": "steps/s", "cores": threads,
                "kind": kind,
                "sample": f"cpu_ispa_force (src/isspa_force.omp.c) on the same {N}-atom geometry, "
                          f"nMC={s.n_mc}, one whole step per call ({pairs:.3e} point-atom pairs); "
                          f"median of {trials} call(s) at {threads} threads, one call at 1 thread; "
                          "LJ-only single-pass model",
                "seconds_per_step": round(t_cpu, 3),
                "pair_evals_per_s": round(pairs / t_cpu, 1),
                "single_thread": {"seconds_per_step": round(t_one, 3),
                                  "pair_evals_per_s": round(pairs / t_one, 1),
                                  "speedup_at_cores": round(t_one / t_cpu, 2)},
                "cpu_model": cpu_model(),
                "nproc": os.cpu_count(),
            }
            out["gpu_over_cpu_step_ratio"] = round(t_cpu / (ms_step * 1e-3), 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

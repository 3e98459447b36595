"""Batched umbrella windows on the GPU: one batched model over 64 windows of
uneven size against 64 plain single-window runs (window w run alone with
Philox system id w) and against the FP64 oracle per window.

With the same explicit tiling a batched window runs exactly the arithmetic of
its plain run (same points, same per-split and per-chunk reduction order), so
IS-SPA forces, bonded forces and restraint forces agree bit for bit.  The
nonbond kernel picks its partner chunking from the launch's total work, so
nonbond forces agree to FP32 rounding (relative RMS <= 1e-6).
"""
import numpy as np
import pytest
import torch

import pyoracle as po
from isspa_amd import batch
from isspa_amd.system import max_abs_rel_rms

pytestmark = pytest.mark.gpu


def _gpu(system, **kw):
    from isspa_amd.engine import GpuSystem
    return GpuSystem(system, "cuda:0", **kw)


@pytest.fixture(scope="module")
def windows():
    return batch.umbrella_windows()


@pytest.fixture(scope="module")
def merged(windows):
    return batch.merge_windows(windows)


def _one_of_each(windows):
    seen, out = set(), []
    for w, s in enumerate(windows):
        kind = s.name.split("_")[1]
        if kind not in seen:
            seen.add(kind)
            out.append(w)
    return out


def test_batch_isspa_matches_plain_windows_bitwise(windows, merged):
    print(f"batch of {len(windows)} windows, {merged.n_atoms} atoms", flush=True)
    g = _gpu(merged, nonbond=False, bonded=False, us=False)
    g.set_tiling(2, 3)
    comp = {k: v.cpu().numpy() for k, v in g.isspa_components(step=5).items()}
    pts = g.field()["pts"].cpu().numpy()
    g.close()
    sl = batch.window_slices(merged)
    for w, (s, (a, b)) in enumerate(zip(windows, sl)):
        p = _gpu(s, nonbond=False, bonded=False, us=False, system_id=w)
        p.set_tiling(2, 3)
        pc = p.isspa_components(step=5)
        pp = p.field()["pts"].cpu().numpy()
        assert np.array_equal(pts[a * s.n_mc:b * s.n_mc], pp), f"window {w}: points differ"
        for k in ("lj", "pol", "cdd"):
            assert np.array_equal(comp[k][a:b], pc[k].cpu().numpy()), f"window {w} {k}"
        p.close()


def test_batch_isspa_matches_oracle(windows, merged):
    g = _gpu(merged, nonbond=False, bonded=False, us=False)
    comp = {k: v.cpu().numpy() for k, v in g.isspa_components(step=9).items()}
    g.close()
    sl = batch.window_slices(merged)
    for w in _one_of_each(windows):
        s, (a, b) = windows[w], sl[w]
        ref = po.isspa_total(s, step=9, sys_id=w)
        for k in ("lj", "pol", "cdd"):
            err = max_abs_rel_rms(comp[k][a:b], ref[k])
            print(f"window {w} ({s.name}) {k}: rel RMS {err:.2e}", flush=True)
            assert err <= 1e-5, (w, k, err)


def test_batch_nonbond_bonded_us_match_windows(windows, merged):
    g = _gpu(merged)
    nb = {k: v.cpu().numpy() for k, v in g.nonbond_components().items()}
    g.zero_forces()
    g.bonded_force()
    fb = g.f.cpu().numpy().copy()
    g.zero_forces()
    _, com = g.us_force()
    fu = g.f.cpu().numpy().copy()
    com = com.cpu().numpy().reshape(-1, 6)
    g.close()
    sl = batch.window_slices(merged)
    for w, (s, (a, b)) in enumerate(zip(windows, sl)):
        p = _gpu(s, isspa=False, system_id=w)
        pn = p.nonbond_components()
        for k in ("lj", "coul", "pair"):
            ref = pn[k].cpu().numpy()
            if np.any(ref):
                assert max_abs_rel_rms(nb[k][a:b], ref) <= 1e-6, (w, k)
            else:
                assert not np.any(nb[k][a:b]), (w, k)
        if p.bonded is not None:
            p.zero_forces()
            p.bonded_force()
            assert np.array_equal(fb[a:b], p.f.cpu().numpy()), f"window {w} bonded"
        p.zero_forces()
        _, pc = p.us_force()
        assert np.array_equal(fu[a:b], p.f.cpu().numpy()), f"window {w} restraint"
        assert np.array_equal(com[w], pc.cpu().numpy()), f"window {w} com"
        p.close()
    # nonbond against the oracle for one window of each kind
    for w in _one_of_each(windows):
        a, b = sl[w]
        ref = po.nonbond(windows[w])
        for k in ("lj", "coul", "pair"):
            if np.any(ref[k]):
                assert max_abs_rel_rms(nb[k][a:b], ref[k]) <= 1e-5, (w, k)


def test_batch_engine_matches_plain_engines(windows, merged):
    """Forces and one integration step of the batched engine against each
    window's own engine (same seed, system id w)."""
    g = _gpu(merged)
    g.set_tiling(2, 3)
    g.set_step(3)
    f = g.forces().cpu().numpy().copy()
    g.step(1)
    x = g.x.cpu().numpy()
    v = g.v.cpu().numpy()
    g.close()
    sl = batch.window_slices(merged)
    for w, (s, (a, b)) in enumerate(zip(windows, sl)):
        p = _gpu(s, system_id=w)
        p.set_tiling(2, 3)
        p.set_step(3)
        pf = p.forces().cpu().numpy()
        assert max_abs_rel_rms(f[a:b, :3], pf[:, :3]) <= 1e-6, f"window {w} forces"
        p.step(1)
        assert np.abs(x[a:b] - p.x.cpu().numpy()).max() <= 1e-4, f"window {w} positions"
        assert np.array_equal(x[a:b, 3], p.x.cpu().numpy()[:, 3])
        assert max_abs_rel_rms(v[a:b, :3], p.v.cpu().numpy()[:, :3]) <= 1e-5, f"window {w} vel"
        p.close()


def test_batch_graph_replay_matches_eager(merged):
    eager = _gpu(merged)
    graph = _gpu(merged)
    eager.set_step(0)
    graph.set_step(0)
    graph.capture(2)
    for r in range(3):
        eager.step(2)
        graph.replay(1)
        torch.cuda.synchronize()
        junk = torch.randn(1000, device="cuda:0").sum().item()
        assert np.isfinite(junk)
        assert np.array_equal(eager.x.cpu().numpy(), graph.x.cpu().numpy()), f"replay {r}"
        assert np.array_equal(eager.v.cpu().numpy(), graph.v.cpu().numpy()), f"replay {r}"
    eager.close()
    graph.close()


def test_autotune_engine_step_restores_state(merged):
    """Tuning with the graph-replayed step as the confirmation runs MD steps
    on the system; positions, velocities and the step counter are restored
    after, so a run that follows equals a run of an untuned copy at the chosen
    geometry, bit for bit."""
    from isspa_amd import autotune
    tuned = _gpu(merged)
    x0, v0 = tuned.x.clone(), tuned.v.clone()
    r = autotune.tune(tuned, reps=1, waves=(2048, 8192), tilings=[(2, 3), (3, 5)],
                      confirm=autotune.engine_step_ms)
    print(f"case autotune by engine step: {r.kind} {r.value} {r.ms:.4f} ms", flush=True)
    assert r.measure == "engine step" and np.isfinite(r.ms) and r.ms > 0
    assert torch.equal(tuned.x, x0) and torch.equal(tuned.v, v0)
    ref = _gpu(merged)
    autotune._apply(ref, r.kind, r.value)
    assert ref.tiling() == tuned.tiling()
    for g in (tuned, ref):
        g.set_step(0)
        g.capture(1)
        g.replay(3)
    torch.cuda.synchronize()
    assert np.array_equal(tuned.x.cpu().numpy(), ref.x.cpu().numpy())
    assert np.array_equal(tuned.v.cpu().numpy(), ref.v.cpu().numpy())
    tuned.close()
    ref.close()


def _mixed_box_windows():
    """Four windows whose boxes make the minimum image fire in two of them:
    a 22-atom chain in an 18 A box, 10 chains (220 atoms) in a 40 A box, and
    a two-atom LJ pair and a chain in 200 A boxes."""
    import copy
    from isspa_amd import synth
    from isspa_amd.system import UmbrellaRestraint
    out = []
    a = synth.make_chain()
    a.lbox = 18.0
    a.xyzq[:, :3] = (a.xyzq[:, :3] - a.xyzq[:, :3].mean(0) + 9.0).astype(np.float32)
    b = synth.make_lattice(10, 40.0, 3, 40.0, seed=31)
    c = synth.two_atom_lj(sep=5.0)
    d = synth.make_chain()
    d.xyzq[:, :3] += np.float32(100.0)
    for s, (g1, g2) in zip((a, b, c, d), (([0], [21]), ([0], [22]), ([0], [1]), ([0], [21]))):
        s.us = UmbrellaRestraint.from_groups(g1, g2, s.true_mass, 10.0, 6.0)
        s.n_mc, s.temperature_k, s.pnu, s.seed = 8, 298.0, 0.01, 777
        out.append(s)
    return out


def test_batch_mixed_boxes_minimum_image():
    """Per-window boxes in one batch: windows whose minimum image fires sit
    next to windows whose does not.  IS-SPA forces equal the plain per-window
    runs bit for bit (same tiling) and the oracle to 1e-5; nonbond forces
    match the oracle per window to 1e-5."""
    wins = _mixed_box_windows()
    m = batch.merge_windows(wins)
    sl = batch.window_slices(m)
    g = _gpu(m, bonded=False, us=False)
    g.set_tiling(2, 3)
    comp = {k: v.cpu().numpy() for k, v in g.isspa_components(step=4).items()}
    nb = g.nonbond_components()
    nbt = {k: v.cpu().numpy() for k, v in nb.items()}
    g.close()
    for w, (s, (a, b)) in enumerate(zip(wins, sl)):
        print(f"window {w}: N={s.n_atoms} box {s.lbox}", flush=True)
        p = _gpu(s, nonbond=False, bonded=False, us=False, system_id=w)
        p.set_tiling(2, 3)
        pc = p.isspa_components(step=4)
        for k in ("lj", "pol", "cdd"):
            assert np.array_equal(comp[k][a:b], pc[k].cpu().numpy()), (w, k)
        p.close()
        ref = po.isspa_total(s, step=4, sys_id=w)
        for k in ("lj", "pol", "cdd"):
            if np.abs(ref[k]).max() == 0.0:
                assert np.abs(comp[k][a:b]).max() == 0.0
                continue
            err = max_abs_rel_rms(comp[k][a:b], ref[k])
            print(f"  isspa {k}: rel RMS {err:.2e}", flush=True)
            assert err <= 1e-5, (w, k, err)
        if s.n_atoms > 1:
            nref = po.nonbond(s)
            for k in ("lj", "coul", "pair"):
                if np.abs(nref[k]).max() == 0.0:
                    continue
                err = max_abs_rel_rms(nbt[k][a:b], nref[k])
                print(f"  nonbond {k}: rel RMS {err:.2e}", flush=True)
                assert err <= 1e-5, (w, k, err)

"""IS-SPA solvent kernels on the GPU against the FP64 oracle (same Philox points).

Tolerance: per-component relative RMS over atoms <= 1e-5 (FP32 compute, FP64
accumulation), the north-star bar.  MC points are compared bit for bit, and
the in-sphere counts exactly (both sides take the same FP32 branch).
"""
import numpy as np
import pytest

import pyoracle as po
from isspa_amd.system import max_abs_rel_rms

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _gpu(system, **kw):
    from isspa_amd.engine import GpuSystem
    return GpuSystem(system, "cuda:0", **kw)


@pytest.fixture(scope="module")
def systems(cfg0, cfg1, adi2):
    return {"cfg0": cfg0, "cfg1": cfg1, "adi2": adi2}


@pytest.mark.parametrize("name,step", [("cfg0", 0), ("cfg0", 7), ("cfg1", 3), ("adi2", 1)])
def test_isspa_components_match_oracle(systems, name, step):
    s = systems[name]
    print(f"case {name} step {step} N={s.n_atoms} nMC={s.n_mc}", flush=True)
    g = _gpu(s, nonbond=False, bonded=False)
    comp = g.isspa_components(step=step)
    fld = {k: v.cpu().numpy() for k, v in g.field().items()}
    ref_pts = po.mc_points(s, step=step)
    # MC points bit for bit
    assert np.array_equal(fld["pts"][:, :3].view(np.uint32), ref_pts[:, :3].view(np.uint32))
    rf = po.isspa_field(s, ref_pts)
    assert np.array_equal(fld["cnt"], rf["cnt"])
    assert max_abs_rel_rms(fld["g"], rf["g"]) < 1e-6
    assert max_abs_rel_rms(fld["E"], rf["E"]) < 1e-6
    assert max_abs_rel_rms(fld["E0"], rf["E0"]) < 1e-6
    ref = po.isspa_force(s, ref_pts, rf)
    for k in ("lj", "pol", "cdd"):
        err = max_abs_rel_rms(comp[k].cpu().numpy(), ref[k])
        print(f"  {k}: rel RMS {err:.3e}", flush=True)
        assert err <= TOL, (k, err)
    g.close()


def test_isspa_total_adds_into_f(cfg1):
    s = cfg1
    g = _gpu(s, nonbond=False, bonded=False)
    import torch
    base = torch.randn_like(g.f) * 0.1
    g.f.copy_(base)
    g.isspa_force(step=5)
    got = (g.f - base)[:, :3].double().cpu().numpy()
    ref = po.isspa_total(s, step=5)["total"]
    assert max_abs_rel_rms(got, ref) <= TOL
    g.close()


def test_isspa_tiling_invariance_and_determinism(cfg1):
    s = cfg1
    g = _gpu(s, nonbond=False, bonded=False)
    base = g.isspa_components(step=2)
    again = g.isspa_components(step=2)
    for k in base:
        assert np.array_equal(base[k].cpu().numpy(), again[k].cpu().numpy()), "not deterministic"
    for splits, chunks in ((1, 1), (3, 7), (17, 64)):
        g.set_tiling(splits, chunks)
        other = g.isspa_components(step=2)
        for k in base:
            assert max_abs_rel_rms(other[k].cpu().numpy(), base[k].cpu().numpy()) < 1e-6
    g.close()


def test_isspa_pbc_path(cfg0):
    """A box small enough that minimum-image shifts fire on many pairs."""
    import copy
    s = copy.deepcopy(cfg0)
    s.lbox = 18.0
    s.xyzq[:, :3] = (s.xyzq[:, :3] - s.xyzq[:, :3].mean(0) + 9.0).astype(np.float32)
    g = _gpu(s, nonbond=False, bonded=False)
    comp = g.isspa_components(step=0)
    ref = po.isspa_total(s, step=0)
    for k in ("lj", "pol", "cdd"):
        assert max_abs_rel_rms(comp[k].cpu().numpy(), ref[k]) <= TOL, k
    g.close()


def test_isspa_large_full_field_and_sampled_forces():
    """19,998 atoms, nMC=20 (the benchmark workload), against oracle-only
    inputs: all 399,960 MC points bit for bit, the FP64 oracle field of every
    point (8.0e9 point-atom pairs on the host cores) against the device's,
    and the forces on 48 sampled receivers computed by the oracle from its own
    points and field."""
    import time
    from isspa_amd import synth
    s = synth.workload(2)
    g = _gpu(s, nonbond=False, bonded=False)
    comp = g.isspa_components(step=0)
    fld = {k: v.cpu().numpy() for k, v in g.field().items()}
    ref_pts = po.mc_points(s, step=0)
    assert np.array_equal(fld["pts"][:, :3].view(np.uint32), ref_pts[:, :3].view(np.uint32))
    t0 = time.perf_counter()
    rf = po.isspa_field(s, ref_pts)
    print(f"  oracle field of {s.n_points} points: {time.perf_counter() - t0:.1f} s", flush=True)
    assert np.array_equal(fld["cnt"], rf["cnt"])
    for k in ("g", "E", "E0"):
        err = max_abs_rel_rms(fld[k], rf[k])
        print(f"  field {k}: rel RMS {err:.3e} over all points", flush=True)
        assert err < 1e-6, k
    rng = np.random.default_rng(0)
    atoms = np.sort(rng.choice(s.n_atoms, 48, replace=False))
    refs = {k: [] for k in ("lj", "pol", "cdd")}
    for a in atoms:
        ref = po.isspa_force(s, ref_pts, rf, atoms=(int(a), int(a) + 1))
        for k in refs:
            refs[k].append(ref[k][a])
    for k in refs:
        got = comp[k].cpu().numpy()[atoms]
        err = max_abs_rel_rms(got, np.array(refs[k]))
        print(f"  sampled {k}: rel RMS {err:.3e}", flush=True)
        assert err <= TOL, (k, err)
    g.close()


def test_isspa_mean_over_10k_steps_matches_reference_curand_frames(adi2):
    """The device path's mean IS-SPA components over 10,000 Philox steps
    (ADI2, nMC=1000) against the reference's own curand output, 1000 frames
    of timing_tests/ADI2/10000MC/compare/ISSPA_force_{LJsolv,Csolv,Csolv_CDD}_10000.xyz
    (mean and standard error per atom and axis in tests/golden/adi2_isspa_stats.npz).
    z = (mean_ours - mean_ref)/sqrt(se_ours^2 + se_ref^2).  Random halvings of
    the fixture give z-RMS up to ~1.46, so the bar is z-RMS < 1.8; every one
    of the 132 values of each component must lie within 3 sigma (the draws
    are fixed by the seed, so the outcome is deterministic)."""
    import os
    import torch
    from conftest import GOLDEN
    st = np.load(os.path.join(GOLDEN, "adi2_isspa_stats.npz"))
    g = _gpu(adi2, nonbond=False, bonded=False)
    K = 10000
    s1 = {k: torch.zeros((adi2.n_atoms, 3), dtype=torch.float64, device="cuda:0")
          for k in ("lj", "pol", "cdd")}
    s2 = {k: torch.zeros_like(v) for k, v in s1.items()}
    for f in range(K):
        comp = g.isspa_components(step=f)
        for k in s1:
            v = comp[k].double()
            s1[k] += v
            s2[k] += v * v
        if f % 2000 == 0:
            print(f"  step {f}", flush=True)
    for k, name in (("lj", "LJsolv"), ("pol", "Csolv"), ("cdd", "Csolv_CDD")):
        m = (s1[k] / K).cpu().numpy()
        var = (s2[k] / K).cpu().numpy() - m * m
        se = np.sqrt(np.maximum(var, 0.0) * K / (K - 1) / K)
        z = (m - st[f"{name}_mean"]) / np.sqrt(se ** 2 + st[f"{name}_se"] ** 2)
        zrms = float(np.sqrt((z ** 2).mean()))
        print(f"{name}: z-RMS {zrms:.2f} max|z| {np.abs(z).max():.2f} "
              f"(|z| > 3 for {int((np.abs(z) > 3).sum())} of {z.size})", flush=True)
        assert zrms < 1.8 and np.abs(z).max() < 3.0, name
    g.close()

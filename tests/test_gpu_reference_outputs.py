"""The MD engine's step-0 forces against the reference's own output frames.

The engine's force evaluation (md_engine_forces: bonded + IS-SPA + all-pairs
nonbond, the force the reference's driver writes before integrating,
src/total_force_cuda.cpp:109-150), averaged over K Philox draws, against
timing_tests/ADI2/10000MC/ADI.2.forces.xyz and frame 0 of
timing_tests/ADI25/ADI.25.forces.xyz.  Each reference frame is one curand
draw, so z = (F_ref - mean) / (sd sqrt(v + 1/K)) with sd the engine's own
one-draw spread and v the variance factor of the reference's float point
index (1 for ADI2; 2.5 for ADI25's receivers 305..549, refout.py).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from refout import float_index_weights, read_xyz_frames

pytestmark = pytest.mark.gpu


def _engine_draws(system, K):
    import torch
    from isspa_amd.engine import GpuSystem
    g = GpuSystem(system, "cuda:0")
    n = system.n_atoms
    s1 = torch.zeros((n, 3), dtype=torch.float64, device="cuda:0")
    s2 = torch.zeros_like(s1)
    for k in range(K):
        g.set_step(k)
        f = g.forces()[:, :3].double()
        s1 += f
        s2 += f * f
    m = (s1 / K).cpu().numpy()
    var = (s2 / K).cpu().numpy() - m * m
    g.close()
    return m, np.sqrt(np.maximum(var, 0.0) * K / (K - 1))


def _check(ref, m, sd, v, K, label, bar_max):
    z = (ref - m) / (sd * np.sqrt(v[:, None] + 1.0 / K))
    zrms, zmax = float(np.sqrt((z ** 2).mean())), float(np.abs(z).max())
    print(f"{label}: z-RMS {zrms:.3f}, max |z| {zmax:.2f} over {z.size} values", flush=True)
    assert 0.8 < zrms < 1.3 and zmax < bar_max, label


def test_engine_step0_forces_match_adi2_reference_output(adi2):
    print("case ADI2 44 atoms nMC=1000, 2000 draws", flush=True)
    ref = read_xyz_frames(os.path.join(GOLDEN, "adi2", "ADI.2.forces.xyz"), 44)[0]
    K = 2000
    m, sd = _engine_draws(adi2, K)
    _check(ref, m, sd, np.ones(44), K, "ADI2 step 0", 4.0)


def test_engine_step0_forces_match_adi25_reference_output():
    from isspa_amd import io
    d = os.path.join(GOLDEN, "adi25")
    s = io.load_system(os.path.join(d, "ADI25.prmtop"), os.path.join(d, "ADI25.isspatop"),
                       os.path.join(d, "ADI25.rst7"), lbox=200.0, n_mc=100, seed=12345,
                       name="ADI25")
    print(f"case ADI25 {s.n_atoms} atoms nMC=100, 1000 draws", flush=True)
    ref = read_xyz_frames(os.path.join(d, "ADI.25.forces.frame0.xyz"), s.n_atoms)[0]
    K = 1000
    m, sd = _engine_draws(s, K)
    v = float_index_weights(s.n_atoms, s.n_mc)
    for lo, hi in ((0, 304), (305, s.n_atoms)):
        _check(ref[lo:hi], m[lo:hi], sd[lo:hi], v[lo:hi], K, f"ADI25 atoms {lo}..{hi - 1}", 6.0)

"""The oracle against the reference's own step-0 total forces (no GPU).

The reference's driver writes f_d at the input coordinates before the first
integration step (src/total_force_cuda.cpp:109-150): bonded + IS-SPA +
nonbond, with the IS-SPA part one Monte Carlo draw.  Two such frames ship:

* timing_tests/ADI2/10000MC/ADI.2.forces.xyz (44 atoms, nMC = 1000, seed
  8622; tests/golden/adi2/config.txt is that run's deck);
* frame 0 of timing_tests/ADI25/ADI.25.forces.xyz (550 atoms, nMC = 100;
  tests/golden/adi25/config.txt).

They are the only reference outputs that constrain the bonded terms.  The
deterministic part (oracle bonded + nonbond) plus the IS-SPA mean must match
each frame within one draw's IS-SPA spread: z = (F_ref - F_det - mean) / sd
per atom and axis, with z-RMS near 1.  A wrong bonded term moves z by orders
of magnitude (checked below by planting one in the oracle's output).

ADI25's frame carries the reference's float point index (quirk Q2): past
thread index 2^24 its force kernel uses some MC points two or three times
and others not at all, which inflates the variance of receivers 305..549 by
2.5 (refout.float_index_weights, derived from src/isspa_force_cuda.cu:237).
The test uses that variance model and also shows the inflation itself.
"""
import copy
import os

import numpy as np

import pyoracle as po
from conftest import GOLDEN
from isspa_amd import io
from refout import float_index_weights, read_xyz_frames


def _draws(system, K):
    """K independent Philox draws of the oracle's IS-SPA total and components"""
    n = system.n_atoms
    tot = np.zeros((K, n, 3))
    comp = {k: np.zeros((K, n, 3)) for k in ("lj", "pol", "cdd")}
    for k in range(K):
        r = po.isspa_total(system, step=k)
        tot[k] = r["total"]
        for c in comp:
            comp[c][k] = r[c]
    return tot, comp


def _zstats(z):
    return float(np.sqrt((z ** 2).mean())), float(np.abs(z).max())


def test_adi2_step0_total_force_matches_reference_output(adi2):
    """ADI.2.forces.xyz against oracle bonded + nonbond + the IS-SPA mean of
    the reference's own 1000 curand frames (tests/golden/adi2_isspa_stats.npz,
    nMC = 1000 as in this run).  The one-draw spread is the oracle's, over 100
    Philox draws; the oracle's per-component spreads are checked against the
    fixture frames' (standard error x sqrt(1000)) first."""
    ref = read_xyz_frames(os.path.join(GOLDEN, "adi2", "ADI.2.forces.xyz"), 44)
    assert len(ref) == 1
    ref = ref[0]
    st = np.load(os.path.join(GOLDEN, "adi2_isspa_stats.npz"))
    tot, comp = _draws(adi2, 100)
    for c, name in (("lj", "LJsolv"), ("pol", "Csolv"), ("cdd", "Csolv_CDD")):
        ratio = np.sqrt((comp[c].var(0, ddof=1)).mean() /
                        ((st[f"{name}_se"] ** 2) * 1000).mean())
        print(f"{name}: oracle/reference one-draw spread {ratio:.3f}")
        assert 0.85 < ratio < 1.15, name
    sd = tot.std(0, ddof=1)
    det = po.bonded(adi2)["total"] + po.nonbond(adi2)["total"]
    mean = st["LJsolv_mean"] + st["Csolv_mean"] + st["Csolv_CDD_mean"]
    z = (ref - det - mean) / sd
    zrms, zmax = _zstats(z)
    print(f"ADI2 step 0: z-RMS {zrms:.3f}, max |z| {zmax:.2f} over {z.size} values "
          f"(residual RMS {np.sqrt(((ref - det - mean) ** 2).mean()):.4f} of {np.sqrt((ref ** 2).mean()):.2f})")
    assert 0.75 < zrms < 1.3 and zmax < 4.0
    # a dropped angle term or a sign error in the dihedrals is seen at once
    b = po.bonded(adi2)
    for wrong in (det - b["angle"], det - 2 * b["dih"], det - 0.5 * b["bond"]):
        assert _zstats((ref - wrong - mean) / sd)[0] > 50


def test_adi25_frame0_total_force_matches_reference_output():
    """Frame 0 of ADI.25.forces.xyz (the ADI25 deck, seed 12345, nMC = 100)
    against oracle bonded + nonbond + the oracle's IS-SPA mean over 48 Philox
    draws, with the variance of the reference's float point index."""
    d = os.path.join(GOLDEN, "adi25")
    s = io.load_system(os.path.join(d, "ADI25.prmtop"), os.path.join(d, "ADI25.isspatop"),
                       os.path.join(d, "ADI25.rst7"), lbox=200.0, n_mc=100, seed=12345,
                       name="ADI25")
    n = s.n_atoms
    assert n == 550
    ref = read_xyz_frames(os.path.join(d, "ADI.25.forces.frame0.xyz"), n)[0]
    K = 48
    tot, _ = _draws(s, K)
    m, sd = tot.mean(0), tot.std(0, ddof=1)
    b = po.bonded(s)
    det = b["total"] + po.nonbond(s)["total"]
    v = float_index_weights(n, s.n_mc)
    assert np.all(v[:304] == 1.0) and np.allclose(v[305:], 2.5, atol=1e-3)
    z = (ref - det - m) / (sd * np.sqrt(v[:, None] + 1.0 / K))
    for lo, hi in ((0, 304), (305, n)):
        zrms, zmax = _zstats(z[lo:hi])
        print(f"ADI25 atoms {lo}..{hi - 1}: z-RMS {zrms:.3f}, max |z| {zmax:.2f}")
        assert 0.8 < zrms < 1.3 and zmax < 6.0
    # without the float-index variance the affected receivers are over-dispersed
    z_plain = (ref - det - m)[305:] / (sd[305:] * np.sqrt(1.0 + 1.0 / K))
    assert _zstats(z_plain)[0] > 1.4
    for wrong in (det - b["angle"], det - b["dih"]):
        assert _zstats((ref - wrong - m) / sd)[0] > 10


def test_float_index_weights_closed_form():
    """Below 2^24 thread indices every point is used once; between 2^24 and
    2^25 the float grid has spacing 2 and ties go to multiples of 4, so a
    receiver whose indices all lie there uses points 4k three times and 4k+2
    once (variance factor (9 + 1) / 4 = 2.5)."""
    assert np.all(float_index_weights(44, 1000) == 1.0)   # ADI2: 44 x 44000 < 2^24
    v = float_index_weights(550, 100)
    assert abs(v[400] - 2.5) < 1e-3
    # 1,100 atoms with nMC = 20 (24.2M thread indices): the tail is affected too
    v = float_index_weights(1100, 20)
    assert v[0] == 1.0 and abs(v[-1] - 2.5) < 1e-3

"""The CPU oracle pinned against the reference's golden vectors, known-answer
tests, statistical fixtures and closed forms (no GPU)."""
import json
import os

import numpy as np
import pytest

import pyoracle as po
from conftest import GOLDEN
from isspa_amd import io
from isspa_amd.system import max_abs_rel_rms


# ---------------------------------------------------------------- Philox ----
def test_philox_published_kat():
    """Random123 kat_vectors for philox4x32_10 (Salmon et al., SC'11)."""
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for c, k, want in kat:
        assert list(po.philox(c, k)) == want


def test_philox_matches_rocrand():
    """52 vectors from rocRAND's ten_rounds (tests/golden/make_philox_kat.py)."""
    d = json.load(open(os.path.join(GOLDEN, "philox_kat.json")))
    for v in d["vectors"]:
        assert list(po.philox(v["ctr"], v["key"])) == v["out"], v


# ------------------------------------------------------- golden nonbond -----
def _golden_system(adi2):
    """ADI2 with FP64 charges and LJ coefficients as the golden generator used."""
    import copy
    s = copy.deepcopy(adi2)
    sec = io._sections(os.path.join(GOLDEN, "adi2", "ADI2.prmtop"))
    q64 = io._read_array(sec["CHARGE"], s.n_atoms, "f")
    ntri = s.n_lj_types * (s.n_lj_types + 1) // 2
    s.lj_a = io._read_array(sec["LENNARD_JONES_ACOEF"], ntri, "f")
    s.lj_b = io._read_array(sec["LENNARD_JONES_BCOEF"], ntri, "f")
    return s, q64


@pytest.mark.parametrize("comp,name,rtol", [("lj", "LJsolu", 1e-13), ("coul", "Csolu", 1e-13),
                                            ("pair", "Csolv_pair", 2e-7)])
def test_nonbond_golden(adi2, comp, name, rtol):
    """timing_tests/ADI2/10000MC/compare/ADI2.multi.frc.*.dat (FP64 golden).
    LJ and Coulomb reproduce to round-off; the IS-SPA pair term carries the
    FP32 constants ep = 2.3473f and rmax = 11.95f (relative 6e-8)."""
    s, q64 = _golden_system(adi2)
    ref = np.loadtxt(os.path.join(GOLDEN, "adi2", f"ADI2.multi.frc.{name}.dat"))
    x64 = np.concatenate([ref[:, :3], q64[:, None]], axis=1)
    got = po.nonbond(s, x64)[comp]
    assert max_abs_rel_rms(got, ref[:, 3:]) <= rtol


def test_nonbond_golden_detects_exclusion_mistake(adi2):
    """Dropping the exclusion list must break the LJ golden comparison."""
    s, q64 = _golden_system(adi2)
    ref = np.loadtxt(os.path.join(GOLDEN, "adi2", "ADI2.multi.frc.LJsolu.dat"))
    x64 = np.concatenate([ref[:, :3], q64[:, None]], axis=1)
    s.excl_list = np.full_like(s.excl_list, -1)
    assert max_abs_rel_rms(po.nonbond(s, x64)["lj"], ref[:, 3:]) > 1e-2


def test_reference_gpu_frames_match_golden():
    """The reference's FP32 GPU single frames vs the FP64 golden files: the
    <=2.8e-5 bar the GPU parity test inherits."""
    st = np.load(os.path.join(GOLDEN, "adi2_isspa_stats.npz"))
    for name in ("LJsolu", "Csolu", "Csolv_pair"):
        ref = np.loadtxt(os.path.join(GOLDEN, "adi2", f"ADI2.multi.frc.{name}.dat"))
        assert np.abs(st[f"{name}_gpu_fp32"] - ref[:, 3:]).max() <= 2.8e-5


# --------------------------------------------------- statistical fixture ----
def test_isspa_statistics_match_reference_fixture(adi2):
    """Mean IS-SPA components over independent Philox streams against the
    reference's 1000 curand frames (nMC=1000) of
    timing_tests/ADI2/10000MC/compare/ISSPA_force_{LJsolv,Csolv,Csolv_CDD}_10000.xyz.
    z = (mean_ours - mean_ref)/sqrt(se_ours^2 + se_ref^2) per atom and axis.
    Random halvings of the fixture itself give z-RMS up to ~1.46, so the bar
    is z-RMS < 1.8 and max |z| < 5 (396 correlated values)."""
    st = np.load(os.path.join(GOLDEN, "adi2_isspa_stats.npz"))
    K = 60
    acc = {k: [] for k in ("lj", "pol", "cdd")}
    for f in range(K):
        o = po.isspa_total(adi2, step=f, seed=4242)
        for k in acc:
            acc[k].append(o[k])
    for k, name in (("lj", "LJsolv"), ("pol", "Csolv"), ("cdd", "Csolv_CDD")):
        a = np.array(acc[k])
        m, se = a.mean(0), a.std(0, ddof=1) / np.sqrt(K)
        z = (m - st[f"{name}_mean"]) / np.sqrt(se ** 2 + st[f"{name}_se"] ** 2)
        zrms = float(np.sqrt((z ** 2).mean()))
        print(f"{name}: z-RMS {zrms:.2f} max|z| {np.abs(z).max():.2f}")
        assert zrms < 1.8 and np.abs(z).max() < 5.0, name


def test_statistical_test_detects_wrong_sign():
    """The same test rejects an oracle with the CDD sign flipped."""
    st = np.load(os.path.join(GOLDEN, "adi2_isspa_stats.npz"))
    fake = -st["Csolv_CDD_mean"]
    z = (fake - st["Csolv_CDD_mean"]) / (np.sqrt(2) * st["Csolv_CDD_se"])
    assert np.sqrt((z ** 2).mean()) > 10


# --------------------------------------------------------- closed forms -----
MPMATH_MOMENTS = {  # coth x - 1/x, 1 - 3M1/x, M1 - 5M2/x with mpmath at 120 digits
    1e-6: (3.333333333333111111111e-7, 6.666666666666031746032e-14, 9.523809523808465608466e-21),
    1e-3: (0.000333333311111113227513, 6.666666031746095238089e-8, 9.52380846560857463523e-12),
    0.1: (0.03331113225398961014527, 0.0006660323803116956418852, 0.00000951323840482805100977),
    0.25: (0.08298816507359656826221, 0.004142019116841180853517, 0.0001477827367729511918602),
    0.5: (0.16395341373865284877, 0.01627951756808290737998, 0.001158238057823774970245),
    1.0: (0.3130352854993313036362, 0.06089414350200608909152, 0.00856456798930085817858),
    2.0: (0.5373147207275480958778, 0.1940279189086778561833, 0.0522449234558534554196),
    3.0: (0.6716364899803558377598, 0.3283635100196441622402, 0.1243639732809489006928),
    4.999: (0.8000509777783823293101, 0.5198733880105727169574, 0.2800735942911121583185),
    5.001: (0.8001306145489213046715, 0.5200176277450981975576, 0.2802169695328263261096),
    10.0: (0.9000000041223072533738, 0.7299999987633078239879, 0.5350000047406533413799),
    40.0: (0.975, 0.926875, 0.859140625),
    200.0: (0.995, 0.985075, 0.970373125),
}


@pytest.mark.parametrize("x", sorted(MPMATH_MOMENTS))
def test_langevin_moments(x):
    got = po.langevin(x)
    want = np.array(MPMATH_MOMENTS[x])
    assert np.all(np.abs(got - want) <= 1e-13 * np.abs(want))


def test_langevin_branch_continuity():
    lo, hi = po.langevin(5.0 - 1e-12), po.langevin(5.0 + 1e-12)
    assert np.all(np.abs(lo - hi) <= 1e-12 * np.abs(hi))


def test_mc_points_in_ball(cfg0):
    """Points lie within sqrt(0.99) rmax of their owner, fill the ball
    uniformly (E|u|^2 = 3/5 of 0.99 rmax^2), and change with the step."""
    s = cfg0
    pts = po.mc_points(s, step=0)
    own = np.repeat(np.arange(s.n_atoms), s.n_mc)
    R = s.isspa.rmax[s.isspa.isspa_type[own]].astype(np.float64)
    u = (pts[:, :3].astype(np.float64) - s.xyzq[own, :3]) / R[:, None]
    r2 = (u ** 2).sum(1)
    assert r2.max() < 0.99 + 1e-6
    big = np.concatenate([(((po.mc_points(s, step=k)[:, :3] - s.xyzq[own, :3]) / R[:, None]) ** 2)
                          .sum(1) for k in range(200)])
    assert abs(big.mean() / 0.99 - 0.6) < 0.01
    assert not np.array_equal(pts, po.mc_points(s, step=1))


# ------------------------------------------------ bonded / nonbond forces ---
def _fd_check(system, energy_fn, force, h=1e-5, atoms=None):
    x = system.xyzq.astype(np.float64)
    atoms = range(system.n_atoms) if atoms is None else atoms
    for i in atoms:
        for a in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, a] += h
            xm[i, a] -= h
            fd = -(energy_fn(xp) - energy_fn(xm)) / (2 * h)
            assert abs(fd - force[i, a]) <= 1e-5 * max(1.0, abs(force[i, a])), (i, a, fd, force[i, a])


def test_bonded_forces_are_minus_gradient(cfg0):
    s = cfg0
    x = s.xyzq.astype(np.float64)
    b = po.bonded(s, x)
    e = lambda y: sum(po.bonded(s, y)[k] for k in ("e_bond", "e_angle", "e_dih", "e_14coul", "e_14lj"))  # noqa: E731
    _fd_check(s, e, b["total"])
    # Newton's third law and translation invariance
    assert np.abs(b["total"].sum(0)).max() < 1e-9
    shifted = x.copy()
    shifted[:, :3] += 3.25
    assert np.allclose(po.bonded(s, shifted)["total"], b["total"], atol=1e-9)


def test_nonbond_lj_coulomb_minus_gradient(cfg0):
    s = cfg0
    x = s.xyzq.astype(np.float64)
    nb = po.nonbond(s, x)
    e = lambda y: po.nonbond(s, y)["e_lj"] + po.nonbond(s, y)["e_coul"]  # noqa: E731
    _fd_check(s, e, nb["lj"] + nb["coul"], atoms=range(0, s.n_atoms, 3))
    assert np.abs((nb["lj"] + nb["coul"]).sum(0)).max() < 1e-9


def test_us_minus_gradient(adi2):
    from isspa_amd.system import UmbrellaRestraint
    import copy
    s = copy.deepcopy(adi2)
    s.us = UmbrellaRestraint.from_groups(range(22), range(22, 44), s.true_mass, 20.0, 6.5)
    u = po.us_force(s)
    _fd_check(s, lambda y: po.us_force(s, y)["energy"], u["f"], atoms=range(0, 44, 5))
    com = u["com"]
    x = s.xyzq[:, :3].astype(np.float64)
    m = s.true_mass.astype(np.float64)
    assert np.allclose(com[0], (m[:22, None] * x[:22]).sum(0) / m[:22].sum(), atol=1e-4)


def test_leapfrog_closed_form_and_andersen(cfg1):
    import copy
    s = copy.deepcopy(cfg1)
    f = np.random.default_rng(0).normal(size=(s.n_atoms, 3))
    s.pnu = 0.0
    x, v, hit = po.leapfrog(s, f, step=0)
    m = s.vel[:, 3:4].astype(np.float64)
    v0 = s.vel[:, :3].astype(np.float64) + f * s.dt / m
    assert not hit.any()
    assert np.allclose(v, v0, rtol=1e-6, atol=1e-9)
    assert np.allclose(x, s.xyzq[:, :3] + v0 * s.dt, atol=1e-4)
    # collisions: fraction pnu, Maxwell-Boltzmann velocities
    s.pnu = 0.5
    hits, vs = [], []
    for step in range(40):
        _, v, hit = po.leapfrog(s, np.zeros_like(f), step=step)
        hits.append(hit.mean())
        vs.append(v[hit.astype(bool)] / np.sqrt(s.kT / m[hit.astype(bool)]))
    assert abs(np.mean(hits) - 0.5) < 0.02
    z = np.concatenate(vs).ravel()
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1.0) < 0.02


# ------------------------------------------------- CPU baseline pinning -----
def test_cpu_baseline_restatement_matches_reference_fixture():
    """oracle's restatement of cpu_ispa_force vs the reference file's own
    output (oracle/_ref build of src/isspa_force.omp.c, srand(1), 1 thread)."""
    d = np.load(os.path.join(GOLDEN, "omp_ref_fixture.npz"))
    f = po.cpu_ispa_force(po.lib().orc_cpu_ispa_force, d["xyz"], d["w"], d["x0"], d["g0"], d["gr2"],
                          d["alpha"], d["lj_A"], d["lj_B"], d["ityp"], int(d["n_mc"]), 1, seed=1)
    assert np.array_equal(f, d["forces"])


def test_cpu_baseline_reference_library_if_built():
    ref = po.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    d = np.load(os.path.join(GOLDEN, "omp_ref_fixture.npz"))
    f = po.cpu_ispa_force(ref.cpu_ispa_force, d["xyz"], d["w"], d["x0"], d["g0"], d["gr2"], d["alpha"],
                          d["lj_A"], d["lj_B"], d["ityp"], int(d["n_mc"]), 1, seed=1)
    assert np.array_equal(f, d["forces"])

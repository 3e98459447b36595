"""The FP64 oracle of the differentiable model (CPU).

What pins it:
  * the analytic g, E, F equal the table generator's closed forms wherever
    the switch is 1 (d >= r_lo), and the plain table-path oracle (itself
    pinned to the reference's fixtures) run on finely sampled tables of the
    analytic functions reproduces the analytic forward;
  * every hand-written derivative (per-type functions, Langevin moments, the
    whole adjoint) equals FP64 central differences with h = 1e-4.
"""
import copy

import numpy as np
import pytest

import pyoracle as po
from isspa_amd import synth
from isspa_amd.diff import DiffParams
from isspa_amd.system import IsspaTables, max_abs_rel_rms


@pytest.fixture(scope="module")
def small():
    """three chains (66 atoms), nMC=6: every term and branch, in seconds"""
    s = synth.make_lattice(3, 16.0, 2, 200.0, seed=5)
    s.n_mc = 6
    s.seed = 77
    return s


def test_analytic_functions_match_generator_tables(cfg0):
    s = cfg0
    prm = DiffParams.from_system(s)
    r = 0.05 + 0.1 * np.arange(120)
    for t in range(prm.n_types):
        p = s.type_params[t]
        for b, d in enumerate(r):
            if d < p["r_lo"] or d > 11.9:
                continue
            v, _ = po.diff_type_fn(s, prm, t, d)
            assert abs(v[0] - s.isspa.g[t, b]) <= 2e-6 * max(1.0, abs(v[0])), (t, d)
            assert abs(v[1] - s.isspa.e[t, b]) <= 2e-6 * max(1e-3, abs(v[1])), (t, d)
            assert abs(v[2] - s.isspa.f[t, b]) <= 2e-6 * max(1.0, abs(v[2])), (t, d)


def test_type_function_derivatives(cfg0):
    s = cfg0
    prm = DiffParams.from_system(s)
    h = 1e-6
    rng = np.random.default_rng(0)
    for t in range(prm.n_types):
        p = s.type_params[t]
        # points in each branch: switch ramp, parabola, beyond r_hi
        ds = [p["r_lo"] - 0.1, p["r_lo"] + 0.05, 0.5 * (p["r_lo"] + p["r_hi"]), p["r_hi"] + 0.3,
              float(rng.uniform(p["r_lo"], 11.0))]
        th = prm.theta
        for d in ds:
            _, dv = po.diff_type_fn(s, prm, t, d)
            for k in range(5):
                tp, tm = th.copy(), th.copy()
                tp[t, k] += h
                tm[t, k] -= h
                vp, _ = po.diff_type_fn(s, prm.with_theta(tp), t, d)
                vm, _ = po.diff_type_fn(s, prm.with_theta(tm), t, d)
                fd = (vp - vm) / (2 * h)
                for f in range(3):
                    assert abs(fd[f] - dv[f, k]) <= 1e-5 * max(1.0, abs(fd[f])), (t, d, f, k)


@pytest.mark.parametrize("x", [1e-3, 0.3, 2.0, 4.999, 5.001, 12.0, 60.0])
def test_langevin_derivatives(x):
    h = 1e-6 * max(1.0, x)
    m, dm = po.langevin_derivs(x)
    mp, _ = po.langevin_derivs(x + h)
    mm, _ = po.langevin_derivs(x - h)
    fd = (mp - mm) / (2 * h)
    assert np.allclose(dm, fd, rtol=1e-6, atol=1e-9)
    m_ref = np.zeros(3)
    po.lib().orc_langevin_moments(x, m_ref)
    assert np.allclose(m, m_ref, rtol=1e-14, atol=1e-300)


def test_analytic_forward_matches_table_path_on_fine_tables(small):
    """Sample the analytic g, E, F on a 0.002 A grid, run the plain oracle
    (field, force, nonbond with Lorentz-Berthelot tables) and compare with the
    analytic forward: they differ only by linear interpolation."""
    s = copy.deepcopy(small)
    prm = DiffParams.from_system(s)
    dr, r0 = 0.002, 0.001
    rows = int(12.2 / dr)
    T = prm.n_types
    G = np.zeros((T, rows), np.float32)
    E = np.zeros((T, rows), np.float32)
    F = np.zeros((T, rows), np.float32)
    for t in range(T):
        for b in range(rows):
            v, _ = po.diff_type_fn(s, prm, t, r0 + dr * b)
            G[t, b], E[t, b], F[t, b] = v
    s.isspa = IsspaTables(isspa_type=s.isspa.isspa_type, rmax=s.isspa.rmax, g=G, e=E, f=F,
                          g_grid=(np.float32(r0), np.float32(dr)),
                          e_grid=(np.float32(r0), np.float32(dr)),
                          f_grid=(np.float32(r0), np.float32(dr)))
    nbparm, a, b = prm.lj_tables()
    s.nbparm, s.lj_a, s.lj_b = nbparm, a.astype(np.float32), b.astype(np.float32)
    pts = po.mc_points(s, step=3)
    fwd = po.diff_forward(s, prm, pts)
    fld = po.isspa_field(s, pts)
    comp = po.isspa_force(s, pts, fld)
    tab = comp["lj"] + comp["pol"] + comp["cdd"]
    err = max_abs_rel_rms(fwd["isspa"], tab)
    print(f"IS-SPA analytic vs fine tables: rel RMS {err:.2e}")
    assert err < 2e-4
    nb = po.nonbond(s)["total"]
    assert max_abs_rel_rms(fwd["nb"], nb) < 1e-6


def test_adjoint_matches_central_differences(small):
    s = small
    prm = DiffParams.from_system(s)
    pts = po.mc_points(s, step=1)
    rng = np.random.default_rng(4)
    lam = rng.normal(size=(s.n_atoms, 3))
    gq, gt = po.diff_vjp(s, prm, pts, lam)
    h = 1e-4

    def L(p):
        return float(np.sum(lam * po.diff_forward(s, p, pts)["total"]))

    th = prm.theta
    scale_t = np.abs(gt).max()
    for t in range(prm.n_types):
        for k in range(5):
            tp, tm = th.copy(), th.copy()
            tp[t, k] += h
            tm[t, k] -= h
            fd = (L(prm.with_theta(tp)) - L(prm.with_theta(tm))) / (2 * h)
            assert abs(fd - gt[t, k]) <= 1e-5 * max(abs(fd), 1e-3 * scale_t), (t, k, fd, gt[t, k])
    scale_q = np.abs(gq).max()
    for i in rng.choice(s.n_atoms, size=12, replace=False):
        qp, qm = prm.q.copy(), prm.q.copy()
        qp[i] += h
        qm[i] -= h
        fd = (L(prm.with_theta(th, qp)) - L(prm.with_theta(th, qm))) / (2 * h)
        assert abs(fd - gq[i]) <= 1e-5 * max(abs(fd), 1e-3 * scale_q), (i, fd, gq[i])


def test_forcematch_perturbation_and_frame_batch(small):
    """Perturbation ranges of the force-matching workload, and frames of one
    topology batched as windows sharing one type block."""
    from isspa_amd import forcematch
    prm = DiffParams.from_system(small)
    rng = np.random.default_rng(0)
    th, q = forcematch.perturb(prm.theta, prm.q, rng)
    r = th / prm.theta - 1.0
    assert np.all(np.abs(r[:, 0]) <= 0.05) and np.all(np.abs(r[:, 1]) <= 0.10)
    assert np.all(r[:, 2] == 0.0)
    assert np.all(np.abs(r[:, 3:]) <= 0.10 + 1e-12) or np.all(th[:, 3] >= 1.05)
    assert np.all(np.abs(q - prm.q) <= 0.05 * forcematch.AMBER_CHARGE)
    frames = [small.xyzq[:, :3] + np.float32(0.01 * k) for k in range(3)]
    b = forcematch.frame_batch(small, frames)
    assert list(b.windows) == [small.n_atoms] * 3
    assert b.isspa.n_types == prm.n_types and b.n_lj_types == small.n_lj_types
    assert np.array_equal(b.xyzq[small.n_atoms:2 * small.n_atoms, :3], frames[1])

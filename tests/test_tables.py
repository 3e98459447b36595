"""IS-SPA table generator and .isspatop writer (CPU)."""
import math
import os

import numpy as np

from conftest import GOLDEN
from isspa_amd import io, synth, tables


def test_writer_round_trips_the_shipped_adi2_tables(tmp_path):
    """Read the reference's ADI2.isspatop, write it, read it back: every array
    and grid identical, and the POINTERS line in the writer's column order."""
    src = os.path.join(GOLDEN, "adi2", "ADI2.isspatop")
    a = io.read_isspatop(src)
    out = tmp_path / "copy.isspatop"
    tables.write_isspatop(str(out), a)
    b = io.read_isspatop(str(out))
    for k in ("isspa_type", "rmax", "g", "e", "f"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    for k in ("g_grid", "e_grid", "f_grid"):
        assert getattr(a, k) == getattr(b, k), k
    with open(src) as fh:
        ref_ptr = fh.read().split("%FLAG POINTERS")[1].splitlines()[2]
    with open(out) as fh:
        our_ptr = fh.read().split("%FLAG POINTERS")[1].splitlines()[2]
    assert ref_ptr.split() == our_ptr.split() == ["44", "16", "120", "120", "120"]


def test_parabola_tables_closed_form():
    """g is the parabola between its roots and 1 beyond r_hi (continuous
    there); F is -dU/dr of the mixed LJ pair (checked against a central
    difference of U); E is the Gaussian bump; everything is 0 below r_lo."""
    sig, eps = [1.2, 3.4], [0.05, 0.18]
    x0 = [s + 0.5 for s in sig]
    g0, al, amp = [1.7, 1.3], [1.9, 0.6], [0.02, -0.04]
    G, E, F, par = tables.parabola_tables(sig, eps, x0, g0, al, amp)
    r = 0.05 + 0.1 * np.arange(120)
    for t in range(2):
        p = par[t]
        s_ts, e_ts = 0.5 * (sig[t] + 3.15), math.sqrt(eps[t] * 0.152)
        assert abs(p["sigma_ts"] - s_ts) < 1e-12 and abs(p["eps_ts"] - e_ts) < 1e-12
        # g(r_hi) = 1 from the parabola side
        assert abs(g0[t] - al[t] * (p["r_hi"] - x0[t]) ** 2 - 1.0) < 1e-12
        U = lambda rr: 4 * e_ts * ((s_ts / rr) ** 12 - (s_ts / rr) ** 6)
        for b, rb in enumerate(r):
            if rb < p["r_lo"]:
                assert G[t, b] == 0 and F[t, b] == 0 and E[t, b] == 0
                continue
            h = 1e-6
            dudr = (U(rb + h) - U(rb - h)) / (2 * h)
            assert abs(F[t, b] + dudr) <= 1e-5 * max(1.0, abs(dudr))
            want_g = g0[t] - al[t] * (rb - x0[t]) ** 2 if rb < p["r_hi"] else 1.0
            assert abs(G[t, b] - want_g) <= 1e-6
            assert abs(E[t, b] - amp[t] * math.exp(-0.5 * (rb - x0[t]) ** 2)) <= 1e-7


def test_chain_tables_come_from_the_generator(tmp_path):
    """The workload chain's tables are the generator's; they survive the
    writer and reader unchanged."""
    s = synth.make_chain()
    out = tmp_path / "chain.isspatop"
    tables.write_isspatop(str(out), s.isspa)
    b = io.read_isspatop(str(out))
    assert np.array_equal(b.g, s.isspa.g) and np.array_equal(b.f, s.isspa.f)
    assert np.array_equal(b.isspa_type, s.isspa.isspa_type)


def test_isspa_domain_rule_and_fit_readers(tmp_path):
    """The reference writer's domain rule (first/last r with g|F| > 0.1) and
    its regression-file conventions: exp() of the log column, zero density
    where the count column is not positive."""
    r = 0.05 + 0.1 * np.arange(10)
    G = np.array([[0, 0, 0.5, 1, 1, 1, 1, 1, 1, 1.0]])
    F = np.array([[9, 9, 9, 2, 1, 0.5, 0.2, 0.05, 0.01, 0.0]])
    dom = tables.isspa_domain(r, G, F)
    assert np.allclose(dom, [[r[2], r[6]]])
    T, rows = 2, 4
    rr = 0.05 + 0.1 * np.arange(rows)
    dens = np.zeros((T * rows, 4))
    force = np.zeros((T * rows, 4))
    rng = np.random.default_rng(0)
    for t in range(T):
        for b in range(rows):
            k = t * rows + b
            dens[k] = (rr[b], rng.normal(), t + 1, 0 if b == 0 else 5)
            force[k] = (rr[b], rng.normal(), 0.0, t + 1)
    np.savetxt(tmp_path / "d.dat", dens)
    np.savetxt(tmp_path / "f.dat", force)
    rd, g = tables.read_density_fit(str(tmp_path / "d.dat"))
    assert np.allclose(rd, rr)
    assert np.all(g[:, 0] == 0)
    assert np.allclose(g[:, 1:], np.exp(dens[:, 1].reshape(T, rows)[:, 1:]))
    rf, f = tables.read_force_fit(str(tmp_path / "f.dat"))
    assert np.allclose(f, np.exp(force[:, 1].reshape(T, rows)))
    re, e = tables.read_efield_fit(str(tmp_path / "d.dat"))
    assert np.allclose(e, dens[:, 1].reshape(T, rows))

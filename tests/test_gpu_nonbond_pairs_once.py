"""The all-pairs total force of large windows, each unordered pair once.

Windows of 4,096 atoms or more take their total all-pairs force from 256 x
256-atom super-tiles (every pair of two blocks at least two apart, evaluated
once and added to both atoms) plus ordered blocks for the pairs of a block
with itself and with its two neighbours.  The split components keep
every pair in both orders.  These tests check the total against the FP64
oracle (src/nonbond_force_cuda.cu:31-174 restated), in the open and in a
periodic box, at 4,400 atoms (a ragged last block of 48) and at the 19,998-atom
benchmark workload; against the component path; and run to run bit for bit.
"""
import copy

import numpy as np
import pytest

import pyoracle as po
from isspa_amd.system import max_abs_rel_rms

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _gpu(system, **kw):
    from isspa_amd.engine import GpuSystem
    return GpuSystem(system, "cuda:0", **kw)


def _lattice(pbc):
    from isspa_amd import synth
    s = synth.make_lattice(200, 120.0, 6, 200.0, seed=91)
    if pbc:
        s = copy.deepcopy(s)
        s.xyzq[:, :3] = (s.xyzq[:, :3] - s.xyzq[:, :3].min(0)).astype(np.float32)
        s.lbox = float(np.ceil(s.xyzq[:, :3].max())) + 4.0
    return s


def _near_form_pairs_across_blocks(s):
    """Pairs of atoms in 256-atom blocks at least two apart that lie inside
    the partner's IS-SPA sphere (d^2 <= 4 R_j^2): the super-tiles' rare
    branch.  FP64 minimum image, as a count for the test's own coverage."""
    x = s.xyzq[:, :3].astype(np.float64)
    r = s.isspa.rmax[s.isspa.isspa_type].astype(np.float64)
    n = len(x)
    blk = np.arange(n) // 256
    count = 0
    for i0 in range(0, n, 512):
        d = x[i0:i0 + 512, None, :] - x[None, :, :]
        if s.lbox < 1e4:
            d -= s.lbox * np.round(d / s.lbox)
        d2 = (d * d).sum(-1)
        far_blocks = np.abs(blk[i0:i0 + 512, None] - blk[None, :]) >= 2
        count += int(np.count_nonzero(far_blocks & (d2 <= 4.0 * r[None, :] ** 2)))
    return count


@pytest.mark.parametrize("pbc", [False, True])
def test_pairs_once_matches_oracle_and_components(pbc):
    s = _lattice(pbc)
    assert s.n_atoms >= 4096 and s.n_atoms % 256 != 0
    near = _near_form_pairs_across_blocks(s)
    print(f"case pairs-once: N={s.n_atoms} box={s.lbox} pbc={pbc} near-form pairs across "
          f"blocks={near}", flush=True)
    assert near > 0
    ref = po.nonbond(s)
    g = _gpu(s, isspa=False, bonded=False)
    g.zero_forces()
    g.nonbond_force()
    tot = g.f[:, :3].clone()
    err = max_abs_rel_rms(tot.double().cpu().numpy(), ref["total"])
    print(f"  total vs oracle: rel RMS {err:.3e}", flush=True)
    assert err <= TOL
    # the component path evaluates every pair in both orders
    comp = {k: v.cpu().numpy() for k, v in g.nonbond_components().items()}
    for k in ("lj", "coul", "pair"):
        assert max_abs_rel_rms(comp[k], ref[k]) <= TOL, k
    both = comp["lj"] + comp["coul"] + comp["pair"]
    assert max_abs_rel_rms(tot.double().cpu().numpy(), both) <= 1e-6
    # run to run, bit for bit
    g.zero_forces()
    g.nonbond_force()
    assert bool((g.f[:, :3] == tot).all())
    g.close()


def test_pairs_once_full_size():
    """19,998 atoms (79 blocks, 3,003 super-tiles): the total against the
    oracle."""
    from isspa_amd import synth
    s = synth.workload(2)
    print(f"case pairs-once full size: N={s.n_atoms}", flush=True)
    ref = po.nonbond(s)
    g = _gpu(s, isspa=False, bonded=False)
    g.zero_forces()
    g.nonbond_force()
    tot = g.f[:, :3].double().cpu().numpy()
    err = max_abs_rel_rms(tot, ref["total"])
    print(f"  total vs oracle: rel RMS {err:.3e}", flush=True)
    assert err <= TOL
    g.close()


def test_pairs_once_window_in_a_batch():
    """A batch of a 4,400-atom window (super-tiles) and two small windows
    (partner chunks): each window's total against the oracle, and the large
    window bit for bit equal to its plain run (its blocks depend on the window
    alone)."""
    from isspa_amd import batch, synth
    big = _lattice(False)
    chain = synth.make_chain()
    chain.xyzq[:, :3] += np.float32(100.0)
    mid = synth.make_lattice(10, 40.0, 3, 200.0, seed=7)
    for s in (big, chain, mid):
        s.lbox, s.n_mc, s.temperature_k, s.pnu = 200.0, 20, 298.0, 0.01
    wins = [chain, big, mid]
    merged = batch.merge_windows(wins)
    print(f"case pairs-once batch: windows {[s.n_atoms for s in wins]}", flush=True)
    g = _gpu(merged, isspa=False, bonded=False, us=False)
    g.zero_forces()
    g.nonbond_force()
    tot = g.f[:, :3].clone()
    g.close()
    for s, (a, b) in zip(wins, batch.window_slices(merged)):
        err = max_abs_rel_rms(tot[a:b].double().cpu().numpy(), po.nonbond(s)["total"])
        print(f"  window of {s.n_atoms}: rel RMS {err:.3e}", flush=True)
        assert err <= TOL
    p = _gpu(big, isspa=False, bonded=False, us=False)
    p.zero_forces()
    p.nonbond_force()
    a, b = batch.window_slices(merged)[1]
    assert bool((p.f[:, :3] == tot[a:b]).all())
    p.close()


def _many_lj_types(s, T=89):
    """The same system with its LJ types spread over T types of identical
    parameters (type t becomes t + base * k): the forces are unchanged, and
    the T x T table with a super-tile's partner sums exceeds a block's 64 KB
    of LDS, so the total force keeps the ordered path."""
    s = copy.deepcopy(s)
    base = s.n_lj_types
    reps = (T - 1) // base
    it = s.ityp.astype(np.int64)
    new_t = it + base * (np.arange(s.n_atoms) % reps)

    def tri(a, b):
        lo, hi = min(a, b), max(a, b)
        return hi * (hi + 1) // 2 + lo

    nbparm = np.empty(T * T, np.int32)
    lj_a = np.zeros(T * (T + 1) // 2, np.float32)
    lj_b = np.zeros(T * (T + 1) // 2, np.float32)
    for a in range(T):
        for b in range(T):
            k = tri(a, b)
            nbparm[a * T + b] = k
            src = s.nbparm[(a % base) * base + (b % base)]
            lj_a[k], lj_b[k] = s.lj_a[src], s.lj_b[src]
    s.ityp = new_t.astype(s.ityp.dtype)
    s.n_lj_types, s.nbparm, s.lj_a, s.lj_b = T, nbparm, lj_a, lj_b
    return s


def test_many_lj_types_keep_the_ordered_path():
    s = _many_lj_types(_lattice(False))
    print(f"case many LJ types: N={s.n_atoms} T={s.n_lj_types}", flush=True)
    ref = po.nonbond(s)
    g = _gpu(s, isspa=False, bonded=False)
    g.zero_forces()
    g.nonbond_force()
    err = max_abs_rel_rms(g.f[:, :3].double().cpu().numpy(), ref["total"])
    print(f"  total vs oracle: rel RMS {err:.3e}", flush=True)
    assert err <= TOL
    g.close()

"""Edge cases of the HIP path against the FP64 oracle: a single atom, nMC = 1,
systems one atom past a wave (65+) and past a block (257+), tilings with
more chunks and splits than there is work, and the all-pairs and bonded terms
on those sizes.  Same tolerances as the main parity tests: IS-SPA components
<= 1e-5 relative RMS, nonbond and bonded <= 1e-5."""
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


def one_atom(n_mc):
    from isspa_amd import synth
    s = copy.deepcopy(synth.two_atom_lj(sep=4.0))
    for name in ("xyzq", "vel", "true_mass", "ityp"):
        setattr(s, name, getattr(s, name)[:1].copy())
    s.xyzq[0, 3] = 0.3
    s.nexcl_prefix = np.array([1], np.int32)
    s.excl_list = np.array([-1], np.int32)
    s.isspa.isspa_type = s.isspa.isspa_type[:1].copy()
    s.us = None
    s.n_mc = n_mc
    s.name = f"one_atom_nmc{n_mc}"
    s.check()
    return s


def two_charged(n_mc):
    from isspa_amd import synth
    s = copy.deepcopy(synth.two_atom_lj(sep=3.1))
    s.xyzq[:, 3] = np.array([0.4, -0.25], np.float32)
    s.us = None
    s.n_mc = n_mc
    s.name = f"two_atoms_nmc{n_mc}"
    return s


def lattice(chains, n_mc):
    from isspa_amd import synth
    s = synth.make_lattice(chains, 30.0 + 6 * chains ** (1 / 3), 4, 200.0, seed=77 + chains)
    s.n_mc = n_mc
    s.name = f"lattice{chains}_nmc{n_mc}"
    return s


CASES = {
    "one_atom_nmc1": lambda: one_atom(1),
    "one_atom_nmc7": lambda: one_atom(7),
    "two_atoms_nmc3": lambda: two_charged(3),
    "66_atoms_nmc2": lambda: lattice(3, 2),     # one wave and two atoms
    "264_atoms_nmc1": lambda: lattice(12, 1),   # one block and eight atoms
}


@pytest.mark.parametrize("case", list(CASES))
def test_isspa_edge_sizes_match_oracle(case):
    s = CASES[case]()
    print(f"case {case}: N={s.n_atoms} nMC={s.n_mc}", flush=True)
    g = _gpu(s, nonbond=False, bonded=False)
    for step in (0, 3):
        comp = g.isspa_components(step=step)
        ref = po.isspa_total(s, step=step)
        for k in ("lj", "pol", "cdd"):
            got = comp[k].cpu().numpy()
            assert np.all(np.isfinite(got)), (k, step)
            if np.abs(ref[k]).max() == 0.0:
                assert np.abs(got).max() == 0.0, (k, step)
                continue
            err = max_abs_rel_rms(got, ref[k])
            print(f"  step {step} {k}: rel RMS {err:.3e}", flush=True)
            assert err <= TOL, (k, step, err)
    g.close()


@pytest.mark.parametrize("case", ["one_atom_nmc7", "66_atoms_nmc2"])
def test_isspa_oversized_tilings_are_invariant(case):
    """More field splits than atoms and more force chunks than points: the
    extra splits and chunks are empty and the result does not change."""
    s = CASES[case]()
    g = _gpu(s, nonbond=False, bonded=False)
    base = {k: v.cpu().numpy() for k, v in g.isspa_components(step=1).items()}
    for splits, chunks in ((s.n_atoms + 5, s.n_points + 9), (1, 1), (64, 1000)):
        g.set_tiling(splits, chunks)
        other = g.isspa_components(step=1)
        for k in base:
            o = other[k].cpu().numpy()
            if np.abs(base[k]).max() == 0.0:
                assert np.abs(o).max() == 0.0
            else:
                assert max_abs_rel_rms(o, base[k]) < 1e-6, (splits, chunks, k)
    g.close()


@pytest.mark.parametrize("case", ["one_atom_nmc7", "264_atoms_nmc1"])
def test_force_keep_matches_force_and_keeps_the_isspa_part(case):
    """isspa_model_force_keep adds into f exactly what isspa_model_force adds
    and stores that same FP32 force alone (w = 0) in the second array."""
    import torch
    s = CASES[case]()
    g = _gpu(s, nonbond=False, bonded=False)
    gen = torch.Generator().manual_seed(5)
    f0 = torch.randn((s.n_atoms, 4), generator=gen).to("cuda:0")
    ref = g.isspa_force(step=2, f=f0.clone())
    alone = torch.zeros_like(f0)
    g.isspa_force(step=2, f=alone)
    f = f0.clone()
    kept = torch.full_like(f0, 7.0)
    g.isspa_force_keep(f, kept, step=2)
    torch.cuda.synchronize()
    assert torch.equal(f, ref)
    assert torch.equal(kept[:, :3], alone[:, :3])
    assert torch.all(kept[:, 3] == 0.0)
    g.close()


@pytest.mark.parametrize("case", ["two_atoms_nmc3", "66_atoms_nmc2", "264_atoms_nmc1"])
def test_nonbond_and_bonded_edge_sizes(case):
    s = CASES[case]()
    g = _gpu(s, isspa=False)
    got = g.nonbond_force()[:, :3].double().cpu().numpy()
    ref = po.nonbond(s)["total"]
    assert max_abs_rel_rms(got, ref) <= TOL
    if len(s.bond_atoms):
        g.zero_forces()
        got = g.bonded_force()[0][:, :3].double().cpu().numpy()
        assert max_abs_rel_rms(got, po.bonded(s)["total"]) <= TOL
    g.close()


def test_single_atom_engine_steps():
    """The whole step on one atom: finite, and the graph replay equals the
    eager step bit for bit."""
    import torch
    s = one_atom(5)
    a, b = _gpu(s), _gpu(s)
    a.set_step(0)
    b.set_step(0)
    b.capture(1)
    a.step(4)
    b.replay(4)
    torch.cuda.synchronize()
    assert torch.isfinite(a.x).all() and torch.isfinite(a.v).all()
    assert torch.equal(a.x, b.x) and torch.equal(a.v, b.v)
    a.close()
    b.close()

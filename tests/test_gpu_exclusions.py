"""The nonbond exclusion list beyond the +-64-index bit windows, on the GPU.

Exclusions within 64 indices of the receiver are bits of a per-atom 128-bit
window; any farther one turns on a per-atom list that the all-pairs kernel
and the nonbond adjoint scan as well (the reference keeps the whole list in
shared memory, src/nonbond_force_cuda.cu:34,57-108).  Relabelled systems
(tests/relabel.py) put excluded pairs up to 80 and more indices apart:

* ADI2 spread over 352 atoms (its atoms at every 8th index, inert fillers
  between them) against the reference's FP64 golden files
  ADI2.multi.frc.{LJsolu,Csolu,Csolv_pair}.dat, at the bar the reference's
  own FP32 output meets (<= 2.8e-5 absolute);
* the 1,100-atom lattice in a random atom order against the oracle and
  against the same lattice in its own order;
* the differentiable forward and adjoint on a permuted lattice against the
  FP64 oracle adjoint.
"""
import os

import numpy as np
import pytest
import torch

import pyoracle as po
from conftest import GOLDEN
from isspa_amd.system import max_abs_rel_rms
from relabel import max_exclusion_span, relabel, spread_with_fillers

pytestmark = pytest.mark.gpu


def _nonbond(s):
    from isspa_amd.engine import GpuSystem
    g = GpuSystem(s, "cuda:0", isspa=False, bonded=False)
    comp = {k: v.cpu().numpy() for k, v in g.nonbond_components().items()}
    g.zero_forces()
    g.nonbond_force()
    tot = g.f[:, :3].double().cpu().numpy()
    g.close()
    return comp, tot


def test_exclusion_list_spread_adi2_golden(adi2):
    t = spread_with_fillers(adi2, stride=8)
    print(f"case spread ADI2: {t.n_atoms} atoms, exclusions up to {max_exclusion_span(t)} apart",
          flush=True)
    comp, tot = _nonbond(t)
    for k, name in (("lj", "LJsolu"), ("coul", "Csolu"), ("pair", "Csolv_pair")):
        ref = np.loadtxt(os.path.join(GOLDEN, "adi2", f"ADI2.multi.frc.{name}.dat"))[:, 3:]
        got = comp[k][0::8]
        err = np.abs(got - ref).max()
        print(f"  {name}: max abs {err:.2e}, rel RMS {max_abs_rel_rms(got, ref):.2e}", flush=True)
        assert err <= 2.8e-5 and max_abs_rel_rms(got, ref) <= 1e-5
    filler = np.ones(t.n_atoms, bool)
    filler[0::8] = False
    assert all(np.all(comp[k][filler] == 0.0) for k in comp)
    ref = po.nonbond(t)["total"]
    assert max_abs_rel_rms(tot, ref) <= 1e-5


def test_exclusion_list_permuted_lattice(cfg1):
    perm = np.random.default_rng(3).permutation(cfg1.n_atoms)
    t = relabel(cfg1, perm, cfg1.n_atoms)
    print(f"case permuted lattice: {t.n_atoms} atoms, exclusions up to "
          f"{max_exclusion_span(t)} apart", flush=True)
    comp, tot = _nonbond(t)
    ref = po.nonbond(t)
    for k in ("lj", "coul", "pair"):
        assert max_abs_rel_rms(comp[k], ref[k]) <= 1e-5, k
    assert max_abs_rel_rms(tot, ref["total"]) <= 1e-5
    base, _ = _nonbond(cfg1)
    for k in ("lj", "coul", "pair"):
        assert max_abs_rel_rms(comp[k][perm], base[k]) <= 1e-5, k


def test_exclusion_list_adjoint_permuted_lattice():
    from isspa_amd import synth
    from isspa_amd.diff import DiffForces, DiffParams
    s = synth.make_lattice(4, 20.0, 2, 200.0, seed=5)
    s.n_mc = 6
    s.seed = 77
    perm = np.random.default_rng(8).permutation(s.n_atoms)
    t = relabel(s, perm, s.n_atoms)
    t.type_params = s.type_params
    print(f"case adjoint permuted lattice: {t.n_atoms} atoms, exclusions up to "
          f"{max_exclusion_span(t)} apart", flush=True)
    assert max_exclusion_span(t) > 64
    prm = DiffParams.from_system(t)
    m = DiffForces(t, prm, step=2)
    q = torch.as_tensor(np.asarray(prm.q, np.float64), device="cuda:0")
    th = torch.as_tensor(np.asarray(prm.theta, np.float64), device="cuda:0")
    f = m.forces(q, th).cpu().numpy()
    pts = po.mc_points(t, step=2)
    ref = po.diff_forward(t, prm, pts)
    assert max_abs_rel_rms(f, ref["total"]) <= 1e-5
    lam = np.random.default_rng(2).normal(size=(t.n_atoms, 3))
    gq, gt = m.vjp(q, th, torch.as_tensor(lam, device="cuda:0"))
    rq, rt = po.diff_vjp(t, prm, pts, lam)
    eq = max_abs_rel_rms(gq.cpu().numpy(), rq)
    gt = gt.cpu().numpy()
    et = np.abs(gt - rt).max() / np.abs(rt).max()
    print(f"  dL/dq rel RMS {eq:.2e}; dL/dtheta max rel {et:.2e}", flush=True)
    assert eq <= 1e-4 and et <= 1e-4

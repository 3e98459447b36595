"""Receiver replicas of the force pass on the GPU.

Windows of at most 64 atoms split each chunk's points over the force block's
four waves (4 replicas of the receivers), windows of 65-128 atoms over two
pairs of waves (2 replicas); larger windows have none.  These tests cover each
regime and both boundaries (44, 66, 110 and 132 atoms: 2, 3, 5 and 6 chains of
22) against the FP64 oracle with the same Philox points, component by
component (the split-mode kernel sums nine components across replicas), and
check that the engine's forces -- the epilogue inside the force launch for
these small windows -- equal the separate calls bit for bit.
"""
import numpy as np
import pytest
import torch

import pyoracle as po
from isspa_amd.system import max_abs_rel_rms

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _system(chains):
    from isspa_amd import synth
    s = synth.make_lattice(chains, 30.0, 3, 200.0, seed=40 + chains)
    s.n_mc = 20
    s.seed = 1234
    return s


def _gpu(system, **kw):
    from isspa_amd.engine import GpuSystem
    return GpuSystem(system, "cuda:0", **kw)


@pytest.mark.parametrize("chains,tiling", [(2, None), (3, None), (3, (3, 7)), (5, None),
                                           (5, (4, 50)), (6, None)])
def test_replica_force_components_match_oracle(chains, tiling):
    s = _system(chains)
    g = _gpu(s, nonbond=False, bonded=False)
    if tiling is not None:
        g.set_tiling(*tiling)
    print(f"case replicas: N={s.n_atoms} nMC={s.n_mc} tiling {g.tiling()}", flush=True)
    comp = g.isspa_components(step=3)
    pts = po.mc_points(s, step=3)
    rf = po.isspa_field(s, pts)
    ref = po.isspa_force(s, pts, rf)
    for k in ("lj", "pol", "cdd"):
        err = max_abs_rel_rms(comp[k].cpu().numpy(), ref[k])
        print(f"  {k}: rel RMS {err:.3e}", flush=True)
        assert err <= TOL, (k, err)
    g.zero_forces()
    g.isspa_force(step=3)
    tot = g.f[:, :3].double().cpu().numpy()
    assert max_abs_rel_rms(tot, po.isspa_total(s, step=3)["total"]) <= TOL
    g.close()


@pytest.mark.parametrize("chains", [2, 3, 5, 6])
def test_replica_engine_equals_separate_calls_bitwise(chains):
    s = _system(chains)
    g = _gpu(s, us=False)
    print(f"case replica engine: N={s.n_atoms} tiling {g.tiling()}", flush=True)
    g.set_step(5)
    got = g.forces().clone()
    g.zero_forces()
    g.bonded_force()
    g.isspa_force(step=5)
    g.nonbond_force()
    torch.cuda.synchronize()
    assert torch.equal(got, g.f)
    # and a second evaluation gives the same bits (fixed replica order)
    g.set_step(5)
    again = g.forces().clone()
    torch.cuda.synchronize()
    assert torch.equal(got, again)
    g.close()

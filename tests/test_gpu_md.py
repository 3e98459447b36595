"""Nonbond, bonded, integrator, restraint and whole-step kernels on the GPU
against the oracle and the reference's golden vectors."""
import copy
import os

import numpy as np
import pytest
import torch

import pyoracle as po
from conftest import GOLDEN
from isspa_amd.system import UmbrellaRestraint, max_abs_rel_rms

pytestmark = pytest.mark.gpu


def _gpu(system, **kw):
    from isspa_amd.engine import GpuSystem
    return GpuSystem(system, "cuda:0", **kw)


def test_nonbond_golden_adi2(adi2):
    """timing_tests/ADI2/10000MC/compare/ADI2.multi.frc.{LJsolu,Csolu,Csolv_pair}.dat:
    FP64 golden per-atom forces.  The reference's own FP32 GPU output matches
    them to <= 2.8e-5 absolute; the HIP path must do at least as well."""
    g = _gpu(adi2, isspa=False, bonded=False)
    comp = g.nonbond_components()
    for k, name in (("lj", "LJsolu"), ("coul", "Csolu"), ("pair", "Csolv_pair")):
        ref = np.loadtxt(os.path.join(GOLDEN, "adi2", f"ADI2.multi.frc.{name}.dat"))
        got = comp[k].cpu().numpy()
        err = np.abs(got - ref[:, 3:]).max()
        print(f"{name}: max abs {err:.2e}  rel RMS {max_abs_rel_rms(got, ref[:, 3:]):.2e}")
        assert err <= 2.8e-5
        assert max_abs_rel_rms(got, ref[:, 3:]) <= 1e-5
    g.close()


@pytest.mark.parametrize("name", ["cfg0", "cfg1"])
def test_nonbond_matches_oracle(name, cfg0, cfg1):
    s = {"cfg0": cfg0, "cfg1": cfg1}[name]
    g = _gpu(s, isspa=False, bonded=False)
    comp = g.nonbond_components()
    ref = po.nonbond(s)
    for k in ("lj", "coul", "pair"):
        assert max_abs_rel_rms(comp[k].cpu().numpy(), ref[k]) <= 1e-5, k
    g.zero_forces()
    g.nonbond_force()
    assert max_abs_rel_rms(g.f[:, :3].double().cpu().numpy(), ref["total"]) <= 1e-5
    g.close()


def test_nonbond_pbc(cfg1):
    s = copy.deepcopy(cfg1)
    s.lbox = 70.0
    s.xyzq[:, :3] = (s.xyzq[:, :3] - s.xyzq[:, :3].min(0)).astype(np.float32)
    g = _gpu(s, isspa=False, bonded=False)
    comp = g.nonbond_components()
    ref = po.nonbond(s)
    for k in ("lj", "coul", "pair"):
        assert max_abs_rel_rms(comp[k].cpu().numpy(), ref[k]) <= 1e-5, k
    g.close()


@pytest.mark.parametrize("name", ["cfg1", "adi2"])
def test_bonded_matches_oracle(name, cfg1, adi2):
    """Positions jittered by 0.05 A: the generator builds bonds at exactly r0,
    where the bond force is FP32 cancellation noise in any implementation."""
    s = copy.deepcopy({"cfg1": cfg1, "adi2": adi2}[name])
    jit = np.random.default_rng(1).normal(scale=0.05, size=(s.n_atoms, 3))
    s.xyzq[:, :3] = (s.xyzq[:, :3] + jit).astype(np.float32)
    g = _gpu(s, isspa=False, nonbond=False)
    g.zero_forces()
    _, e = g.bonded_force()
    ref = po.bonded(s)
    got = g.f[:, :3].double().cpu().numpy()
    assert max_abs_rel_rms(got, ref["total"]) <= 1e-5
    e = e.cpu().numpy()
    for k, key in enumerate(("e_bond", "e_angle", "e_dih", "e_14coul", "e_14lj")):
        assert abs(e[k] - ref[key]) <= 1e-5 * max(1.0, abs(ref[key])), key
    for mask, key in ((1, "bond"), (2, "angle"), (4, "dih")):
        g.zero_forces()
        g.bonded_force(mask)
        assert max_abs_rel_rms(g.f[:, :3].double().cpu().numpy(), ref[key]) <= 1e-5, key
    g.close()


def test_leapfrog_andersen_matches_oracle(cfg1):
    s = copy.deepcopy(cfg1)
    s.pnu = 0.3
    g = _gpu(s, isspa=False, nonbond=False, bonded=False)
    rng = np.random.default_rng(3)
    f = rng.normal(size=(s.n_atoms, 3)) * 5.0
    g.f[:, :3] = torch.from_numpy(f.astype(np.float32)).to(g.dev)
    g.leapfrog(step=11)
    po_x, po_v, hit = po.leapfrog(s, g.f[:, :3].double().cpu().numpy(), step=11)
    x = g.x[:, :3].double().cpu().numpy()
    v = g.v[:, :3].double().cpu().numpy()
    assert 0.2 < hit.mean() < 0.4
    assert max_abs_rel_rms(v, po_v) < 1e-5
    assert np.abs(x - po_x).max() < 1e-4
    # charge and mass columns untouched
    assert np.array_equal(g.x[:, 3].cpu().numpy(), s.xyzq[:, 3])
    assert np.array_equal(g.v[:, 3].cpu().numpy(), s.vel[:, 3])
    g.close()


def test_us_matches_oracle(adi2):
    s = copy.deepcopy(adi2)
    s.us = UmbrellaRestraint.from_groups(range(0, 22), range(22, 44), s.true_mass, 20.0, 6.5)
    g = _gpu(s, isspa=False, nonbond=False, bonded=False)
    g.zero_forces()
    _, com = g.us_force()
    ref = po.us_force(s)
    assert max_abs_rel_rms(g.f[:, :3].double().cpu().numpy(), ref["f"]) <= 1e-5
    assert np.abs(com.cpu().numpy().reshape(2, 3) - ref["com"]).max() < 1e-4
    g.close()


def _oracle_forces(s, xyzq, step):
    it = po.isspa_total(s, step=step, xyzq=xyzq)["total"]
    nb = po.nonbond(s, xyzq)["total"]
    bd = po.bonded(s, xyzq)["total"]
    tot = it + nb + bd
    if s.us is not None:
        tot = tot + po.us_force(s, xyzq)["f"]
    return tot


def test_engine_forces_and_step(cfg0):
    s = copy.deepcopy(cfg0)
    s.us = UmbrellaRestraint.from_groups(range(0, 5), range(17, 22), s.true_mass, 10.0, 8.0)
    g = _gpu(s)
    g.set_step(4)
    g.forces()
    ref = _oracle_forces(s, s.xyzq, 4)
    assert max_abs_rel_rms(g.f[:, :3].double().cpu().numpy(), ref) <= 1e-5
    g.step(1)  # recomputes the same forces (step 4), then integrates
    x, v, _ = po.leapfrog(s, g.f[:, :3].double().cpu().numpy(), step=4)
    assert np.abs(g.x[:, :3].double().cpu().numpy() - x).max() < 1e-4
    g.close()


@pytest.mark.parametrize("name", ["adi2", "cfg1"])
def test_standalone_calls_equal_the_engine_bitwise(name, adi2, cfg1):
    """Without the side stream: zero + bonded_force + nonbond_force give the
    engine's forces (prologue, all-pairs partials, assembly) bit for bit."""
    s = {"adi2": adi2, "cfg1": cfg1}[name]
    g = _gpu(s, isspa=False, us=False)
    g.set_step(0)
    got = g.forces().clone()
    g.zero_forces()
    g.bonded_force()
    g.nonbond_force()
    torch.cuda.synchronize()
    assert torch.equal(got, g.f)
    g.close()


@pytest.mark.parametrize("name,tiling", [("adi2", (5, 200)), ("cfg0", (2, 10)),
                                         ("cfg0_auto", None)])
def test_engine_fused_force_launch_equals_standalone_calls_bitwise(name, tiling, adi2, cfg0):
    """Systems whose windows fit one receiver block (and whose chunks hold at
    most 256 points, 512 blocks in all) run the epilogue inside the force
    launch, records in LDS.  Their engine forces must equal zero +
    bonded_force + isspa_force (epilogue and force pass as two launches,
    records in memory) + nonbond_force bit for bit.  ADI2 at nMC = 1000 with
    200 chunks has 220 points a chunk; the 22-atom chain at nMC = 20 has 44
    with 10 chunks, 16 with the automatic tiling."""
    s = {"adi2": adi2, "cfg0": cfg0, "cfg0_auto": cfg0}[name]
    g = _gpu(s, us=False)
    if tiling is not None:
        g.set_tiling(*tiling)
    print(f"case fused force launch {name}: N={s.n_atoms} nMC={s.n_mc} tiling {g.tiling()}",
          flush=True)
    assert s.n_atoms <= 256
    g.set_step(2)
    got = g.forces().clone()
    g.zero_forces()
    g.bonded_force()
    g.isspa_force(step=2)
    g.nonbond_force()
    torch.cuda.synchronize()
    assert torch.equal(got, g.f)
    assert torch.all(torch.isfinite(got))
    g.close()


@pytest.mark.parametrize("mode", ["eager", "graph"])
def test_engine_step_counter_advances_once_per_step(cfg1, mode):
    """The engine keeps its step on the device and the leap-frog launch moves
    it on.  After set_step(5) and 4 steps (eager, or two replays of a 2-step
    graph) the next force evaluation draws the MC points of step 9 around the
    current coordinates, bit for bit as the oracle draws them."""
    s = cfg1
    g = _gpu(s)
    g.set_step(5)
    if mode == "eager":
        g.step(4)
    else:
        g.capture(2)
        g.replay(2)
    g.forces()
    torch.cuda.synchronize()
    pts = g.field()["pts"].cpu().numpy()
    x = g.x.cpu().numpy()
    g.close()
    ref = po.mc_points(s, xyzq=x, step=9)
    assert np.array_equal(pts[:, :3].view(np.uint32), ref[:, :3].view(np.uint32))
    wrong = po.mc_points(s, xyzq=x, step=8)
    assert not np.array_equal(pts[:, :3].view(np.uint32), wrong[:, :3].view(np.uint32))


def test_graph_replay_matches_eager(cfg1):
    """Three replays of a captured 2-step graph, with eager work and
    synchronisations between them, against the same steps run eagerly."""
    s = cfg1
    eager = _gpu(s)
    graph = _gpu(s)
    eager.set_step(0)
    graph.set_step(0)
    graph.capture(2)
    for r in range(3):
        eager.step(2)
        graph.replay(1)
        torch.cuda.synchronize()
        junk = torch.randn(1000, device="cuda:0").sum().item()  # eager work in between
        assert np.isfinite(junk)
        xe = eager.x.cpu().numpy()
        xg = graph.x.cpu().numpy()
        assert np.array_equal(xe, xg), f"replay {r} differs"
        assert np.array_equal(eager.v.cpu().numpy(), graph.v.cpu().numpy())
    eager.close()
    graph.close()


def test_engine_side_stream_path():
    """At 4,400 atoms the engine runs the all-pairs kernel on a side stream
    beside the IS-SPA passes.  Its forces equal the separate calls made in the
    same order bit for bit, and two replays of the captured step equal two
    eager steps bit for bit."""
    from isspa_amd import synth
    s = synth.make_lattice(200, 120.0, 6, 200.0, seed=91)
    s.n_mc = 2
    assert s.n_atoms >= 4096
    print(f"N={s.n_atoms} nMC={s.n_mc}", flush=True)
    g = _gpu(s)
    g.set_step(3)
    g.forces()
    torch.cuda.synchronize()
    got = g.f.clone()
    g.zero_forces()
    g.bonded_force()
    g.isspa_force(step=3)
    g.nonbond_force()
    torch.cuda.synchronize()
    assert torch.equal(got, g.f)
    g.close()
    eager, graph = _gpu(s), _gpu(s)
    eager.set_step(0)
    graph.set_step(0)
    graph.capture(1)
    eager.step(2)
    graph.replay(2)
    torch.cuda.synchronize()
    assert torch.equal(eager.x, graph.x) and torch.equal(eager.v, graph.v)
    eager.close()
    graph.close()


def test_full_size_invariants():
    """19,998 atoms, nMC = 20 (the benchmark workload), through properties
    that hold at any size.
    * The LJ and Coulomb pair forces and the bonded forces each sum to zero
      (Newton's third law, up to FP32 rounding).
    * Translation, bit for bit: with every coordinate in [256, 512), where
      FP32 values lie on a 2^-15 grid, a shift by a multiple of 2^-14 is exact,
      the MC points (drawn relative to their owners with the same counters)
      shift exactly too, and every separation is the same float.  So every
      force term must come out identical; a wrong image, index or counter
      would change it."""
    from isspa_amd import synth
    s = copy.deepcopy(synth.workload(2))
    s.xyzq[:, :3] = (s.xyzq[:, :3].astype(np.float64) + 150.0).astype(np.float32)
    assert s.xyzq[:, :3].min() > 256 + 12 and s.xyzq[:, :3].max() < 512 - 12
    print(f"N={s.n_atoms} nMC={s.n_mc}", flush=True)

    def terms(system):
        g = _gpu(system, us=False)
        nb = {k: v.cpu().numpy() for k, v in g.nonbond_components().items()}
        g.zero_forces()
        fb = g.bonded_force()[0][:, :3].double().cpu().numpy()
        out = {**{k: v.cpu().numpy() for k, v in g.isspa_components(step=2).items()},
               "nb_lj": nb["lj"], "nb_coul": nb["coul"], "nb_pair": nb["pair"], "bonded": fb}
        g.close()
        return out

    base = terms(s)
    for k in ("nb_lj", "nb_coul"):
        net = np.abs(base[k].sum(axis=0)).max()
        scale = np.abs(base[k]).sum(axis=0).max()
        print(f"  {k}: |sum F| {net:.3e} of sum |F| {scale:.3e}", flush=True)
        assert net <= 1e-6 * scale, k
    fb = base["bonded"]
    assert np.abs(fb.sum(axis=0)).max() <= 1e-5 * np.abs(fb).sum(axis=0).max()
    t = copy.deepcopy(s)
    shift = np.array([3, -5, 7], np.float32) * np.float32(2.0 ** -10)
    t.xyzq[:, :3] = s.xyzq[:, :3] + shift
    assert np.array_equal(t.xyzq[:, :3] - s.xyzq[:, :3], np.broadcast_to(shift, (s.n_atoms, 3)))
    moved = terms(t)
    for k in base:
        same = np.array_equal(moved[k], base[k])
        print(f"  translated {k}: identical {same}", flush=True)
        assert same, k

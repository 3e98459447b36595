"""Differentiable IS-SPA + nonbond forces on the GPU against the FP64 oracle
(forward and hand-written adjoint; the oracle's adjoint is pinned to central
differences by tests/test_diff_oracle.py).

Tolerances: forces <= 1e-5 relative RMS (FP32 with FP64 accumulation);
gradients <= 1e-4 relative (per-type parameters: per component against the
largest of its kind; charges: relative RMS over atoms)."""
import numpy as np
import pytest
import torch

import pyoracle as po
from isspa_amd import batch, synth
from isspa_amd.diff import DiffForces, DiffParams, isspa_forces
from isspa_amd.system import max_abs_rel_rms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    s = synth.make_lattice(3, 16.0, 2, 200.0, seed=5)
    s.n_mc = 6
    s.seed = 77
    return s


def _t(a):
    return torch.as_tensor(np.asarray(a, np.float64), device="cuda:0")


def _grad_close(got, ref, tol=1e-4):
    got, ref = np.asarray(got), np.asarray(ref)
    scale = np.abs(ref).max(axis=0, keepdims=True) if ref.ndim == 2 else np.abs(ref).max()
    return np.abs(got - ref).max() / max(float(np.max(scale)), 1e-300), np.all(
        np.abs(got - ref) <= tol * np.maximum(scale, 1e-30))


@pytest.mark.parametrize("name,step", [("small", 1), ("cfg1", 4)])
def test_diff_forces_and_adjoint_match_oracle(name, step, small, cfg1):
    s = {"small": small, "cfg1": cfg1}[name]
    prm = DiffParams.from_system(s)
    print(f"case {name} N={s.n_atoms} nMC={s.n_mc} step {step}", flush=True)
    m = DiffForces(s, prm, step=step)
    q, th = _t(prm.q), _t(prm.theta)
    f = m.forces(q, th).cpu().numpy()
    pts = po.mc_points(s, step=step)
    ref = po.diff_forward(s, prm, pts)
    err = max_abs_rel_rms(f, ref["total"])
    print(f"  forces rel RMS {err:.2e}", flush=True)
    assert err <= 1e-5
    rng = np.random.default_rng(2)
    lam = rng.normal(size=(s.n_atoms, 3))
    gq, gt = m.vjp(q, th, _t(lam))
    rq, rt = po.diff_vjp(s, prm, pts, lam)
    eq = max_abs_rel_rms(gq.cpu().numpy(), rq)
    et, ok = _grad_close(gt.cpu().numpy(), rt)
    print(f"  dL/dq rel RMS {eq:.2e}; dL/dtheta max rel {et:.2e}", flush=True)
    assert eq <= 1e-4
    assert ok, (gt.cpu().numpy(), rt)
    m.close()


def test_diff_batched_windows_match_oracle_per_window(small):
    """Four configurations of one topology as windows of one model: each
    window's forces and its share of the adjoint match the oracle run on
    that configuration alone with its own points (system id w)."""
    rng = np.random.default_rng(9)
    confs = []
    for w in range(4):
        c = synth.make_lattice(3, 16.0, 2, 200.0, seed=5)
        c.n_mc, c.seed = 6, 77
        c.xyzq[:, :3] += rng.normal(scale=0.05, size=(c.n_atoms, 3)).astype(np.float32)
        confs.append(c)
    merged = batch.merge_windows(confs)
    prm1 = DiffParams.from_system(confs[0])
    # the windows share one type block: per-type theta unchanged, charges
    # repeated per window
    assert merged.isspa.n_types == prm1.n_types
    prm = DiffParams(sigma=prm1.sigma, eps=prm1.eps, x0=prm1.x0, g0=prm1.g0, alpha=prm1.alpha,
                     camp=prm1.camp, q=merged.xyzq[:, 3].astype(np.float64))
    m = DiffForces(merged, prm, step=2)
    q, th = _t(prm.q), _t(prm.theta)
    f = m.forces(q, th).cpu().numpy()
    lam = rng.normal(size=(merged.n_atoms, 3))
    gq, gt = m.vjp(q, th, _t(lam))
    gq = gq.cpu().numpy()
    rt_sum = np.zeros_like(gt.cpu().numpy())
    for w, (c, (a, b)) in enumerate(zip(confs, batch.window_slices(merged))):
        pts = po.mc_points(c, step=2, sys_id=w)
        ref = po.diff_forward(c, prm1, pts)
        assert max_abs_rel_rms(f[a:b], ref["total"]) <= 1e-5, w
        rq, rt = po.diff_vjp(c, prm1, pts, lam[a:b])
        assert max_abs_rel_rms(gq[a:b], rq) <= 1e-4, w
        rt_sum += rt
    et, ok = _grad_close(gt.cpu().numpy(), rt_sum)
    assert ok, et
    m.close()


def test_autograd_function_and_force_matching_step(small):
    """isspa_forces as a torch op: the gradient of a force-matching loss
    through the Function equals the adjoint applied to dL/dF, and a few
    Adam steps from perturbed parameters reduce the loss."""
    s = small
    prm = DiffParams.from_system(s)
    m = DiffForces(s, prm, step=0)
    th0, q0 = _t(prm.theta), _t(prm.q)
    target = m.forces(q0, th0)
    rng = np.random.default_rng(1)
    pert = prm.theta * (1.0 + rng.uniform(-0.03, 0.03, size=prm.theta.shape))
    th = _t(pert).requires_grad_(True)
    q = (q0 + _t(rng.uniform(-0.5, 0.5, size=s.n_atoms))).requires_grad_(True)
    loss = ((isspa_forces(th, q, m) - target) ** 2).mean()
    loss.backward()
    lam = 2.0 * (m.forces(q.detach(), th.detach()) - target) / target.numel()
    gq, gt = m.vjp(q.detach(), th.detach(), lam)
    assert torch.allclose(th.grad, gt) and torch.allclose(q.grad, gq)
    # a short optimisation in relative units
    rel = torch.zeros_like(th0, requires_grad=True)
    dq = torch.zeros_like(q0, requires_grad=True)
    base_th = _t(pert)
    base_q = q.detach().clone()
    opt = torch.optim.Adam([rel, dq], lr=5e-3)
    losses = []
    for it in range(30):
        opt.zero_grad()
        f = isspa_forces(base_th * (1.0 + rel), base_q + dq, m)
        L = ((f - target) ** 2).mean()
        L.backward()
        opt.step()
        losses.append(L.item())
    print(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
    assert losses[-1] < 0.9 * losses[0]
    m.close()


def test_vjp_after_forces_reuses_the_forward(small):
    """isspa_diff_vjp_after_forces (the adjoint without re-running the
    forward) equals isspa_diff_vjp bit for bit, and refuses when the forward
    state it would reuse is gone: before any forces(), after a plain vjp(),
    and after a tiling change."""
    s = small
    prm = DiffParams.from_system(s)
    m = DiffForces(s, prm, step=2)
    th, q = _t(prm.theta), _t(prm.q)
    lam = _t(np.random.default_rng(5).normal(size=(s.n_atoms, 3)))
    with pytest.raises(RuntimeError):
        m.vjp(q, th, lam, after_forces=True)
    m.forces(q, th)
    gq_r, gt_r, ga_r = m.vjp(q, th, lam, per_atom=True, after_forces=True)
    gq_p, gt_p, ga_p = m.vjp(q, th, lam, per_atom=True)
    torch.cuda.synchronize()
    assert torch.equal(gq_r, gq_p) and torch.equal(gt_r, gt_p) and torch.equal(ga_r, ga_p)
    with pytest.raises(RuntimeError):  # the plain vjp replaced the forward state
        m.vjp(q, th, lam, after_forces=True)
    m.forces(q, th)
    m.set_tiling(2, 3)
    with pytest.raises(RuntimeError):  # launch tables rebuilt since that forward
        m._capi.check(m._capi.lib.isspa_diff_vjp_after_forces(
            m.h, m.x.data_ptr(), m._fwd[0].data_ptr(), m._fwd[1].data_ptr(), lam.data_ptr(),
            torch.empty(s.n_atoms, dtype=torch.float64, device="cuda:0").data_ptr(),
            torch.empty((prm.n_types, 5), dtype=torch.float64, device="cuda:0").data_ptr(),
            None, m._stream()), "isspa_diff_vjp_after_forces")
    m.close()

"""The force-matching workload: frames of the 1,100-atom system (50 chains,
nMC = 20) from its own MD at 298 K, as windows of one differentiable model
with fixed common-random-number MC points (isspa_amd.forcematch).

* 8 frames: every window's forces (<= 1e-5 relative RMS) and its share of
  dL/dq (<= 1e-4) against the FP64 oracle run on that frame alone with its
  own points (system id = window), and dL/dtheta against the sum of the
  oracle's per-frame gradients (<= 1e-4 of each parameter's largest);
* 256 frames (281,600 atoms, 5.6 M points, the full workload): the per-type
  gradient equals the per-type sum of the per-atom contributions the adjoint
  returns, two runs are bit-identical, and 8 windows spread over the batch
  match the oracle as above (forces, dL/dq, and their per-window share of
  dL/dtheta taken from the per-atom contributions).
"""
import numpy as np
import pytest
import torch

import pyoracle as po
from isspa_amd.system import max_abs_rel_rms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def frames():
    from isspa_amd import forcematch, synth
    base = synth.workload(1)
    print("sampling 256 frames of the 1,100-atom system", flush=True)
    return base, forcematch.sample_frames(base, 256, stride=20)


def _model(base, fr):
    import dataclasses
    from isspa_amd import forcematch
    from isspa_amd.diff import DiffForces, DiffParams
    prm = DiffParams.from_system(base)
    sysb = forcematch.frame_batch(base, fr)
    prm_b = dataclasses.replace(prm, q=np.tile(prm.q, len(fr)))
    return prm, sysb, DiffForces(sysb, prm_b, step=0)


def _frame_system(base, x):
    import dataclasses
    c = dataclasses.replace(base)
    c.xyzq = base.xyzq.copy()
    c.xyzq[:, :3] = x
    return c


def _scale_ok(got, ref, tol):
    scale = np.maximum(np.abs(ref).max(axis=0, keepdims=True), 1e-300)
    return float((np.abs(got - ref) / scale).max()) <= tol


def _check_window(base, prm, x, w, f_w, gq_w, lam_w):
    c = _frame_system(base, x)
    pts = po.mc_points(c, step=0, sys_id=w)
    ref = po.diff_forward(c, prm, pts)
    ef = max_abs_rel_rms(f_w, ref["total"])
    rq, rt = po.diff_vjp(c, prm, pts, lam_w)
    eq = max_abs_rel_rms(gq_w, rq)
    print(f"  window {w}: forces {ef:.2e}, dL/dq {eq:.2e}", flush=True)
    assert ef <= 1e-5 and eq <= 1e-4, w
    return rt


def test_forcematch_8_frames_against_oracle(frames):
    base, fr = frames
    fr8 = fr[::32]
    prm, sysb, m = _model(base, fr8)
    print(f"case 8 frames x {base.n_atoms} atoms", flush=True)
    n = base.n_atoms
    q = torch.as_tensor(np.tile(prm.q, 8), device="cuda:0")
    th = torch.as_tensor(prm.theta, device="cuda:0")
    f = m.forces(q, th).cpu().numpy()
    lam = np.random.default_rng(4).normal(size=(8 * n, 3))
    gq, gt = m.vjp(q, th, torch.as_tensor(lam, device="cuda:0"))
    gq, gt = gq.cpu().numpy(), gt.cpu().numpy()
    rt_sum = np.zeros_like(gt)
    for w in range(8):
        sl = slice(w * n, (w + 1) * n)
        rt_sum += _check_window(base, prm, fr8[w], w, f[sl], gq[sl], lam[sl])
    assert _scale_ok(gt, rt_sum, 1e-4)
    m.close()


def test_forcematch_256_frames_properties(frames):
    base, fr = frames
    prm, sysb, m = _model(base, fr)
    W, n = len(fr), base.n_atoms
    print(f"case {W} frames x {n} atoms ({sysb.n_atoms} atoms)", flush=True)
    q = torch.as_tensor(np.tile(prm.q, W), device="cuda:0")
    th = torch.as_tensor(prm.theta, device="cuda:0")
    lam = torch.as_tensor(np.random.default_rng(6).normal(size=(W * n, 3)), device="cuda:0")
    runs = []
    for _ in range(2):
        f = m.forces(q, th)
        gq, gt, ga = m.vjp(q, th, lam, per_atom=True, after_forces=True)
        runs.append([t.cpu().numpy() for t in (f, gq, gt, ga)])
    for a, b in zip(*runs):
        assert np.array_equal(a, b), "two runs differ"
    f, gq, gt, ga = runs[0]
    assert np.all(np.isfinite(f)) and np.all(np.isfinite(gt))
    types = np.tile(base.ityp, W)
    per_type = np.zeros_like(gt)
    np.add.at(per_type, types, ga)
    assert _scale_ok(per_type, gt, 1e-12)
    lam = lam.cpu().numpy()
    for w in np.linspace(0, W - 1, 8).astype(int):
        sl = slice(w * n, (w + 1) * n)
        rt = _check_window(base, prm, fr[w], int(w), f[sl], gq[sl], lam[sl])
        share = np.zeros_like(gt)
        np.add.at(share, base.ityp, ga[sl])
        assert _scale_ok(share, rt, 1e-4), w
    m.close()

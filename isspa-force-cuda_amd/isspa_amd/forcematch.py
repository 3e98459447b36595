"""Force matching with the differentiable IS-SPA + nonbond model.

Workload: a batch of configurations of one topology (by default 256 frames of
the 1,100-atom, 50-chain system, sampled from its own MD at 298 K), each frame
a window of one DiffForces model with fixed common-random-number MC points.
Target forces come from the model at the true parameters.  The fit starts
from perturbed parameters (sigma +-5 %, eps +-10 %, charges +-0.05 e,
g0 and alpha +-10 %) and runs Adam on

    L = mean_i |F_i(theta, q) - F_i^target|^2 / mean_i |F_i^target|^2

with gradients from the hand-written adjoint (isspa_forces).  Charges are
per topology atom, shared by every frame.

    python -m isspa_amd.forcematch [--frames 256] [--steps 200] [--lr 0.01] [--lr-q 0.1]
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import sys
import time
from typing import List

import numpy as np

from .system import System

AMBER_CHARGE = 18.2223  # e -> AMBER internal charge units


def sample_frames(system: System, n_frames: int, stride: int = 20, warmup: int = 200,
                  device: str = "cuda:0") -> List[np.ndarray]:
    """Positions (float32[N, 3]) every `stride` MD steps after `warmup`
    steps of the plain engine (bonded + IS-SPA + nonbond + Andersen)."""
    import torch

    from .engine import GpuSystem
    g = GpuSystem(system, device)
    g.set_step(0)
    g.capture(1)
    g.replay(warmup)
    frames = []
    for _ in range(n_frames):
        g.replay(stride)
        torch.cuda.synchronize()
        frames.append(g.x[:, :3].cpu().numpy().copy())
    g.close()
    return frames


def frame_batch(system: System, frames: List[np.ndarray]) -> System:
    """One batched System whose windows are the frames (same topology)."""
    from .batch import merge_windows
    ws = []
    for x in frames:
        s = dataclasses.replace(system)
        s.xyzq = system.xyzq.copy()
        s.xyzq[:, :3] = x
        ws.append(s)
    return merge_windows(ws, name=f"{system.name}x{len(frames)}frames")


def perturb(theta: np.ndarray, q: np.ndarray, rng, sigma=0.05, eps=0.10, dq_e=0.05, g0a=0.10):
    th = theta.copy()
    T = th.shape[0]
    th[:, 0] *= 1.0 + rng.uniform(-sigma, sigma, T)
    th[:, 1] *= 1.0 + rng.uniform(-eps, eps, T)
    th[:, 3] *= 1.0 + rng.uniform(-g0a, g0a, T)
    th[:, 4] *= 1.0 + rng.uniform(-g0a, g0a, T)
    # keep g0 > 1 (the parabola must reach g = 1)
    th[:, 3] = np.maximum(th[:, 3], 1.05)
    return th, q + rng.uniform(-dq_e, dq_e, q.shape) * AMBER_CHARGE


def fit(model, n_frames: int, theta0, q0, target, steps: int = 200, lr: float = 0.01,
        lr_q: float = None, cosine: bool = False, lbfgs: bool = False, log=None):
    """Adam in relative units for theta and in 0.05 e units for the charges
    (learning rate lr_q, default lr), optionally with a cosine decay to lr/100.
    Returns (theta, q, history)."""
    import torch

    from .diff import isspa_forces
    dev = target.device
    th_base = torch.as_tensor(theta0, dtype=torch.float64, device=dev)
    q_base = torch.as_tensor(q0, dtype=torch.float64, device=dev)
    rel = torch.zeros_like(th_base, requires_grad=True)
    dq = torch.zeros_like(q_base, requires_grad=True)
    qs = 0.05 * AMBER_CHARGE
    norm = float((target ** 2).mean())
    hist = []
    if lbfgs:
        # reference optimiser for identifiability: L-BFGS with a strong-Wolfe
        # line search, `steps` outer iterations
        opt = torch.optim.LBFGS([rel, dq], lr=1.0, max_iter=20, history_size=20,
                                line_search_fn="strong_wolfe")

        def closure():
            opt.zero_grad()
            f = isspa_forces(th_base * (1.0 + rel), (q_base + qs * dq).repeat(n_frames), model)
            L = ((f - target) ** 2).mean() / norm
            L.backward()
            hist.append(L.item())
            return L

        for it in range(steps):
            opt.step(closure)
            if log:
                log(f"lbfgs iteration {it:4d}  evaluations {len(hist)}  loss {hist[-1]:.4e}")
        with torch.no_grad():
            return (th_base * (1.0 + rel)).cpu().numpy(), (q_base + qs * dq).cpu().numpy(), hist
    opt = torch.optim.Adam([{"params": [rel], "lr": lr},
                            {"params": [dq], "lr": lr if lr_q is None else lr_q}])
    sched = None
    if cosine:
        sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=steps, eta_min=lr / 100)
    for it in range(steps):
        opt.zero_grad()
        th = th_base * (1.0 + rel)
        q = (q_base + qs * dq).repeat(n_frames)
        f = isspa_forces(th, q, model)
        L = ((f - target) ** 2).mean() / norm
        L.backward()
        opt.step()
        if sched is not None:
            sched.step()
        hist.append(L.item())
        if log and (it % 20 == 0 or it == steps - 1):
            log(f"step {it:4d}  loss {hist[-1]:.4e}")
    with torch.no_grad():
        return (th_base * (1.0 + rel)).cpu().numpy(), (q_base + qs * dq).cpu().numpy(), hist


def curvature(base, frames, theta, q, fd_rel_h: float = 1e-2, device: str = "cuda:0"):
    """Gauss-Newton curvature of the force-matching loss along the per-type
    parameters, in relative units (theta_t,k -> theta_t,k (1 + delta)).

    For each frame (one DiffForces model of that frame alone, its MC points
    drawn at system id 0, fixed through the analysis) every row of the Jacobian
    J = d F / d delta comes from the hand-written adjoint with a unit
    cotangent (3N adjoint calls reusing one forward).  The curvature of
    L = mean |F - F_t|^2 / mean |F_t|^2 at the optimum is
    H = 2 J^T J / (n mean |F_t|^2) over the n force components of all frames.
    Central differences of the forward (relative step fd_rel_h) check the
    columns of J.  Returns H, its eigen-decomposition and the check."""
    import dataclasses

    import torch

    from .diff import DiffForces, DiffParams
    prm = DiffParams.from_system(base)
    th0 = torch.as_tensor(theta, dtype=torch.float64, device=device)
    qq = torch.as_tensor(q, dtype=torch.float64, device=device)
    T, K = th0.shape
    n = 3 * base.n_atoms
    JtJ = torch.zeros((T * K, T * K), dtype=torch.float64, device=device)
    fsq, fd_err = 0.0, []
    for w, x in enumerate(frames):
        c = dataclasses.replace(base)
        c.xyzq = base.xyzq.copy()
        c.xyzq[:, :3] = x
        m = DiffForces(c, prm, device, step=0)
        fsq += float((m.forces(qq, th0) ** 2).sum())
        rows = torch.empty((n, T * K), dtype=torch.float64, device=device)
        lam = torch.zeros((base.n_atoms, 3), dtype=torch.float64, device=device)
        for k in range(n):
            lam.zero_()
            lam.view(-1)[k] = 1.0
            _, gt = m.vjp(qq, th0, lam, after_forces=True)
            rows[k] = (gt * th0).reshape(-1)
        JtJ += rows.T @ rows
        if w == 0:
            # central differences of the forward against the adjoint's columns
            for a in range(T * K):
                d = torch.zeros_like(th0)
                d.view(-1)[a] = fd_rel_h * th0.view(-1)[a]
                col = ((m.forces(qq, th0 + d) - m.forces(qq, th0 - d)) / (2 * fd_rel_h)).reshape(-1)
                ref = rows[:, a]
                fd_err.append(float((col - ref).norm() / max(float(ref.norm()), 1e-300)))
        m.close()
    ncomp = n * len(frames)
    norm = fsq / ncomp
    H = 2.0 * JtJ / (ncomp * norm)
    w_, V = torch.linalg.eigh(H)
    return {"H": H.cpu().numpy(), "eigenvalues": w_.cpu().numpy(),
            "eigenvectors": V.cpu().numpy(), "fd_column_rel": fd_err, "force_norm": norm,
            "components": ncomp}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=256)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--lr", type=float, default=0.01)
    ap.add_argument("--lr-q", type=float, default=0.1)
    ap.add_argument("--cosine", action="store_true")
    ap.add_argument("--lbfgs", action="store_true", help="L-BFGS instead of Adam (diagnostic)")
    ap.add_argument("--stride", type=int, default=20)
    ap.add_argument("--seed", type=int, default=2025)
    ap.add_argument("--curvature", action="store_true",
                    help="Gauss-Newton curvature along the per-type parameters instead of a fit")
    args = ap.parse_args(argv)
    import torch

    from . import synth
    from .diff import DiffForces, DiffParams

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    base = synth.workload(1)
    prm = DiffParams.from_system(base)
    t0 = time.perf_counter()
    frames = sample_frames(base, args.frames, stride=args.stride)
    sysb = frame_batch(base, frames)
    log(f"{args.frames} frames of {base.n_atoms} atoms sampled in {time.perf_counter() - t0:.1f}s")
    prm_b = dataclasses.replace(prm, q=np.tile(prm.q, args.frames))
    model = DiffForces(sysb, prm_b, step=0)
    dev = model.dev
    th_true = prm.theta
    q_true = prm.q
    if args.curvature:
        model.close()
        c = curvature(base, frames, th_true, q_true)
        w, V = c["eigenvalues"], c["eigenvectors"]
        T = th_true.shape[0]
        names = [f"{n}[{t}]" for t in range(T) for n in ("sigma", "eps", "x0", "g0", "alpha")]
        diag = np.diag(c["H"])
        per_kind = {n: [float(diag[t * 5 + k]) for t in range(T)]
                    for k, n in enumerate(("sigma", "eps", "x0", "g0", "alpha"))}
        modes = []
        for i in np.argsort(w)[::-1]:
            v = V[:, i]
            top = np.argsort(-np.abs(v))[:3]
            modes.append({"eigenvalue": float(w[i]),
                          "largest": [(names[j], round(float(v[j]), 3)) for j in top]})
        print(json.dumps({"workload": f"{args.frames} frames x {base.n_atoms} atoms, nMC={base.n_mc}",
                          "curvature_diag": per_kind, "modes": modes,
                          "fd_column_rel": c["fd_column_rel"]}), flush=True)
        return
    target = model.forces(torch.as_tensor(np.tile(q_true, args.frames), device=dev),
                          torch.as_tensor(th_true, device=dev))
    rng = np.random.default_rng(args.seed)
    th0, q0 = perturb(th_true, q_true, rng)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    th, q, hist = fit(model, args.frames, th0, q0, target, steps=args.steps, lr=args.lr,
                      lr_q=args.lr_q, cosine=args.cosine, lbfgs=args.lbfgs, log=log)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    rel_err = lambda a, b: float(np.max(np.abs(a - b) / np.abs(b)))
    out = {
        "workload": f"{args.frames} frames x {base.n_atoms} atoms, nMC={base.n_mc}",
        "optimizer": "lbfgs" if args.lbfgs else "adam", "evaluations": len(hist),
        "steps": args.steps, "lr": args.lr, "lr_q": args.lr_q, "cosine": args.cosine,
        "loss_first": hist[0], "loss_last": hist[-1],
        "seconds_per_step": wall / args.steps,
        "seconds_per_evaluation": wall / len(hist),
        "theta_max_rel_err_start": {n: rel_err(th0[:, k], th_true[:, k])
                                    for k, n in enumerate(("sigma", "eps", "x0", "g0", "alpha"))},
        "theta_max_rel_err_end": {n: rel_err(th[:, k], th_true[:, k])
                                  for k, n in enumerate(("sigma", "eps", "x0", "g0", "alpha"))},
        "charge_max_abs_err_e_start": float(np.abs(q0 - q_true).max() / AMBER_CHARGE),
        "charge_max_abs_err_e_end": float(np.abs(q - q_true).max() / AMBER_CHARGE),
    }
    print(json.dumps(out), flush=True)
    model.close()


if __name__ == "__main__":
    main()

"""Differentiable IS-SPA + nonbond forces for force matching.

The differentiable model is the plain path with its per-type tables replaced
by the analytic functions the tables are sampled from (isspa_amd.tables):

    sigma_ts = (sigma + sigma_s)/2,  eps_ts = sqrt(eps eps_s)
    r_lo = max(x0 - sqrt(g0/alpha), core sigma_ts),  r_hi = x0 + sqrt((g0 - 1)/alpha)
    S(d) = smoothstep((d - r_lo)/delta + 1)     0 below r_lo - delta, 1 above r_lo
    g(d) = S(d) max(0, g0 - alpha (d - x0)^2)  for d < r_hi,  1 beyond
    E(d) = c exp(-(d - x0)^2/2) S(d)
    F(d) = 24 eps_ts/d (2 (sigma_ts/d)^12 - (sigma_ts/d)^6) S(d)
    nonbond A_ij = 4 eps_ij sigma_ij^12, B_ij = 4 eps_ij sigma_ij^6 (Lorentz-Berthelot)

The switch S replaces the tables' hard core floor so that g, E and F are
continuous in distance and parameters.  Trainable: per type
theta_t = (sigma, eps, x0, g0, alpha); per atom the charge q_i.  Positions,
MC points (common random numbers: they depend on positions and rmax only),
rmax, the E-field amplitudes c_t and the model constants are fixed.
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

from .system import System

THETA_NAMES = ("sigma", "eps", "x0", "g0", "alpha")


@dataclasses.dataclass
class DiffParams:
    sigma: np.ndarray   # [T]
    eps: np.ndarray     # [T]
    x0: np.ndarray      # [T]
    g0: np.ndarray      # [T]
    alpha: np.ndarray   # [T]
    camp: np.ndarray    # [T] E-field amplitude (fixed)
    q: np.ndarray       # [N] charges (AMBER units, as in positions' .w)
    sig_s: float = 3.15
    eps_s: float = 0.152
    core: float = 0.85
    delta: float = 0.25

    @property
    def n_types(self) -> int:
        return int(len(self.sigma))

    @property
    def theta(self) -> np.ndarray:
        """[T, 5] in THETA_NAMES order"""
        return np.stack([self.sigma, self.eps, self.x0, self.g0, self.alpha], axis=1)

    def with_theta(self, theta: np.ndarray, q: Optional[np.ndarray] = None) -> "DiffParams":
        th = np.asarray(theta, np.float64)
        return dataclasses.replace(self, sigma=th[:, 0].copy(), eps=th[:, 1].copy(),
                                   x0=th[:, 2].copy(), g0=th[:, 3].copy(), alpha=th[:, 4].copy(),
                                   q=self.q.copy() if q is None else np.asarray(q, np.float64))

    @staticmethod
    def from_system(system: System) -> "DiffParams":
        """The parameters a synthetic system's tables were generated from
        (one IS-SPA type per LJ type)."""
        par = system.type_params
        T = len(par)
        if not np.array_equal(system.ityp, system.isspa.isspa_type):
            raise ValueError("the differentiable model needs IS-SPA type == LJ type")
        get = lambda k: np.array([par[t][k] for t in range(T)], np.float64)
        return DiffParams(sigma=get("sigma"), eps=get("eps"), x0=get("x0"), g0=get("g0"),
                          alpha=get("alpha"), camp=get("c"),
                          q=system.xyzq[:, 3].astype(np.float64))

    def lj_tables(self):
        """(nbparm int[T*T], lj_a, lj_b float64[T(T+1)/2]) with Lorentz-Berthelot
        mixing, in the prmtop layout the plain nonbond model takes."""
        T = self.n_types
        ntri = T * (T + 1) // 2
        a = np.zeros(ntri)
        b = np.zeros(ntri)
        nbparm = np.zeros(T * T, np.int32)
        for i in range(T):
            for j in range(T):
                lo, hi = min(i, j), max(i, j)
                k = hi * (hi + 1) // 2 + lo
                nbparm[i * T + j] = k
                s = 0.5 * (self.sigma[i] + self.sigma[j])
                e = np.sqrt(self.eps[i] * self.eps[j])
                a[k] = 4 * e * s ** 12
                b[k] = 4 * e * s ** 6
        return nbparm, a, b


class DiffForces:
    """The differentiable model of one system, or of a batch of windows
    (a `System` from batch.merge_windows), resident on one GPU.

    forces(q, theta) and vjp(q, theta, lam) take and return torch FP64 device
    tensors: q [N], theta [T, 5], forces and lam [N, 3]."""

    def __init__(self, system: System, prm: DiffParams, device: str = "cuda:0", step: int = 0):
        import ctypes as C

        import torch

        from . import capi
        if not torch.cuda.is_available():
            raise RuntimeError("DiffForces needs a GPU")
        self.dev = torch.device(device)
        self.system = system
        self.n, self.T = system.n_atoms, prm.n_types
        c = capi.default_constants()
        d = capi.DiffDesc()
        d.n_atoms, d.n_types, d.n_mc, d.lbox = self.n, self.T, system.n_mc, system.lbox
        d.e0, d.p0, d.q0, d.o0, d.ep = c["e0"], c["p0"], c["q0"], c["o0"], c["ep"]
        d.sig_s, d.eps_s, d.core, d.delta = prm.sig_s, prm.eps_s, prm.core, prm.delta
        d.seed = int(system.seed) & (2 ** 64 - 1)
        d.step = int(step)
        win = None
        self._keep = []
        if getattr(system, "windows", None) is not None:
            arrs = (np.ascontiguousarray(system.windows, np.int32),
                    np.ascontiguousarray(system.window_sys, np.uint32),
                    np.ascontiguousarray(system.window_lbox, np.float32))
            self._keep.append(arrs)
            win = capi.Windows(len(arrs[0]), *(a.ctypes.data for a in arrs))
        it = np.ascontiguousarray(system.isspa.isspa_type, np.int32)
        rmax = np.ascontiguousarray(system.isspa.rmax, np.float32)
        camp = np.ascontiguousarray(prm.camp, np.float64)
        pre = np.ascontiguousarray(system.nexcl_prefix, np.int32)
        lst = np.ascontiguousarray(system.excl_list, np.int32)
        h = C.c_void_p()
        capi.check(capi.lib.isspa_diff_create(C.byref(d), C.byref(win) if win else None,
                                              it.ctypes.data, rmax.ctypes.data, camp.ctypes.data,
                                              pre.ctypes.data, lst.ctypes.data, len(lst),
                                              C.byref(h)), "isspa_diff_create")
        self.h = h
        self.x = torch.from_numpy(np.ascontiguousarray(system.xyzq, np.float32)).to(self.dev)
        self._capi = capi
        # the (q, theta) buffers of the last forces() call, kept alive so the
        # adjoint can reuse that forward (isspa_diff_vjp_after_forces), and a
        # token replaced by every forces() and plain vjp() call
        self._fwd = None
        self.state = object()

    def _stream(self):
        import torch
        return torch.cuda.current_stream(self.dev).cuda_stream

    def set_tiling(self, splits: int, chunks: int):
        self._capi.check(self._capi.lib.isspa_diff_set_tiling(self.h, splits, chunks))

    def forces(self, q, theta):
        import torch
        q = q.detach().to(self.dev, torch.float64).contiguous()
        theta = theta.detach().to(self.dev, torch.float64).contiguous()
        f = torch.empty((self.n, 3), dtype=torch.float64, device=self.dev)
        self._capi.check(self._capi.lib.isspa_diff_forces(
            self.h, self.x.data_ptr(), q.data_ptr(), theta.data_ptr(), f.data_ptr(),
            self._stream()), "isspa_diff_forces")
        self._fwd = (q, theta)
        self.state = object()
        return f

    def vjp(self, q, theta, lam, per_atom: bool = False, after_forces: bool = False):
        """dL/dq and dL/dtheta for lam = dL/dF.  With after_forces the last
        forces() call's forward is reused (q and theta are then those of that
        call; the C ABI refuses if its state is gone)."""
        import torch
        if after_forces:
            if self._fwd is None:
                raise RuntimeError("vjp(after_forces=True) needs a preceding forces() call")
            q, theta = self._fwd
            fn, name = self._capi.lib.isspa_diff_vjp_after_forces, "isspa_diff_vjp_after_forces"
        else:
            q = q.detach().to(self.dev, torch.float64).contiguous()
            theta = theta.detach().to(self.dev, torch.float64).contiguous()
            fn, name = self._capi.lib.isspa_diff_vjp, "isspa_diff_vjp"
            self._fwd = None
            self.state = object()
        lam = lam.detach().to(self.dev, torch.float64).contiguous()
        gq = torch.empty(self.n, dtype=torch.float64, device=self.dev)
        gt = torch.empty((self.T, 5), dtype=torch.float64, device=self.dev)
        ga = torch.empty((self.n, 5), dtype=torch.float64, device=self.dev) if per_atom else None
        self._capi.check(fn(self.h, self.x.data_ptr(), q.data_ptr(), theta.data_ptr(),
                            lam.data_ptr(), gq.data_ptr(), gt.data_ptr(),
                            ga.data_ptr() if per_atom else None, self._stream()), name)
        return (gq, gt, ga) if per_atom else (gq, gt)

    def close(self):
        if getattr(self, "h", None) is not None:
            self._capi.lib.isspa_diff_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _autograd_function():
    import torch

    class IsspaForces(torch.autograd.Function):
        """F(theta, q) of a DiffForces model; backward is the hand-written
        adjoint (isspa_diff_vjp), not autograd through the kernels.  When no
        other forces() or vjp() call ran on the model since this forward, the
        backward reuses the forward's points and records
        (isspa_diff_vjp_after_forces) instead of recomputing them."""

        @staticmethod
        def forward(ctx, theta, q, model):
            ctx.model = model
            ctx.save_for_backward(theta, q)
            f = model.forces(q, theta)
            ctx.state = model.state
            return f

        @staticmethod
        def backward(ctx, gf):
            theta, q = ctx.saved_tensors  # raises if modified in place since forward
            reuse = ctx.model.state is ctx.state
            gq, gt = ctx.model.vjp(q, theta, gf, after_forces=reuse)
            return gt.to(theta.dtype), gq.to(q.dtype), None

    return IsspaForces


_FN = None


def isspa_forces(theta, q, model: DiffForces):
    """Differentiable IS-SPA + nonbond forces [N, 3] (FP64) for per-type
    parameters theta [T, 5] and per-atom charges q [N] (torch tensors)."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(theta, q, model)

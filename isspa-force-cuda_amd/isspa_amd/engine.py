"""Device front end over the C ABI: uploads a `System` and runs the HIP kernels.

PyTorch is used only for device memory and streams; every force, point and
integration step is computed by libisspa_amd.so.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np
import torch

from . import capi
from .capi import check, ptr
from .system import System


def _np(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class GpuSystem:
    """One system resident on one GPU, with the models of each force term."""

    def __init__(self, system: System, device: str = "cuda:0", *, isspa: bool = True,
                 nonbond: bool = True, bonded: bool = True, us: bool = True,
                 system_id: int = 0):
        if not torch.cuda.is_available():
            raise RuntimeError("GpuSystem needs a GPU (torch.cuda.is_available() is False)")
        self.system = system
        self.dev = torch.device(device)
        torch.cuda.set_device(self.dev)
        self.system_id = int(system_id)
        self.windows = getattr(system, "windows", None)
        self._win = None
        if self.windows is not None:
            # isspa_windows points into these arrays; keep them alive with the object
            self._win_arrays = (_np(system.windows, np.int32), _np(system.window_sys, np.uint32),
                                _np(system.window_lbox, np.float32))
            self._win = capi.Windows(len(self.windows), *(ptr(a) for a in self._win_arrays))
        n = system.n_atoms
        self.x = torch.from_numpy(_np(system.xyzq, np.float32)).to(self.dev)
        self.v = torch.from_numpy(_np(system.vel, np.float32)).to(self.dev)
        self.f = torch.zeros((n, 4), dtype=torch.float32, device=self.dev)
        self._keep = []
        self.isspa = self._make_isspa() if isspa else None
        self.nonbond = self._make_nonbond() if nonbond else None
        self.bonded = self._make_bonded() if bonded and self._has_bonded() else None
        self.restraints = self._restraints()
        self.us = self._make_us() if us and self.restraints else None
        self.engine = None

    # ---- construction -------------------------------------------------------
    def stream(self) -> int:
        return torch.cuda.current_stream(self.dev).cuda_stream

    def _make_isspa(self):
        s = self.system
        it = s.isspa
        c = capi.default_constants()
        d = capi.ModelDesc()
        d.n_atoms, d.n_types, d.n_mc = s.n_atoms, it.n_types, s.n_mc
        d.n_g, d.n_e, d.n_f = it.g.shape[1], it.e.shape[1], it.f.shape[1]
        d.g_r0, d.g_dr = it.g_grid
        d.e_r0, d.e_dr = it.e_grid
        d.f_r0, d.f_dr = it.f_grid
        d.lbox = s.lbox
        d.e0, d.p0, d.q0, d.o0, d.ep = c["e0"], c["p0"], c["q0"], c["o0"], c["ep"]
        d.seed = int(s.seed) & (2 ** 64 - 1)
        d.system_id = self.system_id
        h = C.c_void_p()
        arrs = [_np(it.isspa_type, np.int32), _np(it.rmax, np.float32), _np(it.g, np.float32),
                _np(it.e, np.float32), _np(it.f, np.float32)]
        if self._win is None:
            check(capi.lib.isspa_model_create(C.byref(d), *(ptr(a) for a in arrs), C.byref(h)),
                  "isspa_model_create")
        else:
            check(capi.lib.isspa_model_create_batch(C.byref(d), C.byref(self._win),
                                                    *(ptr(a) for a in arrs), C.byref(h)),
                  "isspa_model_create_batch")
        self.desc = d
        return h

    def _make_nonbond(self):
        s = self.system
        c = capi.default_constants()
        h = C.c_void_p()
        args = [_np(s.ityp, np.int32), _np(s.nbparm, np.int32), _np(s.lj_a, np.float32),
                _np(s.lj_b, np.float32), _np(s.nexcl_prefix, np.int32), _np(s.excl_list, np.int32),
                _np(s.isspa.isspa_type, np.int32), _np(s.isspa.rmax, np.float32)]
        common = (s.n_atoms, s.n_lj_types, ptr(args[0]), ptr(args[1]), ptr(args[2]), ptr(args[3]),
                  ptr(args[4]), ptr(args[5]), int(len(s.excl_list)), ptr(args[6]),
                  s.isspa.n_types, ptr(args[7]), c["ep"], s.lbox)
        if self._win is None:
            check(capi.lib.nonbond_model_create(*common, C.byref(h)), "nonbond_model_create")
        else:
            check(capi.lib.nonbond_model_create_batch(*common, C.byref(self._win), C.byref(h)),
                  "nonbond_model_create_batch")
        return h

    def _has_bonded(self):
        s = self.system
        return len(s.bond_atoms) + len(s.angle_atoms) + len(s.dih_atoms) > 0

    def _make_bonded(self):
        s = self.system
        h = C.c_void_p()
        A = [_np(s.bond_atoms, np.int32), _np(s.bond_params, np.float32),
             _np(s.angle_atoms, np.int32), _np(s.angle_params, np.float32),
             _np(s.dih_atoms, np.int32), _np(s.dih_types, np.int32),
             _np(s.dih_params, np.float32), _np(s.sc14, np.float32), _np(s.ityp, np.int32),
             _np(s.nbparm, np.int32), _np(s.lj_a, np.float32), _np(s.lj_b, np.float32)]
        check(capi.lib.bonded_model_create(
            s.n_atoms, len(A[0]), ptr(A[0]), len(A[1]), ptr(A[1]), len(A[2]), ptr(A[2]),
            len(A[3]), ptr(A[3]), len(A[4]), ptr(A[4]), ptr(A[5]), len(A[6]), ptr(A[6]),
            ptr(A[7]), s.n_lj_types, ptr(A[8]), ptr(A[9]), ptr(A[10]), ptr(A[11]), s.lbox,
            C.byref(h)), "bonded_model_create")
        return h

    def _restraints(self):
        """[(restraint, box)] in global atom indices: the system's one
        restraint, or one per window of a batch (windows without one skipped)."""
        s = self.system
        if self.windows is None:
            return [(s.us, s.lbox)] if s.us is not None else []
        return [(r, float(L)) for r, L in zip(s.restraints, s.window_lbox) if r is not None]

    def _make_us(self):
        s = self.system
        h = C.c_void_p()
        if self.windows is None:
            us = s.us
            al = _np(us.atom_list, np.int32)
            ms = _np(us.mass, np.float32)
            kumb = us.kumb
            check(capi.lib.us_model_create(s.n_atoms, len(al), ptr(al), ptr(ms), ptr(kumb), us.x0,
                                           s.lbox, C.byref(h)), "us_model_create")
            return h
        rs = self.restraints
        boff = _np(np.concatenate([[0], np.cumsum([len(r.atom_list) for r, _ in rs])]), np.int32)
        al = _np(np.concatenate([r.atom_list for r, _ in rs]), np.int32)
        ms = _np(np.concatenate([r.mass for r, _ in rs]), np.float32)
        k = _np([r.k for r, _ in rs], np.float32)
        x0 = _np([r.x0 for r, _ in rs], np.float32)
        L = _np([b for _, b in rs], np.float32)
        check(capi.lib.us_model_create_batch(s.n_atoms, len(rs), ptr(boff), ptr(al), ptr(ms),
                                             ptr(k), ptr(x0), ptr(L), C.byref(h)),
              "us_model_create_batch")
        return h

    # ---- single terms (reference wrapper semantics: forces ADD into f) -----
    def zero_forces(self):
        check(capi.lib.isspa_zero(ptr(self.f), self.f.numel() * 4, self.stream()))

    def isspa_force(self, step: int = 0, f: Optional[torch.Tensor] = None) -> torch.Tensor:
        f = self.f if f is None else f
        check(capi.lib.isspa_model_force(self.isspa, ptr(self.x), ptr(f), step, None,
                                         self.stream()), "isspa_model_force")
        return f

    def isspa_force_keep(self, f: torch.Tensor, f_isspa: torch.Tensor, step: int = 0):
        """Adds the IS-SPA force into f and stores it alone in f_isspa, in one
        reduction (the reference wrapper's isspaf_d argument)."""
        check(capi.lib.isspa_model_force_keep(self.isspa, ptr(self.x), ptr(f), ptr(f_isspa), step,
                                              None, self.stream()), "isspa_model_force_keep")
        return f, f_isspa

    def isspa_components(self, step: int = 0):
        n = self.system.n_atoms
        out = [torch.zeros((n, 3), dtype=torch.float64, device=self.dev) for _ in range(3)]
        check(capi.lib.isspa_model_force_components(self.isspa, ptr(self.x), ptr(out[0]),
                                                    ptr(out[1]), ptr(out[2]), step, self.stream()),
              "isspa_model_force_components")
        return dict(lj=out[0], pol=out[1], cdd=out[2])

    def field(self):
        P = self.system.n_points
        o = dict(pts=torch.zeros((P, 4), dtype=torch.float32, device=self.dev),
                 g=torch.zeros(P, dtype=torch.float64, device=self.dev),
                 cnt=torch.zeros(P, dtype=torch.int32, device=self.dev),
                 E=torch.zeros((P, 3), dtype=torch.float64, device=self.dev),
                 E0=torch.zeros((P, 3), dtype=torch.float64, device=self.dev))
        check(capi.lib.isspa_model_get_field(self.isspa, ptr(o["pts"]), ptr(o["g"]), ptr(o["cnt"]),
                                             ptr(o["E"]), ptr(o["E0"]), self.stream()))
        return o

    def profile(self, reps: int = 5):
        """Mean ms per launch of {points, field, epilogue, force, reduce}."""
        ms = (C.c_float * 5)()
        check(capi.lib.isspa_model_profile(self.isspa, ptr(self.x), reps, ms, self.stream()),
              "isspa_model_profile")
        return dict(zip(("points", "field", "epilogue", "force", "reduce"), list(ms)))

    def set_tiling(self, field_splits: int, force_chunks: int):
        check(capi.lib.isspa_model_set_tiling(self.isspa, field_splits, force_chunks))

    def set_target_waves(self, waves: int):
        check(capi.lib.isspa_model_set_target_waves(self.isspa, int(waves)))

    def tiling(self):
        a, b = C.c_int(), C.c_int()
        check(capi.lib.isspa_model_get_tiling(self.isspa, C.byref(a), C.byref(b)))
        return a.value, b.value

    def nonbond_force(self, f: Optional[torch.Tensor] = None) -> torch.Tensor:
        f = self.f if f is None else f
        check(capi.lib.nonbond_force(self.nonbond, ptr(self.x), ptr(f), self.stream()),
              "nonbond_force")
        return f

    def nonbond_components(self):
        n = self.system.n_atoms
        out = [torch.zeros((n, 3), dtype=torch.float64, device=self.dev) for _ in range(3)]
        check(capi.lib.nonbond_force_components(self.nonbond, ptr(self.x), ptr(out[0]),
                                                ptr(out[1]), ptr(out[2]), self.stream()),
              "nonbond_force_components")
        return dict(lj=out[0], coul=out[1], pair=out[2])

    def bonded_force(self, terms: int = capi.TERM_ALL, f: Optional[torch.Tensor] = None):
        f = self.f if f is None else f
        e = torch.zeros(5, dtype=torch.float64, device=self.dev)
        check(capi.lib.bonded_force(self.bonded, terms, ptr(self.x), ptr(f), ptr(e),
                                    self.stream()), "bonded_force")
        return f, e

    def us_force(self, f: Optional[torch.Tensor] = None):
        f = self.f if f is None else f
        com = torch.zeros(6 * len(self.restraints), dtype=torch.float64, device=self.dev)
        check(capi.lib.us_force(self.us, ptr(self.x), ptr(f), ptr(com), self.stream()), "us_force")
        return f, com

    def leapfrog(self, step: int = 0, f: Optional[torch.Tensor] = None):
        """Standalone integrator call: one leapfrog_andersen per window (the
        whole system when not batched), each with its own box and system id."""
        s = self.system
        f = self.f if f is None else f
        if self.windows is None:
            parts = [(0, s.n_atoms, s.lbox, self.system_id)]
        else:
            a0 = np.concatenate([[0], np.cumsum(s.windows)[:-1]])
            parts = list(zip(a0.tolist(), s.windows.tolist(), s.window_lbox.tolist(),
                             s.window_sys.tolist()))
        for a0, n, L, sid in parts:
            off = 16 * int(a0)
            check(capi.lib.leapfrog_andersen(int(n), ptr(self.x) + off, ptr(self.v) + off,
                                             ptr(f) + off, s.kT, s.dt, s.pnu, float(L),
                                             int(s.seed) & (2 ** 64 - 1), int(sid), step, None,
                                             self.stream()), "leapfrog_andersen")

    # ---- whole step --------------------------------------------------------
    def make_engine(self):
        if self.engine is None:
            s = self.system
            h = C.c_void_p()
            check(capi.lib.md_engine_create(self.isspa, self.nonbond, self.bonded, self.us,
                                            s.n_atoms, ptr(self.x), ptr(self.v), ptr(self.f), s.kT,
                                            s.dt, s.pnu, s.lbox, int(s.seed) & (2 ** 64 - 1),
                                            self.system_id, C.byref(h)), "md_engine_create")
            self.engine = h
        return self.engine

    def set_step(self, step: int):
        check(capi.lib.md_engine_set_step(self.make_engine(), step, self.stream()))

    def step(self, n: int = 1):
        e = self.make_engine()
        for _ in range(n):
            check(capi.lib.md_engine_step(e, self.stream()), "md_engine_step")

    def forces(self):
        check(capi.lib.md_engine_forces(self.make_engine(), self.stream()), "md_engine_forces")
        return self.f

    def com(self):
        """restraint centres of mass of the last engine force evaluation"""
        out = torch.zeros(6 * max(1, len(self.restraints)), dtype=torch.float64, device=self.dev)
        check(capi.lib.md_engine_get_com(self.make_engine(), ptr(out), self.stream()))
        return out

    def _side(self):
        # graph capture needs a stream other than the null stream; default priority
        if getattr(self, "_side_stream", None) is None:
            self._side_stream = torch.cuda.Stream(self.dev)
        return self._side_stream

    def capture(self, steps_per_graph: int = 1):
        side = self._side()
        side.wait_stream(torch.cuda.current_stream(self.dev))
        check(capi.lib.md_engine_capture(self.make_engine(), steps_per_graph, side.cuda_stream),
              "md_engine_capture")
        torch.cuda.current_stream(self.dev).wait_stream(side)

    def replay(self, replays: int = 1):
        side = self._side()
        side.wait_stream(torch.cuda.current_stream(self.dev))
        check(capi.lib.md_engine_replay(self.make_engine(), replays, side.cuda_stream),
              "md_engine_replay")
        torch.cuda.current_stream(self.dev).wait_stream(side)

    def sync(self):
        torch.cuda.synchronize(self.dev)

    def close(self):
        if self.engine is not None:
            capi.lib.md_engine_destroy(self.engine)
            self.engine = None
        for h, fn in ((self.isspa, capi.lib.isspa_model_destroy),
                      (self.nonbond, capi.lib.nonbond_model_destroy),
                      (self.bonded, capi.lib.bonded_model_destroy),
                      (self.us, capi.lib.us_model_destroy)):
            if h is not None:
                fn(h)
        self.isspa = self.nonbond = self.bonded = self.us = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

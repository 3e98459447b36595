"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It loads oracle/liboracle.so (built by oracle/Makefile) and,
for the CPU baseline, oracle/_ref/libisspa_omp_ref.so (the reference's own
src/isspa_force.omp.c compiled where it lies).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")


class Consts(C.Structure):
    _fields_ = [("e0", C.c_float), ("p0", C.c_float), ("q0", C.c_float), ("o0", C.c_float),
                ("ep", C.c_float)]


class Tables(C.Structure):
    _fields_ = [("n_types", C.c_int), ("n_g", C.c_int), ("n_e", C.c_int), ("n_f", C.c_int),
                ("g_r0", C.c_float), ("g_dr", C.c_float), ("e_r0", C.c_float),
                ("e_dr", C.c_float), ("f_r0", C.c_float), ("f_dr", C.c_float),
                ("g", C.c_void_p), ("e", C.c_void_p), ("f", C.c_void_p),
                ("rmax", C.c_void_p), ("vtot", C.c_void_p)]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_default_consts.argtypes = [C.POINTER(Consts)]
        L.orc_mc_points.argtypes = [C.c_int, C.c_int, f32p, i32p, f32p, C.c_uint64, C.c_uint32,
                                    C.c_uint32, f32p]
        L.orc_isspa_field.argtypes = [C.c_int, C.c_int, f32p, i32p, C.POINTER(Tables), f32p,
                                      C.c_float, C.POINTER(Consts), f64p, i32p, f64p, f64p]
        L.orc_isspa_field_subset.argtypes = [C.c_int, C.c_long, C.c_int, i32p, f32p, i32p,
                                             C.POINTER(Tables), f32p, C.c_float,
                                             C.POINTER(Consts), f64p, i32p, f64p, f64p]
        L.orc_isspa_force.argtypes = [C.c_int, C.c_int, f32p, i32p, C.POINTER(Tables), f32p, f64p,
                                      i32p, f64p, f64p, C.c_float, C.POINTER(Consts), C.c_int,
                                      C.c_int, f64p, f64p, f64p]
        L.orc_langevin_moments.argtypes = [C.c_double, f64p]
        L.orc_nonbond.argtypes = [C.c_int, f64p, C.c_int, i32p, i32p, f64p, f64p, i32p, i32p, i32p,
                                  f32p, C.c_float, C.POINTER(Consts), f64p, f64p, f64p, f64p]
        L.orc_bonds.argtypes = [C.c_int, i32p, f32p, f64p, C.c_float, f64p, f64p]
        L.orc_angles.argtypes = [C.c_int, i32p, f32p, f64p, C.c_float, f64p, f64p]
        L.orc_dihedrals.argtypes = [C.c_int, i32p, i32p, f32p, f32p, f64p, C.c_int, i32p, i32p,
                                    f64p, f64p, C.c_float, f64p, f64p]
        L.orc_leapfrog.argtypes = [C.c_int, f32p, f32p, f64p, C.c_float, C.c_float, C.c_float,
                                   C.c_float, C.c_uint64, C.c_uint32, C.c_uint32, f64p, f64p, i32p]
        L.orc_us_force.argtypes = [C.c_int, i32p, f32p, f32p, C.c_float, f64p, C.c_float, f64p,
                                   f64p, f64p]
        L.orc_cpu_ispa_force.argtypes = [f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p,
                                         i32p, C.c_int, C.c_int, C.c_int]
        L.orc_srand.argtypes = [C.c_uint]
        _lib = L
    return _lib


def ref_lib():
    """oracle/_ref/libisspa_omp_ref.so, or None if it was never built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libisspa_omp_ref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.cpu_ispa_force.argtypes = [f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p, i32p,
                                     C.c_int, C.c_int, C.c_int]
        _ref = R
    return _ref


def consts() -> Consts:
    c = Consts()
    lib().orc_default_consts(C.byref(c))
    return c


def philox(ctr, key) -> np.ndarray:
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def langevin(x: float) -> np.ndarray:
    m = np.zeros(3)
    lib().orc_langevin_moments(float(x), m)
    return m


class _TabHold:
    """Keeps the numpy arrays behind a Tables struct alive."""

    def __init__(self, system):
        it = system.isspa
        self.g = np.ascontiguousarray(it.g, np.float32)
        self.e = np.ascontiguousarray(it.e, np.float32)
        self.f = np.ascontiguousarray(it.f, np.float32)
        self.rmax = np.ascontiguousarray(it.rmax, np.float32)
        self.vtot = it.vtot(system.n_mc)
        t = Tables()
        t.n_types = it.n_types
        t.n_g, t.n_e, t.n_f = self.g.shape[1], self.e.shape[1], self.f.shape[1]
        t.g_r0, t.g_dr = it.g_grid
        t.e_r0, t.e_dr = it.e_grid
        t.f_r0, t.f_dr = it.f_grid
        t.g = self.g.ctypes.data
        t.e = self.e.ctypes.data
        t.f = self.f.ctypes.data
        t.rmax = self.rmax.ctypes.data
        t.vtot = self.vtot.ctypes.data
        self.t = t


def mc_points(system, xyzq: Optional[np.ndarray] = None, step: int = 0, sys_id: int = 0,
              seed: Optional[int] = None) -> np.ndarray:
    xyzq = np.ascontiguousarray(system.xyzq if xyzq is None else xyzq, np.float32)
    P = system.n_points
    pts = np.zeros((P, 4), np.float32)
    lib().orc_mc_points(system.n_atoms, system.n_mc, xyzq, system.isspa.isspa_type,
                        np.ascontiguousarray(system.isspa.rmax, np.float32),
                        int(system.seed if seed is None else seed) & (2**64 - 1), step, sys_id, pts)
    return pts


def isspa_field(system, pts: np.ndarray, xyzq: Optional[np.ndarray] = None):
    xyzq = np.ascontiguousarray(system.xyzq if xyzq is None else xyzq, np.float32)
    P = system.n_points
    th = _TabHold(system)
    g = np.zeros(P)
    cnt = np.zeros(P, np.int32)
    E = np.zeros((P, 3))
    E0 = np.zeros((P, 3))
    c = consts()
    lib().orc_isspa_field(system.n_atoms, system.n_mc, xyzq, system.isspa.isspa_type, C.byref(th.t),
                          np.ascontiguousarray(pts, np.float32), system.lbox, C.byref(c), g, cnt, E, E0)
    return dict(g=g, cnt=cnt, E=E, E0=E0)


def isspa_field_subset(system, pts, owner, xyzq=None):
    """Field of the points `pts` (float32[S,4]) owned by atoms `owner`."""
    xyzq = np.ascontiguousarray(system.xyzq if xyzq is None else xyzq, np.float32)
    S = len(owner)
    th = _TabHold(system)
    g = np.zeros(S)
    cnt = np.zeros(S, np.int32)
    E = np.zeros((S, 3))
    E0 = np.zeros((S, 3))
    c = consts()
    lib().orc_isspa_field_subset(system.n_atoms, S, system.n_mc,
                                 np.ascontiguousarray(owner, np.int32), xyzq,
                                 system.isspa.isspa_type, C.byref(th.t),
                                 np.ascontiguousarray(pts, np.float32), system.lbox, C.byref(c),
                                 g, cnt, E, E0)
    return dict(g=g, cnt=cnt, E=E, E0=E0)


def isspa_force(system, pts, field, xyzq=None, atoms=None):
    """Per-component IS-SPA forces (lj, pol, cdd), each float64[N,3].  `atoms`
    = (i0, i1) restricts the receivers (others stay zero)."""
    xyzq = np.ascontiguousarray(system.xyzq if xyzq is None else xyzq, np.float32)
    n = system.n_atoms
    th = _TabHold(system)
    out = {k: np.zeros((n, 3)) for k in ("lj", "pol", "cdd")}
    i0, i1 = atoms if atoms is not None else (0, n)
    c = consts()
    lib().orc_isspa_force(n, system.n_mc, xyzq, system.isspa.isspa_type, C.byref(th.t),
                          np.ascontiguousarray(pts, np.float32),
                          np.ascontiguousarray(field["g"], np.float64),
                          np.ascontiguousarray(field["cnt"], np.int32),
                          np.ascontiguousarray(field["E"], np.float64),
                          np.ascontiguousarray(field["E0"], np.float64), system.lbox, C.byref(c),
                          int(i0), int(i1), out["lj"], out["pol"], out["cdd"])
    return out


def isspa_total(system, step=0, sys_id=0, xyzq=None, seed=None):
    pts = mc_points(system, xyzq, step, sys_id, seed)
    fld = isspa_field(system, pts, xyzq)
    comp = isspa_force(system, pts, fld, xyzq)
    comp["total"] = comp["lj"] + comp["pol"] + comp["cdd"]
    comp["pts"] = pts
    comp["field"] = fld
    return comp


def _pos64(system, xyzq):
    """positions as FP64 [N,4]; FP32 inputs widen exactly."""
    return np.ascontiguousarray(system.xyzq if xyzq is None else xyzq, np.float64)


def nonbond(system, xyzq=None):
    xyzq = _pos64(system, xyzq)
    n = system.n_atoms
    out = {k: np.zeros((n, 3)) for k in ("lj", "coul", "pair")}
    en = np.zeros(2)
    c = consts()
    lib().orc_nonbond(n, xyzq, system.n_lj_types, system.ityp, system.nbparm,
                      np.ascontiguousarray(system.lj_a, np.float64),
                      np.ascontiguousarray(system.lj_b, np.float64), system.nexcl_prefix, system.excl_list, system.isspa.isspa_type,
                      np.ascontiguousarray(system.isspa.rmax, np.float32), system.lbox, C.byref(c),
                      out["lj"], out["coul"], out["pair"], en)
    out["total"] = out["lj"] + out["coul"] + out["pair"]
    out["e_lj"], out["e_coul"] = float(en[0]), float(en[1])
    return out


def bonded(system, xyzq=None):
    xyzq = _pos64(system, xyzq)
    n = system.n_atoms
    fb, fa, fd = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    eb, ea, ed = np.zeros(1), np.zeros(1), np.zeros(3)
    L = lib()
    L.orc_bonds(len(system.bond_atoms), np.ascontiguousarray(system.bond_atoms), system.bond_params,
                xyzq, system.lbox, fb, eb)
    L.orc_angles(len(system.angle_atoms), np.ascontiguousarray(system.angle_atoms),
                 system.angle_params, xyzq, system.lbox, fa, ea)
    L.orc_dihedrals(len(system.dih_atoms), np.ascontiguousarray(system.dih_atoms), system.dih_types,
                    np.ascontiguousarray(system.dih_params), np.ascontiguousarray(system.sc14), xyzq,
                    system.n_lj_types, system.ityp, system.nbparm,
                    np.ascontiguousarray(system.lj_a, np.float64),
                    np.ascontiguousarray(system.lj_b, np.float64), system.lbox, fd, ed)
    return dict(bond=fb, angle=fa, dih=fd, e_bond=float(eb[0]), e_angle=float(ea[0]),
                e_dih=float(ed[0]), e_14coul=float(ed[1]), e_14lj=float(ed[2]),
                total=fb + fa + fd)


def us_force(system, xyzq=None):
    xyzq = _pos64(system, xyzq)
    us = system.us
    f = np.zeros((system.n_atoms, 3))
    com = np.zeros(6)
    e = np.zeros(1)
    lib().orc_us_force(len(us.atom_list), np.ascontiguousarray(us.atom_list, np.int32),
                       np.ascontiguousarray(us.mass, np.float32), us.kumb, us.x0, xyzq, system.lbox,
                       f, com, e)
    return dict(f=f, com=com.reshape(2, 3), energy=float(e[0]))


def leapfrog(system, force, pos=None, vel=None, step=0, sys_id=0):
    pos = np.ascontiguousarray(system.xyzq if pos is None else pos, np.float32)
    vel = np.ascontiguousarray(system.vel if vel is None else vel, np.float32)
    n = system.n_atoms
    po, vo = np.zeros((n, 3)), np.zeros((n, 3))
    hit = np.zeros(n, np.int32)
    lib().orc_leapfrog(n, pos, vel, np.ascontiguousarray(force, np.float64), system.kT, system.dt,
                       system.pnu, system.lbox, int(system.seed) & (2**64 - 1), step, sys_id, po,
                       vo, hit)
    return po, vo, hit


def cpu_ispa_force(fn, xyz, w, x0, g0, gr2, alpha, lj_A, lj_B, ityp, n_mc, n_threads, seed=None):
    """Call a cpu_ispa_force-compatible function (the reference build or the
    oracle restatement).  Returns the accumulated float32[N*3] forces."""
    n = len(ityp)
    f = np.zeros(3 * n, np.float32)
    if seed is not None:
        lib().orc_srand(seed)  # the same glibc rand() state both libraries draw from
    fn(np.ascontiguousarray(xyz, np.float32), f, *(np.ascontiguousarray(a, np.float32) for a in
       (w, x0, g0, gr2, alpha, lj_A, lj_B)), np.ascontiguousarray(ityp, np.int32), n, n_mc,
       n_threads)
    return f


# ---- differentiable model (oracle_diff.c) --------------------------------
class DParams(C.Structure):
    _fields_ = [("n_types", C.c_int), ("sigma", C.c_void_p), ("eps", C.c_void_p),
                ("x0", C.c_void_p), ("g0", C.c_void_p), ("alpha", C.c_void_p),
                ("camp", C.c_void_p), ("rmax", C.c_void_p), ("vtot", C.c_void_p),
                ("sig_s", C.c_double), ("eps_s", C.c_double), ("core", C.c_double),
                ("delta", C.c_double), ("e0", C.c_double), ("p0", C.c_double),
                ("q0", C.c_double), ("o0", C.c_double), ("ep", C.c_double)]


class _DHold:
    """keeps the arrays behind a DParams alive"""

    def __init__(self, system, prm):
        c = consts()
        self.arrs = [np.ascontiguousarray(getattr(prm, k), np.float64)
                     for k in ("sigma", "eps", "x0", "g0", "alpha", "camp")]
        self.rmax = np.ascontiguousarray(system.isspa.rmax, np.float32)
        self.vtot = system.isspa.vtot(system.n_mc).astype(np.float64)
        p = DParams()
        p.n_types = len(self.arrs[0])
        p.sigma, p.eps, p.x0, p.g0, p.alpha, p.camp = (a.ctypes.data for a in self.arrs)
        p.rmax, p.vtot = self.rmax.ctypes.data, self.vtot.ctypes.data
        p.sig_s, p.eps_s, p.core, p.delta = prm.sig_s, prm.eps_s, prm.core, prm.delta
        p.e0, p.p0, p.q0, p.o0, p.ep = c.e0, c.p0, c.q0, c.o0, c.ep
        self.p = p


def _diff_sigs(L):
    if getattr(L, "_diff_ready", False):
        return
    L.orc_diff_type_fn.argtypes = [C.POINTER(DParams), C.c_int, C.c_double, f64p, f64p]
    L.orc_langevin_derivs.argtypes = [C.c_double, f64p, f64p]
    L.orc_diff_forward.argtypes = [C.c_int, C.c_int, f32p, f64p, i32p, f32p, C.c_float,
                                   C.POINTER(DParams), i32p, i32p, f64p, f64p, f64p]
    L.orc_diff_vjp.argtypes = [C.c_int, C.c_int, f32p, f64p, i32p, f32p, C.c_float,
                               C.POINTER(DParams), i32p, i32p, f64p, f64p, f64p]
    L._diff_ready = True


def diff_type_fn(system, prm, t: int, d: float):
    """(val[3] = g, E, F; dval[3, 5]) of the analytic model at distance d"""
    L = lib()
    _diff_sigs(L)
    h = _DHold(system, prm)
    v, dv = np.zeros(3), np.zeros(15)
    L.orc_diff_type_fn(C.byref(h.p), t, d, v, dv)
    return v, dv.reshape(3, 5)


def langevin_derivs(x: float):
    L = lib()
    _diff_sigs(L)
    m, dm = np.zeros(3), np.zeros(3)
    L.orc_langevin_derivs(x, m, dm)
    return m, dm


def _diff_inputs(system, pts, xyzq):
    xyzq = np.ascontiguousarray(system.xyzq if xyzq is None else xyzq, np.float32)
    return (xyzq, np.ascontiguousarray(system.isspa.isspa_type, np.int32),
            np.ascontiguousarray(pts, np.float32), np.ascontiguousarray(system.nexcl_prefix, np.int32),
            np.ascontiguousarray(system.excl_list, np.int32))


def diff_forward(system, prm, pts, xyzq=None):
    """{'isspa', 'nb', 'total': double[N,3], 'psi': double[P,8]} of the
    analytic model; charges are prm.q."""
    L = lib()
    _diff_sigs(L)
    h = _DHold(system, prm)
    x, it, p, pre, lst = _diff_inputs(system, pts, xyzq)
    n = system.n_atoms
    fi, fn = np.zeros((n, 3)), np.zeros((n, 3))
    psi = np.zeros((n * system.n_mc, 8))
    q = np.ascontiguousarray(prm.q, np.float64)
    L.orc_diff_forward(n, system.n_mc, x, q, it, p, system.lbox, C.byref(h.p), pre, lst, fi, fn,
                       psi)
    return {"isspa": fi, "nb": fn, "total": fi + fn, "psi": psi}


def diff_vjp(system, prm, pts, lam, xyzq=None):
    """(dL/dq double[N], dL/dtheta double[T,5]) for cotangent lam = dL/dF."""
    L = lib()
    _diff_sigs(L)
    h = _DHold(system, prm)
    x, it, p, pre, lst = _diff_inputs(system, pts, xyzq)
    n = system.n_atoms
    gq = np.zeros(n)
    gt = np.zeros((h.p.n_types, 5))
    q = np.ascontiguousarray(prm.q, np.float64)
    L.orc_diff_vjp(n, system.n_mc, x, q, it, p, system.lbox, C.byref(h.p), pre, lst,
                   np.ascontiguousarray(lam, np.float64), gq, gt)
    return gq, gt

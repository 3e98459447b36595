"""Batched umbrella windows: many independent systems in one model.

The reference runs one system per process (src/total_force_cuda.cpp:36-100);
an umbrella-sampling campaign is many such runs.  Here W windows of uneven
size are concatenated into one `System` whose `windows` field lists the atom
count of each window; the C ABI's *_create_batch entry points then run every
window in the same launches, each window seeing only its own atoms, with its
own box and Philox system id.

Type tables are shared: windows built from the same parameter set (same LJ
table, same IS-SPA tables, same bonded parameters) reuse one block of types,
so the merged LJ table stays small enough for the nonbond kernel's LDS.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence

import numpy as np

from .system import IsspaTables, System, UmbrellaRestraint


def _key(*arrays) -> bytes:
    return b"|".join(np.ascontiguousarray(a).tobytes() + str(a.shape).encode() for a in arrays)


class _Blocks:
    """Deduplicated concatenation of per-window parameter blocks."""

    def __init__(self):
        self.index = {}
        self.parts = []
        self.count = 0

    def add(self, key: bytes, part, size: int) -> int:
        if key not in self.index:
            self.index[key] = self.count
            self.parts.append(part)
            self.count += size
        return self.index[key]


def merge_windows(windows: Sequence[System], system_ids: Optional[Sequence[int]] = None,
                  name: str = "batch") -> System:
    """Concatenate systems into one batched `System` (atoms in window order).

    All windows must share n_mc, temperature, pnu and the IS-SPA table grids.
    The merged system's seed is the first window's; window w keeps its own
    Monte-Carlo stream through system id `system_ids[w]` (default w)."""
    if not windows:
        raise ValueError("no windows")
    w0 = windows[0]
    for s in windows:
        if (s.n_mc, s.temperature_k, s.pnu) != (w0.n_mc, w0.temperature_k, w0.pnu):
            raise ValueError(f"window {s.name}: n_mc/temperature/pnu differ from {w0.name}")
        it, i0 = s.isspa, w0.isspa
        if (it.g_grid, it.e_grid, it.f_grid) != (i0.g_grid, i0.e_grid, i0.f_grid) or \
                it.g.shape[1] != i0.g.shape[1] or it.e.shape[1] != i0.e.shape[1] or \
                it.f.shape[1] != i0.f.shape[1]:
            raise ValueError(f"window {s.name}: IS-SPA table grids differ")
    W = len(windows)
    sys_ids = np.arange(W, dtype=np.uint32) if system_ids is None else \
        np.asarray(system_ids, np.uint32)

    # LJ type blocks and IS-SPA type blocks
    lj_blocks, is_blocks = _Blocks(), _Blocks()
    bond_b, ang_b, dih_b = _Blocks(), _Blocks(), _Blocks()
    lj_off, is_off, b_off, a_off, d_off = [], [], [], [], []
    for s in windows:
        lj_off.append(lj_blocks.add(_key(s.nbparm, s.lj_a, s.lj_b), s, s.n_lj_types))
        it = s.isspa
        is_off.append(is_blocks.add(_key(it.rmax, it.g, it.e, it.f), it, it.n_types))
        b_off.append(bond_b.add(_key(s.bond_params), s.bond_params, len(s.bond_params)))
        a_off.append(ang_b.add(_key(s.angle_params), s.angle_params, len(s.angle_params)))
        d_off.append(dih_b.add(_key(s.dih_params, s.sc14), (s.dih_params, s.sc14),
                               len(s.dih_params)))
    T = lj_blocks.count
    ntri = T * (T + 1) // 2
    lj_a = np.zeros(ntri, np.float32)
    lj_b = np.zeros(ntri, np.float32)
    nbparm = np.zeros((T, T), np.int32)
    for a in range(T):
        for b in range(T):
            lo, hi = min(a, b), max(a, b)
            nbparm[a, b] = lo * T - lo * (lo - 1) // 2 + (hi - lo)
    # fill same-block pairs from each block's own table
    for off, s in zip(lj_blocks.index.values(), lj_blocks.parts):
        Tb = s.n_lj_types
        nbp = np.asarray(s.nbparm).reshape(Tb, Tb)
        for a in range(Tb):
            for b in range(Tb):
                k = nbparm[off + a, off + b]
                lj_a[k] = s.lj_a[nbp[a, b]]
                lj_b[k] = s.lj_b[nbp[a, b]]
    tabs = is_blocks.parts
    merged_tab = IsspaTables(
        isspa_type=np.concatenate([s.isspa.isspa_type + o for s, o in zip(windows, is_off)])
        .astype(np.int32),
        rmax=np.concatenate([t.rmax for t in tabs]).astype(np.float32),
        g=np.concatenate([t.g for t in tabs]).astype(np.float32),
        e=np.concatenate([t.e for t in tabs]).astype(np.float32),
        f=np.concatenate([t.f for t in tabs]).astype(np.float32),
        g_grid=w0.isspa.g_grid, e_grid=w0.isspa.e_grid, f_grid=w0.isspa.f_grid)

    xs, vs, ms, ity = [], [], [], []
    pre, lst = [], []
    bonds, angles, dihs, dtypes = [], [], [], []
    restraints: List[UmbrellaRestraint] = []
    base = 0
    lbase = 0
    for w, s in enumerate(windows):
        n = s.n_atoms
        xs.append(s.xyzq)
        vs.append(s.vel)
        ms.append(s.true_mass)
        ity.append(s.ityp + lj_off[w])
        pre.append(s.nexcl_prefix.astype(np.int64) + lbase)
        lst.append(np.where(s.excl_list >= 0, s.excl_list + base, s.excl_list))
        lbase += len(s.excl_list)
        if len(s.bond_atoms):
            b = s.bond_atoms.astype(np.int64).copy()
            b[:, :2] += base
            b[:, 2] += b_off[w]
            bonds.append(b)
        if len(s.angle_atoms):
            a = s.angle_atoms.astype(np.int64).copy()
            a[:, :3] += base
            a[:, 3] += a_off[w]
            angles.append(a)
        if len(s.dih_atoms):
            dihs.append(_shift_dih(s.dih_atoms, base))
            dtypes.append(s.dih_types + d_off[w])
        if s.us is not None:
            al = s.us.atom_list.copy()
            al[:, 0] += base
            restraints.append(UmbrellaRestraint(al, s.us.mass.copy(), s.us.k, s.us.x0))
        else:
            restraints.append(None)
        base += n

    def cat(parts, shape, dt):
        return np.concatenate(parts).astype(dt) if parts else np.zeros(shape, dt)

    dpar = dih_b.parts
    out = System(
        xyzq=np.concatenate(xs).astype(np.float32),
        vel=np.concatenate(vs).astype(np.float32),
        true_mass=np.concatenate(ms).astype(np.float32),
        ityp=np.concatenate(ity).astype(np.int32),
        n_lj_types=T,
        nbparm=nbparm.reshape(-1),
        lj_a=lj_a,
        lj_b=lj_b,
        nexcl_prefix=np.concatenate(pre).astype(np.int32),
        excl_list=np.concatenate(lst).astype(np.int32),
        bond_atoms=cat(bonds, (0, 4), np.int32),
        bond_params=cat(bond_b.parts, (0, 2), np.float32),
        angle_atoms=cat(angles, (0, 4), np.int32),
        angle_params=cat(ang_b.parts, (0, 2), np.float32),
        dih_atoms=cat(dihs, (0, 4), np.int32),
        dih_types=cat(dtypes, (0,), np.int32),
        dih_params=cat([p[0] for p in dpar], (0, 4), np.float32),
        sc14=cat([p[1] for p in dpar], (0, 2), np.float32),
        isspa=merged_tab,
        lbox=float(max(s.lbox for s in windows)),
        n_mc=w0.n_mc,
        temperature_k=w0.temperature_k,
        pnu=w0.pnu,
        seed=w0.seed,
        us=None,
        name=name,
    )
    out.windows = np.array([s.n_atoms for s in windows], np.int32)
    out.window_lbox = np.array([s.lbox for s in windows], np.float32)
    out.window_sys = sys_ids
    out.restraints = restraints
    out.check()
    return out


def _shift_dih(dih: np.ndarray, base: int) -> np.ndarray:
    """Shift signed dihedral atom indices by `base`; a negative (flagged) third
    or fourth index stays negative (src/dih_force_cuda.cu:60,83-84)."""
    a = dih.astype(np.int64).copy()
    for col in range(4):
        v = a[:, col]
        a[:, col] = np.where(v < 0, v - base, v + base)
    return a.astype(np.int32)


def window_slices(system: System):
    """(start, stop) atom range of every window of a batched system."""
    n = np.asarray(system.windows, np.int64)
    a0 = np.concatenate([[0], np.cumsum(n)[:-1]])
    return [(int(a), int(a + k)) for a, k in zip(a0, n)]


def umbrella_windows(n_ljpair: int = 32, n_chain: int = 24, n_big: int = 8,
                     k: float = 10.0, x_lo: float = 3.0, x_hi: float = 15.0,
                     n_mc: int = 20, lbox: float = 200.0, seed: int = 12348) -> List[System]:
    """The batched umbrella-sampling workload: 64 windows of uneven size --
    32 two-atom LJ pairs (sigma 3.8 A, eps 0.24, uncharged), 24 copies of the
    22-atom chain and 8 systems of 10 chains (220 atoms) -- each with a
    harmonic distance restraint (k kcal/mol/A^2) between two designated atoms,
    centres evenly spaced over [x_lo, x_hi] in window order.  Window kinds are
    interleaved so each kind spans the whole range of centres.

    Designated atoms: the two atoms of an LJ pair; the first and last atom of a
    chain; the first atoms of the first two chains of a 10-chain system."""
    from .synth import make_chain, make_lattice, maxwell_velocities, two_atom_lj

    W = n_ljpair + n_chain + n_big
    counts = {"ljpair": n_ljpair, "chain": n_chain, "big": n_big}
    # interleave: spread each kind evenly over the window order
    order = []
    for kind, c in counts.items():
        order += [((i + 0.5) / c, kind) for i in range(c)]
    order.sort()
    kinds = [kd for _, kd in order]
    centres = np.linspace(x_lo, x_hi, W)
    out = []
    for w, (kind, x0) in enumerate(zip(kinds, centres)):
        x0 = float(x0)
        if kind == "ljpair":
            s = two_atom_lj(sep=x0, lbox=lbox)
            s.us = UmbrellaRestraint.from_groups([0], [1], s.true_mass, k, x0)
        elif kind == "chain":
            s = make_chain()
            s.xyzq[:, :3] += np.float32(0.5 * lbox)
            s.lbox = lbox
            s.us = UmbrellaRestraint.from_groups([0], [s.n_atoms - 1], s.true_mass, k, x0)
        else:
            s = make_lattice(10, 40.0, 3, lbox, seed=seed + w)
            s.us = UmbrellaRestraint.from_groups([0], [22], s.true_mass, k, x0)
        s.n_mc, s.temperature_k, s.pnu, s.seed = n_mc, 298.0, 0.01, seed
        maxwell_velocities(s, seed=1000 + w)
        s.name = f"w{w:02d}_{kind}"
        out.append(s)
    return out

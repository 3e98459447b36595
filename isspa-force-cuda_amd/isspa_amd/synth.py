"""Seeded synthetic IS-SPA systems of the benchmark workloads.

The reference ships only the ADI2/ADI25/ADI50/LJ decks; the benchmark
workloads (a 22-atom chain and lattices of 50 and 909 copies of it) are
generated here.  Every choice without a reference counterpart is frozen in
``SPEC`` below and recorded in the bench output.

Chain topology (22 atoms, 21 bonds, 36 angles, 41 dihedral terms): a
backbone B0..B8 with 3,2,0,2,1,0,2,0,3 terminal atoms, atoms numbered
depth-first so exclusions stay near the diagonal; 32 proper dihedral
quadruples over the backbone bonds plus 9 extra periodicity terms on the
first 9 quadruples (third index negated, so their 1-4 pair is not counted
twice -- the AMBER convention the reference honours at
src/dih_force_cuda.cu:60,83-84).

IS-SPA tables per type t (rows r_b = 0.05 + 0.1 b, b < 120, rmax = 12 A):
  x0 = sigma_t + 0.5, w = sqrt(g0/alpha), r_hi = x0 + sqrt((g0-1)/alpha)
  r_lo = max(x0 - w, 0.85 sigma_ts)        (sigma_ts = (sigma_t + 3.15)/2)
  G(r) = 0 (r < r_lo); g0 - alpha (r-x0)^2 (r_lo <= r < r_hi); 1 (r >= r_hi)
  F(r) = -dU/dr of 4 eps_ts ((sigma_ts/r)^12 - (sigma_ts/r)^6), eps_ts =
         sqrt(eps_t * 0.152), for r >= r_lo; 0 below
  E(r) = c_t exp(-(r - x0)^2 / 2) for r >= r_lo, c_t ~ U[-0.05, 0.05]
The r_lo floor keeps MC points out of the steep LJ core of small atoms,
where x0 - w < 0 would put g > 0 at the nucleus.
"""
from __future__ import annotations

import math

import numpy as np

from .system import IsspaTables, System, UmbrellaRestraint
from .tables import parabola_tables

SPEC = {
    "chain_seed": 2024,
    "n_lj_types": 6,
    "sigma_range": (1.0, 3.8),
    "eps_range": (0.01, 0.2),
    "charge_range": (-0.5, 0.5),
    "amber_charge_scale": 18.2223,
    "bond_r0": (1.0, 1.5), "bond_k": (300.0, 500.0),
    "angle_theta0_deg": (105.0, 120.0), "angle_k": (50.0, 80.0),
    "dih_n": (1, 3), "dih_V": (0.0, 2.0), "dih_phase": (0.0, math.pi),
    "scee": 1.2, "scnb": 2.0,
    "g0_range": (1.2, 2.0), "alpha_range": (0.5, 2.0),
    "solvent_sigma": 3.15, "solvent_eps": 0.152,
    "rmax": 12.0, "table_r0": 0.05, "table_dr": 0.1, "table_rows": 120,
    "core_frac": 0.85, "efield_amp": 0.05,
    "min_sep": 1.0,
    "masses_backbone": (12.01, 14.01, 16.00, 12.01),
    "masses_terminal": (1.008, 1.008),
}

BACKBONE_LEAVES = (3, 2, 0, 2, 1, 0, 2, 0, 3)


def chain_topology():
    """Atom list in depth-first order: (is_backbone, parent) and the bond list."""
    atoms, bonds, backbone = [], [], []
    for k, nleaf in enumerate(BACKBONE_LEAVES):
        b = len(atoms)
        atoms.append((True, backbone[-1] if backbone else -1))
        if backbone:
            bonds.append((backbone[-1], b))
        backbone.append(b)
        for _ in range(nleaf):
            atoms.append((False, b))
            bonds.append((b, len(atoms) - 1))
    return atoms, bonds, backbone


def _neighbours(n, bonds):
    nb = [[] for _ in range(n)]
    for a, b in bonds:
        nb[a].append(b)
        nb[b].append(a)
    return nb


def _tables(sig, eps, rng):
    """Draw the per-type parabola and E-field parameters and build the tables
    (isspa_amd.tables.parabola_tables)."""
    S = SPEC
    T = len(sig)
    g0 = rng.uniform(*S["g0_range"], size=T)
    al = rng.uniform(*S["alpha_range"], size=T)
    ce = rng.uniform(-S["efield_amp"], S["efield_amp"], size=T)
    x0 = [sig[t] + 0.5 for t in range(T)]
    return parabola_tables(sig, eps, x0, g0, al, ce, r0=S["table_r0"], dr=S["table_dr"],
                           rows=S["table_rows"], core=S["core_frac"],
                           solvent_sigma=S["solvent_sigma"], solvent_eps=S["solvent_eps"])


def make_chain(seed: int = SPEC["chain_seed"]) -> System:
    """The 22-atom chain, centred at the origin."""
    S = SPEC
    rng = np.random.default_rng(seed)
    atoms, bonds, backbone = chain_topology()
    n = len(atoms)
    nbr = _neighbours(n, bonds)
    T = S["n_lj_types"]
    sig = rng.uniform(*S["sigma_range"], size=T)
    eps = rng.uniform(*S["eps_range"], size=T)
    ityp = np.array([rng.integers(0, 4) if bb else rng.integers(4, 6) for bb, _ in atoms], np.int32)
    mass = np.array([(S["masses_backbone"] + S["masses_terminal"])[t] for t in ityp], np.float32)
    q = rng.uniform(*S["charge_range"], size=n)
    q = (q - q.mean()) * S["amber_charge_scale"]

    # LJ tables in prmtop layout (Lorentz-Berthelot mixing)
    ntri = T * (T + 1) // 2
    lj_a = np.zeros(ntri, np.float32)
    lj_b = np.zeros(ntri, np.float32)
    nbparm = np.zeros(T * T, np.int32)
    for i in range(T):
        for j in range(T):
            a, b = min(i, j), max(i, j)
            idx = b * (b + 1) // 2 + a
            nbparm[i * T + j] = idx
            s_ij = 0.5 * (sig[i] + sig[j])
            e_ij = math.sqrt(eps[i] * eps[j])
            lj_a[idx] = 4 * e_ij * s_ij ** 12
            lj_b[idx] = 4 * e_ij * s_ij ** 6

    # bonded terms, one parameter type per term
    bond_atoms = np.zeros((len(bonds), 4), np.int32)
    bond_params = np.zeros((len(bonds), 2), np.float32)
    r0 = {}
    for k, (a, b) in enumerate(bonds):
        rr = rng.uniform(*S["bond_r0"])
        kk = rng.uniform(*S["bond_k"])
        bond_atoms[k] = (a, b, k, 0)
        bond_params[k] = (2.0 * kk, rr)
        r0[(a, b)] = r0[(b, a)] = rr
    angles = [(a, c, b) for c in range(n) for ia, a in enumerate(nbr[c]) for b in nbr[c][ia + 1:]]
    angle_atoms = np.zeros((len(angles), 4), np.int32)
    angle_params = np.zeros((len(angles), 2), np.float32)
    th0 = {}
    for k, (a, c, b) in enumerate(angles):
        t0 = math.radians(rng.uniform(*S["angle_theta0_deg"]))
        angle_atoms[k] = (a, c, b, k)
        angle_params[k] = (2.0 * rng.uniform(*S["angle_k"]), t0)
        th0[(a, c, b)] = th0[(b, c, a)] = t0
    quads = []
    for k in range(len(backbone) - 1):
        j, l = backbone[k], backbone[k + 1]
        for i in nbr[j]:
            if i == l:
                continue
            for m in nbr[l]:
                if m != j:
                    quads.append((i, j, l, m))
    extra = quads[:9]
    dih_atoms = np.array([q4 for q4 in quads] + [(i, j, -l, m) for i, j, l, m in extra], np.int32)
    nd = len(dih_atoms)
    dih_types = np.arange(nd, dtype=np.int32)
    dih_params = np.zeros((nd, 4), np.float32)
    for k in range(nd):
        dih_params[k] = (rng.integers(S["dih_n"][0], S["dih_n"][1] + 1), rng.uniform(*S["dih_V"]),
                         S["dih_phase"][rng.integers(0, 2)], 0.0)
    sc14 = np.tile(np.array([[S["scee"], S["scnb"]]], np.float32), (nd, 1))

    # exclusions (1-2, 1-3, 1-4), AMBER layout: partners with larger index,
    # atoms without any get one placeholder entry (-1 after the 1-based shift)
    excl = [set() for _ in range(n)]
    for a, b in bonds:
        excl[min(a, b)].add(max(a, b))
    for a, c, b in angles:
        excl[min(a, b)].add(max(a, b))
    for i, j, l, m in quads:
        excl[min(i, m)].add(max(i, m))
    counts, lst = [], []
    for i in range(n):
        e = sorted(excl[i])
        if not e:
            e = [-1]
        counts.append(len(e))
        lst.extend(e)

    # IS-SPA tables: one IS-SPA type per LJ type
    G, E, F, par = _tables(sig, eps, rng)
    rmax = np.full(T, S["rmax"], np.float32)
    tab = IsspaTables(isspa_type=ityp.copy(), rmax=rmax, g=G, e=E, f=F,
                      g_grid=(np.float32(S["table_r0"]), np.float32(S["table_dr"])),
                      e_grid=(np.float32(S["table_r0"]), np.float32(S["table_dr"])),
                      f_grid=(np.float32(S["table_r0"]), np.float32(S["table_dr"])))

    xyz = _build_geometry(n, atoms, nbr, r0, th0, rng)
    xyzq = np.zeros((n, 4), np.float32)
    xyzq[:, :3] = xyz - xyz.mean(axis=0)
    xyzq[:, 3] = q
    vel = np.zeros((n, 4), np.float32)
    vel[:, 3] = np.where(mass < 2.0, np.float32(12.0), mass)  # src/read_prmtop.cpp:144-145
    sysm = System(xyzq=xyzq, vel=vel, true_mass=mass, ityp=ityp, n_lj_types=T, nbparm=nbparm,
                  lj_a=lj_a, lj_b=lj_b, nexcl_prefix=np.cumsum(counts).astype(np.int32),
                  excl_list=np.array(lst, np.int32), bond_atoms=bond_atoms,
                  bond_params=bond_params, angle_atoms=angle_atoms, angle_params=angle_params,
                  dih_atoms=dih_atoms, dih_types=dih_types, dih_params=dih_params, sc14=sc14,
                  isspa=tab, name="chain22")
    sysm.type_params = par
    sysm.sigma, sysm.epsilon = sig, eps
    return sysm


def _build_geometry(n, atoms, nbr, r0, th0, rng):
    """Place atoms in index order: bond length r0, angle theta0 to the
    grandparent, random azimuth; reject placements closer than min_sep."""
    for _attempt in range(1000):
        pos = np.zeros((n, 3))
        ok = True
        for i in range(1, n):
            par = atoms[i][1]
            gp = atoms[par][1] if par >= 0 else -1
            if gp < 0:
                # grandparent: any already placed neighbour of the parent
                cand = [a for a in nbr[par] if a < i and a != i]
                gp = cand[0] if cand else -1
            placed = False
            for _ in range(200):
                if gp < 0:
                    d = rng.normal(size=3)
                    d /= np.linalg.norm(d)
                else:
                    t0 = th0.get((gp, par, i), math.radians(110.0))
                    u = pos[gp] - pos[par]
                    u /= np.linalg.norm(u)
                    v = rng.normal(size=3)
                    v -= v.dot(u) * u
                    v /= np.linalg.norm(v)
                    d = math.cos(t0) * u + math.sin(t0) * v
                p = pos[par] + r0[(par, i)] * d
                if i == 1 or np.min(np.linalg.norm(pos[:i] - p, axis=1)) >= SPEC["min_sep"]:
                    pos[i] = p
                    placed = True
                    break
            if not placed:
                ok = False
                break
        if ok:
            return pos
    raise RuntimeError("could not build a self-avoiding chain")


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def make_lattice(copies: int, region: float, grid: int, lbox: float, seed: int,
                 chain_seed: int = SPEC["chain_seed"]) -> System:
    """`copies` randomly rotated chains on distinct sites of a grid^3 lattice
    filling a `region`-wide cube centred in the box."""
    base = make_chain(chain_seed)
    rng = np.random.default_rng(seed)
    if copies > grid ** 3:
        raise ValueError("more copies than lattice sites")
    sites = rng.permutation(grid ** 3)[:copies]
    spacing = region / grid
    ctr = 0.5 * lbox - 0.5 * region + 0.5 * spacing
    x0 = base.xyzq[:, :3].astype(np.float64)
    offsets, rots, placed = [], [], []
    for s in sites:
        ix, iy, iz = s // (grid * grid), (s // grid) % grid, s % grid
        off = ctr + spacing * np.array([ix, iy, iz], np.float64)
        for _ in range(100):
            R = _random_rotation(rng)
            x = x0 @ R.T + off
            if not placed or _min_dist(x, placed) >= SPEC["min_sep"]:
                break
        placed.append(x)
        offsets.append(off)
        rots.append(R)
    sysm = base.replicate(copies, np.array(offsets), np.array(rots), name=f"chain22x{copies}")
    sysm.lbox = lbox
    sysm.type_params = base.type_params
    sysm.sigma, sysm.epsilon = base.sigma, base.epsilon
    return sysm


def _min_dist(x, placed):
    near = [p for p in placed[-64:]]
    allp = np.concatenate(near)
    d = np.linalg.norm(x[:, None, :] - allp[None, :, :], axis=2)
    return float(d.min())


def maxwell_velocities(system: System, seed: int) -> None:
    """Maxwell-Boltzmann velocities at system.temperature_k with the dynamic
    masses (src/atom_class.cpp:53-63 draws the same distribution)."""
    rng = np.random.default_rng(seed)
    m = system.vel[:, 3].astype(np.float64)
    s = np.sqrt(system.kT / m)
    system.vel[:, :3] = (rng.normal(size=(system.n_atoms, 3)) * s[:, None]).astype(np.float32)


WORKLOAD_NAMES = ("chain22_nmc10", "chains50_1100atoms_nmc20", "chains909_19998atoms_nmc20")


_LATTICE_SEEDS = {1: 11, 2: 12}


def workload_seeds(index: int) -> dict:
    """Every seed the plain workload `index` is built from: the chain
    generator, the lattice placement (none for the single chain), the Philox
    key of MC points and Andersen draws, and the Maxwell velocities."""
    out = {"chain": SPEC["chain_seed"], "philox": 12345 + index, "velocities": 100 + index}
    if index in _LATTICE_SEEDS:
        out["lattice"] = _LATTICE_SEEDS[index]
    return out


def workload(index: int) -> System:
    """The plain workloads, by index:
    0: one chain, nMC=10, box 200 A;
    1: 50 chains (1,100 atoms) on a 4^3 lattice in a 60 A region, nMC=20, box 200 A;
    2: 909 chains (19,998 atoms) on a 10^3 lattice in a 160 A region, nMC=20,
       box 400 A (the minimum image never fires)."""
    if index == 0:
        s = make_chain()
        s.lbox, s.n_mc = 200.0, 10
        s.xyzq[:, :3] += np.float32(100.0)
    elif index == 1:
        s = make_lattice(50, 60.0, 4, 200.0, seed=_LATTICE_SEEDS[1])
        s.n_mc = 20
    elif index == 2:
        s = make_lattice(909, 160.0, 10, 400.0, seed=_LATTICE_SEEDS[2])
        s.n_mc = 20
    else:
        raise ValueError(f"no plain workload {index}")
    seeds = workload_seeds(index)
    s.temperature_k, s.pnu, s.seed = 298.0, 0.01, seeds["philox"]
    maxwell_velocities(s, seed=seeds["velocities"])
    s.name = WORKLOAD_NAMES[index]
    s.check()
    return s


def two_atom_lj(sep: float = 4.0, lbox: float = 200.0) -> System:
    """Two uncharged LJ atoms (sigma 3.8 A, eps 0.24), the LJ-pair window
    system of the batched workload."""
    base = make_chain()
    sig, eps = 3.8, 0.24
    A, B = 4 * eps * sig ** 12, 4 * eps * sig ** 6
    xyzq = np.array([[0.5 * lbox, 0.5 * lbox, 0.5 * lbox, 0.0],
                     [0.5 * lbox + sep, 0.5 * lbox, 0.5 * lbox, 0.0]], np.float32)
    G, E, F, par = _tables(np.array([sig]), np.array([eps]), np.random.default_rng(5))
    tab = IsspaTables(isspa_type=np.zeros(2, np.int32), rmax=np.full(1, SPEC["rmax"], np.float32),
                      g=G, e=E, f=F, g_grid=base.isspa.g_grid, e_grid=base.isspa.e_grid,
                      f_grid=base.isspa.f_grid)
    vel = np.zeros((2, 4), np.float32)
    vel[:, 3] = 39.948
    s = System(xyzq=xyzq, vel=vel, true_mass=np.full(2, 39.948, np.float32),
               ityp=np.zeros(2, np.int32), n_lj_types=1, nbparm=np.zeros(1, np.int32),
               lj_a=np.array([A], np.float32), lj_b=np.array([B], np.float32),
               nexcl_prefix=np.array([1, 2], np.int32), excl_list=np.array([-1, -1], np.int32),
               bond_atoms=np.zeros((0, 4), np.int32), bond_params=np.zeros((0, 2), np.float32),
               angle_atoms=np.zeros((0, 4), np.int32), angle_params=np.zeros((0, 2), np.float32),
               dih_atoms=np.zeros((0, 4), np.int32), dih_types=np.zeros(0, np.int32),
               dih_params=np.zeros((0, 4), np.float32), sc14=np.zeros((0, 2), np.float32),
               isspa=tab, lbox=lbox, name="ljpair")
    s.type_params = par
    s.us = UmbrellaRestraint.from_groups([0], [1], s.true_mass, 10.0, sep)
    return s

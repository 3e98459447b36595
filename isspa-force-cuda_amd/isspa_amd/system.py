"""In-memory description of an IS-SPA MD system, in the device layouts.

Every array here has the dtype and layout the C-ABI expects, so a `System`
can be uploaded as is.  The layouts follow the reference's device arrays
(src/atom_class.h:31-49, src/isspa_class.h:14-42, src/bond_class.h,
src/angle_class.h, src/dih_class.h, src/us_class.h:121-131):

* positions  float32[N,4]  (x, y, z, charge)            -- pos_d
* velocities float32[N,4]  (vx, vy, vz, dynamic mass)   -- vel_d
* bonds      int32[nb,4]   (i, j, type, 0) + float32[nbt,2] (2k, r0)
* angles     int32[na,4]   (i, j, k, type) + float32[nat,2] (2k, theta0)
* dihedrals  int32[nd,4]   (i, j, +-k, +-l) + int32[nd] type
             + float32[ndt,4] (n, V, phase, 0) + float32[ndt,2] (scee, scnb)
* exclusions int32[N] inclusive prefix sum + int32[L] list (0-based)
* IS-SPA     int32[N] type, float32[T] rmax, tables float32[T, n] on r0 + b*dr
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

# src/constants.h:6 (PI as the reference writes it, a float literal)
PI_F32 = float(np.float32(3.1415926535))
# src/config_class.cpp:123-125
DT_PS = 0.002
AMBER_TIME = 20.455
KB_KCAL = 0.00198717


@dataclasses.dataclass
class IsspaTables:
    """Per-type IS-SPA tables (src/isspa_class.cpp:193-241)."""

    isspa_type: np.ndarray  # int32[N]
    rmax: np.ndarray  # float32[T]
    g: np.ndarray  # float32[T, n_g]  ISSPA_DENSITIES
    e: np.ndarray  # float32[T, n_e]  ISSPA_EFIELD
    f: np.ndarray  # float32[T, n_f]  ISSPA_FORCES
    g_grid: tuple  # (r0, dr) as float32
    e_grid: tuple
    f_grid: tuple

    @property
    def n_types(self) -> int:
        return int(self.rmax.shape[0])

    def vtot(self, n_mc: int) -> np.ndarray:
        """Monte-Carlo normalisation, src/isspa_class.cpp:110-112, evaluated in
        double with the reference's operand order and stored as float."""
        out = np.empty(self.n_types, np.float32)
        for t in range(self.n_types):
            r = float(self.rmax[t])
            out[t] = np.float32(4.0 / 3.0 * PI_F32 * r * r * r * 0.0074 / float(np.float32(n_mc)))
        return out


@dataclasses.dataclass
class UmbrellaRestraint:
    """Harmonic COM-distance restraint (src/us_class.cpp:13-157)."""

    atom_list: np.ndarray  # int32[nbias, 2] (atom, group)
    mass: np.ndarray  # float32[nbias] mass / group mass
    k: float
    x0: float

    @property
    def kumb(self) -> np.ndarray:
        return np.array([-self.k, self.k], np.float32)

    @staticmethod
    def from_groups(group1, group2, true_mass, k, x0) -> "UmbrellaRestraint":
        """populate_mass (src/us_class.cpp:136-157): masses normalised by the
        FP32 group total, true (not hydrogen-reweighted) masses."""
        atoms = [(int(a), 0) for a in group1] + [(int(a), 1) for a in group2]
        al = np.array(atoms, np.int32).reshape(-1, 2)
        m = np.array([true_mass[a] for a, _ in atoms], np.float32)
        tot = [np.float32(0.0), np.float32(0.0)]
        for (a, gidx), mm in zip(atoms, m):
            tot[gidx] = np.float32(tot[gidx] + mm)
        mn = np.array([mm / tot[gidx] for (a, gidx), mm in zip(atoms, m)], np.float32)
        return UmbrellaRestraint(al, mn, float(k), float(x0))


@dataclasses.dataclass
class System:
    xyzq: np.ndarray
    vel: np.ndarray
    true_mass: np.ndarray
    ityp: np.ndarray
    n_lj_types: int
    nbparm: np.ndarray
    lj_a: np.ndarray
    lj_b: np.ndarray
    nexcl_prefix: np.ndarray
    excl_list: np.ndarray
    bond_atoms: np.ndarray
    bond_params: np.ndarray
    angle_atoms: np.ndarray
    angle_params: np.ndarray
    dih_atoms: np.ndarray
    dih_types: np.ndarray
    dih_params: np.ndarray
    sc14: np.ndarray
    isspa: IsspaTables
    lbox: float = 200.0
    n_mc: int = 20
    temperature_k: float = 298.0
    pnu: float = 0.01
    seed: int = 12345
    us: Optional[UmbrellaRestraint] = None
    name: str = "system"

    @property
    def n_atoms(self) -> int:
        return int(self.xyzq.shape[0])

    @property
    def n_points(self) -> int:
        return self.n_atoms * self.n_mc

    @property
    def kT(self) -> float:
        """temperature in kcal/mol as the reference stores it (float)."""
        return float(np.float32(np.float32(self.temperature_k) * np.float32(KB_KCAL)))

    @property
    def dt(self) -> float:
        return float(np.float32(DT_PS * AMBER_TIME))

    def check(self) -> None:
        n = self.n_atoms
        assert self.xyzq.dtype == np.float32 and self.xyzq.shape == (n, 4)
        assert self.vel.dtype == np.float32 and self.vel.shape == (n, 4)
        assert self.ityp.shape == (n,) and self.isspa.isspa_type.shape == (n,)
        assert self.nexcl_prefix.shape == (n,)
        assert int(self.nexcl_prefix[-1]) == self.excl_list.shape[0] if n else True
        for a in (self.bond_atoms, self.angle_atoms, self.dih_atoms):
            assert a.dtype == np.int32 and (a.size == 0 or a.shape[1] == 4)
        assert np.all(self.isspa.isspa_type < self.isspa.n_types)
        assert np.all(self.ityp < self.n_lj_types)

    def replicate(self, copies: int, offsets: np.ndarray, rotations=None, name=None) -> "System":
        """Concatenate `copies` translated/rotated copies of this system."""
        n = self.n_atoms
        xs, vs = [], []
        for c in range(copies):
            x = self.xyzq[:, :3].astype(np.float64)
            if rotations is not None:
                ctr = x.mean(axis=0)
                x = (x - ctr) @ rotations[c].T + ctr
            x = x + offsets[c]
            xs.append(np.concatenate([x.astype(np.float32), self.xyzq[:, 3:4]], axis=1))
            vs.append(self.vel.copy())

        def rep_idx(arr, ncols_atoms):
            out = []
            for c in range(copies):
                a = arr.copy()
                for col in range(ncols_atoms):
                    v = a[:, col]
                    sign = np.where(v < 0, -1, 1)
                    a[:, col] = sign * (np.abs(v) + c * n)
                out.append(a)
            return np.concatenate(out) if out else arr

        # exclusions: list entries shift by c*n, but the reference's "0" (-1)
        # placeholder entries stay negative
        pre, lst = [], []
        base = 0
        for c in range(copies):
            counts = np.diff(np.concatenate([[0], self.nexcl_prefix]))
            l2 = np.where(self.excl_list >= 0, self.excl_list + c * n, self.excl_list)
            lst.append(l2)
            pre.append(np.cumsum(counts) + base)
            base += len(self.excl_list)
        us = None
        return System(
            xyzq=np.concatenate(xs).astype(np.float32),
            vel=np.concatenate(vs).astype(np.float32),
            true_mass=np.tile(self.true_mass, copies),
            ityp=np.tile(self.ityp, copies),
            n_lj_types=self.n_lj_types,
            nbparm=self.nbparm,
            lj_a=self.lj_a,
            lj_b=self.lj_b,
            nexcl_prefix=np.concatenate(pre).astype(np.int32),
            excl_list=np.concatenate(lst).astype(np.int32),
            bond_atoms=rep_idx(self.bond_atoms, 2).astype(np.int32),
            bond_params=self.bond_params,
            angle_atoms=rep_idx(self.angle_atoms, 3).astype(np.int32),
            angle_params=self.angle_params,
            dih_atoms=rep_idx_dih(self.dih_atoms, copies, n),
            dih_types=np.tile(self.dih_types, copies),
            dih_params=self.dih_params,
            sc14=self.sc14,
            isspa=dataclasses.replace(self.isspa, isspa_type=np.tile(self.isspa.isspa_type, copies)),
            lbox=self.lbox,
            n_mc=self.n_mc,
            temperature_k=self.temperature_k,
            pnu=self.pnu,
            seed=self.seed,
            us=us,
            name=name or f"{self.name}x{copies}",
        )


def rep_idx_dih(dih: np.ndarray, copies: int, n: int) -> np.ndarray:
    """Shift signed dihedral atom indices; the sign flags survive (a negative
    third/fourth index stays negative, src/dih_force_cuda.cu:60,83-84)."""
    out = []
    for c in range(copies):
        a = dih.astype(np.int64).copy()
        for col in range(4):
            v = a[:, col]
            a[:, col] = np.where(v < 0, v - c * n, v + c * n)
        out.append(a)
    return np.concatenate(out).astype(np.int32) if out else dih


def max_abs_rel_rms(a: np.ndarray, b: np.ndarray) -> float:
    """relative RMS of a against b: sqrt(sum|a-b|^2 / sum|b|^2)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = float(np.sum(b * b))
    return math.sqrt(float(np.sum((a - b) ** 2)) / den) if den > 0 else float(np.sqrt(np.sum(a * a)))

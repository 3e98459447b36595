"""Readers for the reference's input formats.

* AMBER prmtop (subset), following the conversions of src/read_prmtop.cpp:10-593
  - charges divided by sqrt(dielectric)                      (:125)
  - dynamic mass 12 for atoms lighter than 2 (hydrogens)      (:141-146)
  - exclusion counts turned into an inclusive prefix sum      (:177-183)
  - coordinate indices divided by 3, type indices minus 1     (:405 ff.)
  - bond and angle force constants doubled                    (:233, :267)
* IS-SPA topology (.isspatop), src/isspa_class.cpp:115-248
  - POINTERS: nAtoms, nTypes, nGRs, nERs, nRs at columns 0/8/16/24/32
* AMBER restart (.rst7), src/atom_class.cpp:76-156 (6 values of width 12 a line)
* run deck (`key = value`, `#` comments), src/config_class.cpp:14-133
* umbrella file, src/us_class.cpp:13-104

Fixed-width fields are cut by the width the section's %FORMAT declares, as
the reference cuts them by its hard-wired widths; for the shipped files the
two agree.
"""
from __future__ import annotations

import math
import re
from typing import Dict, List

import numpy as np

from .system import IsspaTables, System, UmbrellaRestraint

_FMT = re.compile(r"\((\d+)([aAiIeEfF])(\d+)(?:\.(\d+))?\)")


def _sections(path: str) -> Dict[str, List[str]]:
    """Split an AMBER-style file into {flag: [format, data lines...]}."""
    out: Dict[str, List[str]] = {}
    cur = None
    with open(path, "r") as fh:
        for line in fh:
            line = line.rstrip("\n")
            if line.startswith("%FLAG"):
                cur = line.split()[1]
                out[cur] = []
            elif cur is not None:
                out[cur].append(line)
    return out


def _fields(lines: List[str], count: int, width: int, per_line: int) -> List[str]:
    vals: List[str] = []
    for line in lines:
        for c in range(per_line):
            if len(vals) >= count:
                return vals
            vals.append(line[c * width:(c + 1) * width])
        if len(vals) >= count:
            break
    return vals


def _read_array(sec: List[str], count: int, kind: str):
    m = _FMT.search(sec[0])
    if not m:
        raise ValueError(f"bad %FORMAT line: {sec[0]!r}")
    per_line, width = int(m.group(1)), int(m.group(3))
    raw = _fields(sec[1:], count, width, per_line)
    if len(raw) < count:
        raise ValueError(f"section truncated: wanted {count} values, found {len(raw)}")
    if kind == "i":
        return np.array([int(s) for s in raw], np.int64)
    return np.array([float(s) for s in raw], np.float64)


def read_prmtop(path: str, dielectric: float = 1.0) -> dict:
    """Return the topology part of a System as a dict of arrays."""
    s = _sections(path)
    if "POINTERS" not in s:
        raise ValueError(f"{path}: no %FLAG POINTERS")
    ptr = _read_array(s["POINTERS"], 18, "i")
    n, ntypes = int(ptr[0]), int(ptr[1])
    nbonh, nbona, nthetah, ntheta, nphih, nphia = (int(x) for x in ptr[2:8])
    nnb, nbt, nat, ndt = int(ptr[10]), int(ptr[15]), int(ptr[16]), int(ptr[17])
    sqrt_eps = np.float32(math.sqrt(dielectric))

    q = _read_array(s["CHARGE"], n, "f").astype(np.float32)
    xyzq = np.zeros((n, 4), np.float32)
    xyzq[:, 3] = q / sqrt_eps
    mass = _read_array(s["MASS"], n, "f").astype(np.float32)
    vel = np.zeros((n, 4), np.float32)
    vel[:, 3] = np.where(mass < 2.0, np.float32(12.0), mass)
    ityp = (_read_array(s["ATOM_TYPE_INDEX"], n, "i") - 1).astype(np.int32)
    nex = _read_array(s["NUMBER_EXCLUDED_ATOMS"], n, "i")
    nbparm = (_read_array(s["NONBONDED_PARM_INDEX"], ntypes * ntypes, "i") - 1).astype(np.int32)
    ntri = ntypes * (ntypes + 1) // 2
    lj_a = _read_array(s["LENNARD_JONES_ACOEF"], ntri, "f").astype(np.float32)
    lj_b = _read_array(s["LENNARD_JONES_BCOEF"], ntri, "f").astype(np.float32)

    bond_params = np.zeros((nbt, 2), np.float32)
    bond_params[:, 0] = _read_array(s["BOND_FORCE_CONSTANT"], nbt, "f").astype(np.float32) * np.float32(2.0)
    bond_params[:, 1] = _read_array(s["BOND_EQUIL_VALUE"], nbt, "f")
    angle_params = np.zeros((nat, 2), np.float32)
    angle_params[:, 0] = _read_array(s["ANGLE_FORCE_CONSTANT"], nat, "f").astype(np.float32) * np.float32(2.0)
    angle_params[:, 1] = _read_array(s["ANGLE_EQUIL_VALUE"], nat, "f")
    dih_params = np.zeros((ndt, 4), np.float32)
    dih_params[:, 0] = _read_array(s["DIHEDRAL_PERIODICITY"], ndt, "f")
    dih_params[:, 1] = _read_array(s["DIHEDRAL_FORCE_CONSTANT"], ndt, "f")
    dih_params[:, 2] = _read_array(s["DIHEDRAL_PHASE"], ndt, "f")
    sc14 = np.zeros((ndt, 2), np.float32)
    sc14[:, 0] = _read_array(s["SCEE_SCALE_FACTOR"], ndt, "f")
    sc14[:, 1] = _read_array(s["SCNB_SCALE_FACTOR"], ndt, "f")

    def c_div3(v):  # C integer division (truncation toward zero)
        return np.where(v < 0, -((-v) // 3), v // 3)

    def bonds(flag, cnt):
        raw = _read_array(s[flag], 3 * cnt, "i").reshape(cnt, 3) if cnt else np.zeros((0, 3), np.int64)
        out = np.zeros((cnt, 4), np.int32)
        out[:, 0] = c_div3(raw[:, 0]); out[:, 1] = c_div3(raw[:, 1]); out[:, 2] = raw[:, 2] - 1
        return out

    def angles(flag, cnt):
        raw = _read_array(s[flag], 4 * cnt, "i").reshape(cnt, 4) if cnt else np.zeros((0, 4), np.int64)
        out = np.zeros((cnt, 4), np.int32)
        out[:, :3] = c_div3(raw[:, :3]); out[:, 3] = raw[:, 3] - 1
        return out

    def dihs(flag, cnt):
        raw = _read_array(s[flag], 5 * cnt, "i").reshape(cnt, 5) if cnt else np.zeros((0, 5), np.int64)
        return c_div3(raw[:, :4]).astype(np.int32), (raw[:, 4] - 1).astype(np.int32)

    bond_atoms = np.concatenate([bonds("BONDS_INC_HYDROGEN", nbonh), bonds("BONDS_WITHOUT_HYDROGEN", nbona)])
    angle_atoms = np.concatenate([angles("ANGLES_INC_HYDROGEN", nthetah), angles("ANGLES_WITHOUT_HYDROGEN", ntheta)])
    d1, t1 = dihs("DIHEDRALS_INC_HYDROGEN", nphih)
    d2, t2 = dihs("DIHEDRALS_WITHOUT_HYDROGEN", nphia)
    excl = (_read_array(s["EXCLUDED_ATOMS_LIST"], nnb, "i") - 1).astype(np.int32)
    return dict(
        xyzq=xyzq, vel=vel, true_mass=mass, ityp=ityp, n_lj_types=ntypes, nbparm=nbparm,
        lj_a=lj_a, lj_b=lj_b, nexcl_prefix=np.cumsum(nex).astype(np.int32), excl_list=excl,
        bond_atoms=bond_atoms, bond_params=bond_params, angle_atoms=angle_atoms,
        angle_params=angle_params, dih_atoms=np.concatenate([d1, d2]),
        dih_types=np.concatenate([t1, t2]), dih_params=dih_params, sc14=sc14,
    )


def read_isspatop(path: str) -> IsspaTables:
    """src/isspa_class.cpp:115-248."""
    s = _sections(path)
    ptr = s["POINTERS"][1]
    n = int(ptr[0:8])
    nt = int(ptr[8:16])
    n_g = int(ptr[16:24])
    n_e = int(ptr[24:32])
    n_f = int(ptr[32:40])
    it = (_read_array(s["ISSPA_TYPE_INDEX"], n, "i") - 1).astype(np.int32)
    rmax = _read_array(s["ISSPA_RMAX"], nt, "f").astype(np.float32)

    def table(flag, rows):
        r = np.zeros(rows, np.float32)
        tab = np.zeros((nt, rows), np.float32)
        for b, line in enumerate(s[flag][1:1 + rows]):
            r[b] = float(line[0:16])
            for t in range(nt):
                tab[t, b] = float(line[(t + 1) * 16:(t + 2) * 16])
        grid = (np.float32(r[0]), np.float32(r[1] - r[0]))
        return tab, grid

    g, gg = table("ISSPA_DENSITIES", n_g)
    e, eg = table("ISSPA_EFIELD", n_e)
    f, fg = table("ISSPA_FORCES", n_f)
    return IsspaTables(isspa_type=it, rmax=rmax, g=g, e=e, f=f, g_grid=gg, e_grid=eg, f_grid=fg)


def read_rst7(path: str, n_atoms: int) -> np.ndarray:
    """Positions (or velocities) from an AMBER restart, float32[N,3]
    (src/atom_class.cpp:76-115)."""
    with open(path, "r") as fh:
        lines = fh.readlines()
    vals = []
    for line in lines[2:]:
        for c in range(6):
            if len(vals) >= 3 * n_atoms:
                break
            vals.append(float(line[c * 12:(c + 1) * 12]))
        if len(vals) >= 3 * n_atoms:
            break
    if len(vals) < 3 * n_atoms:
        raise ValueError(f"{path}: expected {3 * n_atoms} values, found {len(vals)}")
    return np.array(vals, np.float32).reshape(n_atoms, 3)


def write_rst7(path: str, xyz: np.ndarray, lbox: float, title: str = "Restart from isspa_amd") -> None:
    """src/atom_class.cpp:292-327 layout (%12.7f, 6 per line, then a box line)."""
    n = xyz.shape[0]
    with open(path, "w") as fh:
        fh.write(title + "\n")
        fh.write(f"{n}\n")
        for i in range(0, n, 2):
            fh.write("%12.7f%12.7f%12.7f" % tuple(xyz[i, :3]))
            if i + 1 < n:
                fh.write("%12.7f%12.7f%12.7f\n" % tuple(xyz[i + 1, :3]))
            else:
                fh.write("\n")
        fh.write("%12.7f%12.7f%12.7f%12.7f%12.7f%12.7f\n" % (lbox, lbox, lbox, 90.0, 90.0, 90.0))


def read_config(path: str) -> dict:
    """`key = value` run deck, src/config_class.cpp:14-133.  Keys are matched
    by prefix in the reference's order, which this reproduces."""
    keys = ["prmtop", "isspaPrmtop", "posFile", "posRstFile", "velFile", "velRstFile", "forFile",
            "inputCoord", "nMC", "temperature", "seed", "cutoff", "dielectric", "nSteps",
            "deltaWrite", "boxLength", "pnu", "velRst", "inputVel", "US", "usCfgFile"]
    ints = {"nMC", "nSteps", "deltaWrite", "velRst", "US"}
    floats = {"temperature", "cutoff", "dielectric", "boxLength", "pnu"}
    cfg: dict = {}
    with open(path) as fh:
        for line in fh:
            line = line.split("#", 1)[0]
            if "=" not in line:
                continue
            tok, val = line.split("=", 1)
            tok, val = tok.strip(), val.strip()
            for k in keys:
                if tok.startswith(k):
                    if k in ints:
                        cfg[k] = int(val)
                    elif k == "seed":
                        cfg[k] = int(float(val))
                    elif k in floats:
                        cfg[k] = float(val)
                    else:
                        cfg[k] = val
                    break
    return cfg


def read_us_config(path: str) -> dict:
    """src/us_class.cpp:13-104 (atom lists are 1-based in the file)."""
    out: dict = {}
    with open(path) as fh:
        for line in fh:
            line = line.split("#", 1)[0]
            if "=" not in line:
                continue
            tok, val = (x.strip() for x in line.split("=", 1))
            if tok.startswith("k"):
                out["k"] = float(val)
            elif tok.startswith("x0"):
                out["x0"] = float(val)
            elif tok.startswith("cvFile"):
                out["cvFile"] = val
            elif tok.startswith("atomSelect1"):
                out["group1"] = [int(a) - 1 for a in val.split(",") if a.strip()]
            elif tok.startswith("atomSelect2"):
                out["group2"] = [int(a) - 1 for a in val.split(",") if a.strip()]
    return out


def load_system(prmtop: str, isspatop: str, rst7: str, *, lbox: float = 200.0, n_mc: int = 20,
                dielectric: float = 1.0, temperature_k: float = 298.0, pnu: float = 0.0,
                seed: int = 12345, name: str = "system") -> System:
    topo = read_prmtop(prmtop, dielectric)
    tabs = read_isspatop(isspatop)
    n = topo["xyzq"].shape[0]
    if tabs.isspa_type.shape[0] != n:
        raise ValueError(f"atom count mismatch: prmtop {n}, isspatop {tabs.isspa_type.shape[0]}")
    topo["xyzq"][:, :3] = read_rst7(rst7, n)
    sysm = System(isspa=tabs, lbox=lbox, n_mc=n_mc, temperature_k=temperature_k, pnu=pnu,
                  seed=seed, name=name, **topo)
    sysm.check()
    return sysm


def load_us(us_cfg: str, system: System) -> UmbrellaRestraint:
    c = read_us_config(us_cfg)
    return UmbrellaRestraint.from_groups(c["group1"], c["group2"], system.true_mass, c["k"], c["x0"])

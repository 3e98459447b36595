"""IS-SPA table generator and .isspatop writer.

Counterpart of the reference's scripts/isspa_prmtop/write_isspa_prmtop.py.
The writer produces the file that src/isspa_class.cpp:45-178 reads.

Tables per type t on a uniform grid r_b = r0 + b*dr:

  ISSPA_DENSITIES  g_t(r): the parabola model
        g = 0                        r < r_lo
        g = g0 - alpha (r - x0)^2    r_lo <= r < r_hi
        g = 1                        r >= r_hi
    with r_hi = x0 + sqrt((g0 - 1)/alpha) and r_lo = max(x0 - sqrt(g0/alpha),
    core * sigma_ts).  The core floor keeps points out of the steep LJ core
    of small atoms, where x0 - w < 0 would put g > 0 at the nucleus.
  ISSPA_FORCES     F_t(r) = -dU/dr of the solute-solvent LJ pair
        U = 4 eps_ts ((sigma_ts/r)^12 - (sigma_ts/r)^6)
    using Lorentz-Berthelot mixing with the solvent sphere
    (sigma_ts = (sigma_t + sigma_s)/2, eps_ts = sqrt(eps_t eps_s)), for
    r >= r_lo, and 0 below.
  ISSPA_EFIELD     E_t(r) = c_t exp(-(r - x0)^2 / 2) for r >= r_lo, and 0
    below (the E-table policy: a smooth, zero-mean-free bump centred on the
    contact distance; c_t = 0 switches the term off).

Other table sources use the reference writer's fit-file readers below:
`read_density_fit` (g = exp(column 1) where the count column is positive),
`read_efield_fit` (column 1 as is) and `read_force_fit` (exp(column 1)).

One deliberate difference from the reference writer: it stores log(F) in the
ISSPA_FORCES section (write_isspa_prmtop.py:262-267).  The reader uses the
column as is (src/isspa_class.cpp:150-171), and the shipped ADI2 tables hold
raw forces, zeros included.  This writer therefore stores F itself.
"""
from __future__ import annotations

import math
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from .system import IsspaTables

SOLVENT_SIGMA = 3.15   # A, the LJ solvent sphere
SOLVENT_EPS = 0.152    # kcal/mol


def lj_solvent_mixing(sigma: float, eps: float, solvent_sigma: float = SOLVENT_SIGMA,
                      solvent_eps: float = SOLVENT_EPS) -> Tuple[float, float]:
    """Lorentz-Berthelot mixing of a solute type with the solvent sphere."""
    return 0.5 * (sigma + solvent_sigma), math.sqrt(eps * solvent_eps)


def parabola_domain(x0: float, g0: float, alpha: float, sigma_ts: float,
                    core: float = 0.85) -> Tuple[float, float]:
    """(r_lo, r_hi) of the parabola model."""
    w = math.sqrt(g0 / alpha)
    r_hi = x0 + math.sqrt((g0 - 1.0) / alpha)
    return max(x0 - w, core * sigma_ts), r_hi


def parabola_tables(sigma: Sequence[float], eps: Sequence[float], x0: Sequence[float],
                    g0: Sequence[float], alpha: Sequence[float],
                    efield_amp: Optional[Sequence[float]] = None, r0: float = 0.05,
                    dr: float = 0.1, rows: int = 120, core: float = 0.85,
                    solvent_sigma: float = SOLVENT_SIGMA, solvent_eps: float = SOLVENT_EPS):
    """G, E, F float32[T, rows] and the per-type parameters used."""
    T = len(sigma)
    r = (r0 + dr * np.arange(rows)).astype(np.float64)
    G = np.zeros((T, rows), np.float32)
    E = np.zeros((T, rows), np.float32)
    F = np.zeros((T, rows), np.float32)
    amp = np.zeros(T) if efield_amp is None else np.asarray(efield_amp, np.float64)
    par: Dict[int, dict] = {}
    for t in range(T):
        s_ts, e_ts = lj_solvent_mixing(sigma[t], eps[t], solvent_sigma, solvent_eps)
        r_lo, r_hi = parabola_domain(x0[t], g0[t], alpha[t], s_ts, core)
        for b, rb in enumerate(r):
            if rb < r_lo:
                continue
            G[t, b] = g0[t] - alpha[t] * (rb - x0[t]) ** 2 if rb < r_hi else 1.0
            sr6 = (s_ts / rb) ** 6
            F[t, b] = 24.0 * e_ts / rb * (2.0 * sr6 * sr6 - sr6)
            E[t, b] = amp[t] * math.exp(-0.5 * (rb - x0[t]) ** 2)
        par[t] = dict(x0=x0[t], g0=g0[t], alpha=alpha[t], w=math.sqrt(g0[t] / alpha[t]),
                      r_lo=r_lo, r_hi=r_hi, sigma_ts=s_ts, eps_ts=e_ts, c=float(amp[t]),
                      sigma=float(sigma[t]), eps=float(eps[t]),
                      lj_A=4 * e_ts * s_ts ** 12, lj_B=4 * e_ts * s_ts ** 6)
    return G, E, F, par


def isspa_domain(r: np.ndarray, G: np.ndarray, F: np.ndarray, thresh: float = 0.1):
    """(mcmin, rmax) per type: the first and last grid radius where
    g |F| > thresh, the reference writer's rule (write_isspa_prmtop.py:190-200).
    G and F are [T, rows]."""
    n = min(G.shape[1], F.shape[1])
    fg = G[:, :n].astype(np.float64) * np.abs(F[:, :n].astype(np.float64))
    out = np.zeros((G.shape[0], 2))
    for t in range(G.shape[0]):
        idx = np.nonzero(fg[t] > thresh)[0]
        if idx.size == 0:
            raise ValueError(f"type {t}: g|F| never exceeds {thresh}")
        out[t] = (r[idx.min()], r[idx.max()])
    return out


def _fit_rows(path: str, type_col: int):
    data = np.loadtxt(path)
    n_types = int(data[-1, type_col])
    rows = data.shape[0] // n_types
    return data, n_types, rows


def read_density_fit(path: str):
    """(r, g[T, rows]) from a regression file with columns r, log g, type,
    count: g = exp(log g) where count > 0, else 0
    (write_isspa_prmtop.py:133-149)."""
    data, T, rows = _fit_rows(path, 2)
    g = np.where(data[:, 3] > 0, np.exp(data[:, 1]), 0.0).reshape(T, rows)
    return data[:rows, 0].copy(), g


def read_efield_fit(path: str):
    """(r, E[T, rows]), column 1 as is (write_isspa_prmtop.py:151-164)."""
    data, T, rows = _fit_rows(path, 2)
    return data[:rows, 0].copy(), data[:, 1].reshape(T, rows).copy()


def read_force_fit(path: str):
    """(r, F[T, rows]) = exp(column 1), type in column 3
    (write_isspa_prmtop.py:166-181)."""
    data, T, rows = _fit_rows(path, 3)
    return data[:rows, 0].copy(), np.exp(data[:, 1]).reshape(T, rows)


def _write_ints(fh, vals):
    for k, v in enumerate(vals):
        if k and k % 10 == 0:
            fh.write("\n")
        fh.write("%8d" % int(v))
    fh.write("\n")


def _write_reals5(fh, vals):
    for k, v in enumerate(vals):
        if k and k % 5 == 0:
            fh.write("\n")
        fh.write("%16.8e" % float(v))
    fh.write("\n")


def write_isspatop(path: str, tables: IsspaTables, mcmin: Optional[Sequence[float]] = None,
                   title: str = "PDI") -> None:
    """Write an .isspatop file.  isspa_type is 0-based here and written
    1-based.  POINTERS holds nAtoms, nTypes, nGRs, nERs, nFRs in the
    writer's column order (write_isspa_prmtop.py:212-218), the one the reader
    slices at columns 0/8/16/24/32 (src/isspa_class.cpp:80-88)."""
    T = tables.n_types
    n = int(tables.isspa_type.shape[0])
    mcmin = np.zeros(T) if mcmin is None else np.asarray(mcmin, np.float64)
    with open(path, "w") as fh:
        fh.write("%VERSION \n%FLAG TITLE\n%FORMAT(20a4)\n")
        fh.write(title + "\n")
        fh.write("%FLAG POINTERS\n%FORMAT(10I8)\n")
        fh.write("%8d%8d%8d%8d%8d\n" % (n, T, tables.g.shape[1], tables.e.shape[1],
                                        tables.f.shape[1]))
        fh.write("%FLAG ISSPA_TYPE_INDEX\n%FORMAT(10I8)\n")
        _write_ints(fh, np.asarray(tables.isspa_type) + 1)
        fh.write("%FLAG ISSPA_MCMIN\n%FORMAT(5E16.8)\n")
        _write_reals5(fh, mcmin)
        fh.write("%FLAG ISSPA_RMAX\n%FORMAT(5E16.8)\n")
        _write_reals5(fh, tables.rmax)
        for flag, tab, (r0, dr) in (("ISSPA_DENSITIES", tables.g, tables.g_grid),
                                    ("ISSPA_EFIELD", tables.e, tables.e_grid),
                                    ("ISSPA_FORCES", tables.f, tables.f_grid)):
            fh.write(f"%FLAG {flag}\n%FORMAT({T + 1}E16.8)\n")
            for b in range(tab.shape[1]):
                fh.write("%16.8e" % (float(r0) + float(dr) * b))
                for t in range(T):
                    fh.write("%16.8e" % float(tab[t, b]))
                fh.write("\n")

"""Helpers for comparing with the reference's own output files (test
infrastructure): frame readers and the variance model of the reference's
force-kernel point index."""
from __future__ import annotations

import math

import numpy as np


def read_xyz_frames(path, n_atoms):
    """Frames of the reference's xyz writer (src/atom_class.cpp:248-256):
    two count lines, then one 'C x y z' line per atom."""
    lines = open(path).read().split("\n")
    out, i = [], 0
    while i + 2 + n_atoms <= len(lines) and lines[i].strip():
        out.append(np.array([[float(v) for v in l.split()[1:4]]
                             for l in lines[i + 2:i + 2 + n_atoms]]))
        i += 2 + n_atoms
    return out


def float_index_weights(n_atoms, n_mc):
    """How often the reference's IS-SPA force kernel uses each MC point for
    each receiving atom (src/isspa_force_cuda.cu:208-239, 376-385).

    The launch has forceThreads = 32*ceil(N*nMC/32) threads per receiver;
    thread `index` takes receiver atom = index / forceThreads and point
    MC = int(index - atom*forceThreads) evaluated in float, so once
    index >= 2^24 the index rounds to the float grid and some points are used
    two or three times and others not at all (quirk Q2; this build fixes it).
    Returns w[N, P] as a generator of (atom, weights) for the atoms whose
    weights are not all one, and the per-atom variance factor
    sum_m w_m^2 / P (1 where the arithmetic is exact)."""
    P = n_atoms * n_mc
    T = 32 * math.ceil(P / 32)
    fac = np.ones(n_atoms)
    for a in range(n_atoms):
        if (a + 1) * T <= 2 ** 24:
            continue
        i = a * T + np.arange(T, dtype=np.int64)
        fi = i.astype(np.float32)                       # unsigned -> float, nearest even
        prod = np.float32(a) * np.float32(T)            # int * float in float
        mc = (fi - prod).astype(np.float32).astype(np.int64)  # float subtraction, int()
        ok = (mc >= 0) & (mc < P)
        w = np.bincount(mc[ok], minlength=P)[:P]
        fac[a] = float((w.astype(np.float64) ** 2).sum() / P)
    return fac

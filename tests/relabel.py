"""Atom relabelling for tests (test infrastructure): a system with its atoms
in another order, optionally with inert filler atoms between them, so that
excluded pairs end up far apart in index (the nonbond kernels keep
exclusions within +-64 indices in per-atom bit windows and fall back to a
per-atom list beyond that)."""
from __future__ import annotations

import copy

import numpy as np


def exclusion_pairs(s):
    """{(i, j), i < j} of the reference's per-atom lists: the entries of atom
    i's segment of excl_list above i (src/nonbond_force_cuda.cu:71-108; -1
    marks an atom with none)"""
    pairs = set()
    lo = 0
    for i in range(s.n_atoms):
        hi = int(s.nexcl_prefix[i])
        for j in s.excl_list[lo:hi]:
            if int(j) > i:
                pairs.add((i, int(j)))
        lo = hi
    return pairs


def _lists(n, pairs):
    """AMBER-style lists: for atom i the partners above i, ascending, or a
    single -1; returned as (inclusive prefix sum, list)"""
    part = [[] for _ in range(n)]
    for i, j in pairs:
        a, b = min(i, j), max(i, j)
        part[a].append(b)
    pre, lst = [], []
    for i in range(n):
        lst.extend(sorted(part[i]) if part[i] else [-1])
        pre.append(len(lst))
    return np.array(pre, np.int32), np.array(lst, np.int32)


def max_exclusion_span(s):
    return max((j - i for i, j in exclusion_pairs(s)), default=0)


def relabel(s, new_of_old, n_new, filler_xyz=None):
    """The nonbond/IS-SPA view of system s with old atom k at index
    new_of_old[k] of an n_new-atom system.  Indices not taken are filler
    atoms at filler_xyz (one row each, in index order): zero charge, an LJ
    type whose A and B coefficients are 0 with every type, IS-SPA type 0.
    Bonded terms are dropped (the tests using this look at the nonbond and
    IS-SPA terms)."""
    new_of_old = np.asarray(new_of_old, np.int64)
    n = s.n_atoms
    assert len(new_of_old) == n and len(set(new_of_old.tolist())) == n
    t = copy.deepcopy(s)
    free = np.setdiff1d(np.arange(n_new), new_of_old)
    T = s.n_lj_types
    xyzq = np.zeros((n_new, 4), np.float32)
    xyzq[new_of_old] = s.xyzq
    vel = np.zeros((n_new, 4), np.float32)
    vel[new_of_old] = s.vel
    vel[free, 3] = 12.0
    mass = np.full(n_new, 12.0, np.float32)
    mass[new_of_old] = s.true_mass
    ityp = np.full(n_new, T, np.int32)
    ityp[new_of_old] = s.ityp
    itp = np.zeros(n_new, np.int32)
    itp[new_of_old] = s.isspa.isspa_type
    if len(free):
        assert filler_xyz is not None and len(filler_xyz) == len(free)
        xyzq[free, :3] = filler_xyz
        # one more LJ type, zero coefficients against every type
        nb = np.full((T + 1, T + 1), len(s.lj_a), np.int32)
        nb[:T, :T] = np.asarray(s.nbparm, np.int32).reshape(T, T)
        t.nbparm = nb.ravel()
        t.lj_a = np.concatenate([s.lj_a, [0.0]]).astype(s.lj_a.dtype)
        t.lj_b = np.concatenate([s.lj_b, [0.0]]).astype(s.lj_b.dtype)
        t.n_lj_types = T + 1
    pairs = {(int(new_of_old[i]), int(new_of_old[j])) for i, j in exclusion_pairs(s)}
    t.nexcl_prefix, t.excl_list = _lists(n_new, pairs)
    t.xyzq, t.vel, t.true_mass, t.ityp = xyzq, vel, mass, ityp
    t.isspa = copy.deepcopy(s.isspa)
    t.isspa.isspa_type = itp
    t.bond_atoms = np.zeros((0, 4), np.int32)
    t.bond_params = np.zeros((0, 2), np.float32)
    t.angle_atoms = np.zeros((0, 4), np.int32)
    t.angle_params = np.zeros((0, 2), np.float32)
    t.dih_atoms = np.zeros((0, 4), np.int32)
    t.dih_types = np.zeros(0, np.int32)
    t.dih_params = np.zeros((0, 4), np.float32)
    t.sc14 = np.zeros((0, 2), np.float32)
    t.us = None
    t.name = s.name + "_relabelled"
    t.check()
    return t


def spread_with_fillers(s, stride=3, offset=(60.0, 0.0, 0.0), spacing=4.0):
    """Old atom k at index stride*k; the stride-1 filler atoms after each sit
    on a cubic grid of the given spacing, shifted by `offset` from the
    system's lowest corner, well away from every atom"""
    n = s.n_atoms
    n_new = stride * n
    nf = n_new - n
    side = int(np.ceil(nf ** (1.0 / 3.0)))
    g = np.stack(np.meshgrid(*(np.arange(side),) * 3, indexing="ij"), -1).reshape(-1, 3)[:nf]
    lo = s.xyzq[:, :3].min(0)
    fill = (lo + np.asarray(offset) + spacing * g).astype(np.float32)
    return relabel(s, stride * np.arange(n), n_new, fill)

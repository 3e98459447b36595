"""Relabelled systems for the nonbond exclusion paths, checked on the CPU:
the oracle's nonbond forces of a relabelled system are those of the
original, moved with the atoms (FP64, exact up to summation order), and the
relabelling puts excluded pairs more than 64 indices apart."""
import numpy as np

import pyoracle as po
from relabel import exclusion_pairs, max_exclusion_span, relabel, spread_with_fillers


def test_spread_adi2_keeps_oracle_forces(adi2):
    t = spread_with_fillers(adi2, stride=8)
    assert t.n_atoms == 352 and max_exclusion_span(t) == 80
    assert len(exclusion_pairs(t)) == len(exclusion_pairs(adi2))
    a, b = po.nonbond(adi2), po.nonbond(t)
    for k in ("lj", "coul", "pair"):
        assert np.abs(b[k][0::8] - a[k]).max() <= 1e-12 * np.abs(a[k]).max()
        filler = np.ones(t.n_atoms, bool)
        filler[0::8] = False
        assert np.all(b[k][filler] == 0.0)


def test_permuted_lattice_keeps_oracle_forces(cfg1):
    perm = np.random.default_rng(3).permutation(cfg1.n_atoms)
    t = relabel(cfg1, perm, cfg1.n_atoms)
    assert max_exclusion_span(t) > 64
    a, b = po.nonbond(cfg1), po.nonbond(t)
    for k in ("lj", "coul", "pair"):
        assert np.abs(b[k][perm] - a[k]).max() <= 1e-11 * np.abs(a[k]).max()

"""Batched umbrella windows on the CPU: the window merge and the C ABI's
argument checks for the batch entry points (no device call is made)."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
from isspa_amd import batch


@pytest.fixture(scope="module")
def windows():
    return batch.umbrella_windows()


@pytest.fixture(scope="module")
def merged(windows):
    return batch.merge_windows(windows)


def test_workload_shape(windows, merged):
    kinds = [w.name.split("_")[1] for w in windows]
    assert len(windows) == 64
    assert kinds.count("ljpair") == 32 and kinds.count("chain") == 24 and kinds.count("big") == 8
    assert merged.n_atoms == 32 * 2 + 24 * 22 + 8 * 220
    assert list(merged.windows) == [w.n_atoms for w in windows]
    # restraint centres evenly spaced over 3..15 A in window order, k = 10
    x0 = [r.x0 for r in merged.restraints]
    assert np.allclose(x0, np.linspace(3.0, 15.0, 64))
    assert all(r.k == 10.0 for r in merged.restraints)
    # the chain and the 10-chain systems share one parameter set: 6 + 1 LJ types
    assert merged.n_lj_types == 7 and merged.isspa.n_types == 7


def test_merge_keeps_every_index_inside_its_window(merged):
    sl = batch.window_slices(merged)
    win = np.repeat(np.arange(len(sl)), merged.windows)
    for r, (a, b) in zip(merged.restraints, sl):
        assert np.all((r.atom_list[:, 0] >= a) & (r.atom_list[:, 0] < b))
    for arr, cols in ((merged.bond_atoms, 2), (merged.angle_atoms, 3), (merged.dih_atoms, 4)):
        at = np.abs(arr[:, :cols])
        assert np.all(win[at] == win[at[:, :1]])
    pre = merged.nexcl_prefix
    for i in range(merged.n_atoms):
        lo = pre[i - 1] if i else 0
        ex = merged.excl_list[lo:pre[i]]
        ex = ex[ex >= 0]
        assert np.all(win[ex] == win[i])


def test_merge_preserves_pair_parameters(windows, merged):
    """Every same-window LJ pair reads the window's own A, B; every atom keeps
    its IS-SPA table rows."""
    sl = batch.window_slices(merged)
    T = merged.n_lj_types
    nbp = merged.nbparm.reshape(T, T)
    for w, (a, b) in zip(windows, sl):
        Tw = w.n_lj_types
        wnb = w.nbparm.reshape(Tw, Tw)
        ti, tj = np.meshgrid(w.ityp, w.ityp, indexing="ij")
        mi, mj = np.meshgrid(merged.ityp[a:b], merged.ityp[a:b], indexing="ij")
        assert np.array_equal(w.lj_a[wnb[ti, tj]], merged.lj_a[nbp[mi, mj]])
        assert np.array_equal(w.lj_b[wnb[ti, tj]], merged.lj_b[nbp[mi, mj]])
        mt = merged.isspa.isspa_type[a:b]
        assert np.array_equal(merged.isspa.g[mt], w.isspa.g[w.isspa.isspa_type])
        assert np.array_equal(merged.isspa.rmax[mt], w.isspa.rmax[w.isspa.isspa_type])


def test_merged_bonded_oracle_is_the_windows_oracle(windows, merged):
    """Bonded terms never cross windows: the FP64 oracle over the merged
    system equals the concatenation of the per-window oracles."""
    ref = po.bonded(merged)
    parts = [po.bonded(w)["total"] for w in windows]
    assert np.allclose(ref["total"], np.concatenate(parts), rtol=0, atol=1e-12)
    for key in ("e_bond", "e_angle", "e_dih", "e_14coul", "e_14lj"):
        tot = sum(po.bonded(w)[key] for w in windows)
        assert abs(ref[key] - tot) <= 1e-9 * max(1.0, abs(tot)), key


def test_batch_abi_rejects_bad_windows_without_gpu(merged):
    from isspa_amd import capi
    from isspa_amd.engine import _np
    d = capi.ModelDesc()
    it = merged.isspa
    d.n_atoms, d.n_types, d.n_mc = merged.n_atoms, it.n_types, 20
    d.n_g, d.n_e, d.n_f = it.g.shape[1], it.e.shape[1], it.f.shape[1]
    d.g_r0, d.g_dr = it.g_grid
    d.e_r0, d.e_dr = it.e_grid
    d.f_r0, d.f_dr = it.f_grid
    d.lbox = 200.0
    arrs = [_np(it.isspa_type, np.int32), _np(it.rmax, np.float32), _np(it.g, np.float32),
            _np(it.e, np.float32), _np(it.f, np.float32)]
    bad = _np(merged.windows, np.int32).copy()
    bad[0] += 1  # windows no longer sum to n_atoms
    win = capi.Windows(len(bad), bad.ctypes.data, None, None)
    h = C.c_void_p()
    rc = capi.lib.isspa_model_create_batch(C.byref(d), C.byref(win),
                                           *(a.ctypes.data for a in arrs), C.byref(h))
    assert rc == 1 and b"sum to" in capi.lib.isspa_last_error()
    # an exclusion that crosses windows
    two = np.array([1, 1], np.int32)
    win2 = capi.Windows(2, two.ctypes.data, None, None)
    ity = np.zeros(2, np.int32)
    nbparm = np.zeros(1, np.int32)
    ab = np.ones(1, np.float32)
    pre = np.array([1, 1], np.int32)
    lst = np.array([1], np.int32)
    rmax = np.full(1, 12.0, np.float32)
    rc = capi.lib.nonbond_model_create_batch(2, 1, ity.ctypes.data, nbparm.ctypes.data,
                                             ab.ctypes.data, ab.ctypes.data, pre.ctypes.data,
                                             lst.ctypes.data, 1, ity.ctypes.data, 1,
                                             rmax.ctypes.data, 2.3473, 200.0, C.byref(win2),
                                             C.byref(h))
    assert rc == 1 and b"crosses windows" in capi.lib.isspa_last_error()
    # a restraint with an atom twice
    boff = np.array([0, 3], np.int32)
    al = np.array([[0, 0], [1, 1], [0, 1]], np.int32)
    m = np.ones(3, np.float32)
    k = np.ones(1, np.float32)
    rc = capi.lib.us_model_create_batch(2, 1, boff.ctypes.data, al.ctypes.data, m.ctypes.data,
                                        k.ctypes.data, k.ctypes.data, rmax.ctypes.data, C.byref(h))
    assert rc == 1 and b"twice" in capi.lib.isspa_last_error()
    # us_model_create keeps the reference's {-k, +k} convention
    kumb = np.array([-1.0, 2.0], np.float32)
    rc = capi.lib.us_model_create(2, 2, al.ctypes.data, m.ctypes.data, kumb.ctypes.data, 1.0,
                                  10.0, C.byref(h))
    assert rc == 5

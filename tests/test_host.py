"""Host-side checks without a GPU: readers, workload generator, C-ABI exports."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from isspa_amd import io, synth


# ------------------------------------------------------------- readers ------
def test_adi2_counts_match_reference_log(adi2):
    """Counts printed by the reference for this deck
    (timing_tests/ADI2_US/window_6.5/ADI2.run0.log:31-61)."""
    s = adi2
    assert s.n_atoms == 44 and s.n_lj_types == 7
    assert len(s.bond_atoms) == 24 + 18
    assert len(s.angle_atoms) == 50 + 22
    assert len(s.dih_atoms) == 86 + 48
    assert len(s.excl_list) == 198
    assert s.isspa.n_types == 16
    assert s.isspa.g.shape == (16, 120) and s.isspa.e.shape == (16, 120) and s.isspa.f.shape == (16, 120)
    assert len(s.bond_params) == 11 and len(s.angle_params) == 21 and len(s.dih_params) == 20
    # group masses of the umbrella log (:40-41): 144.175995 each, true masses
    assert abs(float(s.true_mass[:22].sum()) - 144.175995) < 1e-3


def test_prmtop_conventions(adi2):
    s = adi2
    # hydrogens move with mass 12 (src/read_prmtop.cpp:144-145), true masses kept
    light = s.true_mass < 2.0
    assert light.any() and np.all(s.vel[light, 3] == 12.0)
    assert np.all(s.vel[~light, 3] == s.true_mass[~light])
    # prefix-summed exclusions, 0-based indices, doubled force constants
    assert np.all(np.diff(s.nexcl_prefix) >= 0) and s.nexcl_prefix[-1] == len(s.excl_list)
    assert s.bond_atoms[:, :2].max() < 44 and s.bond_atoms[:, :2].min() >= 0
    sec = io._sections(os.path.join(GOLDEN, "adi2", "ADI2.prmtop"))
    k = io._read_array(sec["BOND_FORCE_CONSTANT"], 11, "f")
    assert np.allclose(s.bond_params[:, 0], 2 * k)
    # signed dihedral indices survive (negative third/fourth atom)
    assert (s.dih_atoms[:, 2] < 0).any()
    # charge scaled by 1/sqrt(dielectric)
    s4 = io.load_system(*(os.path.join(GOLDEN, "adi2", f) for f in ("ADI2.prmtop", "ADI2.isspatop", "ADI2.rst7")),
                        dielectric=4.0)
    assert np.allclose(s4.xyzq[:, 3], s.xyzq[:, 3] / 2.0)


def test_isspatop_tables(adi2):
    it = adi2.isspa
    assert np.allclose(it.rmax, 11.95)
    assert abs(float(it.g_grid[0]) - 0.05) < 1e-7 and abs(float(it.g_grid[1]) - 0.1) < 1e-6
    # first rows are zero (excluded volume), tail of g near 1
    assert np.all(it.g[:, 0] == 0) and np.all(np.abs(it.g[:, -1] - 1) < 0.05)
    assert it.isspa_type.min() == 0 and it.isspa_type.max() == 15
    # vtot (src/isspa_class.cpp:110-112)
    vt = it.vtot(100)
    assert abs(vt[0] - 4 / 3 * np.pi * 11.95 ** 3 * 0.0074 / 100) < 1e-5


def test_rst7_round_trip(tmp_path, adi2):
    p = tmp_path / "x.rst7"
    io.write_rst7(str(p), adi2.xyzq[:, :3], 200.0)
    back = io.read_rst7(str(p), adi2.n_atoms)
    assert np.abs(back - adi2.xyzq[:, :3]).max() < 1e-6


def test_reader_errors(tmp_path):
    bad = tmp_path / "bad.prmtop"
    bad.write_text("%FLAG TITLE\n%FORMAT(20a4)\nX\n")
    with pytest.raises(ValueError):
        io.read_prmtop(str(bad))
    short = tmp_path / "short.rst7"
    short.write_text("t\n44\n   1.0000000   2.0000000\n")
    with pytest.raises(ValueError):
        io.read_rst7(str(short), 44)


def test_config_reader():
    p = os.path.join(ROOT, "tests", "golden", "run_deck.txt")
    cfg = io.read_config(p)
    assert cfg["nMC"] == 100 and cfg["boxLength"] == 200.0 and cfg["velRst"] == 0
    assert cfg["prmtop"] == "ADI2.prmtop" and cfg["seed"] == 12345 and cfg["US"] == 0


# ----------------------------------------------------------- workloads ------
def test_chain_topology_counts():
    c = synth.make_chain()
    assert c.n_atoms == 22 and len(c.bond_atoms) == 21 and len(c.angle_atoms) == 36
    assert len(c.dih_atoms) == 41 and c.n_lj_types == 6
    assert abs(float(c.xyzq[:, 3].sum())) < 1e-3
    assert (c.dih_atoms[:, 2] < 0).sum() == 9
    d = np.linalg.norm(c.xyzq[:, None, :3] - c.xyzq[None, :, :3], axis=2)
    assert d[np.triu_indices(22, 1)].min() >= 1.0 - 1e-5


def test_workloads_deterministic_and_sized():
    a, b = synth.workload(1), synth.workload(1)
    assert np.array_equal(a.xyzq, b.xyzq) and np.array_equal(a.vel, b.vel)
    assert a.n_atoms == 1100 and a.n_mc == 20
    ext = a.xyzq[:, :3].max(0) - a.xyzq[:, :3].min(0)
    assert np.all(ext + 12.0 < a.lbox / 2)  # no minimum-image shift ever needed


def test_parabola_tables_closed_form():
    c = synth.make_chain()
    r = 0.05 + 0.1 * np.arange(120)
    for t, p in c.type_params.items():
        g = c.isspa.g[t]
        inside = (r >= p["r_lo"]) & (r < p["r_hi"])
        assert np.allclose(g[inside], p["g0"] - p["alpha"] * (r[inside] - p["x0"]) ** 2, atol=1e-6)
        assert np.all(g[r < p["r_lo"]] == 0) and np.all(g[r >= p["r_hi"]] == 1)
        # LJ-solvent force table is -dU/dr
        rr = r[inside][:3]
        h = 1e-6
        U = lambda x: 4 * p["eps_ts"] * ((p["sigma_ts"] / x) ** 12 - (p["sigma_ts"] / x) ** 6)  # noqa: E731
        assert np.allclose(c.isspa.f[t][inside][:3], -(U(rr + h) - U(rr - h)) / (2 * h), rtol=1e-4)


# ------------------------------------------------------------ C ABI ---------
def _declared_functions():
    txt = open(os.path.join(ROOT, "include", "isspa_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b([a-z_0-9]+)\s*\(", txt)) - {"if", "sizeof"})


def test_abi_library_exports_every_declared_symbol():
    lib = os.path.join(ROOT, "isspa-force-cuda_amd", "libisspa_amd.so")
    if not os.path.exists(lib):
        pytest.fail("libisspa_amd.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = [f for f in _declared_functions() if f not in exported]
    assert not missing, missing


def test_abi_python_binding_loads_and_covers_header():
    from isspa_amd import capi
    declared = set(_declared_functions())
    assert declared <= set(capi.EXPORTED), declared - set(capi.EXPORTED)
    c = capi.default_constants()
    e0 = 3.0 / (4.0 * np.pi * 0.0075 * 4.173705894203977) * (1.0 - 1.0 / 2.3473)  # src/constants.h:13
    assert abs(c["p0"] - 4.173705894203977) < 1e-6 and abs(c["e0"] - e0) < 1e-5 * e0


def test_abi_rejects_bad_arguments_without_gpu():
    """Argument validation happens before any device call."""
    import ctypes as C
    from isspa_amd import capi
    d = capi.ModelDesc()
    d.n_atoms, d.n_types, d.n_mc = 0, 1, 1
    h = C.c_void_p()
    z = np.zeros(4, np.float32)
    zi = np.zeros(4, np.int32)
    rc = capi.lib.isspa_model_create(C.byref(d), zi.ctypes.data, z.ctypes.data, z.ctypes.data,
                                     z.ctypes.data, z.ctypes.data, C.byref(h))
    assert rc == 1 and b"n_atoms" in capi.lib.isspa_last_error()
    rc = capi.lib.us_model_create(10, 1, zi.ctypes.data, z.ctypes.data, z.ctypes.data, 1.0, 10.0,
                                  C.byref(h))
    assert rc == 1  # one group empty


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus 2 exits with a message and status 2
    before touching a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""

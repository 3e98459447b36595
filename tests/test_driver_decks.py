"""The reference's own run decks through the host driver (isspa_md.x).

timing_tests/{ADI50,ADI25,LJ}/config.txt are run unchanged from a copy of
their input files (tests/golden/{adi50,adi25,lj2}): the counts the reader
prints, then (on the GPU) the whole run, with frame 0 of the force file --
the total force at the input coordinates, written before the first
integration step (src/total_force_cuda.cpp:132-150) -- against the oracle
with the same Philox points (seed and nMC of the deck).  Frames are printed
with 6 decimals, as the reference prints them.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from refout import read_xyz_frames

EXE = os.path.join(ROOT, "isspa-force-cuda_amd", "host", "isspa_md.x")

DECKS = {
    # name: (fixture dir, prmtop, isspatop, rst7, counts the reader prints)
    "ADI50": ("adi50", "ADI50.prmtop", "ADI50.isspatop", "ADI50.rst7",
              ("Number of atoms from prmtop file: 1100", "Length of excluded atoms list: 4950",
               "Number of ISSPA types: 16")),
    "ADI25": ("adi25", "ADI25.prmtop", "ADI25.isspatop", "ADI25.rst7",
              ("Number of atoms from prmtop file: 550", "Length of excluded atoms list: 2475",
               "Number of ISSPA types: 16")),
    "LJ": ("lj2", "LJ2.prmtop", "LJ2.isspatop", "LJ2.rst7",
           ("Number of atoms from prmtop file: 2", "Length of excluded atoms list: 2",
            "Number of ISSPA types: 2")),
}


def _deck_dir(tmp_path, name):
    d = os.path.join(GOLDEN, DECKS[name][0])
    for f in os.listdir(d):
        shutil.copy(os.path.join(d, f), tmp_path / f)
    return tmp_path / "config.txt"


def _deck_values(cfg):
    vals = {}
    for line in open(cfg):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    return vals


@pytest.mark.parametrize("name", sorted(DECKS))
def test_reference_deck_counts(tmp_path, name):
    cfg = _deck_dir(tmp_path, name)
    r = subprocess.run([EXE, str(cfg), "--check"], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    for want in DECKS[name][4]:
        assert want in r.stdout, (want, r.stdout)


@pytest.mark.gpu
@pytest.mark.parametrize("name,mode", [("ADI50", "eager"), ("ADI25", "eager"), ("ADI25", "graph"),
                                       ("LJ", "eager")])
def test_reference_deck_runs_and_frame0_matches_oracle(tmp_path, name, mode):
    import pyoracle as po
    from isspa_amd import io
    from isspa_amd.system import max_abs_rel_rms
    cfg = _deck_dir(tmp_path, name)
    v = _deck_values(cfg)
    print(f"case deck {name} ({mode}): nMC {v['nMC']}, nSteps {v['nSteps']}, "
          f"deltaWrite {v['deltaWrite']}", flush=True)
    args = [EXE, str(cfg)] + (["--graph"] if mode == "graph" else [])
    r = subprocess.run(args, capture_output=True, text=True, cwd=tmp_path, timeout=300)
    print(r.stdout[-600:])
    assert r.returncode == 0, r.stderr
    assert "Average ns/day" in r.stdout
    _, prm, itop, rst, _ = DECKS[name]
    s = io.load_system(str(tmp_path / prm), str(tmp_path / itop), str(tmp_path / rst),
                       lbox=float(v["boxLength"]), n_mc=int(v["nMC"]), seed=int(v["seed"]),
                       name=name)
    n_frames = (int(v["nSteps"]) + int(v["deltaWrite"]) - 1) // int(v["deltaWrite"])
    frc = read_xyz_frames(tmp_path / v["forFile"], s.n_atoms)
    pos = read_xyz_frames(tmp_path / v["posFile"], s.n_atoms)
    assert len(frc) == len(pos) == n_frames
    assert all(np.all(np.isfinite(f)) for f in frc + pos)
    assert os.path.exists(tmp_path / v["posRstFile"]) and os.path.exists(tmp_path / v["velRstFile"])
    ref = po.isspa_total(s, step=0)["total"] + po.nonbond(s)["total"]
    if len(s.bond_atoms) + len(s.angle_atoms) + len(s.dih_atoms):
        ref = ref + po.bonded(s)["total"]
    err = max_abs_rel_rms(frc[0], ref)
    print(f"{name} frame 0 force rel RMS vs oracle: {err:.2e}", flush=True)
    assert err < 2e-5


# ---- one umbrella window through the reference's equil -> run0 -> run1 chain
def _window_decks(root, x0="6.5", n_equil=40, n_run0=30, n_run1=20, delta=10, run1_seed=None):
    """timing_tests/ADI2_US as submit_US_equil.sh and submit_US_run.sh lay it
    out: topology files one level up, one directory per window, XXX -> the
    window centre.  Changed: the step counts and write interval (production
    lengths are 5e5 and 2.5e7 steps), and the equilibration's input, which in
    the reference is the previous window's equilibrated restart (not shipped):
    here ADI2.rst7."""
    src = os.path.join(GOLDEN, "adi2_us")
    root.mkdir(parents=True, exist_ok=True)
    for f in ("ADI2.prmtop", "ADI2.isspatop", "ADI2.rst7"):
        shutil.copy(os.path.join(GOLDEN, "adi2", f), root / f)
    w = root / f"window_{x0}"
    w.mkdir()

    def sub(name, out, steps=None, extra=None):
        txt = open(os.path.join(src, name)).read().replace("XXX", x0)
        lines = []
        for line in txt.split("\n"):
            key = line.split("=", 1)[0].strip() if "=" in line else ""
            if key == "nSteps" and steps is not None:
                line = f"nSteps = {steps}"
            elif key == "deltaWrite" and steps is not None:
                line = f"deltaWrite = {delta}"
            elif key == "inputCoord" and "YYY" in line:
                line = "inputCoord = ../ADI2.rst7"
            elif extra and key in extra:
                line = f"{key} = {extra[key]}"
            lines.append(line)
        (w / out).write_text("\n".join(lines))

    sub("ADI2.us", "us_params.dat")
    sub("ADI2.run0.us", "us_params.run0.dat")
    sub("ADI2.run1.us", "us_params.run1.dat")
    sub("equil_config.txt", "equil_config.dat", n_equil)
    sub("run0_config.txt", "run0_config.dat", n_run0)
    sub("run1_config.txt", "run1_config.dat", n_run1,
        {"seed": run1_seed} if run1_seed is not None else None)
    return w


def _run(w, deck, mode):
    args = [EXE, deck] + (["--graph"] if mode == "graph" else [])
    r = subprocess.run(args, capture_output=True, text=True, cwd=w, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["eager", "graph"])
def test_umbrella_window_chain_and_exact_restart(tmp_path, mode):
    """The window at x0 = 6.5 A: equilibration, production run0 from the
    equilibrated positions, then run1 resumed with velRst = 1 from run0's
    restart files (timing_tests/ADI2_US/run1_config.txt:2-6;
    src/atom_class.cpp:292-327 writes them).  With run1 given run0's seed,
    run0 (30 steps) + run1 (20 steps) must equal one 50-step run0 bit for
    bit: the restart keeps the exact float state beside the text file and the
    step count, and run1 continues the Philox counter at step 30."""
    print(f"case umbrella window chain ({mode})", flush=True)
    w = _window_decks(tmp_path / "chain", run1_seed="34669")
    _run(w, "equil_config.dat", mode)
    assert (w / "ADI2.equil.window.6.5.positions.rst").exists()
    _run(w, "run0_config.dat", mode)
    out = _run(w, "run1_config.dat", mode)
    assert "Resuming at step 30" in out
    cv0 = np.loadtxt(w / "ADI2.run0.window.6.5.dat")
    cv1 = np.loadtxt(w / "ADI2.run1.window.6.5.dat")
    assert cv0.shape == (3, 2) and cv1.shape == (2, 2)
    # the same window, run0 for 50 steps without a restart
    u = _window_decks(tmp_path / "whole", n_run0=50)
    _run(u, "equil_config.dat", mode)
    _run(u, "run0_config.dat", mode)
    whole = np.loadtxt(u / "ADI2.run0.window.6.5.dat")
    assert np.array_equal(whole[:, 1], np.concatenate([cv0[:, 1], cv1[:, 1]]))
    for kind in ("positions", "velocities"):
        a = (w / f"ADI2.run1.window.6.5.{kind}.rst").read_text()
        b = (u / f"ADI2.run0.window.6.5.{kind}.rst").read_text()
        assert a.split("\n", 1)[1] == b.split("\n", 1)[1], kind  # all but the title line
        assert a.split("\n", 1)[0].endswith("step 50") and b.split("\n", 1)[0].endswith("step 50")
        ea = (w / f"ADI2.run1.window.6.5.{kind}.rst.exact").read_bytes()
        eb = (u / f"ADI2.run0.window.6.5.{kind}.rst.exact").read_bytes()
        assert ea == eb, kind

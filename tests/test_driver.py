"""The C++ host layer and step driver (isspa_md.x) on the reference's ADI2 deck."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

EXE = os.path.join(ROOT, "isspa-force-cuda_amd", "host", "isspa_md.x")


def _deck(tmp_path, n_steps=20, delta=10, extra=""):
    d = os.path.join(GOLDEN, "adi2")
    txt = f"""velRst = 0
prmtop = {d}/ADI2.prmtop
inputCoord = {d}/ADI2.rst7
isspaPrmtop = {d}/ADI2.isspatop
forFile = frc.xyz
posFile = pos.xyz
velFile = vel.xyz
posRstFile = pos.rst
velRstFile = vel.rst
nMC = 100
nSteps = {n_steps}
deltaWrite = {delta}
temperature = 298.00
seed = 12345
cutoff = 12.0
dielectric = 1.0
boxLength = 200.0
pnu = 0.01
US = 0
{extra}"""
    p = tmp_path / "config.txt"
    p.write_text(txt)
    return p


def _frames(path, n):
    lines = open(path).read().split("\n")
    out = []
    i = 0
    while i + 2 + n <= len(lines) and lines[i].strip():
        out.append(np.array([[float(v) for v in l.split()[1:4]] for l in lines[i + 2:i + 2 + n]]))
        i += 2 + n
    return out


def test_driver_parses_reference_deck(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("isspa_md.x not built")
    cfg = _deck(tmp_path)
    r = subprocess.run([EXE, str(cfg), "--check"], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    for want in ("Number of atoms from prmtop file: 44", "Number of bonds: 42",
                 "Number of angles: 72", "Number of dihedrals: 134",
                 "Length of excluded atoms list: 198", "Number of ISSPA types: 16"):
        assert want in r.stdout, want


def test_driver_reports_missing_files(tmp_path):
    cfg = tmp_path / "c.txt"
    cfg.write_text("prmtop = nope.prmtop\nisspaPrmtop = nope.isspatop\nnMC = 10\n")
    r = subprocess.run([EXE, str(cfg), "--check"], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode != 0 and "nope.prmtop" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["eager", "graph"])
def test_driver_runs_deck_on_gpu(tmp_path, mode, adi2):
    import copy
    import pyoracle as po
    from isspa_amd.system import max_abs_rel_rms
    cfg = _deck(tmp_path)
    args = [EXE, str(cfg)] + (["--graph"] if mode == "graph" else [])
    r = subprocess.run(args, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    print(r.stdout[-800:])
    assert r.returncode == 0, r.stderr
    assert "Average ns/day" in r.stdout
    frc = _frames(tmp_path / "frc.xyz", 44)
    pos = _frames(tmp_path / "pos.xyz", 44)
    assert len(frc) == 2 and len(pos) == 2
    assert np.all(np.isfinite(frc[1])) and np.all(np.isfinite(pos[1]))
    # frame 0: forces at the input coordinates, MC points of step 0
    s = copy.deepcopy(adi2)
    s.n_mc = 100
    ref = (po.isspa_total(s, step=0, seed=12345)["total"] + po.nonbond(s)["total"] +
           po.bonded(s)["total"])
    err = max_abs_rel_rms(frc[0], ref)
    print(f"frame 0 force rel RMS vs oracle: {err:.2e} (frames printed with 6 decimals)")
    assert err < 2e-5
    assert os.path.exists(tmp_path / "pos.rst") and os.path.exists(tmp_path / "vel.rst")


@pytest.mark.gpu
def test_driver_umbrella_cv_in_eager_and_graph_modes(tmp_path):
    """An ADI2 run with a centre-of-mass distance restraint: the CV file
    written in graph mode (centres of mass from the engine) matches the eager
    run's (from the wrapper), and its first value is the input geometry's."""
    cvs = {}
    for mode in ("eager", "graph"):
        d = tmp_path / mode
        d.mkdir()
        (d / "us.cfg").write_text("k = 10.0\nx0 = 8.0\ncvFile = cv.dat\n"
                                  "atomSelect1 = 1,2,3,4,5\natomSelect2 = 23,24,25,26,27\n")
        cfg = _deck(d, n_steps=40, delta=10, extra="US = 1\nusCfgFile = us.cfg\n")
        txt = cfg.read_text().replace("US = 0\n", "")
        cfg.write_text(txt)
        args = [EXE, str(cfg)] + (["--graph"] if mode == "graph" else [])
        r = subprocess.run(args, capture_output=True, text=True, cwd=d, timeout=120)
        assert r.returncode == 0, r.stderr
        cvs[mode] = np.loadtxt(d / "cv.dat")
    e, g = cvs["eager"], cvs["graph"]
    print("eager", e.tolist(), "graph", g.tolist())
    assert e.shape == g.shape == (4, 2)
    assert np.array_equal(e[:, 0], g[:, 0])
    assert abs(e[0, 1] - g[0, 1]) < 1e-3  # same input coordinates
    assert np.abs(e[:, 1] - g[:, 1]).max() < 0.05  # FP32 summation order only


@pytest.mark.gpu
@pytest.mark.parametrize("n_steps,delta_ten", [(23, 10), (15, 20), (53, 100)])
def test_driver_graph_steps_per_replay_leave_the_trajectory_unchanged(tmp_path, n_steps,
                                                                       delta_ten):
    """Graph mode replays the largest of 50, 20, 10, 5, 2 steps per graph that
    divides the write interval and fits the run, and the last steps
    (23 = 2 x 10 + 3, 15 = 10 + 5, 53 = 50 + 3) with a one-step graph.  With
    15 steps and writes every 20 the one-step graph is captured right after
    the 10-step replay is launched, with no write (and so no synchronisation)
    between them: the recapture must wait for that replay before it frees its
    graph.  The trajectory equals, digit for digit, the run that replays one
    step per graph (write interval 23: no divisor among them)."""
    out = {}
    for name, delta in (("ten", delta_ten), ("one", 23)):
        d = tmp_path / name
        d.mkdir()
        cfg = _deck(d, n_steps=n_steps, delta=delta)
        r = subprocess.run([EXE, str(cfg), "--graph"], capture_output=True, text=True, cwd=d,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        out[name] = ((d / "pos.rst").read_text(), (d / "vel.rst").read_text())
    assert out["ten"][0] == out["one"][0]
    assert out["ten"][1] == out["one"][1]

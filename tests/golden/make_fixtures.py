"""Writes the golden fixtures under tests/golden/ from the reference's own test data.

Run once in a container that has the reference tree (it is not needed at test
time; the outputs are committed):

    python tests/golden/make_fixtures.py /root/reference

Outputs
* adi2/ADI2.{prmtop,isspatop,rst7}: the 44-atom ADI2 deck of
  timing_tests/ADI2/10000MC/ (input data files, copied verbatim);
* adi2/ADI2.multi.frc.{LJsolu,Csolu,Csolv_pair}.dat: FP64 per-atom golden
  forces (x y z Fx Fy Fz) for the deterministic nonbond components;
* adi2_isspa_stats.npz: per-atom mean and standard error over the 1000
  independent-seed frames of timing_tests/ADI2/10000MC/compare/
  ISSPA_force_{LJsolv,Csolv,Csolv_CDD}_10000.xyz (curand, nMC=1000), plus the
  single-frame FP32 GPU outputs of the three deterministic components.
  Frames are parsed whole (all 44 atoms each).
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def read_frames(path):
    with open(path) as fh:
        lines = fh.readlines()
    frames, i = [], 0
    while i < len(lines):
        n = int(lines[i].split()[0])
        rows = [list(map(float, l.split()[1:7])) for l in lines[i + 2:i + 2 + n]]
        frames.append(rows)
        i += 2 + n
    return np.array(frames)  # [F, N, 6]


def main(ref):
    src = os.path.join(ref, "timing_tests", "ADI2", "10000MC")
    out = os.path.join(HERE, "adi2")
    os.makedirs(out, exist_ok=True)
    for f in ("ADI2.prmtop", "ADI2.isspatop", "ADI2.rst7"):
        shutil.copyfile(os.path.join(src, f), os.path.join(out, f))
    for comp in ("LJsolu", "Csolu", "Csolv_pair"):
        shutil.copyfile(os.path.join(src, "compare", f"ADI2.multi.frc.{comp}.dat"),
                        os.path.join(out, f"ADI2.multi.frc.{comp}.dat"))
    stats = {}
    for comp in ("LJsolv", "Csolv", "Csolv_CDD"):
        fr = read_frames(os.path.join(src, "compare", f"ISSPA_force_{comp}_10000.xyz"))
        f = fr[:, :, 3:6]
        stats[f"{comp}_mean"] = f.mean(axis=0)
        stats[f"{comp}_se"] = f.std(axis=0, ddof=1) / np.sqrt(f.shape[0])
        stats[f"{comp}_frames"] = np.array(f.shape[0])
        stats["xyz"] = fr[0, :, 0:3]
    for comp in ("LJsolu", "Csolu", "Csolv_pair"):
        fr = read_frames(os.path.join(src, "compare", f"ISSPA_force_{comp}_10000.xyz"))
        stats[f"{comp}_gpu_fp32"] = fr[0, :, 3:6]
    np.savez_compressed(os.path.join(HERE, "adi2_isspa_stats.npz"), **stats)
    print("wrote", sorted(stats))
    omp_fixture()


def omp_inputs(n=50, n_mc=10):
    """Inputs shaped like src/isspa_force.omp.c:157-170 (atoms on a line 7 A
    apart, one type with the file's own parabola and LJ constants)."""
    xyz = np.zeros(3 * n, np.float32)
    xyz[0::3] = np.arange(n, dtype=np.float32) * np.float32(7.0)
    one = lambda v: np.array([v], np.float32)  # noqa: E731
    return dict(xyz=xyz, w=one(0.801), x0=one(4.118), g0=one(1.714),
                gr2=np.array([11.002, 21.478], np.float32), alpha=one(2.674),
                lj_A=one(6.669e7), lj_B=one(1.103e4), ityp=np.zeros(n, np.int32), n_mc=n_mc)


def omp_fixture():
    """Forces of the reference's own cpu_ispa_force (oracle/_ref build of
    src/isspa_force.omp.c) at one thread after srand(1)."""
    sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
    import pyoracle as po
    ref = po.ref_lib()
    if ref is None:
        print("oracle/_ref not built: skipping the OpenMP fixture")
        return
    a = omp_inputs()
    f = po.cpu_ispa_force(ref.cpu_ispa_force, a["xyz"], a["w"], a["x0"], a["g0"], a["gr2"],
                          a["alpha"], a["lj_A"], a["lj_B"], a["ityp"], a["n_mc"], 1, seed=1)
    np.savez_compressed(os.path.join(HERE, "omp_ref_fixture.npz"), forces=f, seed=1, threads=1,
                        **{k: v for k, v in a.items()})
    print("wrote omp_ref_fixture.npz", float(np.abs(f).sum()))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

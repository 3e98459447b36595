"""The HIP runtime a Python process binds the library to.

PyTorch's ROCm wheel ships its own libamdhip64.so.7; isspa_amd.capi imports
PyTorch before loading libisspa_amd.so so that the library's dependency
resolves to that same runtime.  A fresh process that imports the binding
first (before anything else touches PyTorch) must compute normally."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

PROG = """
import sys
sys.path[:0] = [{pkg!r}]
from isspa_amd import capi          # the library first
from isspa_amd import synth
from isspa_amd.engine import GpuSystem
g = GpuSystem(synth.workload(0), "cuda:0")
g.isspa_force(step=1)
g.sync()
import numpy as np
f = g.f.cpu().numpy()
assert np.all(np.isfinite(f)) and np.abs(f).max() > 0
print("binding-first process ok")
"""


def test_binding_imported_first_in_a_fresh_process():
    pkg = os.path.join(ROOT, "isspa-force-cuda_amd")
    r = subprocess.run([sys.executable, "-c", PROG.format(pkg=pkg)], capture_output=True,
                       text=True, timeout=240)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0 and "binding-first process ok" in r.stdout

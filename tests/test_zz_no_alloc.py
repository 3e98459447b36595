"""isspa_model_force_noalloc against isspa_model_force, through the C ABI.

The no-allocation entry point runs the same kernels with its per-call
buffers in the library's static device array, so its forces must equal the
allocating entry point's bit for bit.  Two calls in a row, the second on a
smaller system, check that what the first call leaves in the shared array
does not leak into the second.  A model whose tiling needs more than the
array holds is refused with an error naming the sizes.
"""
import numpy as np
import pytest
import torch  # noqa: F401  (the process's HIP runtime; see isspa_amd.capi)

pytestmark = pytest.mark.gpu


def _forces(g, fn, step):
    import torch
    from isspa_amd import capi
    from isspa_amd.capi import ptr
    base = torch.linspace(-1.0, 1.0, g.f.numel(), device="cuda:0").reshape(g.f.shape)
    f = base.clone()
    rc = fn(g.isspa, ptr(g.x), ptr(f), step, None, g.stream())
    capi.check(rc, fn.__name__)
    torch.cuda.synchronize()
    return f.cpu().numpy(), base.cpu().numpy()


def test_noalloc_equals_model_force_on_two_calls(cfg1, cfg0):
    from isspa_amd import capi
    from isspa_amd.engine import GpuSystem
    lib = capi.lib
    for name, s, step in (("1,100 atoms nMC 20", cfg1, 3), ("22 atoms nMC 10", cfg0, 5)):
        print(f"case noalloc {name} step {step}", flush=True)
        g = GpuSystem(s, "cuda:0", nonbond=False, bonded=False)
        got, base = _forces(g, lib.isspa_model_force_noalloc, step)
        want, _ = _forces(g, lib.isspa_model_force, step)
        assert np.array_equal(got, want), name
        assert np.abs(got - base).max() > 0  # the force was added
        g.close()


def test_noalloc_refuses_models_beyond_its_scratch():
    from isspa_amd import capi, synth
    from isspa_amd.capi import ptr
    from isspa_amd.engine import GpuSystem
    s = synth.workload(2)  # 19,998 atoms: hundreds of MB of partials
    print("case noalloc refusal 19,998 atoms", flush=True)
    g = GpuSystem(s, "cuda:0", nonbond=False, bonded=False)
    rc = capi.lib.isspa_model_force_noalloc(g.isspa, ptr(g.x), ptr(g.f), 0, None, g.stream())
    assert rc == 1  # ISSPA_ERR_ARG
    msg = capi.lib.isspa_last_error().decode()
    assert "bytes of scratch" in msg, msg
    g.close()

"""Host logic of the launch-geometry autotuner (no GPU)."""
import math

from isspa_amd import autotune


def test_effective_pieces_follow_the_library_rule():
    # isspa_model_set_tiling: pieces of ceil(n / p) items, clamped to [1, n]
    for n in (1, 2, 7, 64, 1000, 19998, 399960):
        for p in (1, 2, 3, 10, 63, 206, 5000, 10 ** 6):
            per = math.ceil(n / max(1, min(p, n)))
            assert autotune.effective_pieces(n, p) == math.ceil(n / per)
            assert 1 <= autotune.effective_pieces(n, p) <= n


def test_round_fill():
    # 2048 blocks x 4 waves = 8192 waves: exactly one round
    assert autotune.round_fill(2048, 1) == 1.0
    assert autotune.round_fill(1024, 1) == 0.5
    assert abs(autotune.round_fill(1563, 10) - 0.6318) < 1e-3   # 62520 waves = 7.63 rounds


def test_round_aligned_candidates():
    n, P = 19998, 19998 * 20
    fb, pb = math.ceil(P / 256), math.ceil(n / 256)
    s = autotune.round_aligned(n, fb, 10)
    c = autotune.round_aligned(P, pb, 206)
    assert 10 in s and 206 in c
    for v in s:
        assert v == 10 or autotune.round_fill(fb, v) >= 0.85
        assert 5 <= v <= 20
    for v in c:
        assert v == 206 or autotune.round_fill(pb, v) >= 0.85
    assert 9 in s and 13 in s          # 6.87 and 9.92 rounds
    assert len(s) <= 7 and len(c) <= 7


class _FakeModel:
    """Stands in for GpuSystem: stage times are a fixed function of the
    tiling, so the tuner's choice can be checked without a GPU."""

    class _Sys:
        n_atoms, n_points = 1000, 20000

    windows = None
    system = _Sys()

    def __init__(self):
        self.t = (0, 0)
        self.waves = 32768

    def set_tiling(self, s, c):
        self.t = (s, c)

    def set_target_waves(self, w):
        self.waves = w

    def tiling(self):
        if self.t != (0, 0):
            return self.t
        return (max(1, self.waves // 4096), max(1, self.waves // 512))

    def profile(self, reps):
        s, c = self.tiling()
        # the field stages prefer 8 splits, the force stages 64 chunks
        return {"points": 0.0, "field": 1.0 + abs(s - 8) * 0.01, "epilogue": 0.1,
                "force": 2.0 + abs(c - 64) * 0.001, "reduce": 0.1}


def test_tune_picks_fastest_pass_and_confirm_overrides():
    g = _FakeModel()
    r = autotune.tune(g, reps=1)
    assert r.measure == "IS-SPA pass"
    assert g.tiling() == (8, 64) and abs(r.ms - 3.2) < 1e-9
    # the five leaders include waves 32768 and tiling (8, 64) (both 3.2 ms)
    # and tiling (8, 32) (3.232 ms); a confirmation that prefers (8, 32) decides
    calls = []

    def confirm(model, reps):
        calls.append(model.tiling())
        return 1.0 if model.tiling() == (8, 32) else 5.0

    confirm.measure = "engine step"
    g2 = _FakeModel()
    r2 = autotune.tune(g2, reps=1, confirm=confirm)
    assert r2.measure == "engine step"
    assert len(calls) == 5 and (8, 32) in calls and calls.count((8, 64)) == 2
    assert r2.kind == "tiling" and r2.value == (8, 32)
    assert g2.tiling() == (8, 32) and r2.ms == 1.0

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

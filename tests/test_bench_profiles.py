"""bench.py's use of the committed counter profiles (CPU only).

The k_force traffic it reports is read from profiles/, so the profile must
describe the kernel's algorithm: WRITE_SIZE per chunk is the 6 FP64 partials
of 19,998 atoms, the fetched bytes are at least the point records, and the
interpolation reproduces every profiled point.  The VALU counter summary
must be internally consistent (executed FLOP from the instruction counts).
"""
import os
import sys

import pytest

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402

N_ATOMS, N_POINTS = 19998, 399960
TRAFFIC = os.path.join(ROOT, "profiles", bench.TRAFFIC_JSON)
COUNTERS = os.path.join(ROOT, "profiles", bench.COUNTERS_JSON)


def _points():
    import json
    with open(TRAFFIC) as fh:
        tr = json.load(fh)
    return tr, sorted(tr["points"], key=lambda p: p["chunks"])


def test_traffic_profile_matches_the_algorithm():
    tr, pts = _points()
    assert len(pts) >= 2 and tr["write_scale"] == 1.0 and 1.99 < tr["fetch_scale"] < 2.01
    partial_kib = 48 * N_ATOMS / 1024          # 6 FP64 partials per atom and chunk
    record_bytes = 80 * N_POINTS               # 64-byte record + 16-byte aux per point
    # WRITE_SIZE is linear in the chunk count with the partials' slope (the
    # intercept, other writes, is under 1 MiB)
    slope = (pts[-1]["write_kib"] - pts[0]["write_kib"]) / (pts[-1]["chunks"] - pts[0]["chunks"])
    assert abs(slope - partial_kib) <= 2e-3 * partial_kib
    for p in pts:
        assert p["grid"] == p["chunks"] * 79 * 256
        assert abs(p["write_kib"] - partial_kib * p["chunks"]) <= 1024, p
        fetched = p["fetch_kib"] * tr["fetch_scale"] * 1024
        assert fetched >= record_bytes, p
        alg = record_bytes + 48 * N_ATOMS * p["chunks"]
        assert p["traffic_bytes"] >= 0.99 * alg, p


def test_traffic_interpolation():
    tr, pts = _points()
    for p in pts:
        b, note, src = bench.traffic_at(p["chunks"], TRAFFIC)
        assert abs(b - p["traffic_bytes"]) <= 1024, p
        assert src == tr["summary"] and "interpolated" in note
    lo, hi = pts[-3], pts[-2]
    mid = (lo["chunks"] + hi["chunks"]) // 2
    b, _, _ = bench.traffic_at(mid, TRAFFIC)
    assert min(lo["traffic_bytes"], hi["traffic_bytes"]) <= b <= max(lo["traffic_bytes"],
                                                                      hi["traffic_bytes"])
    for c in (pts[0]["chunks"] - 1, pts[-1]["chunks"] + 1):
        b, note, _ = bench.traffic_at(c, TRAFFIC)
        assert b is None and "outside" in note
    assert bench.traffic_at(207, os.path.join(ROOT, "profiles", "missing.json")) is None


@pytest.mark.parametrize("kernel,pairs", [("k_force", N_ATOMS * N_POINTS),
                                          ("k_field", N_ATOMS * N_POINTS)])
def test_counter_summary_is_consistent(kernel, pairs):
    import json
    with open(COUNTERS) as fh:
        vc = json.load(fh)
    k = vc[kernel]
    assert k["pairs"] == pairs
    wave_pairs = pairs / 64
    assert abs(k["SQ_INSTS_VALU"] / wave_pairs - k["valu_per_wave_pair"]) < 0.01
    executed = (2 * k["FMA_F32"] + k["ADD_F32"] + k["MUL_F32"] + k["TRANS_F32"]) / wave_pairs
    assert abs(executed - k["executed_flop_per_pair"]) < 0.01
    assert abs(executed / k["algorithmic_flop_per_pair"] - k["executed_over_algorithmic"]) < 1e-3
    # one v_rsq per pair
    assert abs(k["TRANS_F32"] / wave_pairs - 1.0) < 0.01
    cc = bench.counter_check(COUNTERS)
    assert cc[kernel]["valu_per_wave_pair"] == k["valu_per_wave_pair"]
    assert cc["source"] == vc["summary"]

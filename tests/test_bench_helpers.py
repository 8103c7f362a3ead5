"""Host-side helpers of bench.py: the counted HBM traffic it reports comes from the committed
counter summary, for the operator form that ran, and only for the steps that summary covers."""
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import bench  # noqa: E402


def test_traffic_window_is_the_summary_title():
    assert bench.traffic_window() == [500, 519]


def test_traffic_per_operator_form():
    streamed = bench.measured_traffic("k_cg_apply<1>", [500, 519], "false")
    recomputed = bench.measured_traffic("k_cg_apply<1>", [500, 519], "true")
    assert streamed is not None and recomputed is not None
    # recomputing b_ij streams 4 bytes per pair less than streaming (a_ij, b_ij)
    assert 0.6 * streamed < recomputed < 0.9 * streamed
    assert bench.measured_traffic("k_cg_update3", [500, 519]) is not None


def test_no_traffic_for_another_window():
    assert bench.measured_traffic("k_cg_apply<1>", [500, 1499], "false") is None
    assert bench.measured_traffic("k_cg_apply<1>", [0, 19], "false") is None

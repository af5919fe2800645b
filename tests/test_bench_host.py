"""bench.py's host-side arithmetic and contract, on the CPU: the one-GPU guard,
the byte count of the row-pair format, the roofline object and the counter
summary it reads."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


class _FakePair:
    """a matrix whose FP32 variant is the row-pair layout"""

    def __init__(self, lanes, layout):
        self._lanes, self._layout = lanes, layout

    def lanes(self):
        return self._lanes

    def pair_layout_info(self):
        return dict(self._layout)


def test_format_step_bytes_counts_values_metadata_and_twenty_vector_passes():
    lay = {"slices": 10, "slots": 100, "stored_pairs": 7, "stored_real_pairs": 11, "meta_bytes": 4000,
           "stored_real_singles": 13, "sep_slots": 2}
    rows = 1280
    b, got = bench.format_step_bytes(_FakePair((129, 143), lay), rows)
    assert b == 4 * (7 * 16 + 11 * 8 + 13 * 4 + 4000) + 20 * rows * 8
    assert got["kind"].startswith("row-pair")


def test_roofline_object_fractions():
    rf = bench.roofline_obj("k", "w", 8e6, 1e-3, 1.1e-3, 16e6, "src", 32e6)
    assert rf["achieved"] == pytest.approx(8000.0)          # 8 MB in 1 us = 8 TB/s
    assert rf["frac"] == pytest.approx(1.0)
    assert rf["traffic_over_format_bytes"] == pytest.approx(2.0)
    assert rf["traffic_gbs"] == pytest.approx(16000.0)
    assert rf["csr_model"]["over_peak"] == pytest.approx(4.0)


def test_committed_traffic_summary_is_read():
    t, src = bench.read_traffic(os.path.join(ROOT, "profiles", "k_ks2_stage_chain10_traffic.txt"))
    assert src == os.path.join("profiles", "k_ks2_stage_chain10_traffic.txt")
    assert 5e7 < t < 2e8


def test_kernel_names_cover_every_variant_code():
    for g in list(range(128, 137)) + list(range(140, 145)):
        assert "stage" in bench.kernel_name(g, "float")

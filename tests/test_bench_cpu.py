"""The bench line's contract that does not need a GPU: the counter figures
it quotes come from the committed rocprofv3 summaries in profiles/, and a
mode-product launch's traffic scales with the fields it moves (a carried
step's launches move 14 and 13 fields, one that transforms u 26 and 13)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402  (module level holds only helpers; main() runs under __main__)


@pytest.mark.parametrize("name,bytes_per_field", [
    ("k_modeprod_f16_192cube_hbm_counters.md", 4 * 2 * 196 ** 3),   # FP32 in + out
    ("k_modeprod_dl_192cube_13fields_hbm_counters.md", 8 * 2 * 196 ** 3),
    ("k_modeprod_split_192cube_hbm_counters.md", 4 * 2 * 196 ** 3),
])
def test_mode_product_traffic_per_field_pass_is_committed(name, bytes_per_field):
    v = bench.measured_traffic(name, "traffic_bytes_per_field_pass")
    assert v is not None, name
    # at least the algorithmic bytes, and not wildly above them
    assert bytes_per_field <= v < 1.5 * bytes_per_field
    assert bench.traffic_per_launch(name, 13.5) == pytest.approx(13.5 * v)


def test_fft_pass_traffic_is_committed():
    v = bench.measured_traffic("k_fft16_f32_192cube_hbm_counters.md")
    alg = 2 * 6 * 192 ** 3 * 8.0  # read + write six packed complex FP32 fields
    assert v is not None and alg <= v < 1.1 * alg


def test_missing_summary_gives_none():
    assert bench.measured_traffic("no_such_summary.md") is None
    assert bench.traffic_per_launch("no_such_summary.md", 13.0) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "not supported" in r.stderr

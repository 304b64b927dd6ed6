"""The benchmark's CPU-side contract: the bytes the roofline counts and the
one-GPU guard (no GPU needed).  Algorithmic bytes per point and stage are
w * (4 nf + G) (+ the plastic words per step spread over the stages); the
bytes a stage needs with the df skip average (3 + 4 (s - 2) + 3) / s words
per field over an s-stage step, since every low-storage scheme of
rk.cpp:42-91 has A[0] = 0."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


@pytest.mark.parametrize("nf,w,kw,want", [
    (9, 8, {}, 288.), (9, 4, {}, 144.),                       # 3-D planar
    (5, 8, {"ndim": 2}, 160.), (5, 4, {"ndim": 2}, 80.),      # 2-D planar, ensemble
    (6, 4, {"ndim": 2, "curvilinear": True}, 116.),           # 2-D curvilinear plastic, bulk launch
    (6, 4, {"ndim": 2, "curvilinear": True, "plastic_words": 12}, 125.6),
    (9, 8, {"needed": True}, 259.2), (9, 4, {"needed": True}, 129.6),
    (5, 4, {"ndim": 2, "needed": True}, 72.),
])
def test_bytes_per_point(nf, w, kw, want):
    assert bench.bytes_per_point(nf, w, **kw) == pytest.approx(want)


def test_stage_field_words():
    # brute force over the stages: f read and written always, df read where
    # A != 0 (every stage but the first), df written except on the last
    for s in range(1, 7):
        words = [2 + (1 if k > 0 else 0) + (1 if k < s - 1 else 0) for k in range(s)]
        assert bench.stage_field_words(s) == pytest.approx(sum(words) / s), s
    assert bench.stage_field_words(5) == pytest.approx(3.6)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr

"""benchmark_acceleration (src/acceleration.c:369-507): the printed text
against the reference's own (golden F11, timings masked) and the
machine-readable records of gs_benchmark_acceleration."""
import ctypes
import os
import re
import sys
import tempfile

import numpy as np
import pytest

from grav_sim_amd import lib as L

pytestmark = pytest.mark.gpu

# the parameter sets the golden text was printed for (tests/golden/make_golden.py)
BENCH_SETS = ((1, 0.0, 1e-3, -1, 2), (2, 0.0, 1e-3, -1, 1), (3, 0.5, 1e-3, 1, 2), (1, 0.0, 1e-3, -1, 0))


def capture_stdout(fn):
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    with tempfile.TemporaryFile(mode="w+b") as tmp:
        saved = os.dup(1)
        os.dup2(tmp.fileno(), 1)
        try:
            out = fn()
            libc.fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tmp.seek(0)
        return out, tmp.read().decode()


def mask(text):
    """Timings always differ; so does the Barnes-Hut MAE (the reference's
    walk drops leaf monopoles).  Everything else must be the same text."""
    text = re.sub(r"Avg time: \S+ \(\+- \S+\) s", "Avg time: T (+- S) s", text)
    return re.sub(r"(Method: Barnes-Hut\n    Number of times: \d+\n    Avg time: T \(\+- S\) s\n    MAE: )\S+",
                  r"\1M", text)


def test_benchmark_text_matches_reference(golden):
    f = golden("f11_benchmark.npz")
    recs, text = capture_stdout(lambda: L.benchmark_acceleration(f["x"], f["m"], BENCH_SETS))
    assert mask(text) == mask(str(f["text"]))
    assert "MAE: 0\n" in text  # the massless form with every particle massive equals the pairwise sum
    assert recs[3]["num_times"] == 0 and recs[3]["mean_s"] == 0.0


def test_benchmark_records_every_method():
    from grav_sim_amd import ic
    x, v, m = ic.two_spheres(4096, 12)
    m = m.copy()
    m[::4] = 0.0  # a quarter massless
    L.learned_set_params(np.zeros(L.learned_num_params(), np.float32))
    try:
        sets = [(1, 0.0, 1e-3, -1, 1), (2, 0.0, 1e-3, -1, 1), (16, 0.0, 1e-3, -1, 2), (17, 0.0, 1e-3, -1, 1),
                (3, 0.5, 1e-3, 1, 1), (18, 0.8, 1e-3, 1, 1), (19, 0.5, 1e-3, 1, 1), (20, 0.5, 1e-3, 1, 1),
                (21, 0.5, 1e-3, 1, 1), (22, 0.5, 1e-3, 1, 1), (16, 0.0, 1e-3, -1, 0)]
        recs, text = capture_stdout(lambda: L.benchmark_acceleration(x, m, sets))
    finally:
        L.learned_reset()
    names = re.findall(r"Method: (.*)\n", text)
    assert names == ["Pairwise", "Massless", "Pairwise (FP32 split)", "Massless (FP32 split)", "Barnes-Hut",
                     "Barnes-Hut (learned correction)", "Barnes-Hut (quadrupole)", "Barnes-Hut (group walk)",
                     "Barnes-Hut (group walk, quadrupole)", "Barnes-Hut (group walk, local expansions)"]
    n = len(m)
    massive = np.count_nonzero(m)
    info = L.device_info()
    for r, s in zip(recs, sets):
        assert r["method"] == s[0] and r["num_times"] == s[4]
        if s[4] == 0:
            assert r["mean_s"] == 0.0 and r["interactions"] == 0.0
            continue
        want = {1: n * (n - 1.0), 16: n * (n - 1.0), 2: massive * (n - 1.0), 17: massive * (n - 1.0)}
        if s[0] in want:
            assert r["interactions"] == want[s[0]]
        else:
            assert 0 < r["interactions"] < n * (n - 1.0)
        assert r["mean_s"] > 0 and r["interactions_per_s"] == pytest.approx(r["interactions"] / r["mean_s"])
        assert r["particles_per_s"] == pytest.approx(n / r["mean_s"])
        assert r["flops_per_s"] == pytest.approx(20.0 * r["interactions_per_s"])
        peak = info.fp64_vector_tflops if s[0] in (1, 2) else info.fp32_vector_tflops
        assert r["peak_flops_per_s"] == pytest.approx(peak * 1e12)
        assert r["roofline_frac"] == pytest.approx(r["flops_per_s"] / r["peak_flops_per_s"])
    # MAE against test 0 (FP64 pairwise): the massless form with massless particles differs
    # only in what massless sources add (nothing) -> 0; FP32 within its bound
    assert recs[1]["mae"] == 0.0
    assert 0.0 < recs[2]["mae"] < 1e-4 * np.mean(np.abs(L.acceleration(x, m, 1, eps=1e-3)).sum(axis=1))

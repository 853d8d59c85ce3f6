"""CPU checks of the serving workload generators and latency summary."""
import numpy as np

from radeonrays_sdk_amd import scenes, serving


def test_poisson_arrivals_rate_and_order():
    a = scenes.poisson_arrivals(20000, 10_000.0)
    assert a[0] == 0 and (np.diff(a) >= 0).all()
    rate = (a.size - 1) / (a[-1] / 1e9)
    assert abs(rate / 10_000.0 - 1) < 0.03
    gaps = np.diff(a) / 1e9
    # Exponential gaps: coefficient of variation 1.
    assert abs(gaps.std() / gaps.mean() - 1) < 0.05


def test_serving_requests_shape():
    fake = [scenes.ServingScene("soup" if k % 2 == 0 else "terrain", None, None) for k in range(32)]
    pool = 1 << 20
    reqs = scenes.serving_requests(5000, fake, pool)
    counts = np.array([q.count for q in reqs])
    assert counts.min() >= 1000 and counts.max() <= pool
    # log-uniform in [1e3, 1e6]: median near sqrt(1e9) = 31.6k.
    assert 25_000 < np.median(counts) < 40_000
    for q in reqs:
        assert 0 <= q.offset and q.offset + q.count <= pool
        assert (q.pool < 2) == (fake[q.scene].kind == "soup")
    closest = np.mean([q.query == "closest" for q in reqs])
    assert abs(closest - 0.7) < 0.03
    assert len({q.scene for q in reqs}) == 32


def test_serving_scene_sizes():
    sc = scenes.serving_scenes(n_scenes=6, lo=1000, hi=20000)
    assert [s.kind for s in sc] == ["soup", "terrain"] * 3
    for s in sc:
        assert 500 <= s.indices.shape[0] <= 25000
        assert s.indices.max() < s.vertices.shape[0]


def test_latency_summary():
    t = np.zeros(4, dtype=serving.TIMING_DTYPE)
    t["arrival_ns"] = [0, 1_000_000, 2_000_000, 3_000_000]
    t["complete_ns"] = [1_000_000, 3_000_000, 3_000_000, 4_000_000]
    t["batch_size"] = [1, 2, 2, 1]
    s = serving.latency_summary(t, [1000, 1000, 1000, 1000])
    assert s["p50_ms"] == 1.0 and abs(s["p99_ms"] - 1.97) < 1e-9
    assert s["mrays_s"] == 1.0 and s["mean_batch"] == 1.5

"""Host-side parts of the serving loop (no GPU): the request stream, the HBM footprint estimate
used for admission, and the summary statistics."""
import numpy as np

from taichi_dem_julia_port_amd import serving


def test_stream_is_seeded_and_within_its_ranges():
    a = serving.make_stream(40, rate=20.0, seed=5)
    b = serving.make_stream(40, rate=20.0, seed=5)
    assert [(r.arrival, r.n, r.steps, r.kind) for r in a] == [(r.arrival, r.n, r.steps, r.kind) for r in b]
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0)
    assert all(1000 <= r.n <= 200_000 and 2000 <= r.steps <= 50_000 for r in a)
    assert [r.kind for r in a[:4]] == ["cube", "pour", "cube", "pour"]
    assert all(r.bonded == (r.kind == "cube") for r in a)


def test_footprint_grows_with_grains_slots_and_walls():
    f = serving.footprint
    assert f(2000, 64) == 2 * f(1000, 64)
    assert f(1000, 64) > f(1000, 32) > 0
    assert f(1000, 64, nwall=6) > f(1000, 64, nwall=1)
    # a million grains with 64 slots fits many times in a 288 GB device
    assert f(1_000_000, 64) < 30e9


def test_summary_statistics():
    res = [serving.Result(k, arrival=0.1 * k, admitted=0.1 * k, finished=1.0 + 0.2 * k, n=100,
                          steps=10) for k in range(5)]
    s = serving.summarize(res, makespan=2.0)
    lat = [1.0 + 0.1 * k for k in range(5)]
    assert s["requests"] == 5 and s["scenes_per_s"] == 2.5
    assert s["particle_steps_per_s"] == 5 * 100 * 10 / 2.0
    assert abs(s["latency_p50_s"] - np.percentile(lat, 50)) < 1e-9
    assert abs(s["latency_max_s"] - max(lat)) < 1e-9

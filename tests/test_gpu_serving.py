"""Continuous batching: requests admitted while others run give the same result as alone.

A short stream with staggered arrivals and different step counts forces several re-packs
(admissions and retirements); every scene's final grains, contacts and bond breaks must equal a
one-scene-at-a-time run bit for bit (FP64) -- and within the mixed-precision tolerance of the FP64
path for the mixed server.
"""
import numpy as np
import pytest

from helpers import pair_keys

pytestmark = pytest.mark.gpu


def _stream():
    from taichi_dem_julia_port_amd.serving import Request
    reqs = [Request(0, 0.00, "cube", 700, 240, 11), Request(1, 0.00, "pour", 1500, 100, 12),
            Request(2, 0.05, "cube", 300, 170, 13), Request(3, 0.10, "pour", 900, 310, 14),
            Request(4, 0.10, "cube", 1200, 60, 15)]
    return reqs


def test_served_results_equal_alone_fp64():
    from taichi_dem_julia_port_amd.serving import Server, serve_one_at_a_time
    print("[case] serving fp64 stream of 5", flush=True)
    srv = Server(precision="fp64", epoch_steps=50, keep_results=True)
    served, makespan = srv.run(_stream())
    alone, _ = serve_one_at_a_time(_stream(), precision="fp64", realtime=False, keep_results=True)
    assert srv.repacks >= 4
    assert len(served) == len(alone) == 5
    for a, b in zip(served, alone):
        assert a.rid == b.rid
        for f in ("k", "v", "w", "q"):
            assert np.array_equal(a.grains[f].view(np.uint64), b.grains[f].view(np.uint64)), (a.rid, f)
        ca, ba = a.contacts
        cb, bb = b.contacts
        assert np.array_equal(pair_keys(ca), pair_keys(cb)), a.rid
        assert np.array_equal(ba, bb), a.rid
        for f in ("F", "s", "ti", "tj"):
            assert np.array_equal(ca[f].view(np.uint64), cb[f].view(np.uint64)), (a.rid, f)
        assert a.bonds_broken == b.bonds_broken


def test_served_mixed_results_close_to_fp64_alone():
    from taichi_dem_julia_port_amd.serving import Server, serve_one_at_a_time
    print("[case] serving mixed stream of 5", flush=True)
    served, _ = Server(precision="mixed", epoch_steps=50, keep_results=True).run(_stream())
    alone, _ = serve_one_at_a_time(_stream(), precision="fp64", realtime=False, keep_results=True)
    for a, b in zip(served, alone):
        rel = np.linalg.norm(a.grains["k"] - b.grains["k"], axis=1) / np.linalg.norm(b.grains["k"], axis=1)
        assert rel.max() <= 1e-6, (a.rid, rel.max())

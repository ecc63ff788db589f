"""The serving stream at its real scene sizes, with step counts scaled down.

Twelve requests with Poisson arrivals at 40/s, grains log-uniform in 1k-200k, half bonded cubes
(bond strength, impact speed and seed drawn per request as in the benchmark stream) and half
loose pours, each asking for 100-400 steps (the benchmark asks for 2k-50k).  The continuous-
batching server admits, re-packs and retires them as they arrive and finish:

* FP64 server: every scene's final grains, contacts and bond-break count equal the same request
  run alone, bit for bit;
* one of them (the smallest bonded cube that meets the floor during its run) equals the oracle
  run from the same initial state for the same steps, at the FP64 bar of the parity tests;
* mixed-precision server: every scene's positions agree with the FP64 alone run within 1e-6 of
  the scene's largest |k|.
"""
import numpy as np
import pytest

from helpers import compare_state, pair_keys

pytestmark = pytest.mark.gpu


def _stream():
    from taichi_dem_julia_port_amd.serving import make_stream
    return make_stream(12, rate=40.0, seed=25, smin=100, smax=400)


def _alone():
    from taichi_dem_julia_port_amd.serving import serve_one_at_a_time
    alone, _ = serve_one_at_a_time(_stream(), precision="fp64", realtime=False, keep_results=True)
    return alone


def test_stream_sizes_cover_the_range():
    reqs = _stream()
    n = sorted(r.n for r in reqs)
    assert n[0] < 5_000 and n[-1] > 100_000, n
    assert {r.kind for r in reqs} == {"cube", "pour"}


def test_served_stream_equals_alone_and_one_scene_equals_the_oracle():
    from taichi_dem_julia_port_amd.serving import Server
    print("[case] 12-request stream at real sizes, FP64 server", flush=True)
    srv = Server(precision="fp64", epoch_steps=80, poll_steps=20, min_epoch_s=0.0,
                 keep_results=True)
    served, _ = srv.run(_stream())
    alone = _alone()
    assert srv.repacks >= 6, srv.repacks
    assert len(served) == len(alone) == 12
    for a, b in zip(served, alone):
        assert a.rid == b.rid
        for f in ("k", "v", "w", "q", "a", "dw"):
            assert np.array_equal(a.grains[f].view(np.uint64), b.grains[f].view(np.uint64)), (a.rid, f)
        ca, ba = a.contacts
        cb, bb = b.contacts
        assert np.array_equal(pair_keys(ca), pair_keys(cb)), a.rid
        assert np.array_equal(ba, bb), a.rid
        for f in ("k", "F", "s", "ti", "tj"):
            assert np.array_equal(ca[f].view(np.uint64), cb[f].view(np.uint64)), (a.rid, f)
        assert a.bonds_broken == b.bonds_broken, a.rid

    # the smallest bonded cube against the oracle from its initial state
    from oracle.oracle import OracleScene
    from taichi_dem_julia_port_amd.executor import Scene
    # (the smallest one whose grains touch the floor at the end of its run)
    reqs = {r.rid: r for r in _stream()}
    hit = [r for r in sorted(reqs.values(), key=lambda r: r.n) if r.kind == "cube" and
           np.any(next(b for b in alone if b.rid == r.rid).grains["k"][:, 2] <
                  next(b for b in alone if b.rid == r.rid).grains["r"])]
    assert hit, "no cube of the stream reaches the floor"
    cube = hit[0]
    cfg, g, w = cube.build()
    print(f"[case] oracle: cube rid {cube.rid}, {cube.n} grains, {cube.steps} steps", flush=True)
    with Scene(cfg, g, w) as s:
        s.initialize()
        s.step(cube.steps)
        got = next(r for r in served if r.rid == cube.rid)
        assert np.array_equal(s.read_grains()["k"], got.grains["k"])
        o = OracleScene(cfg, g, w, parallel=True)
        o.initialize()
        o.step(cube.steps - 1)
        o.track_force_scale()
        o.step(1)
        compare_state(s, o)
        o.close()


def test_mixed_server_close_to_fp64_alone():
    from taichi_dem_julia_port_amd.serving import Server
    print("[case] 12-request stream at real sizes, mixed server", flush=True)
    served, _ = Server(precision="mixed", epoch_steps=80, poll_steps=20, min_epoch_s=0.0,
                       keep_results=True).run(_stream())
    alone = _alone()
    worst = 0.0
    for a, b in zip(served, alone):
        assert a.rid == b.rid
        scale = np.abs(b.grains["k"]).max()
        err = np.abs(a.grains["k"] - b.grains["k"]).max() / scale
        worst = max(worst, err)
        assert err <= 1e-6, (a.rid, err)
    print(f"[mixed server] worst position error {worst:.3e} of the scene's largest |k|", flush=True)

"""Batched uneven clips (vipe_amd.batched): the batch, eager and replayed from
one captured graph, equals one ba_extended call per clip; the tile autotuner
picks a feasible tile and caches it; the tiled plan takes every tile when the
only high-degree frame is a wide one."""
import pytest
import torch

from vipe_amd import synth
from vipe_amd.batched import BatchedBA, Clip, PixelTileTuner

pytestmark = pytest.mark.gpu

SMALL_RES = ((24, 32, 30.0, 30.0, 16.0, 12.0), (16, 40, 20.0, 20.0, 20.0, 8.0), (32, 32, 30.0, 30.0, 16.0, 16.0))
ITERS = 4


def _clips(n=5, seed=3):
    gs = synth.make_batch(n, seed, windows=(8, 12, 16), resolutions=SMALL_RES)
    return [Clip.from_graph(g, "cuda") for g in gs]


def _single(clip, iters, tile=0):
    from vipe_amd.ext import slam_ext

    N, H, W = clip.disps.shape
    p, d = clip.poses.clone(), clip.disps.clone()
    plan = slam_ext.get_plan(clip.ii, clip.jj, N, H, W, clip.t0, clip.t1, tile)
    ws = torch.empty((plan.ws_bytes,), dtype=torch.uint8, device="cuda")
    dx = torch.empty((clip.t1 - clip.t0, 6), device="cuda")
    dz = torch.empty((plan.K, H * W), device="cuda")
    slam_ext.run_planned(plan, ws, p, d, clip.intrinsics, clip.disps_sens, clip.targets, clip.weights, clip.eta,
                         clip.depth_active, iters, 1e-4, 0.1, 0.05, dx, dz)
    return p, d, dx, dz


def test_run_planned_equals_ba_extended():
    from vipe_amd.ext import slam_ext

    c = _clips(1)[0]
    p, d, dx, dz = _single(c, ITERS)
    p2, d2 = c.poses.clone(), c.disps.clone()
    dx2, dz2, _ = slam_ext.ba_extended(p2, d2, c.intrinsics, c.disps_sens, c.targets, c.weights, c.eta, c.ii, c.jj,
                                       c.depth_active, c.t0, c.t1, ITERS, 1e-4, 0.1, False, 0.05, False, 1e-6, 1e-6,
                                       1.0, False)
    torch.cuda.synchronize()
    assert torch.equal(p, p2) and torch.equal(d, d2) and torch.equal(dx, dx2) and torch.equal(dz, dz2)


def test_batch_eager_equals_single_clips():
    clips = _clips()
    expect = [_single(c, ITERS) for c in clips]
    b = BatchedBA(clips, iterations=ITERS, n_streams=3)
    b.run()
    torch.cuda.synchronize()
    for n, (c, (p, d, dx, dz)) in enumerate(zip(clips, expect)):
        assert torch.equal(c.poses, p), n
        assert torch.equal(c.disps, d), n
        assert torch.equal(b.dx[n], dx), n
        assert torch.equal(b.dz[n], dz), n


def test_batch_graph_replays_equal_eager():
    clips = _clips(seed=5)
    start = [c.state() for c in clips]
    b = BatchedBA(clips, iterations=ITERS, n_streams=4)
    b.run()
    torch.cuda.synchronize()
    eager = [(c.poses.clone(), c.disps.clone(), dx.clone()) for c, dx in zip(clips, b.dx)]
    for c, s in zip(clips, start):
        c.restore(s)
    b.capture()
    other = _clips(1, seed=11)[0]
    for rep in range(3):
        for c, s in zip(clips, start):
            c.restore(s)
        torch.cuda.synchronize()
        b.replay()
        torch.cuda.synchronize()
        for n, (c, (p, d, dx)) in enumerate(zip(clips, eager)):
            assert torch.equal(c.poses, p), (rep, n)
            assert torch.equal(c.disps, d), (rep, n)
            assert torch.equal(b.dx[n], dx), (rep, n)
        # eager work of the program between replays
        _single(other, 2)
        torch.cuda.synchronize()


def test_tuner_picks_feasible_tile_and_caches():
    clips = _clips(3, seed=7)
    t = PixelTileTuner(iterations=1, repeats=2)
    for c in clips:
        start = c.state()
        tile = t.choose(c)
        assert tile in (64, 128, 256)
        assert torch.equal(c.poses, start[0]) and torch.equal(c.disps, start[1])  # state restored
        n_times = len(t.times)
        assert t.choose(c) == tile and len(t.times) == n_times  # cached
    b = BatchedBA(clips, iterations=ITERS, tuner=t)
    assert [r.plan.pixel_tile for r in b.runs] == [t.choose(c) for c in clips]
    expect = [_single(c, ITERS, t.choose(c)) for c in clips]
    b.run()
    torch.cuda.synchronize()
    for c, (p, d, dx, dz) in zip(clips, expect):
        assert torch.equal(c.poses, p) and torch.equal(c.disps, d)


def test_tuned_tile_matches_default_tile_closely():
    c = _clips(1, seed=9)[0]
    outs = [_single(c, ITERS, tile) for tile in (64, 128, 256)]
    torch.cuda.synchronize()
    for p, d, dx, dz in outs[1:]:
        assert (p - outs[0][0]).abs().max().item() < 1e-5
        assert (d - outs[0][1]).abs().max().item() < 1e-4


def test_tiled_plan_takes_every_tile_with_a_wide_frame():
    from vipe_amd.ext import slam_ext

    # frame 1 (inside the window [1, 24)) with 20 outgoing edges: more than the
    # 18 whose E rows stay in LDS, so it takes the wide-frame kernel and the
    # frame kernel's tile is chosen for the other frames alone
    N = 24
    ii = torch.ones(20, dtype=torch.long)
    jj = torch.arange(2, 22)
    for tile in (64, 128, 256):
        assert slam_ext.get_plan(ii.cuda(), jj.cuda(), N, 16, 32, 1, N, tile).pixel_tile == tile
    with pytest.raises(RuntimeError, match="pixel_tile"):
        slam_ext.get_plan(ii.cuda(), jj.cuda(), N, 16, 32, 1, N, 96)

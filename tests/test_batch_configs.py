"""The batched uneven workload's clip generator (host only)."""
from vipe_amd import synth


def test_batch_configs_follow_the_workload_definition():
    cfgs = synth.batch_configs(16, 0)
    assert len(cfgs) == 16
    res = {(r[0], r[1]) for r in synth.BATCH_RESOLUTIONS}
    for c in cfgs:
        assert c["n_frames"] in synth.BATCH_WINDOWS
        assert (c["H"], c["W"]) in res
        n = c["n_frames"]
        assert 1 <= c["window"] <= n - 1
        # window band + long-range edges, never more than a directed complete graph
        assert 0 < c["n_edges"] <= n * (n - 1)
        assert c["n_edges"] <= n * synth.BATCH_DEGREES[1]
    assert synth.batch_configs(16, 0) == cfgs  # seeded


def test_make_batch_small_clips_are_valid_graphs():
    gs = synth.make_batch(4, 1, windows=(8, 12), resolutions=((16, 24, 20.0, 20.0, 12.0, 8.0),))
    for g in gs:
        N = g.disps.shape[0]
        E = g.ii.numel()
        assert g.targets.shape == (E, 2, 16, 24) and g.weights.shape == (E, 2, 16, 24)
        assert bool(((g.ii >= 0) & (g.ii < N) & (g.jj >= 0) & (g.jj < N) & (g.ii != g.jj)).all())
        pairs = set(zip(g.ii.tolist(), g.jj.tolist()))
        assert len(pairs) == E  # no duplicate edges

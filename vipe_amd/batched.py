"""Independent BA clips in one graph launch, with a per-shape tile autotuner.

A batch holds clips of uneven size: different window lengths, out-degrees and
map resolutions.  Each clip is a full ``ba_extended`` problem (its own poses,
disparity maps, edges and window [t0, t1)).  ``BatchedBA`` prepares every
clip once (plan, workspace, outputs), then launches all clips' iterations with
the clips spread over a few streams.  ``capture()`` records that launch
sequence, forks and joins included, in one HIP graph, so a replay runs the
whole batch as one launch from the host.

``PixelTileTuner`` picks the frame kernel's pixels per pass (64, 128 or 256)
for each clip shape by timing the candidates on the clip itself.  It also
reports the per-clip time, which the batch uses to assign clips to streams
longest-first.  The reference runs one clip per ``ba`` call (geom_kernels.cu:
1626-1847) and has no batched form; the per-clip arithmetic is unchanged.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from .ext import slam_ext


@dataclass
class Clip:
    """One BA problem; tensors are contiguous float32 on the GPU (ii, jj int64)."""
    poses: torch.Tensor
    disps: torch.Tensor
    intrinsics: torch.Tensor
    disps_sens: torch.Tensor
    targets: torch.Tensor
    weights: torch.Tensor
    eta: torch.Tensor
    ii: torch.Tensor
    jj: torch.Tensor
    depth_active: torch.Tensor
    t0: int
    t1: int

    @classmethod
    def from_graph(cls, g, device) -> "Clip":
        return cls(poses=g.poses.to(device), disps=g.disps.to(device), intrinsics=g.intrinsics.to(device),
                   disps_sens=g.disps_sens.to(device), targets=g.targets.to(device), weights=g.weights.to(device),
                   eta=g.eta.to(device), ii=g.ii.to(device), jj=g.jj.to(device),
                   depth_active=torch.ones(g.disps.shape[0], device=device), t0=int(g.t0), t1=int(g.t1))

    def shape_key(self):
        N, H, W = self.disps.shape
        deg = torch.bincount(self.ii.cpu(), minlength=N)
        return (int(N), int(self.ii.numel()), int(H), int(W), self.t0, self.t1, int(deg.max()) if N else 0)

    def plan(self, pixel_tile: int = 0):
        N, H, W = self.disps.shape
        return slam_ext.get_plan(self.ii, self.jj, N, H, W, self.t0, self.t1, pixel_tile)

    def state(self):
        return self.poses.clone(), self.disps.clone()

    def restore(self, state):
        self.poses.copy_(state[0])
        self.disps.copy_(state[1])


class _Run:
    """A clip with its plan, workspace and outputs: what one launch needs."""

    def __init__(self, clip: Clip, pixel_tile: int):
        self.clip = clip
        self.plan = clip.plan(pixel_tile)
        dev = clip.poses.device
        H, W = clip.disps.shape[1:]
        self.ws = torch.empty((max(self.plan.ws_bytes, 1),), dtype=torch.uint8, device=dev)
        self.dx = torch.empty((clip.t1 - clip.t0, 6), dtype=torch.float32, device=dev)
        self.dz = torch.empty((self.plan.K, H * W), dtype=torch.float32, device=dev)

    def launch(self, iterations, lm, ep, alpha):
        c = self.clip
        slam_ext.run_planned(self.plan, self.ws, c.poses, c.disps, c.intrinsics, c.disps_sens, c.targets, c.weights,
                             c.eta, c.depth_active, iterations, lm, ep, alpha, self.dx, self.dz)


class PixelTileTuner:
    """Chooses the frame kernel's pixels per pass for each clip shape by timing
    ``iterations`` GN iterations per candidate (best of ``repeats``) on the clip
    (its poses and disparities are restored afterwards).  Results are cached
    by shape; ``times[key][tile]`` keeps every measurement in ms."""

    def __init__(self, candidates=(64, 128, 256), iterations: int = 2, repeats: int = 3, lm: float = 1e-4,
                 ep: float = 0.1, alpha: float = 0.05):
        self.candidates = tuple(candidates)
        self.iterations, self.repeats = iterations, repeats
        self.lm, self.ep, self.alpha = lm, ep, alpha
        self.best: dict = {}
        self.times: dict = {}

    def choose(self, clip: Clip) -> int:
        key = clip.shape_key()
        if key in self.best:
            return self.best[key]
        saved = clip.state()
        res = {}
        for tile in self.candidates:
            try:
                run = _Run(clip, tile)
            except RuntimeError:  # the tile's E rows do not fit the LDS
                continue
            run.launch(self.iterations, self.lm, self.ep, self.alpha)  # warm
            best = float("inf")
            for _ in range(self.repeats):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                run.launch(self.iterations, self.lm, self.ep, self.alpha)
                e1.record()
                e1.synchronize()
                best = min(best, e0.elapsed_time(e1))
            res[tile] = best / self.iterations
            clip.restore(saved)
        if not res:
            raise RuntimeError(f"no pixel tile fits clip shape {key}")
        self.times[key] = res
        self.best[key] = min(res, key=res.get)
        return self.best[key]

    def ms_per_iteration(self, clip: Clip) -> float:
        key = clip.shape_key()
        return self.times[key][self.best[key]]


class BatchedBA:
    """``iterations`` GN iterations of every clip, the clips spread over
    ``n_streams`` streams (longest first onto the least loaded stream)."""

    def __init__(self, clips, iterations: int = 8, lm: float = 1e-4, ep: float = 0.1, alpha: float = 0.05,
                 n_streams: int = 4, tuner: PixelTileTuner | None = None):
        self.clips = list(clips)
        self.iterations, self.lm, self.ep, self.alpha = iterations, lm, ep, alpha
        tiles = [tuner.choose(c) if tuner else 0 for c in self.clips]
        self.runs = [_Run(c, t) for c, t in zip(self.clips, tiles)]
        cost = [tuner.ms_per_iteration(c) if tuner else float(c.ii.numel() * c.disps[0].numel())
                for c in self.clips]
        self.n_streams = max(1, min(n_streams, len(self.clips)))
        self.streams = [torch.cuda.Stream() for _ in range(self.n_streams)]
        load = [0.0] * self.n_streams
        self.assignment = [[] for _ in range(self.n_streams)]
        for n in sorted(range(len(self.clips)), key=lambda n: -cost[n]):
            s = min(range(self.n_streams), key=lambda i: load[i])
            self.assignment[s].append(n)
            load[s] += cost[n]
        self.graph = None

    def _launch_all(self):
        main = torch.cuda.current_stream()
        for s in self.streams:
            s.wait_stream(main)
        for s, idx in zip(self.streams, self.assignment):
            with torch.cuda.stream(s):
                for n in idx:
                    self.runs[n].launch(self.iterations, self.lm, self.ep, self.alpha)
        for s in self.streams:
            main.wait_stream(s)

    def run(self):
        """Eager launch of the whole batch on the current stream's timeline."""
        self._launch_all()

    def capture(self):
        """Record the batch launch into one graph (run once eagerly first, so that
        every plan and buffer exists)."""
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._launch_all()
        return self.graph

    def replay(self):
        if self.graph is None:
            raise RuntimeError("BatchedBA.replay: capture() first")
        self.graph.replay()

    @property
    def dx(self):
        return [r.dx for r in self.runs]

    @property
    def dz(self):
        return [r.dz for r in self.runs]


def main(argv=None) -> None:
    """Time the 16-clip uneven batch: ``python -m vipe_amd.batched [--clips 16]
    [--iterations 8] [--replays 10] [--streams 4]``.  Prints one JSON line."""
    import argparse
    import json
    import time

    from . import synth

    ap = argparse.ArgumentParser()
    ap.add_argument("--clips", type=int, default=16)
    ap.add_argument("--iterations", type=int, default=8)
    ap.add_argument("--replays", type=int, default=10)
    ap.add_argument("--streams", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    dev = torch.device("cuda", 0)
    t = time.perf_counter()
    graphs = synth.make_batch(a.clips, a.seed)
    clips = [Clip.from_graph(g, dev) for g in graphs]
    t_gen = time.perf_counter() - t
    t = time.perf_counter()
    tuner = PixelTileTuner()
    batch = BatchedBA(clips, iterations=a.iterations, n_streams=a.streams, tuner=tuner)
    torch.cuda.synchronize()
    t_tune = time.perf_counter() - t
    start = [c.state() for c in clips]
    batch.run()  # warm
    torch.cuda.synchronize()

    def timed(fn, n):
        ms = []
        for _ in range(n):
            for c, s in zip(clips, start):
                c.restore(s)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            ms.append(e0.elapsed_time(e1))
        ms.sort()
        return ms[len(ms) // 2]

    eager_ms = timed(batch.run, a.replays)
    for c, s in zip(clips, start):
        c.restore(s)
    batch.capture()
    graph_ms = timed(batch.replay, a.replays)
    serial_ms = sum(tuner.ms_per_iteration(c) for c in clips) * a.iterations
    edges = sum(int(c.ii.numel()) for c in clips)
    print(json.dumps({
        "workload": f"{a.clips} clips, windows {sorted({int(c.disps.shape[0]) for c in clips})}, "
                    f"{edges} edges in all, {a.iterations} GN iterations each",
        "clips": [dict(frames=int(c.disps.shape[0]), edges=int(c.ii.numel()), hw=list(c.disps.shape[1:]),
                       tile=r.plan.pixel_tile, ms_per_iter=round(tuner.ms_per_iteration(c), 4))
                  for c, r in zip(clips, batch.runs)],
        "graph_replay_ms": round(graph_ms, 3),
        "eager_multistream_ms": round(eager_ms, 3),
        "sum_of_single_clip_ms": round(serial_ms, 3),
        "clip_iterations_per_s": round(a.clips * a.iterations / (graph_ms * 1e-3), 1),
        "streams": batch.n_streams,
        "generate_s": round(t_gen, 2), "tune_s": round(t_tune, 2),
    }), flush=True)


if __name__ == "__main__":
    main()

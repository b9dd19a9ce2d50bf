"""Pins oracle/geom.py: rotation-matrix forms of the same maps, closed forms,
and a per-pixel loop of depth_filter on small inputs (host only)."""
import math

import pytest
import torch

from oracle import geom as OG
from vipe_amd import synth


def _rotmat(q):
    """Rotation matrix of unit quaternion (x, y, z, w), written out entry by entry."""
    x, y, z, w = [float(v) for v in q]
    return torch.tensor([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                         [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                         [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]], dtype=torch.float64)


def _graph():
    return synth.make_graph("kf8_e24_48x64", seed=4, with_fmaps=False)


def test_relative_pose_composes():
    g = _graph()
    P = g.poses.double()
    p = torch.randn(3, dtype=torch.float64)
    for a, b in ((0, 3), (5, 1), (2, 2)):
        t, q = OG.relative(P[a], P[b])
        Ra, Rb, Rr = _rotmat(P[a, 3:7]), _rotmat(P[b, 3:7]), _rotmat(q)
        lhs = Rr @ (Ra @ p + P[a, :3]) + t
        rhs = Rb @ p + P[b, :3]
        assert torch.allclose(lhs, rhs, atol=1e-6)


def test_iproj_equals_matrix_form():
    g = _graph()
    pts = OG.iproj(g.poses, g.disps, g.intrinsics)
    fx, fy, cx, cy = [float(v) for v in g.intrinsics]
    N, H, W = g.disps.shape
    for n in (0, 5):
        R = _rotmat(g.poses[n, 3:7].double())
        t = g.poses[n, :3].double()
        for (v, u) in ((0, 0), (7, 33), (H - 1, W - 1)):
            d = float(g.disps[n, v, u])
            X = torch.tensor([(u - cx) / fx, (v - cy) / fy, 1.0], dtype=torch.float64)
            expect = (R @ X + t * d) / d
            assert torch.allclose(pts[n, v, u], expect, atol=1e-6)


def test_projmap_equals_matrix_form_and_defaults():
    g = _graph()
    coords, valid = OG.projmap(g.poses, g.disps, g.intrinsics, g.ii, g.jj)
    fx, fy, cx, cy = [float(v) for v in g.intrinsics]
    P = g.poses.double()
    assert torch.all(coords[..., 2] == 0)
    for e in (0, 7, 20):
        a, b = int(g.ii[e]), int(g.jj[e])
        Ra, Rb = _rotmat(P[a, 3:7]), _rotmat(P[b, 3:7])
        R = Rb @ Ra.T
        t = P[b, :3] - R @ P[a, :3]
        for (v, u) in ((3, 4), (20, 60), (47, 0)):
            d = float(g.disps[a, v, u])
            Y = R @ torch.tensor([(u - cx) / fx, (v - cy) / fy, 1.0], dtype=torch.float64) + t * d
            # float32 quaternions are unit only to ~1e-8, which the quaternion and
            # matrix forms absorb differently
            if Y[2] > 0.01:
                assert abs(coords[e, v, u, 0] - (fx * Y[0] / Y[2] + cx)) < 1e-6
                assert abs(coords[e, v, u, 1] - (fy * Y[1] / Y[2] + cy)) < 1e-6
            assert valid[e, v, u, 0] == (1.0 if Y[2] > 0.25 else 0.0)
    # a point pushed behind the camera keeps its own pixel coordinates
    poses = torch.tensor([[0, 0, 0, 0, 0, 0, 1], [0, 0, -5.0, 0, 0, 0, 1]], dtype=torch.float32)
    disps = torch.ones(2, 3, 4)
    c, vld = OG.projmap(poses, disps, torch.tensor([2.0, 2.0, 2.0, 1.5]), torch.tensor([0]), torch.tensor([1]))
    assert torch.equal(c[0, 1, 2, :2], torch.tensor([2.0, 1.0], dtype=torch.float64))
    assert torch.all(vld == 0)


def test_frame_distance_closed_forms():
    H, W = 6, 8
    intr = torch.tensor([[10.0, 10.0, 4.0, 3.0]])
    disps = torch.full((3, H, W), 0.5)
    tx = 0.2
    poses = torch.tensor([[0, 0, 0, 0, 0, 0, 1], [tx, 0, 0, 0, 0, 0, 1], [0, 0, -3.0, 0, 0, 0, 1]],
                         dtype=torch.float32)
    z = torch.zeros(3, dtype=torch.long)
    pi = torch.tensor([0, 0, 0])
    pj = torch.tensor([0, 1, 2])
    d = OG.frame_distance(poses, disps, intr, pi, pj, z, z, z, 0.7)
    assert d[0] == 0.0                                   # same frame: no flow
    assert abs(d[1] - 10.0 * 0.5 * tx) < 1e-6           # pure x translation: fx * d * tx everywhere
    assert d[2] == 1000.0                                # every point behind MIN_DEPTH


def _depth_filter_loop(poses, disps, intr, ix, thresh):
    """Per-pixel loop restatement of the same rule (small inputs only)."""
    P = poses.double()
    D = disps.double()
    N, H, W = D.shape
    fx, fy, cx, cy = [float(v) for v in intr]
    out = torch.zeros(len(ix), H, W, dtype=torch.float64)
    for b, i in enumerate(ix.tolist()):
        for n in range(6):
            j = i - n - 1 if n < 3 else i + n - 2
            if not 0 <= j < N:
                continue
            Ri, Rj = _rotmat(P[i, 3:7]), _rotmat(P[j, 3:7])
            R = Rj @ Ri.T
            t = P[j, :3] - R @ P[i, :3]
            for v in range(H):
                for u in range(W):
                    d = float(D[i, v, u])
                    Y = R @ torch.tensor([(u - cx) / fx, (v - cy) / fy, 1.0], dtype=torch.float64) + t * d
                    uj, vj, dj = fx * Y[0] / Y[2] + cx, fy * Y[1] / Y[2] + cy, d / Y[2]
                    u0, v0 = math.floor(uj), math.floor(vj)
                    if not (0 <= u0 < W - 1 and 0 <= v0 < H - 1):
                        continue
                    corners = (D[j, v0, u0], D[j, v0, u0 + 1], D[j, v0 + 1, u0], D[j, v0 + 1, u0 + 1])
                    if any(abs(1.0 / dj - 1.0 / float(c)) < float(thresh[b]) for c in corners):
                        out[b, v, u] += 1
    return out


def test_depth_filter_constant_scene_counts_neighbours():
    N, H, W = 7, 5, 6
    poses = torch.tensor([[0, 0, 0, 0, 0, 0, 1]] * N, dtype=torch.float32)
    disps = torch.full((N, H, W), 0.4)
    ix = torch.tensor([0, 3, 6])
    cnt = OG.depth_filter(poses, disps, torch.tensor([5.0, 5.0, 3.0, 2.0]), ix, torch.full((3,), 0.01))
    # identical frames: a pixel lands on itself; it counts for every neighbour in
    # range unless it sits in the last row or column
    for b, i in enumerate(ix.tolist()):
        n_in = sum(1 for n in range(6) if 0 <= (i - n - 1 if n < 3 else i + n - 2) < N)
        assert torch.all(cnt[b, :H - 1, :W - 1] == n_in)
        assert torch.all(cnt[b, H - 1, :] == 0) and torch.all(cnt[b, :, W - 1] == 0)


def test_depth_filter_matches_per_pixel_loop():
    g = synth.make_graph("kf8_e24_48x64", seed=6, with_fmaps=False)
    poses = g.poses.clone()
    poses[:, :3] *= 0.2  # small baselines so that reprojections stay on the map
    disps = g.disps[:, ::6, ::6].contiguous()
    intr = g.intrinsics / 6.0
    ix = torch.tensor([0, 4, 7])
    thresh = torch.tensor([0.3, 0.5, 0.8])
    fast = OG.depth_filter(poses, disps, intr, ix, thresh)
    slow = _depth_filter_loop(poses, disps, intr, ix, thresh)
    assert torch.equal(fast, slow)
    assert fast.sum() > 0


def test_geometry_wrappers_reject_bad_inputs_on_host():
    from vipe_amd.ext import slam_ext

    poses = torch.zeros(4, 7)
    disps = torch.ones(4, 6, 8)
    intr = torch.ones(4)
    with pytest.raises(RuntimeError, match="contiguous"):
        slam_ext.iproj(poses, disps.transpose(1, 2), intr)
    with pytest.raises(RuntimeError, match="GPU"):
        slam_ext.iproj(poses, disps, intr)

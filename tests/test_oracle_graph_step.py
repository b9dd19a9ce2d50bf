"""Pins oracle/graph_step.py (host only): the reprojection against the
rotation-matrix form, identity and pure-translation closed forms, the z < 0.1
rule, the clamp of the motion features, and the epilogue's mask."""
import torch

from oracle import graph_step as OG
from vipe_amd.synth import make_graph


def _rotmat(q):
    x, y, z, w = [float(v) for v in q]
    return torch.tensor([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                         [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                         [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]], dtype=torch.float64)


def test_reprojection_equals_matrix_form():
    g = make_graph("kf8_e24_48x64", seed=4, with_fmaps=False)
    c1 = OG.reproject(g.poses, g.disps, g.intrinsics, g.ii, g.jj)
    fx, fy, cx, cy = [float(v) for v in g.intrinsics]
    P = g.poses.double()
    for e in (0, 5, 17):
        a, b = int(g.ii[e]), int(g.jj[e])
        R = _rotmat(P[b, 3:7]) @ _rotmat(P[a, 3:7]).T
        t = P[b, :3] - R @ P[a, :3]
        for (v, u) in ((0, 0), (11, 40), (47, 63)):
            d = float(g.disps[a, v, u])
            X1 = R @ torch.tensor([(u - cx) / fx, (v - cy) / fy, 1.0], dtype=torch.float64) + t * d
            z = X1[2] if X1[2] >= 0.1 else 1.0
            assert abs(c1[e, v, u, 0] - (fx * X1[0] / z + cx)) < 1e-6
            assert abs(c1[e, v, u, 1] - (fy * X1[1] / z + cy)) < 1e-6


def test_closed_forms_and_clamp():
    H, W = 5, 7
    intr = torch.tensor([10.0, 10.0, 3.0, 2.0])
    disps = torch.full((3, H, W), 0.5)
    tx = 0.3
    poses = torch.tensor([[0, 0, 0, 0, 0, 0, 1], [tx, 0, 0, 0, 0, 0, 1], [0, 0, -5.0, 0, 0, 0, 1]], dtype=torch.float32)
    ii, jj = torch.tensor([0, 0, 0]), torch.tensor([0, 1, 2])
    target = torch.zeros(3, H, W, 2)
    target[..., 0] = 1000.0
    c1, motn = OG.motion_features(poses, disps, intr, ii, jj, target)
    v, u = torch.meshgrid(torch.arange(H, dtype=torch.float64), torch.arange(W, dtype=torch.float64), indexing="ij")
    assert torch.allclose(c1[0, ..., 0], u) and torch.allclose(c1[0, ..., 1], v)           # same frame: the grid
    assert torch.allclose(c1[1, ..., 0], u + 10.0 * 0.5 * tx)                                # x shift fx d tx
    # frame 2 is 5 units behind: z = 1 - 2.5 < 0.1 projects with z = 1
    X = (u - 3.0) / 10.0
    assert torch.allclose(c1[2, ..., 0], 10.0 * X + 3.0)
    assert motn.shape == (3, 4, H, W) and motn[:, 2].max() == 64.0                         # clamped
    assert torch.allclose(motn[1, 0], torch.full((H, W), 1.5, dtype=torch.float64))
    t, w = OG.targets(c1, torch.ones_like(c1), torch.ones_like(c1), torch.tensor([[[0, 1] * 3 + [0]] * H, [[0] * W] * H], dtype=torch.uint8), torch.tensor([0, 1, 0]))
    assert torch.equal(t, c1 + 1) and w[0, 0, 1].sum() == 0 and w[1].min() == 1.0

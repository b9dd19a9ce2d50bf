"""gfx950 geometry siblings (iproj, projmap, frame_distance, depth_filter) vs
the FP64 oracle on seeded frame graphs.  Threshold decisions (validity, the
depth_filter hits, the 0.75 rule of frame_distance) may differ from FP64 only
where the FP64 value lies within FP32 rounding of the threshold."""
import pytest
import torch

from oracle import geom as OG
from vipe_amd import synth

pytestmark = pytest.mark.gpu


def _graph(name="kf8_e24_48x64", seed=2):
    return synth.make_graph(name, seed=seed, with_fmaps=False)


def test_iproj_matches_oracle():
    from vipe_amd.ext import slam_ext

    g = _graph()
    out = slam_ext.iproj(g.poses.cuda(), g.disps.cuda(), g.intrinsics.cuda()).cpu()
    ref = OG.iproj(g.poses, g.disps, g.intrinsics)
    assert out.shape == ref.shape
    assert (out.double() - ref).abs().max().item() < 1e-5 * max(1.0, ref.abs().max().item())


@pytest.mark.parametrize("name", ["kf8_e24_48x64", "kf32_e192_48x64"])
def test_projmap_matches_oracle(name):
    from vipe_amd.ext import slam_ext

    g = _graph(name)
    coords, valid = slam_ext.projmap(g.poses.cuda(), g.disps.cuda(), g.intrinsics.cuda(), g.ii.cuda(), g.jj.cuda())
    rc, rv = OG.projmap(g.poses, g.disps, g.intrinsics, g.ii, g.jj)
    coords, valid = coords.cpu().double(), valid.cpu().double()
    assert coords.shape == rc.shape and valid.shape == rv.shape
    assert torch.all(coords[..., 2] == 0)
    # z of every pixel, to locate the threshold cases
    _, z = synth.reproject(g.poses, g.disps, g.intrinsics, g.ii, g.jj)
    near = ((z - 0.25).abs() < 1e-5) | ((z - 0.01).abs() < 1e-5)
    assert torch.equal(valid[..., 0][~near], rv[..., 0][~near])
    # FP32 projection: the relative error of u = fx x / z + cx grows like 1 / z
    far = ~near & (z > 0.1)
    err = (coords[..., :2] - rc[..., :2]).abs()[far]
    scale = rc[..., :2].abs()[far].clamp(min=1.0)
    assert (err / scale).max().item() < 3e-5


def test_frame_distance_matches_oracle():
    from vipe_amd.ext import slam_ext

    g = _graph("kf32_e192_48x64", seed=5)
    N = g.disps.shape[0]
    gen = torch.Generator().manual_seed(0)
    num = 300
    pi = torch.randint(0, N, (num,), generator=gen)
    pj = torch.randint(0, N, (num,), generator=gen)
    di = torch.where(torch.rand(num, generator=gen) < 0.5, pi, pj)
    intr = torch.stack([g.intrinsics, g.intrinsics * torch.tensor([1.1, 1.1, 1.0, 1.0])])
    qi = torch.randint(0, 2, (num,), generator=gen)
    qj = torch.randint(0, 2, (num,), generator=gen)
    # a few pairs far apart push most points behind the camera (the 1000 rule)
    poses = g.poses.clone()
    poses[N - 1, 2] = -4.0
    pj[:10] = N - 1
    d = slam_ext.frame_distance(poses.cuda(), g.disps.cuda(), intr.cuda(), pi.cuda(), pj.cuda(), qi.cuda(),
                                qj.cuda(), di.cuda(), 0.6).cpu().double()
    r = OG.frame_distance(poses, g.disps, intr, pi, pj, qi, qj, di, 0.6)
    big = r == 1000.0
    assert big.any() and (~big).any()
    assert torch.equal(d[big], r[big])
    # relative error, with an absolute floor of 1e-4 px for pairs with (almost) no flow
    assert ((d[~big] - r[~big]).abs() / r[~big].clamp(min=1.0)).max().item() < 1e-4


def test_depth_filter_matches_oracle():
    from vipe_amd.ext import slam_ext

    g = _graph("kf8_e24_48x64", seed=7)
    poses = g.poses.clone()
    poses[:, :3] *= 0.3  # keep reprojections on the map so that most pixels are tested
    ix = torch.tensor([0, 1, 3, 4, 7])
    thresh = torch.tensor([0.05, 0.2, 0.5, 1.0, 2.0])
    out = slam_ext.depth_filter(poses.cuda(), g.disps.cuda(), g.intrinsics.cuda(), ix.cuda(), thresh.cuda()).cpu()
    ref = OG.depth_filter(poses, g.disps, g.intrinsics, ix, thresh)
    assert out.shape == ref.shape
    mism = (out.double() != ref)
    assert ref.sum() > 0 and (ref < ref.max()).any()
    # only threshold-boundary pixels may differ from the FP64 decision
    assert mism.double().mean().item() < 2e-3, mism.sum().item()
    assert (out.double() - ref).abs().max().item() <= 1


def test_geometry_out_of_range_indices_are_flagged_not_read():
    from vipe_amd.ext import slam_ext

    g = _graph()
    N = g.disps.shape[0]
    ii = torch.tensor([0, N + 5], device="cuda")
    jj = torch.tensor([1, 0], device="cuda")
    coords, valid = slam_ext.projmap(g.poses.cuda(), g.disps.cuda(), g.intrinsics.cuda(), ii, jj)
    torch.cuda.synchronize()
    assert torch.isnan(coords[1]).all() and (valid[1] == 0).all()
    assert torch.isfinite(coords[0]).all()
    cnt = slam_ext.depth_filter(g.poses.cuda(), g.disps.cuda(), g.intrinsics.cuda(),
                                torch.tensor([-1, 2], device="cuda"), torch.ones(2, device="cuda"))
    assert (cnt[0] == 0).all()

"""vh_ba_run_static (the allocation-free BA entry point) against vh_ba_run with
a caller-allocated workspace, through the C ABI: two calls in a row, the
second on a smaller graph, each bit-identical to the allocating entry point
(the kernels and their reduction order are the same).  Also the size limit."""
import ctypes as C

import pytest
import torch

from vipe_amd.synth import make_graph

pytestmark = pytest.mark.gpu


def _inputs(name):
    g = make_graph(name, with_fmaps=False)
    N = g.disps.shape[0]
    ds = g.disps.clone() * 1.01
    ds[:, ::2] = 0.0
    d = {k: (v.cuda() if torch.is_tensor(v) else v) for k, v in g.__dict__.items()}
    d["disps_sens"] = ds.cuda()
    d["depth_active"] = torch.ones(N, device="cuda")
    return g, d


def _call(lib, plan, d, iters, flags, workspace):
    """One vh_ba_run (workspace tensor) or vh_ba_run_static (workspace None) call on fresh copies."""
    from vipe_amd import _lib

    P = d["t1"] - d["t0"]
    poses, disps, intr = d["poses"].clone(), d["disps"].clone(), d["intrinsics"].clone()
    dx = torch.full((P, 6), float("nan"), device="cuda")
    dz = torch.full((plan.K, d["disps"].shape[1] * d["disps"].shape[2]), float("nan"), device="cuda")
    en = torch.full((iters,), float("nan"), device="cuda")
    args = [plan.h, _lib.ptr(poses), _lib.ptr(disps), _lib.ptr(intr), _lib.ptr(d["disps_sens"]),
            _lib.ptr(d["targets"]), _lib.ptr(d["weights"]), _lib.ptr(d["eta"]), _lib.ptr(d["depth_active"]),
            iters, 1e-4, 0.1, 0.05, 1e-6, 1e-6, 0.125, flags, _lib.ptr(dx), _lib.ptr(dz), _lib.ptr(en)]
    stream = _lib.stream_of(poses)
    if workspace is None:
        rc = lib.vh_ba_run_static(*args, stream)
    else:
        rc = lib.vh_ba_run(*args, _lib.ptr(workspace), stream)
    _lib.check(rc, "ba_run" if workspace is not None else "ba_run_static")
    torch.cuda.synchronize()
    return poses.cpu(), disps.cpu(), intr.cpu(), dx.cpu(), dz.cpu(), en.cpu()


@pytest.mark.parametrize("flags", [4, 4 | 2])  # energy; energy + intrinsics
def test_static_entry_point_matches_allocating_one(flags):
    from vipe_amd import _lib
    from vipe_amd.ext import slam_ext

    lib = _lib.load()
    # the larger graph first, then a smaller one: the second call must not see
    # anything the first left in the shared array
    for name, iters in (("kf64_e384_64x80", 3), ("kf8_e24_48x64", 2)):
        g, d = _inputs(name)
        print(f"[no-alloc] {name} iters={iters} flags={flags}", flush=True)
        N, H, W = g.disps.shape
        plan = slam_ext.get_plan(g.ii, g.jj, N, H, W, g.t0, g.t1)
        assert plan.ws_bytes <= (64 << 20)
        ws = torch.empty(plan.ws_bytes, dtype=torch.uint8, device="cuda")
        want = _call(lib, plan, d, iters, flags, ws)
        got = _call(lib, plan, d, iters, flags, None)
        for a, b in zip(got, want):
            assert not torch.isnan(b).any()
            assert torch.equal(a, b)
        # and again on the same array, immediately after
        again = _call(lib, plan, d, iters, flags, None)
        for a, b in zip(again, want):
            assert torch.equal(a, b)


def test_static_entry_point_rejects_plans_over_the_limit():
    from vipe_amd import _lib
    from vipe_amd.ext import slam_ext

    lib = _lib.load()
    # 8 frames of 1024 x 1280 pixels: about 130 MB of workspace
    ii = torch.tensor([0, 1, 2, 3, 4, 5, 6, 1, 2, 3])
    jj = torch.tensor([1, 2, 3, 4, 5, 6, 7, 0, 1, 2])
    plan = slam_ext.get_plan(ii, jj, 8, 1024, 1280, 1, 8)
    assert plan.ws_bytes > (64 << 20)
    t = torch.zeros(16, device="cuda")
    p = [_lib.ptr(t)] * 8
    rc = lib.vh_ba_run_static(plan.h, *p, 1, 1e-4, 0.1, 0.05, 1e-6, 1e-6, 1.0, 0, _lib.ptr(t), _lib.ptr(t), None,
                              _lib.stream_of(t))
    assert rc == _lib.VH_ERR_UNSUPPORTED
    assert "static limit" in _lib.last_error()

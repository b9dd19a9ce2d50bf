"""The update operator's HIP kernels against FP64 references.

Layer tests give each kernel f16 operands and compare with an FP64 CPU
convolution of the same f16 values, so only the FP32 accumulation order and
the output rounding differ.  Module tests compare UpdateModule.forward (FP16
activations, as the reference runs under autocast) with oracle/update_op.py in
FP64 on the same inputs and parameters.  Tolerances are stated per test."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import update_op as OU

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _lib():
    from vipe_amd import _lib as L

    return L


def _h(x):
    return x.to(torch.float16)


def _ref_conv(x_nchw, w, b, glo=None):
    """FP64 CPU convolution; x [n, C, H, W], w [Co, C, k, k]."""
    x, w, b = (t.detach().double().cpu() for t in (x_nchw, w, b))
    y = F.conv2d(x, w, b, padding=w.shape[2] // 2)
    if glo is not None:
        y = y + glo.detach().double().cpu()[:, :, None, None]
    return y


def _runner():
    from vipe_amd.update_op import _Runner

    return _Runner(torch.device(DEV))


def _packed(w, b, cin_real, cin_pad):
    from vipe_amd.update_op import _pack_conv

    conv = torch.nn.Conv2d(w.shape[1], w.shape[0], w.shape[2], padding=w.shape[2] // 2)
    conv.weight.data.copy_(w.float())
    conv.bias.data.copy_(b.float())
    return _pack_conv([conv], cin_real, cin_pad, torch.device(DEV))


def _segments(g, nimg, H, W, chans, wide):
    """NHWC f16 segments; with `wide`, each lives at channel offset 16 of a
    buffer 32 channels wider (exercises seg_off / seg_stride)."""
    P = nimg * H * W
    segs, parts = [], []
    for c in chans:
        if wide:
            buf = _h(torch.randn(P, c + 32, generator=g)).to(DEV)
            segs.append((buf, c, 16))
            parts.append(buf[:, 16:16 + c])
        else:
            buf = _h(torch.randn(P, c, generator=g)).to(DEV)
            segs.append((buf, c, 0))
            parts.append(buf)
    x = torch.cat(parts, 1).float().view(nimg, H * W, -1).permute(0, 2, 1).reshape(nimg, -1, H, W)
    return segs, x


CASES = [
    # (kernel, input segment channels, Cout, epilogue name, nimg, H, W, wide segments)
    (1, [224], 128, "relu", 3, 7, 9, False),
    (3, [128], 128, "relu", 2, 12, 20, True),
    (7, [32], 128, "relu", 3, 7, 9, False),
    (3, [128], 64, "relu", 2, 5, 3, False),
    (3, [128, 128, 128, 64], 256, "none", 2, 9, 11, True),
    (3, [128], 256, "sigmoid", 1, 4, 4, False),
    (1, [128], 576, "none", 2, 6, 10, False),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"k{c[0]}_c{'-'.join(map(str, c[1]))}_o{c[2]}_{c[3]}")
def test_conv_layer_matches_fp64(case):
    """Tolerance: |err| <= 2e-3 |ref|max + 1e-3 |ref| (f16 output rounding is
    2^-11 relative; FP32 accumulation of <= 4032 products adds ~1e-6)."""
    L = _lib()
    k, chans, cout, epi, nimg, H, W, wide = case
    print(f"conv case {case}", flush=True)
    g = torch.Generator().manual_seed(sum(chans) + k + cout)
    segs, x = _segments(g, nimg, H, W, chans, wide)
    cin = sum(chans)
    w = _h(torch.randn(cout, cin, k, k, generator=g) / np.sqrt(cin * k * k))
    b = torch.randn(cout, generator=g) * 0.1
    pk = _packed(w, b, chans, chans)
    code = {"none": L.VH_UO_NONE, "relu": L.VH_UO_RELU, "sigmoid": L.VH_UO_SIGMOID}[epi]
    r = _runner()
    out = r.conv(segs, pk, nimg, H, W, code)
    torch.cuda.synchronize()
    ref = _ref_conv(x, w, b)
    if epi == "relu":
        ref = torch.relu(ref)
    elif epi == "sigmoid":
        ref = torch.sigmoid(ref)
    got = out.double().cpu().view(nimg, H * W, cout).permute(0, 2, 1).reshape(nimg, cout, H, W)
    err = (got - ref).abs()
    tol = 2e-3 * ref.abs().max() + 1e-3 * ref.abs()
    print(f"  max err {err.max().item():.3e} (ref max {ref.abs().max().item():.3f})", flush=True)
    assert bool((err <= tol).all())


def test_conv_nchw_output_with_offset():
    """The upmask form: channels written to [img][out_off + n][pixel]."""
    L = _lib()
    g = torch.Generator().manual_seed(11)
    nimg, H, W, cout = 3, 5, 7, 576
    segs, x = _segments(g, nimg, H, W, [128], False)
    w = _h(torch.randn(cout, 128, 1, 1, generator=g) / 12)
    b = torch.randn(cout, generator=g)
    pk = _packed(w, b, [128], [128])
    out = torch.full((nimg, 600, H, W), 7.0, device=DEV, dtype=torch.float16)
    _runner().conv(segs, pk, nimg, H, W, L.VH_UO_NONE | L.VH_UO_NCHW, out=out, out_stride=600, out_off=10)
    torch.cuda.synchronize()
    ref = _ref_conv(x, w, b)
    got = out.double().cpu()
    assert bool((got[:, :10] == 7).all()) and bool((got[:, 586:] == 7).all())
    err = (got[:, 10:586] - ref).abs()
    assert bool((err <= 2e-3 * ref.abs().max() + 1e-3 * ref.abs()).all())


def test_gate_epilogues_match_fp64():
    """SIGMUL (f32 out), GATES (z and r * net) and GRU, with a per-image glo
    addend, as ConvGRU chains them (droid_net.py:413-419)."""
    L = _lib()
    g = torch.Generator().manual_seed(12)
    nimg, H, W = 3, 6, 9
    P, HW = nimg * H * W, H * W
    net = _h(torch.tanh(torch.randn(P, 128, generator=g))).to(DEV)
    segs, x = _segments(g, nimg, H, W, [128, 128, 64], True)
    net_nchw = net.float().view(nimg, HW, 128).permute(0, 2, 1).reshape(nimg, 128, H, W)
    r = _runner()
    # SIGMUL: sigmoid(w net + b) * net, f32
    w1 = _h(torch.randn(128, 128, 1, 1, generator=g) / 11)
    b1 = torch.randn(128, generator=g) * 0.1
    sg = r.conv([(net, 128, 0)], _packed(w1, b1, [128], [128]), nimg, H, W, L.VH_UO_SIGMUL, out_f32=True,
                 aux0=net, aux_stride=128)
    ref = torch.sigmoid(_ref_conv(net_nchw, w1, b1)) * net_nchw.double().cpu()
    got = sg.double().cpu().view(nimg, HW, 128).permute(0, 2, 1).reshape(nimg, 128, H, W)
    assert float((got - ref).abs().max()) < 1e-5
    # GATES: [z; r] from [net, inp, corr, flow] + glo
    wzr = _h(torch.randn(256, 448, 3, 3, generator=g) / 63)
    bzr = torch.randn(256, generator=g) * 0.1
    glo = (torch.randn(nimg, 256, generator=g) * 0.3).to(DEV)
    z = torch.empty(P, 128, device=DEV, dtype=torch.float16)
    rnet = torch.empty(P, 128, device=DEV, dtype=torch.float16)
    r.conv([(net, 128, 0)] + segs, _packed(wzr, bzr, [128, 128, 128, 64], [128, 128, 128, 64]), nimg, H, W,
           L.VH_UO_GATES, out=z, out_stride=128, glo=glo, glo_stride=256, aux0=net, aux_stride=128, out2=rnet)
    xin = torch.cat([net_nchw, x], 1)
    gates = torch.sigmoid(_ref_conv(xin, wzr, bzr, glo))
    zr = gates[:, :128]
    rr = gates[:, 128:] * net_nchw.double().cpu()

    def nchw(t):
        return t.double().cpu().view(nimg, HW, 128).permute(0, 2, 1).reshape(nimg, 128, H, W)

    assert float((nchw(z) - zr).abs().max()) < 2e-3
    assert float((nchw(rnet) - rr).abs().max()) < 2e-3
    # GRU: (1 - z) net + z tanh(conv([r net, inp, corr, flow]) + glo_q), on the kernel's own z and r * net
    wq = _h(torch.randn(128, 448, 3, 3, generator=g) / 63)
    bq = torch.randn(128, generator=g) * 0.1
    gq = (torch.randn(nimg, 128, generator=g) * 0.3).to(DEV)
    out = r.conv([(rnet, 128, 0)] + segs, _packed(wq, bq, [128, 128, 128, 64], [128, 128, 128, 64]), nimg, H, W,
                 L.VH_UO_GRU, glo=gq, glo_stride=128, aux0=z, aux1=net, aux_stride=128)
    torch.cuda.synchronize()
    q = torch.tanh(_ref_conv(torch.cat([nchw(rnet), x.double().cpu()], 1), wq, bq, gq))
    zz = nchw(z)
    ref = (1 - zz) * net_nchw.double().cpu() + zz * q
    assert float((nchw(out) - ref).abs().max()) < 2e-3


SMALL_CASES = [
    # (kernel, cin, in_off, in_stride, Cout, epilogue, nimg, H, W)
    (3, 128, 0, 256, 2, "none", 3, 7, 9),
    (3, 128, 128, 256, 2, "sigmoid", 2, 12, 20),
    (3, 128, 0, 128, 1, "eta", 5, 3, 5),
    (1, 64, 8, 80, 2, "relu", 1, 11, 13),
    (3, 200, 0, 200, 1, "none", 2, 9, 10),
    (3, 256, 0, 256, 2, "none", 1, 64, 80),
]


@pytest.mark.parametrize("wshift", [0, 1], ids=["wvec", "wscalar"])
@pytest.mark.parametrize("case", SMALL_CASES, ids=lambda c: f"k{c[0]}_c{c[1]}@{c[2]}_o{c[4]}_{c[5]}_{c[7]}x{c[8]}")
def test_small_conv_heads_match_fp64(case, wshift):
    """The 1- and 2-channel heads against an FP64 convolution of the same f16
    inputs: only the FP32 accumulation order differs, so |err| <= 1e-5 |ref|max
    + 1e-5 (sigmoid / eta outputs compared after the activation)."""
    L = _lib()
    from vipe_amd.update_op import _vp

    k, cin, off, stride, cout, epi, nimg, H, W = case
    print(f"small conv case {case}", flush=True)
    g = torch.Generator().manual_seed(cin + k + cout + H)
    P = nimg * H * W
    buf = _h(torch.randn(P, stride, generator=g))
    w = torch.randn(cout, cin, k, k, generator=g) / np.sqrt(cin * k * k)
    b = torch.randn(cout, generator=g) * 0.1
    wp = w.permute(2, 3, 1, 0).reshape(k * k, cin, cout).contiguous()
    out = torch.full((P, cout + 1), 9.0, device=DEV)
    code = {"none": L.VH_UO_NONE, "relu": L.VH_UO_RELU, "sigmoid": L.VH_UO_SIGMOID, "eta": L.VH_UO_ETA}[epi]
    r = _runner()
    # wshift 1 moves the weights off 16-byte alignment: the kernel with 4-byte loads
    bd, bb = buf.to(DEV), b.to(DEV)
    wd = torch.zeros(wp.numel() + wshift, device=DEV)[wshift:]
    wd.copy_(wp.reshape(-1).to(DEV))
    L.check(r.lib.vh_uo_conv_small(_vp(bd), stride, off, cin, _vp(wd), _vp(bb), nimg, H, W, k, k, cout, code,
                                   _vp(out), cout + 1, r.stream), "uo_conv_small")
    torch.cuda.synchronize()
    x = buf[:, off:off + cin].double().view(nimg, H * W, cin).permute(0, 2, 1).reshape(nimg, cin, H, W)
    ref = F.conv2d(x, w.double(), b.double(), padding=k // 2)
    if epi == "relu":
        ref = torch.relu(ref)
    elif epi == "sigmoid":
        ref = torch.sigmoid(ref)
    elif epi == "eta":
        ref = 0.01 * F.softplus(ref)
    got = out.double().cpu()
    assert bool((got[:, cout] == 9.0).all())
    got = got[:, :cout].view(nimg, H * W, cout).permute(0, 2, 1).reshape(nimg, cout, H, W)
    err = float((got - ref).abs().max())
    print(f"  max err {err:.3e} (ref max {float(ref.abs().max()):.3f})", flush=True)
    assert err <= 1e-5 * float(ref.abs().max()) + 1e-5


# the 8 x 8 register-transpose form (HW a multiple of 8, 8-channel groups) and
# the per-element form (other shapes), for f32 and f16 tensors on both sides
@pytest.mark.parametrize("dt", [torch.float32, torch.float16])
@pytest.mark.parametrize("C,Cpad,HW", [(196, 224, 95), (4, 32, 64), (128, 128, 130), (300, 320, 7), (128, 128, 320),
                                       (196, 224, 64), (64, 72, 40), (16, 16, 8)])
def test_layout_round_trip(C, Cpad, HW, dt):
    L = _lib()
    from vipe_amd.update_op import _vp

    g = torch.Generator().manual_seed(C)
    nimg = 3
    x = torch.randn(nimg, C, HW, generator=g).to(dt).to(DEV)
    r = _runner()
    h = r.to_nhwc(x, nimg, C, HW, Cpad)
    back = torch.empty(nimg, C, HW, device=DEV, dtype=dt)
    code = L.VH_F32 if dt == torch.float32 else L.VH_F16
    L.check(r.lib.vh_uo_to_nchw(code, _vp(h), nimg, C, HW, Cpad, 0, _vp(back), r.stream), "to_nchw")
    # a channel window [8, 8 + k) of the padded rows
    k = (C - 8) // 8 * 8
    if k > 0:
        win = torch.empty(nimg, k, HW, device=DEV, dtype=dt)
        L.check(r.lib.vh_uo_to_nchw(code, _vp(h), nimg, k, HW, Cpad, 8, _vp(win), r.stream), "to_nchw window")
    torch.cuda.synchronize()
    assert torch.equal(h.view(nimg, HW, Cpad)[:, :, C:].float().cpu(), torch.zeros(nimg, HW, Cpad - C))
    assert torch.equal(back.cpu(), x.half().to(dt).cpu())
    assert torch.equal(h.view(nimg, HW, Cpad)[:, :, :C].cpu(), x.half().permute(0, 2, 1).cpu())
    if k > 0:
        assert torch.equal(win.cpu(), x[:, 8:8 + k].half().to(dt).cpu())


@pytest.mark.parametrize("C,HW", [(128, 5120), (128, 77), (64, 16), (70, 33)])
def test_mean_over_pixels(C, HW):
    """vh_uo_mean: per-image channel means of [nimg][HW][C] f32 (16-byte and
    per-channel forms) against FP64, and the same bits on a second run."""
    L = _lib()
    from vipe_amd.update_op import _vp

    g = torch.Generator().manual_seed(HW)
    nimg = 5
    x = torch.randn(nimg, HW, C, generator=g).to(DEV)
    r = _runner()
    out = torch.empty(nimg, C, device=DEV)
    out2 = torch.empty(nimg, C, device=DEV)
    L.check(r.lib.vh_uo_mean(_vp(x), nimg, HW, C, _vp(out), r.stream), "uo_mean")
    L.check(r.lib.vh_uo_mean(_vp(x), nimg, HW, C, _vp(out2), r.stream), "uo_mean")
    torch.cuda.synchronize()
    ref = x.double().mean(1).cpu()
    assert (out.double().cpu() - ref).abs().max().item() < 1e-6
    assert torch.equal(out, out2)


def _module(seed):
    from vipe_amd.update_op import UpdateModule

    torch.manual_seed(seed)
    m = UpdateModule()
    with torch.no_grad():  # larger biases than the default init, so the gates and heads leave their centres
        for name, p in m.named_parameters():
            if name.endswith("bias"):
                p.uniform_(-0.5, 0.5)
    return m.to(DEV)


def _inputs(seed, B, N, H, W, flow=True):
    g = torch.Generator().manual_seed(seed)
    net = torch.tanh(torch.randn(B, N, 128, H, W, generator=g)).half()
    inp = torch.relu(torch.randn(B, N, 128, H, W, generator=g)).half()
    corr = torch.randn(B, N, 196, H, W, generator=g) * 0.5
    fl = (torch.randn(B, N, 4, H, W, generator=g) * 4).clamp(-64, 64) if flow else None
    return [t.to(DEV) if t is not None else None for t in (net, inp, corr, fl)]


# Module tolerances (FP16 activations against FP64, about 6 f16 ulps at 1.0):
# net and weight are in
# [-1, 1] / [0, 1]; delta, eta and upmask are compared relative to their
# largest magnitude.
NET_TOL, HEAD_REL = 3e-3, 3e-3


def _check(got, ref, rel, what):
    got, ref = got.double().cpu(), ref.double().cpu()
    err = float((got - ref).abs().max())
    scale = float(ref.abs().max())
    print(f"  {what}: max err {err:.3e}, ref max {scale:.3e}", flush=True)
    assert err <= rel * max(scale, 1e-6), what


@pytest.mark.parametrize("B,N,H,W,ix,flow", [
    (1, 5, 12, 20, [0, 1, 1, 2, 0], True),
    (2, 3, 7, 9, [1, 0, 1], True),
    (1, 4, 8, 8, None, False),
    (1, 1, 3, 5, [0], True),
])
def test_update_module_matches_oracle(B, N, H, W, ix, flow):
    print(f"module B={B} N={N} H={H} W={W} ix={ix} flow={flow}", flush=True)
    m = _module(B * 100 + N)
    net, inp, corr, fl = _inputs(N + H, B, N, H, W, flow)
    ixt = None if ix is None else torch.tensor(ix, device=DEV)
    out = m(net, inp, corr, fl, ixt)
    torch.cuda.synchronize()
    ref = OU.update_module(m.state_dict(), net, inp, corr, fl, None if ix is None else torch.tensor(ix))
    assert len(out) == len(ref)
    assert out[0].dtype == torch.float16 and out[0].shape == (B, N, 128, H, W)
    assert out[1].shape == out[2].shape == (B, N, H, W, 2)
    _check(out[0], ref[0], NET_TOL, "net")
    _check(out[1], ref[1], HEAD_REL, "delta")
    _check(out[2], ref[2], NET_TOL, "weight")
    if ix is not None:
        nu = max(ix) + 1
        assert out[3].shape == (B, nu, H, W) and out[4].shape == (B, nu, 576, H, W)
        _check(out[3], ref[3], HEAD_REL, "eta")
        _check(out[4], ref[4], HEAD_REL, "upmask")


def test_cpu_tensors_raise():
    from vipe_amd.update_op import UpdateModule

    m = UpdateModule()
    x = torch.zeros(1, 1, 128, 4, 4)
    with pytest.raises(RuntimeError, match="GPU"):
        m(x, x, torch.zeros(1, 1, 196, 4, 4))


def test_parameters_repack_after_an_update():
    m = _module(3)
    net, inp, corr, fl = _inputs(4, 1, 2, 6, 8)
    a = m(net, inp, corr, fl)[1].clone()
    with torch.no_grad():
        m.delta[2].bias.add_(1.0)
    b = m(net, inp, corr, fl)[1]
    torch.cuda.synchronize()
    assert torch.allclose(b - a, torch.ones_like(a), atol=1e-5)


def test_measured_size_on_sampled_edges():
    """384 edges at 64 x 80, grouped 6 per frame as 64 keyframes: images 0 and
    383 against the oracle (each image's net / delta / weight depends on that
    image alone), and frame 0's eta / upmask against the oracle's GraphAgg on
    the kernel's own new state of edges 0-5."""
    E, H, W = 384, 64, 80
    print(f"measured size E={E} {H}x{W}", flush=True)
    m = _module(7)
    net, inp, corr, fl = _inputs(8, 1, E, H, W)
    ix = torch.arange(E, device=DEV) // 6
    out = m(net, inp, corr, fl, ix)
    torch.cuda.synchronize()
    sd = m.state_dict()
    for e in (0, E - 1):
        ref = OU.update_module(sd, net[:, e:e + 1], inp[:, e:e + 1], corr[:, e:e + 1], fl[:, e:e + 1])
        _check(out[0][:, e:e + 1], ref[0], NET_TOL, f"net[{e}]")
        _check(out[1][:, e:e + 1], ref[1], HEAD_REL, f"delta[{e}]")
        _check(out[2][:, e:e + 1], ref[2], NET_TOL, f"weight[{e}]")
    eta, upmask = OU.graph_agg(sd, out[0][:, 0:6].double().cpu(), torch.zeros(6, dtype=torch.long))
    _check(out[3][:, 0:1], eta, HEAD_REL, "eta[0]")
    _check(out[4][:, 0:1], upmask, HEAD_REL, "upmask[0]")
    assert out[3].shape == (1, 64, H, W) and out[4].shape == (1, 64, 576, H, W)
    assert bool(torch.isfinite(out[0]).all()) and bool(torch.isfinite(out[4]).all())

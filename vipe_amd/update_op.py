"""The DroidNet update operator on gfx950: a drop-in for ``UpdateModule``
(vipe/slam/networks/droid_net.py:452-519) with its ``ConvGRU`` (:393-420) and
``GraphAgg`` (:423-449).

The module keeps the reference's submodules and parameter names
(``corr_encoder.0``, ``gru.convz``, ``agg.upmask.0`` ...), so the reference's
checkpoint entries under ``update.`` load into it unchanged.  ``forward`` does
not run those torch layers: it packs the parameters once (FP16, channels-last,
fused where two layers read the same input) and runs the layers as the HIP
kernels of ``csrc/update_op.hip``:

  corr_encoder   1x1 196->128 ReLU, 3x3 128->128 ReLU            (:456-461)
  flow_encoder   7x7 4->128 ReLU, 3x3 128->64 ReLU                (:463-468)
  gru            glo = mean(sigmoid(w(net)) * net); z and r as one
                 256-channel convolution of [net, inp, corr, flow] with the
                 gate epilogue (writes z and r * net); q with the state update
                 (1 - z) net + z tanh(.) in its epilogue           (:406-420)
  delta, weight  their 3x3 128->128 ReLU layers as one 256-channel
                 convolution; the 2-channel heads on the VALU      (:470-485)
  agg            conv1 ReLU, scatter_mean over the edges' frames (this
                 package's scatter_ext), conv2 ReLU, eta = 0.01 softplus(.),
                 upmask 1x1 128->576 written straight to [B, U, 576, H, W]
                                                                   (:434-449)

Activations are FP16 with FP32 accumulation, as the reference runs the module
under autocast (factor_graph.py:228, :357).  Outputs: net f16
[B, N, 128, H, W]; delta and weight f32 [B, N, H, W, 2] (the reference's are
f16 and its caller converts them to f32, factor_graph.py:272-274); eta f32
[B, U, H, W]; upmask f16 [B, U, 576, H, W].  Inference only: the outputs carry
no autograd graph (the reference never differentiates the update operator).
CPU tensors raise: there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import torch
import torch.nn as nn

from . import _lib
from ._lib import UoConvArgs, check, load

COR_PLANES = 4 * (2 * 3 + 1) ** 2  # 196 (droid_net.py:455)
_CORR_PAD = 224  # COR_PLANES rounded up to the 32-channel K chunk
_FLOW_PAD = 32


def _pad_to(x: int, m: int) -> int:
    return (x + m - 1) // m * m


def _pack_conv(convs, cin_real, cin_pad, device):
    """Weights of the convolutions `convs` (stacked along the output channels)
    as f16 [KH*KW][Cout_pad][sum cin_pad]: input channel groups of cin_real
    channels each zero-padded to cin_pad, output rows past Cout zero."""
    w = torch.cat([c.weight.detach() for c in convs], 0).to(device=device, dtype=torch.float32)
    b = torch.cat([c.bias.detach() for c in convs], 0).to(device=device, dtype=torch.float32).contiguous()
    cout, _, kh, kw = w.shape
    parts, o = [], 0
    for r, p in zip(cin_real, cin_pad):
        part = w[:, o:o + r]
        if p > r:
            part = torch.cat([part, part.new_zeros(cout, p - r, kh, kw)], 1)
        parts.append(part)
        o += r
    assert o == w.shape[1]
    w = torch.cat(parts, 1).permute(2, 3, 0, 1).reshape(kh * kw, cout, sum(cin_pad))
    cout_pad = _pad_to(cout, 128)
    if cout_pad > cout:
        w = torch.cat([w, w.new_zeros(kh * kw, cout_pad - cout, w.shape[2])], 1)
    return dict(w=w.to(torch.float16).contiguous(), b=b, cout=cout, cout_pad=cout_pad, kh=kh, kw=kw,
                cin=sum(cin_pad))


def _pack_small(conv, device):
    """Weights of a 1- or 2-channel head as f32 [KH*KW][Cin][Cout]."""
    w = conv.weight.detach().to(device=device, dtype=torch.float32)
    cout, cin, kh, kw = w.shape
    return dict(w=w.permute(2, 3, 1, 0).reshape(kh * kw, cin, cout).contiguous(),
                b=conv.bias.detach().to(device=device, dtype=torch.float32).contiguous(), cout=cout, cin=cin, kh=kh,
                kw=kw)


def _pack_fc(convs, device):
    w = torch.cat([c.weight.detach() for c in convs], 0).to(device=device, dtype=torch.float32)
    b = torch.cat([c.bias.detach() for c in convs], 0).to(device=device, dtype=torch.float32)
    return dict(w=w.reshape(w.shape[0], w.shape[1]).contiguous(), b=b.contiguous(), cout=w.shape[0],
                cin=w.shape[1])


def _vp(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class _Runner:
    """Launches the layers on one stream; buffers are torch allocations."""

    def __init__(self, device):
        self.lib = load()
        self.device = device
        self.stream = C.c_void_p(torch.cuda.current_stream(device).cuda_stream)

    def empty(self, *shape, dtype=torch.float16):
        return torch.empty(*shape, device=self.device, dtype=dtype)

    def to_nhwc(self, x, nimg, c, hw, cpad):
        out = self.empty(nimg * hw, cpad)
        dt = _lib.VH_F16 if x.dtype == torch.float16 else _lib.VH_F32
        if x.dtype not in (torch.float16, torch.float32):
            x = x.float()
        x = x.contiguous()
        check(self.lib.vh_uo_to_nhwc(dt, _vp(x), nimg, c, hw, cpad, _vp(out), self.stream), "uo_to_nhwc")
        return out

    def conv(self, segs, pk, nimg, H, W, epi, out=None, out_f32=False, out_stride=None, out_off=0, glo=None,
             glo_stride=0, aux0=None, aux1=None, aux_stride=0, out2=None):
        """segs: list of (buffer [P, stride] f16, channels, offset)."""
        a = UoConvArgs()
        for i, (t, c, off) in enumerate(segs):
            a.seg[i] = t.data_ptr()
            a.seg_c[i] = c
            a.seg_stride[i] = t.shape[1]
            a.seg_off[i] = off
        a.nseg = len(segs)
        assert sum(c for _, c, _ in segs) == pk["cin"]
        a.weight = pk["w"].data_ptr()
        a.bias = pk["b"].data_ptr()
        a.nimg, a.H, a.W = nimg, H, W
        a.Cout, a.Cout_pad, a.KH, a.KW = pk["cout"], pk["cout_pad"], pk["kh"], pk["kw"]
        a.epilogue = epi
        if out is None:
            out = self.empty(nimg * H * W, pk["cout"], dtype=torch.float32 if out_f32 else torch.float16)
        a.out = out.data_ptr()
        a.out_f32 = int(out_f32)
        a.out_stride = out.shape[1] if out_stride is None else out_stride
        a.out_off = out_off
        a.glo = None if glo is None else glo.data_ptr()
        a.glo_stride = glo_stride
        a.aux0 = None if aux0 is None else aux0.data_ptr()
        a.aux1 = None if aux1 is None else aux1.data_ptr()
        a.aux_stride = aux_stride
        a.out2 = None if out2 is None else out2.data_ptr()
        check(self.lib.vh_uo_conv(C.byref(a), self.stream), "uo_conv")
        return out

    def conv_small(self, x, off, pk, nimg, H, W, epi):
        out = self.empty(nimg * H * W, pk["cout"], dtype=torch.float32)
        check(self.lib.vh_uo_conv_small(_vp(x), x.shape[1], off, pk["cin"], _vp(pk["w"]), _vp(pk["b"]), nimg, H, W,
                                        pk["kh"], pk["kw"], pk["cout"], epi, _vp(out), pk["cout"], self.stream),
              "uo_conv_small")
        return out


class ConvGRU(nn.Module):
    """Parameter container of droid_net.py:393-404 (same names and shapes)."""

    def __init__(self, h_planes=128, i_planes=128):
        super().__init__()
        self.convz = nn.Conv2d(h_planes + i_planes, h_planes, 3, padding=1)
        self.convr = nn.Conv2d(h_planes + i_planes, h_planes, 3, padding=1)
        self.convq = nn.Conv2d(h_planes + i_planes, h_planes, 3, padding=1)
        self.w = nn.Conv2d(h_planes, h_planes, 1, padding=0)
        self.convz_glo = nn.Conv2d(h_planes, h_planes, 1, padding=0)
        self.convr_glo = nn.Conv2d(h_planes, h_planes, 1, padding=0)
        self.convq_glo = nn.Conv2d(h_planes, h_planes, 1, padding=0)


class GraphAgg(nn.Module):
    """Parameter container of droid_net.py:423-432 (same names and shapes)."""

    def __init__(self):
        super().__init__()
        self.conv1 = nn.Conv2d(128, 128, 3, padding=1)
        self.conv2 = nn.Conv2d(128, 128, 3, padding=1)
        self.relu = nn.ReLU(inplace=True)
        self.eta = nn.Sequential(nn.Conv2d(128, 1, 3, padding=1), nn.Softplus())
        self.upmask = nn.Sequential(nn.Conv2d(128, 8 * 8 * 9, 1, padding=0))


class UpdateModule(nn.Module):
    """UpdateModule (droid_net.py:452-519) whose forward runs on the HIP kernels."""

    def __init__(self):
        super().__init__()
        self.corr_encoder = nn.Sequential(
            nn.Conv2d(COR_PLANES, 128, 1, padding=0), nn.ReLU(inplace=True),
            nn.Conv2d(128, 128, 3, padding=1), nn.ReLU(inplace=True))
        self.flow_encoder = nn.Sequential(
            nn.Conv2d(4, 128, 7, padding=3), nn.ReLU(inplace=True),
            nn.Conv2d(128, 64, 3, padding=1), nn.ReLU(inplace=True))
        self.weight = nn.Sequential(
            nn.Conv2d(128, 128, 3, padding=1), nn.ReLU(inplace=True),
            nn.Conv2d(128, 2, 3, padding=1), nn.Sigmoid())
        self.delta = nn.Sequential(
            nn.Conv2d(128, 128, 3, padding=1), nn.ReLU(inplace=True),
            nn.Conv2d(128, 2, 3, padding=1))
        self.gru = ConvGRU(128, 128 + 128 + 64)
        self.agg = GraphAgg()
        self._packed = None
        self._packed_key = None

    def _key(self, device):
        return (device,) + tuple((p.data_ptr(), p._version) for p in self.parameters())

    def packed(self, device):
        """The packed parameters on `device`, rebuilt when a parameter changed."""
        key = self._key(device)
        if self._packed_key != key:
            g, d = self.gru, device
            self._packed = dict(
                corr0=_pack_conv([self.corr_encoder[0]], [COR_PLANES], [_CORR_PAD], d),
                corr1=_pack_conv([self.corr_encoder[2]], [128], [128], d),
                flow0=_pack_conv([self.flow_encoder[0]], [4], [_FLOW_PAD], d),
                flow1=_pack_conv([self.flow_encoder[2]], [128], [128], d),
                w=_pack_conv([g.w], [128], [128], d),
                zr=_pack_conv([g.convz, g.convr], [128, 128, 128, 64], [128, 128, 128, 64], d),
                q=_pack_conv([g.convq], [128, 128, 128, 64], [128, 128, 128, 64], d),
                zr_glo=_pack_fc([g.convz_glo, g.convr_glo], d),
                q_glo=_pack_fc([g.convq_glo], d),
                heads=_pack_conv([self.delta[0], self.weight[0]], [128], [128], d),
                delta=_pack_small(self.delta[2], d),
                weight=_pack_small(self.weight[2], d),
                agg1=_pack_conv([self.agg.conv1], [128], [128], d),
                agg2=_pack_conv([self.agg.conv2], [128], [128], d),
                eta=_pack_small(self.agg.eta[0], d),
                upmask=_pack_conv([self.agg.upmask[0]], [128], [128], d),
            )
            self._packed_key = key
        return self._packed

    @torch.no_grad()
    def forward(self, net, inp, corr, flow=None, ix=None):
        """RaftSLAM update operator (droid_net.py:487-519)."""
        if net.device.type != "cuda":
            raise RuntimeError("UpdateModule runs on the GPU only (there is no CPU path)")
        batch, num, ch, ht, wd = net.shape
        if ch != 128 or inp.shape[2] != 128 or corr.shape[2] != COR_PLANES:
            raise ValueError(f"UpdateModule: net / inp / corr channels {ch} / {inp.shape[2]} / {corr.shape[2]}, "
                             f"expected 128 / 128 / {COR_PLANES}")
        if flow is not None and flow.shape[2] != 4:
            raise ValueError(f"UpdateModule: flow has {flow.shape[2]} channels, expected 4")
        dev = net.device
        pk = self.packed(dev)
        r = _Runner(dev)
        nimg, HW = batch * num, ht * wd
        P = nimg * HW
        RELU, NONE = _lib.VH_UO_RELU, _lib.VH_UO_NONE

        net_h = r.to_nhwc(net.reshape(nimg, 128, HW), nimg, 128, HW, 128)
        inp_h = r.to_nhwc(inp.reshape(nimg, 128, HW), nimg, 128, HW, 128)
        corr_h = r.to_nhwc(corr.reshape(nimg, COR_PLANES, HW), nimg, COR_PLANES, HW, _CORR_PAD)
        if flow is None:
            flow_h = torch.zeros(P, _FLOW_PAD, device=dev, dtype=torch.float16)
        else:
            flow_h = r.to_nhwc(flow.reshape(nimg, 4, HW), nimg, 4, HW, _FLOW_PAD)

        # encoders (:497-498)
        c1 = r.conv([(corr_h, _CORR_PAD, 0)], pk["corr0"], nimg, ht, wd, RELU)
        corr_e = r.conv([(c1, 128, 0)], pk["corr1"], nimg, ht, wd, RELU)
        f1 = r.conv([(flow_h, _FLOW_PAD, 0)], pk["flow0"], nimg, ht, wd, RELU)
        flow_e = r.conv([(f1, 128, 0)], pk["flow1"], nimg, ht, wd, RELU)

        # ConvGRU (:406-420)
        sg = r.conv([(net_h, 128, 0)], pk["w"], nimg, ht, wd, _lib.VH_UO_SIGMUL, out_f32=True, aux0=net_h,
                    aux_stride=128)
        glo = torch.empty(nimg, 128, device=dev, dtype=torch.float32)
        check(r.lib.vh_uo_mean(_vp(sg), nimg, HW, 128, _vp(glo), r.stream), "uo_mean")
        glo_zr = torch.empty(nimg, 256, device=dev, dtype=torch.float32)
        glo_q = torch.empty(nimg, 128, device=dev, dtype=torch.float32)
        for name, o in (("zr_glo", glo_zr), ("q_glo", glo_q)):
            f = pk[name]
            check(r.lib.vh_uo_fc(_vp(glo), nimg, f["cin"], _vp(f["w"]), _vp(f["b"]), f["cout"], _vp(o), r.stream),
                  "uo_fc")
        z = r.empty(P, 128)
        rnet = r.empty(P, 128)
        ctx = [(inp_h, 128, 0), (corr_e, 128, 0), (flow_e, 64, 0)]
        r.conv([(net_h, 128, 0)] + ctx, pk["zr"], nimg, ht, wd, _lib.VH_UO_GATES, out=z, out_stride=128,
               glo=glo_zr, glo_stride=256, aux0=net_h, aux_stride=128, out2=rnet)
        net_new = r.conv([(rnet, 128, 0)] + ctx, pk["q"], nimg, ht, wd, _lib.VH_UO_GRU, glo=glo_q, glo_stride=128,
                         aux0=z, aux1=net_h, aux_stride=128)

        # delta and weight heads (:501-507)
        hw = r.conv([(net_new, 128, 0)], pk["heads"], nimg, ht, wd, RELU)
        delta = r.conv_small(hw, 0, pk["delta"], nimg, ht, wd, NONE).view(batch, num, ht, wd, 2)
        weight = r.conv_small(hw, 128, pk["weight"], nimg, ht, wd, _lib.VH_UO_SIGMOID).view(batch, num, ht, wd, 2)

        net_out = torch.empty(batch, num, 128, ht, wd, device=dev, dtype=torch.float16)
        check(r.lib.vh_uo_to_nchw(_lib.VH_F16, _vp(net_new), nimg, 128, HW, 128, 0, _vp(net_out), r.stream),
              "uo_to_nchw")
        if ix is None:
            return net_out, delta, weight

        # GraphAgg (:434-449)
        from .ext import scatter_ext

        a1 = r.conv([(net_new, 128, 0)], pk["agg1"], nimg, ht, wd, RELU)
        ix = ix.to(dev)
        am = scatter_ext.scatter_mean(a1.view(batch, num, HW * 128), ix, 1)
        nu = am.shape[1]
        am = am.reshape(batch * nu * HW, 128)
        a2 = r.conv([(am, 128, 0)], pk["agg2"], batch * nu, ht, wd, RELU)
        eta = r.conv_small(a2, 0, pk["eta"], batch * nu, ht, wd, _lib.VH_UO_ETA).view(batch, nu, ht, wd)
        upmask = torch.empty(batch, nu, 576, ht, wd, device=dev, dtype=torch.float16)
        r.conv([(a2, 128, 0)], pk["upmask"], batch * nu, ht, wd, NONE | _lib.VH_UO_NCHW, out=upmask,
               out_stride=576)
        return net_out, delta, weight, eta, upmask

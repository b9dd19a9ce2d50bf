"""Restatement of the DroidNet update operator in FP64 (oracle; test-only).

Follows vipe/slam/networks/droid_net.py:
  conv            nn.Conv2d with 'same' zero padding (cross-correlation, bias)
  conv_gru        ConvGRU.forward (:406-420): inp = cat(inputs); glo = mean
                  over pixels of sigmoid(w(net)) * net; z, r = sigmoid(conv
                  of cat[net, inp] + conv_glo(glo)); q = tanh(convq(cat[r *
                  net, inp]) + convq_glo(glo)); net = (1 - z) net + z q
  graph_agg       GraphAgg.forward (:434-449): relu(conv1), scatter_mean over
                  dim 1 by ix, relu(conv2), eta = 0.01 softplus(eta conv),
                  upmask = 1x1 conv to 576 channels
  update_module   UpdateModule.forward (:487-519): corr and flow encoders,
                  the GRU on (net, inp, corr, flow), delta and weight heads
                  permuted to [B, N, H, W, 2], GraphAgg when ix is given
Parameters come as the module's state_dict (reference names, e.g.
"gru.convz.weight").  Everything runs in float64 on the CPU, with
torch.nn.functional.conv2d as the convolution routine.

Pinned by tests/test_oracle_update_op.py: conv against a brute-force loop of
the cross-correlation definition; the GRU at z = 0 (state kept) and z = 1
(state = tanh of the candidate, with the candidate computed from one named
input block alone, which fixes the concatenation order); the global context
against a hand-computed mean; the heads and eta at zero weights against their
closed forms (bias, sigmoid(bias), 0.01 log(1 + e^b)); scatter_mean against
per-group means; output shapes and layouts.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F


def _p(params, name):
    return params[name].detach().to(torch.float64).cpu()


def conv(x, params, name, act=None):
    """nn.Conv2d(..., padding=k // 2) (droid_net.py:396-404, 426-432, 456-485)."""
    w, b = _p(params, name + ".weight"), _p(params, name + ".bias")
    y = F.conv2d(x, w, b, padding=(w.shape[2] // 2, w.shape[3] // 2))
    if act == "relu":
        y = torch.relu(y)
    elif act == "sigmoid":
        y = torch.sigmoid(y)
    return y


def conv_gru(params, net, *inputs, prefix="gru."):
    """ConvGRU.forward (droid_net.py:406-420)."""
    inp = torch.cat(inputs, dim=1)
    net_inp = torch.cat([net, inp], dim=1)
    b, c, h, w = net.shape
    glo = torch.sigmoid(conv(net, params, prefix + "w")) * net
    glo = glo.view(b, c, h * w).mean(-1).view(b, c, 1, 1)
    z = torch.sigmoid(conv(net_inp, params, prefix + "convz") + conv(glo, params, prefix + "convz_glo"))
    r = torch.sigmoid(conv(net_inp, params, prefix + "convr") + conv(glo, params, prefix + "convr_glo"))
    q = torch.tanh(conv(torch.cat([r * net, inp], dim=1), params, prefix + "convq")
                   + conv(glo, params, prefix + "convq_glo"))
    return (1 - z) * net + z * q


def scatter_mean_dim1(x, ix):
    """scatter_mean(x, ix, dim=1) with dim_size = max(ix) + 1 (vipe/ext/scatter.py)."""
    ix = ix.cpu().long()
    n_out = int(ix.max()) + 1
    out = torch.zeros((x.shape[0], n_out) + tuple(x.shape[2:]), dtype=x.dtype)
    cnt = torch.zeros(n_out, dtype=x.dtype)
    for e in range(x.shape[1]):
        out[:, ix[e]] += x[:, e]
        cnt[ix[e]] += 1
    shape = (1, n_out) + (1,) * (x.dim() - 2)
    return out / cnt.clamp(min=1).view(shape)


def graph_agg(params, net, ix, prefix="agg."):
    """GraphAgg.forward (droid_net.py:434-449); net [B, N, 128, H, W]."""
    batch, num, ch, ht, wd = net.shape
    x = conv(net.reshape(batch * num, ch, ht, wd), params, prefix + "conv1", "relu")
    x = scatter_mean_dim1(x.view(batch, num, 128, ht, wd), ix)
    x = x.view(-1, 128, ht, wd)
    x = conv(x, params, prefix + "conv2", "relu")
    eta = F.softplus(conv(x, params, prefix + "eta.0")).view(batch, -1, ht, wd)
    upmask = conv(x, params, prefix + "upmask.0").view(batch, -1, 8 * 8 * 9, ht, wd)
    return 0.01 * eta, upmask


def update_module(params, net, inp, corr, flow=None, ix=None):
    """UpdateModule.forward (droid_net.py:487-519), all in float64."""
    net, inp, corr = (t.detach().to(torch.float64).cpu() for t in (net, inp, corr))
    batch, num, ch, ht, wd = net.shape
    if flow is None:
        flow = torch.zeros(batch, num, 4, ht, wd, dtype=torch.float64)
    flow = flow.detach().to(torch.float64).cpu()
    output_dim = (batch, num, -1, ht, wd)
    net = net.reshape(batch * num, -1, ht, wd)
    inp = inp.reshape(batch * num, -1, ht, wd)
    corr = corr.reshape(batch * num, -1, ht, wd)
    flow = flow.reshape(batch * num, -1, ht, wd)
    corr = conv(conv(corr, params, "corr_encoder.0", "relu"), params, "corr_encoder.2", "relu")
    flow = conv(conv(flow, params, "flow_encoder.0", "relu"), params, "flow_encoder.2", "relu")
    net = conv_gru(params, net, inp, corr, flow)
    delta = conv(conv(net, params, "delta.0", "relu"), params, "delta.2").view(*output_dim)
    weight = conv(conv(net, params, "weight.0", "relu"), params, "weight.2", "sigmoid").view(*output_dim)
    delta = delta.permute(0, 1, 3, 4, 2)[..., :2].contiguous()
    weight = weight.permute(0, 1, 3, 4, 2)[..., :2].contiguous()
    net = net.view(*output_dim)
    if ix is not None:
        eta, upmask = graph_agg(params, net, ix)
        return net, delta, weight, eta, upmask
    return net, delta, weight

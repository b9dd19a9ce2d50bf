"""Batch broadcasting of group-op inputs (vipe/ext/lietorch/broadcasting.py:12-35)."""
from __future__ import annotations


def check_broadcastable(x, y):
    if len(x.shape) != len(y.shape):
        raise ValueError("group op inputs must have the same number of batch dimensions")
    for n, m in zip(x.shape[:-1], y.shape[:-1]):
        if not (n == m or n == 1 or m == 1):
            raise ValueError(f"batch dimensions {tuple(x.shape[:-1])} and {tuple(y.shape[:-1])} do not broadcast")


def broadcast_inputs(x, y):
    """Flatten the batch dims of x (and y, repeated to a common shape) to [n, d]."""
    if y is None:
        return (x.reshape(-1, x.shape[-1]).contiguous(),), x.shape[:-1]
    check_broadcastable(x, y)
    xs, xd = x.shape[:-1], x.shape[-1]
    ys, yd = y.shape[:-1], y.shape[-1]
    out_shape = tuple(max(n, m) for n, m in zip(xs, ys))
    if tuple(xs) == tuple(ys):
        return (x.reshape(-1, xd).contiguous(), y.reshape(-1, yd).contiguous()), out_shape
    x_rep = [m if n == 1 else 1 for n, m in zip(xs, ys)]
    y_rep = [n if m == 1 else 1 for n, m in zip(xs, ys)]
    x1 = x.repeat(x_rep + [1]).reshape(-1, xd).contiguous()
    y1 = y.repeat(y_rep + [1]).reshape(-1, yd).contiguous()
    return (x1, y1), out_shape

"""Model compression: quintic tables of the se_e2_a embedding nets.

Builds what the reference's ``DescrptSeA.enable_compression`` /
``DPTabulate.build`` (deepmd/pt/model/descriptor/se_a.py:269-314,
deepmd/utils/tabulate.py:67-372,468-486) build: for every embedding net a
table of quintic pieces on the knots

    arange(lower, upper, stride0) ++ arange(upper, extrapolate*upper, stride1)
    ++ [extrapolate*upper]

whose piece i on [x_i, x_i + h_i] matches the net's value, first and second
derivative at both ends (the coefficient formulas of tabulate.py:305-352),
plus the ``table_info`` row [lower, upper, upper*extrapolate, stride0,
stride1, check_frequency] the tabulate op reads.

The net's y, y', y'' at the knots come from a forward-mode pass through the
tanh MLP in FP64 on the model's device (what tabulate_math.py:353-470 does
with hand-written derivative helpers):
    z = x W + b,  t = tanh(z),  y = t*idt + res(x)
    dz = dx W,    dt = (1-t^2) dz,                 dy = dt*idt + res(dx)
    d2z = d2x W,  d2t = (1-t^2) (d2z - 2 t dz^2),  d2y = d2t*idt + res(d2x)
"""

from __future__ import annotations

import math

import numpy as np
import torch


def spline5_switch(x: float, rmin: float, rmax: float) -> float:
    """deepmd/utils/tabulate.py:490-499."""
    if x < rmin:
        return 1.0
    if x < rmax:
        u = (x - rmin) / (rmax - rmin)
        return u * u * u * (-6 * u * u + 15 * u - 10) + 1
    return 0.0


def env_mat_range(davg, dstd, rcut, rcut_smth, min_nbor_dist):
    """Per centre type (lower, upper) of the normalised s = sw/r:
    tabulate.py:468-486 on the serialized [ntypes, nnei, 4] statistics, whose
    ``[:, 0]`` is slot 0's four components, reduced over them
    (tabulate.py:91-97)."""
    sw = spline5_switch(min_nbor_dist, rcut_smth, rcut)
    a = np.asarray(davg)[:, 0, :]
    d = np.asarray(dstd)[:, 0, :]
    lower = np.floor(-a / d).min(axis=1)
    upper = np.ceil(((1.0 / min_nbor_dist) * sw - a) / d).max(axis=1)
    return lower, upper


def table_ranges(davg, dstd, rcut, rcut_smth, sel, type_one_side, min_nbor_dist):
    """(lower, upper) of each embedding net's table (tabulate.py:111-140)."""
    lower, upper = env_mat_range(davg, dstd, rcut, rcut_smth, min_nbor_dist)
    nt = len(sel)
    if type_one_side:
        keep = np.array([s > 0 for s in sel])
        return [(float(lower[keep].min()), float(upper[keep].max()))] * nt
    return [(float(lower[i // nt]), float(upper[i // nt])) for i in range(nt * nt)]


def knots(lower, upper, stride0, stride1, extrapolate):
    """Knot positions (FP64) and the number of pieces (tabulate.py:121-140)."""
    xx = np.arange(lower, upper, stride0, dtype=np.float64)
    xx = np.append(xx, np.arange(upper, extrapolate * upper, stride1, dtype=np.float64))
    xx = np.append(xx, np.array([extrapolate * upper], dtype=np.float64))
    nspline = int((upper - lower) / stride0 + (extrapolate * upper - upper) / stride1)
    return xx, nspline


def _res(nin, nout, resnet):
    if not resnet:
        return 0
    return 1 if nout == nin else (2 if nout == 2 * nin else 0)


def net_value_derivs(layers, xx: torch.Tensor):
    """y, dy/ds, d2y/ds2 [n, L] of a tanh embedding net at scalar inputs."""
    x = xx.view(-1, 1)
    dx = torch.ones_like(x)
    d2x = torch.zeros_like(x)
    for ly in layers:
        w = torch.as_tensor(ly.w, dtype=torch.float64, device=xx.device)
        b = torch.as_tensor(ly.b, dtype=torch.float64, device=xx.device)
        if ly.act != "tanh":
            raise NotImplementedError(f"compression of activation {ly.act!r}")
        z = x @ w + b
        dz = dx @ w
        d2z = d2x @ w
        t = torch.tanh(z)
        g = 1.0 - t * t
        dt = g * dz
        d2t = g * (d2z - 2.0 * t * dz * dz)
        if ly.idt is not None:
            idt = torch.as_tensor(ly.idt, dtype=torch.float64, device=xx.device)
            t, dt, d2t = t * idt, dt * idt, d2t * idt
        r = _res(w.shape[0], w.shape[1], ly.resnet)
        if r == 1:
            t, dt, d2t = t + x, dt + dx, d2t + d2x
        elif r == 2:
            t = t + torch.cat([x, x], dim=1)
            dt = dt + torch.cat([dx, dx], dim=1)
            d2t = d2t + torch.cat([d2x, d2x], dim=1)
        x, dx, d2x = t, dt, d2t
    return x, dx, d2x


def build_table(layers, lower, upper, stride0, stride1, extrapolate, device="cpu"):
    """Table [nspline, 6 L] (FP64) of one net: row i = (a0..a5) per channel."""
    xx, nspline = knots(lower, upper, stride0, stride1, extrapolate)
    vv, dd, d2 = net_value_derivs(layers, torch.tensor(xx, dtype=torch.float64, device=device))
    L = vv.shape[1]
    tt = torch.full((nspline, 1), float(stride1), dtype=torch.float64, device=vv.device)
    tt[: int((upper - lower) / stride0)] = float(stride0)
    y0, y1 = vv[:nspline], vv[1:nspline + 1]
    p0, p1 = dd[:nspline], dd[1:nspline + 1]
    q0, q1 = d2[:nspline], d2[1:nspline + 1]
    hh = y1 - y0
    coef = torch.empty((nspline, L, 6), dtype=torch.float64, device=vv.device)
    coef[:, :, 0] = y0
    coef[:, :, 1] = p0
    coef[:, :, 2] = 0.5 * q0
    coef[:, :, 3] = (1 / (2 * tt ** 3)) * (20 * hh - (8 * p1 + 12 * p0) * tt
                                           - (3 * q0 - q1) * tt * tt)
    coef[:, :, 4] = (1 / (2 * tt ** 4)) * (-30 * hh + (14 * p1 + 16 * p0) * tt
                                           + (3 * q0 - 2 * q1) * tt * tt)
    coef[:, :, 5] = (1 / (2 * tt ** 5)) * (12 * hh - 6 * (p1 + p0) * tt + (q1 - q0) * tt * tt)
    return coef.reshape(nspline, 6 * L)


def table_info(lower, upper, extrapolate, stride0, stride1, check_frequency=-1):
    """se_a.py:725-736 (compress_info of one net)."""
    return np.array([lower, upper, upper * extrapolate, stride0, stride1, check_frequency],
                    dtype=np.float64)


def build_tables(emb_layers, davg, dstd, rcut, rcut_smth, sel, type_one_side, min_nbor_dist,
                 extrapolate=5.0, stride0=0.01, stride1=0.1, check_frequency=-1, device="cpu"):
    """Tables and infos of every embedding net (NetworkCollection order)."""
    if not (min_nbor_dist > 0 and stride0 > 0 and stride1 > 0 and extrapolate > 1):
        raise ValueError("compression needs min_nbor_dist > 0, strides > 0, extrapolate > 1")
    ranges = table_ranges(davg, dstd, rcut, rcut_smth, sel, type_one_side, min_nbor_dist)
    if len(ranges) != len(emb_layers):
        raise ValueError(f"{len(emb_layers)} embedding nets for {len(ranges)} tables")
    tables, infos = [], []
    for layers, (lo, up) in zip(emb_layers, ranges):
        if not math.isfinite(lo) or not math.isfinite(up) or up <= lo:
            raise ValueError(f"empty table range [{lo}, {up}]")
        tables.append(build_table(layers, lo, up, stride0, stride1, extrapolate, device))
        infos.append(table_info(lo, up, extrapolate, stride0, stride1, check_frequency))
    return tables, infos

"""Oracle compression-table builder (quintic Hermite pieces of the embedding net).

Test infrastructure only (see ``oracle/__init__.py``).

Restates ``deepmd/utils/tabulate.py:67-372`` (knot set, ``nspline``, per-knot
value / first / second derivative, piece coefficients) and
``deepmd/utils/tabulate.py:468-486`` (range).  The embedding net's y, y', y''
at the knots come from torch autograd (a library routine) instead of the
reference's hand-written tanh derivatives (``deepmd/utils/tabulate_math.py``);
the coefficients come from solving the 3x3 Hermite end conditions in closed
form (pinned by ``tests/test_oracle_tabulate.py``: every piece reproduces y,
y', y'' at both of its knots).
"""

from __future__ import annotations

import numpy as np
import torch

from . import sea


def spline5_switch(x, rmin, rmax):
    if x < rmin:
        return 1.0
    if x < rmax:
        u = (x - rmin) / (rmax - rmin)
        return u * u * u * (-6 * u * u + 15 * u - 10) + 1
    return 0.0


def env_mat_range(model: sea.SeAModel, min_nbor_dist: float):
    """``tabulate.py:91-97,468-486`` for se_a: ``davg[:, 0]`` / ``dstd[:, 0]``
    of the serialized ``[ntypes, nnei, 4]`` statistics are slot 0's four
    components; floor / ceil per component, then min / max over them."""
    sw = spline5_switch(min_nbor_dist, model.rcut_smth, model.rcut)
    davg = model.davg[:, 0, :].numpy()
    dstd = model.dstd[:, 0, :].numpy()
    lower = np.floor(-davg / dstd).min(axis=1)
    upper = np.ceil(((1.0 / min_nbor_dist) * sw - davg) / dstd).max(axis=1)
    return lower, upper


def net_ranges(model: sea.SeAModel, min_nbor_dist: float):
    """``tabulate.py:111-140``: (lower, upper) of table ii.  type_one_side: the
    extremes over centre types with sel > 0 (no exclusions here); otherwise the
    range of type ii // ntypes."""
    lower, upper = env_mat_range(model, min_nbor_dist)
    nt = model.ntypes
    out = []
    if model.type_one_side:
        idx = np.array([s > 0 for s in model.sel])
        for _ in range(nt):
            out.append((float(np.min(lower[idx])), float(np.max(upper[idx]))))
    else:
        for ii in range(nt * nt):
            out.append((float(lower[ii // nt]), float(upper[ii // nt])))
    return out


def net_derivs(layers, xx):
    """y, dy/dx, d2y/dx2 of an embedding net at scalar inputs xx: [n, L] each."""
    x = torch.tensor(xx, dtype=torch.float64).view(-1, 1).requires_grad_(True)
    y = sea.run_net(x, layers)
    L = y.shape[1]
    d1 = torch.zeros_like(y)
    d2 = torch.zeros_like(y)
    for k in range(L):
        g, = torch.autograd.grad(y[:, k].sum(), x, create_graph=True)
        h, = torch.autograd.grad(g.sum(), x, retain_graph=True)
        d1[:, k] = g[:, 0].detach()
        d2[:, k] = h[:, 0].detach()
    return y.detach().numpy(), d1.numpy(), d2.numpy()


def hermite5(y0, d0, s0, y1, d1, s1, h):
    """Coefficients a0..a5 of p(t), t in [0, h], with p, p', p'' matching
    (y0, d0, s0) at 0 and (y1, d1, s1) at h."""
    e0 = y1 - (y0 + d0 * h + 0.5 * s0 * h * h)
    e1 = d1 - (d0 + s0 * h)
    e2 = s1 - s0
    a3 = (10 * e0 - 4 * e1 * h + 0.5 * e2 * h * h) / h ** 3
    a4 = (-15 * e0 + 7 * e1 * h - e2 * h * h) / h ** 4
    a5 = (6 * e0 - 3 * e1 * h + 0.5 * e2 * h * h) / h ** 5
    return y0, d0, 0.5 * s0, a3, a4, a5


def build_table(layers, lower, upper, stride0, stride1, extrapolate, dtype=np.float64):
    """``tabulate.py:121-143,275-372``: table [nspline, 6 L] and knots."""
    ll, uu = float(lower), float(upper)
    xx = np.arange(ll, uu, stride0, dtype=np.float64)
    xx = np.append(xx, np.arange(uu, extrapolate * uu, stride1, dtype=np.float64))
    xx = np.append(xx, np.array([extrapolate * uu], dtype=np.float64))
    nspline = int((uu - ll) / stride0 + (extrapolate * uu - uu) / stride1)
    vv, dd, d2 = net_derivs(layers, xx)
    L = vv.shape[1]
    tt = np.full((nspline, 1), stride1, dtype=np.float64)
    tt[: int((uu - ll) / stride0)] = stride0
    coef = hermite5(vv[:nspline], dd[:nspline], d2[:nspline],
                    vv[1:nspline + 1], dd[1:nspline + 1], d2[1:nspline + 1], tt)
    table = np.zeros((nspline, L, 6), dtype=np.float64)
    for k in range(6):
        table[:, :, k] = coef[k]
    return table.reshape(nspline, 6 * L).astype(dtype), xx


def table_info(lower, upper, extrapolate, stride0, stride1, check_frequency=-1,
               dtype=np.float64):
    """``se_a.py:725-736``: [lower, upper, upper*extrapolate, stride0, stride1, freq]."""
    return np.array([lower, upper, upper * extrapolate, stride0, stride1, check_frequency],
                    dtype=dtype)

"""Oracle restatement of the reference C++ kernels' semantics (CPU, numpy FP64/FP32).

Test infrastructure only (see ``oracle/__init__.py``).

* :func:`prod_env_mat_a`    -- ``source/lib/src/env_mat.cc:123-221`` (per slot
  value + derivative w.r.t. the centre + rij) and
  ``source/lib/src/prod_env_mat.cc:14-133`` (normalisation by avg/std,
  virtual centres zeroed); switch ``source/lib/include/switcher.h:60-83``.
* :func:`prod_force_a`      -- ``source/lib/src/prod_force.cc:23-73``
* :func:`prod_virial_a`     -- ``source/lib/src/prod_virial.cc:22-69``
* :func:`prod_force_grad_a` -- ``source/lib/src/prod_force_grad.cc:22-76``
* :func:`prod_virial_grad_a`-- ``source/lib/src/prod_virial_grad.cc:22-63``
* :func:`tabulate_fusion_se_a` / ``_grad`` / ``_grad_grad`` --
  ``source/lib/src/tabulate.cc:45-73,122-158,162-447`` (``locate_xx``,
  quintic evaluation, the ``is_sorted`` padding fold).

All functions take and return numpy arrays with the reference's flat layouts
(``nlist`` [nf, nloc, nnei] with -1 padding, ``em`` [nf, nloc, nnei*4],
``em_deriv`` [nf, nloc, nnei*12], ``rij`` [nf, nloc, nnei*3]).
"""

from __future__ import annotations

import numpy as np


def spline5_switch(xx, rmin, rmax):
    """``switcher.h:60-83``: (value, d value / d r)."""
    xx = np.asarray(xx)
    uu = (xx - rmin) / (rmax - rmin)
    du = 1.0 / (rmax - rmin)
    vv = uu * uu * uu * (-6.0 * uu * uu + 15.0 * uu - 10.0) + 1.0
    dd = (3.0 * uu * uu * (-6.0 * uu * uu + 15.0 * uu - 10.0)
          + uu * uu * uu * (-12.0 * uu + 15.0)) * du
    vv = np.where(xx < rmin, 1.0, np.where(xx < rmax, vv, 0.0))
    dd = np.where(xx < rmin, 0.0, np.where(xx < rmax, dd, 0.0))
    return vv, dd


def prod_env_mat_a(coord, atype, nlist, avg, std, rcut, rcut_smth,
                   float32_cutoffs=False):
    """Env matrix with derivative and rij for a formatted neighbour list.

    coord [nf, nall, 3] FP64, atype [nf, nall] (centre types are the first nloc
    rows), nlist [nf, nloc, nnei] (-1 padded), avg/std [ntypes, nnei*4].
    ``float32_cutoffs`` rounds rcut / rcut_smth through float32 as the C++
    signature (``const float rcut``) does.
    """
    coord = np.asarray(coord, dtype=np.float64)
    nf, nloc, nnei = nlist.shape
    if float32_cutoffs:
        rcut = float(np.float32(rcut))
        rcut_smth = float(np.float32(rcut_smth))
    mask = nlist >= 0
    idx = np.where(mask, nlist, 0)
    ci = coord[:, :nloc, :]
    cj = np.take_along_axis(coord, idx.reshape(nf, -1, 1), axis=1).reshape(nf, nloc, nnei, 3)
    rij = np.where(mask[..., None], cj - ci[:, :, None, :], 0.0)
    nr2 = np.einsum("...d,...d->...", rij, rij)
    nr2s = np.where(mask, nr2, 1.0)
    inr = 1.0 / np.sqrt(nr2s)
    nr = nr2s * inr
    inr2 = inr * inr
    inr4 = inr2 * inr2
    inr3 = inr4 * nr
    sw, dsw = spline5_switch(nr, rcut_smth, rcut)
    x, y, z = rij[..., 0], rij[..., 1], rij[..., 2]
    d0 = 1.0 / nr
    dx, dy, dz = x / nr2s, y / nr2s, z / nr2s
    em = np.stack([d0, dx, dy, dz], -1) * sw[..., None]
    r = [x, y, z]
    dv = [d0, dx, dy, dz]
    der = np.zeros(rij.shape[:-1] + (4, 3))
    for dd in range(3):
        der[..., 0, dd] = r[dd] * inr3 * sw - d0 * dsw * r[dd] * inr
    for cc in range(3):
        for dd in range(3):
            t = 2.0 * r[cc] * r[dd] * inr4 - (inr2 if cc == dd else 0.0)
            der[..., cc + 1, dd] = t * sw - dv[cc + 1] * dsw * r[dd] * inr
    em = np.where(mask[..., None], em, 0.0)
    der = np.where(mask[..., None, None], der, 0.0)
    t_c = atype[:, :nloc]
    valid = t_c >= 0
    tc = np.where(valid, t_c, 0)
    a = avg[tc].reshape(nf, nloc, nnei, 4)
    s = std[tc].reshape(nf, nloc, nnei, 4)
    em_n = np.where(valid[:, :, None, None], (em - a) / s, 0.0)
    der_n = np.where(valid[:, :, None, None, None], der / s[..., None], 0.0)
    return (em_n.reshape(nf, nloc, nnei * 4), der_n.reshape(nf, nloc, nnei * 12),
            rij.reshape(nf, nloc, nnei * 3))


def prod_force_a(net_deriv, env_deriv, nlist, nall):
    """force [nf, nall, 3]: centre gets -Σ g·dR, neighbour j gets +g·dR."""
    nf, nloc, nnei = nlist.shape
    g = net_deriv.reshape(nf, nloc, nnei, 4, 1)
    d = env_deriv.reshape(nf, nloc, nnei, 4, 3)
    per = (g * d).sum(3)                      # [nf, nloc, nnei, 3]
    force = np.zeros((nf, nall, 3))
    force[:, :nloc, :] -= per.sum(2)
    for ff in range(nf):
        m = nlist[ff] >= 0
        np.add.at(force[ff], nlist[ff][m], per[ff][m])
    return force


def prod_virial_a(net_deriv, env_deriv, rij, nlist, nall):
    """virial [nf, 9] and atom_virial [nf, nall, 9];
    V[d0][d1] += g·dR[d0] * rij[d1] accumulated on neighbour j."""
    nf, nloc, nnei = nlist.shape
    g = net_deriv.reshape(nf, nloc, nnei, 4, 1)
    d = env_deriv.reshape(nf, nloc, nnei, 4, 3)
    per = (g * d).sum(3)                      # [nf, nloc, nnei, 3]
    r = rij.reshape(nf, nloc, nnei, 3)
    tv = per[..., :, None] * r[..., None, :]  # [.., 3, 3]
    m = nlist >= 0
    tv = np.where(m[..., None, None], tv, 0.0)
    virial = tv.sum((1, 2)).reshape(nf, 9)
    av = np.zeros((nf, nall, 9))
    for ff in range(nf):
        mm = nlist[ff] >= 0
        np.add.at(av[ff], nlist[ff][mm], tv[ff][mm].reshape(-1, 9))
    return virial, av


def prod_force_grad_a(grad, env_deriv, nlist):
    """grad_net [nf, nloc, nnei*4] from grad [nf, nloc, 3] (neighbour index
    taken modulo nloc, as ``prod_force_grad.cc:54-57``)."""
    nf, nloc, nnei = nlist.shape
    d = env_deriv.reshape(nf, nloc, nnei, 4, 3)
    gi = grad.reshape(nf, nloc, 3)
    out = -np.einsum("fid,fijad->fija", gi, d)
    m = nlist >= 0
    j = np.where(m, nlist % nloc, 0)
    gj = np.take_along_axis(gi, j.reshape(nf, -1, 1), axis=1).reshape(nf, nloc, nnei, 3)
    out += np.where(m[..., None], np.einsum("fijd,fijad->fija", gj, d), 0.0)
    return out.reshape(nf, nloc, nnei * 4)


def prod_virial_grad_a(grad, env_deriv, rij, nlist):
    """grad_net [nf, nloc, nnei*4] from grad [nf, 9]."""
    nf, nloc, nnei = nlist.shape
    d = env_deriv.reshape(nf, nloc, nnei, 4, 3)
    r = rij.reshape(nf, nloc, nnei, 3)
    G = grad.reshape(nf, 3, 3)
    out = np.einsum("fxy,fijy,fijax->fija", G, r, d)
    out = np.where((nlist >= 0)[..., None], out, 0.0)
    return out.reshape(nf, nloc, nnei * 4)


# ------------------------------------------------------------- tabulate ----

def locate_xx(lower, upper, vmax, stride0, stride1, xx, dtype=np.float64):
    """``tabulate.cc:45-73`` vectorised: (table_idx, local x, extrapolate_delta).

    Arithmetic is done in ``dtype`` (the table's precision)."""
    T = dtype
    lower, upper, vmax = T(lower), T(upper), T(vmax)
    stride0, stride1 = T(stride0), T(stride1)
    xx = np.asarray(xx, dtype=T)
    first = int((upper - lower) / stride0)
    # high tail: boundary = nextafter(max, lower)
    bxx = np.nextafter(vmax, lower)
    tail_idx = first + int((bxx - upper) / stride1)
    idx = np.zeros(xx.shape, dtype=np.int64)
    loc = np.zeros(xx.shape, dtype=T)
    ext = np.zeros(xx.shape, dtype=T)
    c0 = xx < lower
    c1 = (~c0) & (xx < upper)
    c2 = (~c0) & (~c1) & (xx < vmax)
    c3 = ~(c0 | c1 | c2)
    ext[c0] = xx[c0] - lower
    i1 = ((xx[c1] - lower) / stride0).astype(np.int64)
    idx[c1] = i1
    loc[c1] = xx[c1] - (i1.astype(T) * stride0 + lower)
    i2 = first + ((xx[c2] - upper) / stride1).astype(np.int64)
    idx[c2] = i2
    loc[c2] = xx[c2] - ((i2 - first).astype(T) * stride1 + upper)
    idx[c3] = tail_idx
    loc[c3] = vmax - (T(tail_idx - first) * stride1 + upper)
    ext[c3] = xx[c3] - vmax
    return idx, loc, ext


def _poly(table, idx, x, L):
    """(value, first derivative) of the quintic pieces: [..., L]."""
    c = table.reshape(table.shape[0], L, 6)[idx]          # [..., L, 6]
    x = x[..., None]
    a0, a1, a2, a3, a4, a5 = [c[..., k] for k in range(6)]
    val = a0 + (a1 + (a2 + (a3 + (a4 + a5 * x) * x) * x) * x) * x
    grd = a1 + (2 * a2 + (3 * a3 + (4 * a4 + 5 * a5 * x) * x) * x) * x
    return val, grd


def _fold_weights(em_x, em, is_sorted=True):
    """Padding fold of ``tabulate.cc:189-229``: weight 1 before the first slot j
    whose em_x equals the last slot's and whose em[1:4] are zero, weight
    (nnei - j) at it, 0 after it."""
    nloc, nnei = em_x.shape
    w = np.ones((nloc, nnei), dtype=em.dtype)
    if not is_sorted or nnei == 0:
        return w
    hit = (em_x == em_x[:, -1:]) & (em[:, :, 1] == 0) & (em[:, :, 2] == 0) & (em[:, :, 3] == 0)
    any_hit = hit.any(1)
    first = np.where(any_hit, hit.argmax(1), nnei)
    j = np.arange(nnei)[None, :]
    w = np.where(j < first[:, None], 1.0, np.where(j == first[:, None],
                 (nnei - first)[:, None].astype(np.float64), 0.0)).astype(em.dtype)
    return w


def _prep(table, info, em_x, em):
    T = table.dtype.type
    nloc, nnei, nd = em.shape
    L = table.shape[1] // 6
    xx = em_x.reshape(nloc, nnei)
    idx, loc, ext = locate_xx(info[0], info[1], info[2], info[3], info[4], xx, T)
    val, grd = _poly(table, idx, loc, L)
    val = val + grd * ext[..., None]
    return T, L, xx, val, grd


def tabulate_fusion_se_a(table, info, em_x, em, last_layer_size, is_sorted=True):
    """out [nloc, nd, L] = Σ_j w_j * em[j, m] * g_k(x_j)."""
    T, L, xx, val, _ = _prep(table, info, em_x, em)
    assert L == last_layer_size
    w = _fold_weights(xx, em, is_sorted)
    return np.einsum("ij,ijm,ijk->imk", w, em, val).astype(T)


def tabulate_fusion_se_a_grad(table, info, em_x, em, dy, is_sorted=True):
    """(dy_dem_x [nloc, nnei], dy_dem [nloc, nnei, nd])."""
    T, L, xx, val, grd = _prep(table, info, em_x, em)
    w = _fold_weights(xx, em, is_sorted)
    dot = np.einsum("ijm,imk->ijk", em, dy)              # em_j · dy[:, k]
    dx = w * np.einsum("ijk,ijk->ij", grd, dot)
    dem = w[..., None] * np.einsum("ijk,imk->ijm", val, dy)
    return dx.astype(T), dem.astype(T)


def tabulate_fusion_se_a_grad_grad(table, info, em_x, em, dz_dy_dem_x, dz_dy_dem,
                                   is_sorted=True):
    """dz_dy [nloc, nd, L] = Σ_j w_j (g_k(x_j) hh[j,m] + dz_x[j] g'_k(x_j) em[j,m])."""
    T, L, xx, val, grd = _prep(table, info, em_x, em)
    w = _fold_weights(xx, em, is_sorted)
    dzx = dz_dy_dem_x.reshape(xx.shape)
    out = np.einsum("ij,ijk,ijm->imk", w, val, dz_dy_dem)
    out += np.einsum("ij,ij,ijk,ijm->imk", w, dzx, grd, em)
    return out.astype(T)

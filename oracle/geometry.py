"""Oracle geometry: PBC wrap, periodic images and neighbour lists (CPU, FP64).

Test infrastructure only (see ``oracle/__init__.py``).

Restates, op for op:

* ``deepmd/pt/utils/region.py:5-105``  (phys2inter, inter2phys, to_face_distance,
  normalize_coord)
* ``deepmd/pt/utils/nlist.py:14-199,304-337,424-513``  (neighbour list, type
  partition, ghost images)
* ``deepmd/pt/model/model/make_model.py:541-658``  (format_nlist)

Distances use ``torch.linalg.norm`` on FP64 differences exactly as the
reference does; on x86 CPUs torch evaluates the 3-vector norm as
``sqrt(fma(z, z, fma(y, y, x*x)))`` (pinned by ``tests/test_oracle_geometry.py``),
which is the rounding the HIP kernels reproduce.
"""

from __future__ import annotations

import numpy as np
import torch


# ---------------------------------------------------------------- region ----

def phys2inter(coord: torch.Tensor, cell: torch.Tensor) -> torch.Tensor:
    """``region.py:5-28``: fractional = coord @ inv(cell)."""
    rec_cell, _ = torch.linalg.inv_ex(cell)
    return torch.matmul(coord, rec_cell)


def inter2phys(coord: torch.Tensor, cell: torch.Tensor) -> torch.Tensor:
    """``region.py:31-47``: physical = fractional @ cell."""
    return torch.matmul(coord, cell)


def to_face_distance(cell: torch.Tensor) -> torch.Tensor:
    """``region.py:50-83``: V / |a_j x a_k| for the three cell faces."""
    cshape = cell.shape
    c = cell.view(-1, 3, 3)
    volume = torch.linalg.det(c)
    h_yz = volume / torch.linalg.norm(torch.cross(c[:, 1], c[:, 2], dim=-1), dim=-1)
    h_zx = volume / torch.linalg.norm(torch.cross(c[:, 2], c[:, 0], dim=-1), dim=-1)
    h_xy = volume / torch.linalg.norm(torch.cross(c[:, 0], c[:, 1], dim=-1), dim=-1)
    return torch.stack([h_yz, h_zx, h_xy], dim=1).view(list(cshape[:-2]) + [3])


def normalize_coord(coord: torch.Tensor, cell: torch.Tensor) -> torch.Tensor:
    """``region.py:86-105``: wrap into the cell via remainder(frac, 1)."""
    icoord = phys2inter(coord, cell)
    icoord = torch.remainder(icoord, 1.0)
    return inter2phys(icoord, cell)


# ---------------------------------------------------------------- images ----

def image_nbuff(cell: torch.Tensor, rcut: float) -> torch.Tensor:
    """``nlist.py:474-481``: images per direction, max over frames."""
    to_face = to_face_distance(cell.view(-1, 3, 3))
    nbuff = torch.ceil(rcut / to_face).to(torch.int64)
    return torch.amax(nbuff, dim=0)


def image_shifts(nbuff) -> torch.Tensor:
    """``nlist.py:482-498``: integer shift triples ordered by
    ``torch.argsort(norm)`` (the CPU, non-stable argsort the reference runs
    when its device is the CPU)."""
    nb = [int(x) for x in nbuff]
    xi = torch.arange(-nb[0], nb[0] + 1, 1, dtype=torch.int64)
    yi = torch.arange(-nb[1], nb[1] + 1, 1, dtype=torch.int64)
    zi = torch.arange(-nb[2], nb[2] + 1, 1, dtype=torch.int64)
    eye_3 = torch.eye(3, dtype=torch.float64)
    xyz = xi.view(-1, 1, 1, 1) * eye_3[0]
    xyz = xyz + yi.view(1, -1, 1, 1) * eye_3[1]
    xyz = xyz + zi.view(1, 1, -1, 1) * eye_3[2]
    xyz = xyz.view(-1, 3)
    return xyz[torch.argsort(torch.linalg.norm(xyz, dim=-1))]


def extend_coord_with_ghosts(coord, atype, cell, rcut):
    """``nlist.py:424-513``: image-major extended coordinates.

    Returns ``(extended_coord [nf, nall*3], extended_atype [nf, nall],
    mapping [nf, nall])`` with ``nall = ns * nloc`` and extended index
    ``e = s * nloc + i`` for shift ``s`` in :func:`image_shifts` order.
    """
    nf, nloc = atype.shape
    aidx = torch.tile(torch.arange(nloc, dtype=torch.int64).unsqueeze(0), [nf, 1])
    if cell is None:
        return coord.clone().view(nf, nloc * 3), atype.clone(), aidx
    coord = coord.view(nf, nloc, 3)
    cell = cell.view(nf, 3, 3)
    shift_idx = image_shifts(image_nbuff(cell, rcut))
    ns = shift_idx.shape[0]
    shift_vec = torch.einsum("sd,fdk->fsk", shift_idx, cell)
    ext = coord[:, None, :, :] + shift_vec[:, :, None, :]
    ext_atype = torch.tile(atype.unsqueeze(-2), [1, ns, 1])
    ext_aidx = torch.tile(aidx.unsqueeze(-2), [1, ns, 1])
    return (ext.reshape(nf, ns * nloc * 3), ext_atype.view(nf, ns * nloc),
            ext_aidx.view(nf, ns * nloc))


# ---------------------------------------------------------- neighbour list --

def nlist_distinguish_types(nlist, atype, sel):
    """``nlist.py:304-337``: stable per-type partition, truncated to sel[t]."""
    nf, nloc, nnei = nlist.shape
    ret = []
    tmp_atype = torch.tile(atype.unsqueeze(1), [1, nloc, 1])
    mask = nlist == -1
    tnlist = torch.gather(tmp_atype, 2, nlist.masked_fill(mask, 0))
    tnlist = tnlist.masked_fill(mask, -1)
    for ii, ss in enumerate(sel):
        pick = (tnlist == ii).to(torch.int32)
        pick, imap = torch.sort(pick, dim=-1, descending=True, stable=True)
        inlist = torch.gather(nlist, 2, imap)
        inlist = inlist.masked_fill(~(pick.to(torch.bool)), -1)
        ret.append(inlist[..., :ss])
    return torch.concat(ret, dim=-1)


def _trim_mask(is_vir, rr, nlist, rcut, nsel):
    """``nlist.py:153-199`` (mixed-type branch): pad/trim to nsel and drop
    ``rr > rcut`` and virtual centres."""
    nf, nloc, nnei = rr.shape
    if nsel <= nnei:
        rr = rr[:, :, :nsel]
        nlist = nlist[:, :, :nsel]
    else:
        rr = torch.cat([rr, torch.ones([nf, nloc, nsel - nnei], dtype=rr.dtype) + rcut], -1)
        nlist = torch.cat([nlist, torch.ones([nf, nloc, nsel - nnei], dtype=nlist.dtype)], -1)
    return torch.where(torch.logical_or(rr > rcut, is_vir[:, :nloc, None]), -1, nlist)


def build_neighbor_list_dense(ext_coord, ext_atype, nloc, rcut, sel,
                              distinguish_types=True):
    """``nlist.py:49-150``: dense O(nloc*nall) distances + ``topk``.

    Used for small frames; :func:`build_neighbor_list_cells` gives the same
    list (up to the order of exactly equal distances) for large ones.
    """
    nf = ext_coord.shape[0]
    nall = ext_coord.shape[1] // 3
    if isinstance(sel, int):
        sel = [sel]
    xmax = torch.max(ext_coord) + 2.0 * rcut if ext_coord.numel() > 0 else 2.0 * rcut
    xyz = ext_coord.view(nf, nall, 3)
    is_vir = ext_atype < 0
    vxyz = torch.where(is_vir[:, :, None], xmax, xyz)
    diff = vxyz.unsqueeze(1) - vxyz[:, :nloc, :].unsqueeze(2)
    rr = torch.linalg.norm(diff, dim=-1)
    idx = torch.arange(min(nloc, nall), dtype=torch.int64)
    rr[:, idx, idx] -= 1.0
    nsel = sum(sel)
    top_k = min(nsel + 1, rr.shape[-1])
    rr, nlist = torch.topk(rr, top_k, largest=False)
    rr = rr[:, :, 1:]
    nlist = nlist[:, :, 1:]
    nlist = _trim_mask(is_vir, rr, nlist, rcut, nsel)
    if distinguish_types:
        return nlist_distinguish_types(nlist, ext_atype, sel)
    return nlist


def build_neighbor_list_cells(ext_coord, ext_atype, nloc, rcut, sel,
                              distinguish_types=True):
    """O(N) restatement of :func:`build_neighbor_list_dense` for large frames.

    Same contract as ``nlist.py:49-150``: the Σsel nearest extended atoms with
    ``rr <= rcut`` in ascending FP64 distance (ties broken by extended index),
    then the per-type partition.  Candidates come from a Cartesian cell grid of
    edge ``rcut`` over the real extended atoms; distances are the same
    ``torch.linalg.norm`` of the same FP64 differences.
    """
    if isinstance(sel, int):
        sel = [sel]
    nsel = sum(sel)
    nf = ext_coord.shape[0]
    nall = ext_coord.shape[1] // 3
    out = torch.full((nf, nloc, nsel), -1, dtype=torch.int64)
    for ff in range(nf):
        xyz = ext_coord[ff].view(nall, 3).numpy()
        at = ext_atype[ff].numpy()
        real = np.nonzero(at >= 0)[0]
        if real.size == 0:
            continue
        lo = xyz[real].min(0)
        grid = np.floor((xyz[real] - lo) / rcut).astype(np.int64)
        dims = grid.max(0) + 1
        lin = (grid[:, 0] * dims[1] + grid[:, 1]) * dims[2] + grid[:, 2]
        order = np.argsort(lin, kind="stable")
        sorted_lin = lin[order]
        members = real[order]
        ncell = int(dims.prod())
        start = np.searchsorted(sorted_lin, np.arange(ncell + 1))
        # group local real centres by cell
        cen = real[real < nloc]
        cen_cell = np.floor((xyz[cen] - lo) / rcut).astype(np.int64)
        cen_lin = (cen_cell[:, 0] * dims[1] + cen_cell[:, 1]) * dims[2] + cen_cell[:, 2]
        offs = np.array([(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)])
        for cl in np.unique(cen_lin):
            cc = cen[cen_lin == cl]
            g = cen_cell[np.nonzero(cen_lin == cl)[0][0]]
            nb = g[None, :] + offs
            ok = np.all((nb >= 0) & (nb < dims), axis=1)
            nb = nb[ok]
            nbl = (nb[:, 0] * dims[1] + nb[:, 1]) * dims[2] + nb[:, 2]
            cand = np.concatenate([members[start[c]:start[c + 1]] for c in nbl])
            cand.sort()
            diff = torch.from_numpy(xyz[cand][None, :, :] - xyz[cc][:, None, :])
            rr = torch.linalg.norm(diff, dim=-1).numpy()
            rr[cand[None, :] == cc[:, None]] = -1.0
            for row, ii in enumerate(cc):
                r = rr[row]
                keep = np.nonzero(r <= rcut)[0]
                o = np.lexsort((cand[keep], r[keep]))
                picked = cand[keep[o]][1:nsel + 1]
                out[ff, ii, : picked.size] = torch.from_numpy(picked)
    if distinguish_types:
        return nlist_distinguish_types(out, ext_atype, sel)
    return out


def build_neighbor_list(ext_coord, ext_atype, nloc, rcut, sel, distinguish_types=True,
                        dense_limit=4.0e7):
    """Dense reference algorithm when ``nloc*nall`` is small, cell list otherwise."""
    nall = ext_coord.shape[1] // 3
    if ext_coord.shape[0] * nloc * nall <= dense_limit:
        return build_neighbor_list_dense(ext_coord, ext_atype, nloc, rcut, sel, distinguish_types)
    return build_neighbor_list_cells(ext_coord, ext_atype, nloc, rcut, sel, distinguish_types)


def extend_input_and_build_neighbor_list(coord, atype, rcut, sel, box=None,
                                         mixed_types=True):
    """``nlist.py:14-46``: wrap, add images, build the Σsel nearest list."""
    nf, nloc = atype.shape
    if box is not None:
        cn = normalize_coord(coord.view(nf, nloc, 3), box.reshape(nf, 3, 3))
    else:
        cn = coord.clone()
    ext, ext_atype, mapping = extend_coord_with_ghosts(cn, atype, box, rcut)
    nlist = build_neighbor_list(ext, ext_atype, nloc, rcut, sel,
                                distinguish_types=not mixed_types)
    return ext.view(nf, -1, 3), ext_atype, mapping, nlist


def format_nlist(ext_coord, ext_atype, nlist, rcut, sel, extra_nlist_sort=False,
                 mixed_types=False):
    """``make_model.py:541-658``: pad to Σsel, optionally sort/truncate by
    distance, drop ``rr > rcut``, then the per-type partition."""
    nnei = sum(sel)
    n_nf, n_nloc, n_nnei = nlist.shape
    ext_coord = ext_coord.view(n_nf, -1, 3)
    if n_nnei < nnei:
        nlist = torch.cat([nlist, -1 * torch.ones([n_nf, n_nloc, nnei - n_nnei],
                                                  dtype=nlist.dtype)], dim=-1)
    if n_nnei > nnei or extra_nlist_sort:
        n_nf, n_nloc, n_nnei = nlist.shape
        m_real = nlist >= 0
        nlist = torch.where(m_real, nlist, 0)
        coord0 = ext_coord[:, :n_nloc, :]
        index = nlist.view(n_nf, n_nloc * n_nnei, 1).expand(-1, -1, 3)
        coord1 = torch.gather(ext_coord, 1, index).view(n_nf, n_nloc, n_nnei, 3)
        rr = torch.linalg.norm(coord0[:, :, None, :] - coord1, dim=-1)
        rr = torch.where(m_real, rr, float("inf"))
        rr, nmap = torch.sort(rr, dim=-1)
        nlist = torch.gather(nlist, 2, nmap)
        nlist = torch.where(rr > rcut, -1, nlist)
        nlist = nlist[..., :nnei]
    else:
        m_real = nlist >= 0
        coord0 = ext_coord[:, :n_nloc, :]
        index = torch.where(m_real, nlist, 0).view(n_nf, n_nloc * n_nnei, 1).expand(-1, -1, 3)
        coord1 = torch.gather(ext_coord, 1, index).view(n_nf, n_nloc, n_nnei, 3)
        rr = torch.linalg.norm(coord0[:, :, None, :] - coord1, dim=-1)
        nlist = torch.where(m_real & (rr > rcut), -1, nlist)
    if not mixed_types:
        nlist = nlist_distinguish_types(nlist, ext_atype, sel)
    return nlist

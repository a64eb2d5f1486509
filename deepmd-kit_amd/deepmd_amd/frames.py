"""Host-side set-up of a packed batch of frames for the neighbour search.

For every frame this computes, on the CPU and with the same torch routines the
reference uses, the quantities the HIP kernels need:

* the cell inverse (``torch.linalg.inv_ex``, deepmd/pt/utils/region.py:24),
* face distances and the image count ``nbuff = ceil(rcut / face)``
  (region.py:50-83, nlist.py:474-481; max over frames for uniform batches),
* the shift list ordered by ``torch.argsort`` of the shift norms
  (nlist.py:482-498) and ``shift_vec = shift_idx @ box`` (nlist.py:501),
* a fractional cell grid with cell width > rcut (so neighbours lie in the
  adjacent cells), the per-frame cell offsets and search ranges.

The result (:class:`FrameSet`) is uploaded once and reused by every step on
the same boxes.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from ._lib import Frame

_EPS = 1e-9


def to_face_distance(box: torch.Tensor) -> torch.Tensor:
    """V / |a_j x a_k| per face (deepmd/pt/utils/region.py:71-83), box [3,3]."""
    c = box.view(1, 3, 3)
    v = torch.linalg.det(c)
    h0 = v / torch.linalg.norm(torch.cross(c[:, 1], c[:, 2], dim=-1), dim=-1)
    h1 = v / torch.linalg.norm(torch.cross(c[:, 2], c[:, 0], dim=-1), dim=-1)
    h2 = v / torch.linalg.norm(torch.cross(c[:, 0], c[:, 1], dim=-1), dim=-1)
    return torch.stack([h0, h1, h2], dim=1)[0]


def shift_list(nb) -> torch.Tensor:
    """Integer shifts in [-nb, nb]^3 in the reference's order: x-major grid,
    then ``torch.argsort`` of their norms on the CPU (nlist.py:482-498)."""
    xi = torch.arange(-nb[0], nb[0] + 1, 1, dtype=torch.int64)
    yi = torch.arange(-nb[1], nb[1] + 1, 1, dtype=torch.int64)
    zi = torch.arange(-nb[2], nb[2] + 1, 1, dtype=torch.int64)
    eye = torch.eye(3, dtype=torch.float64)
    xyz = xi.view(-1, 1, 1, 1) * eye[0]
    xyz = xyz + yi.view(1, -1, 1, 1) * eye[1]
    xyz = xyz + zi.view(1, 1, -1, 1) * eye[2]
    xyz = xyz.view(-1, 3)
    return xyz[torch.argsort(torch.linalg.norm(xyz, dim=-1))]


@dataclass
class FrameSet:
    frames_dev: torch.Tensor      # uint8 bytes of dpk_frame[nframes]
    frames_host: list             # Frame structs
    img_pos: torch.Tensor         # int32 device
    shift_vec: torch.Tensor       # float64 device [nimg, 3]
    img_shift: torch.Tensor       # int32 device [nimg, 3]: integer shift of each image
    frame_of: torch.Tensor        # int32 device [natoms]
    atom_off: np.ndarray          # [nframes + 1]
    ncells: int
    nframes: int
    natoms: int
    max_cand: int
    density_max: float


def _cap(n):
    for c in (256, 512, 1024, 2048):
        if n <= c:
            return c
    return 2048


def make_frames(nlocs, boxes, rcut: float, device, coords_host=None,
                shared_images: bool = True) -> FrameSet:
    """Frames of a packed batch.

    nlocs: atoms per frame; boxes: list of [3,3] float64 arrays (rows = cell
    vectors) or None for open boundaries (then ``coords_host`` [natoms,3] is
    needed to size the grid).  ``shared_images`` uses the max image count over
    frames, as the reference does for a uniform batch.
    """
    nframes = len(nlocs)
    atom_off = np.zeros(nframes + 1, dtype=np.int64)
    atom_off[1:] = np.cumsum(nlocs)
    natoms = int(atom_off[-1])
    rc = float(rcut)
    faces, nbs = [], []
    for b in boxes:
        if b is None:
            faces.append(None)
            nbs.append(np.zeros(3, dtype=np.int64))
            continue
        bt = torch.as_tensor(np.asarray(b, dtype=np.float64).reshape(3, 3))
        f = to_face_distance(bt)
        faces.append(f)
        nbs.append(torch.ceil(rc / f).to(torch.int64).numpy())
    if shared_images:
        nb_all = np.max(np.stack(nbs), axis=0)
        nbs = [nb_all if b is not None else np.zeros(3, dtype=np.int64) for b in boxes]
    frames, img_tabs, shift_tabs, int_tabs = [], [], [], []
    img_off, cell_off, dens = 0, 0, 0.0
    for f in range(nframes):
        fr = Frame()
        fr.atom_off = int(atom_off[f])
        fr.out_off = int(atom_off[f])
        fr.nloc = int(nlocs[f])
        nb = [int(x) for x in nbs[f]]
        b = boxes[f]
        if b is not None:
            bt = torch.as_tensor(np.asarray(b, dtype=np.float64).reshape(3, 3))
            rec, _ = torch.linalg.inv_ex(bt)
            face = faces[f].numpy()
            for k in range(9):
                fr.box[k] = float(bt.view(-1)[k])
                fr.rec[k] = float(rec.reshape(-1)[k])
            fr.pbc = 1
            nc = [max(1, int(math.floor(face[d] / (rc * (1 + _EPS))))) for d in range(3)]
            rng = [max(1, int(math.ceil(rc * (1 + _EPS) / (face[d] / nc[d])))) for d in range(3)]
            shifts = shift_list(nb)
            shift_vec = torch.einsum("sd,dk->sk", shifts.to(torch.float64), bt)
            vol = abs(float(torch.linalg.det(bt)))
        else:
            assert coords_host is not None, "open-boundary frames need host coordinates"
            x = np.asarray(coords_host[fr.atom_off:fr.atom_off + fr.nloc], dtype=np.float64)
            lo = x.min(0) if len(x) else np.zeros(3)
            hi = x.max(0) if len(x) else np.ones(3)
            ext = np.maximum(hi - lo, 1e-6) * (1 + 1e-6) + 1e-6
            fr.pbc = 0
            nc = [max(1, int(math.floor(ext[d] / (rc * (1 + _EPS))))) for d in range(3)]
            rng = [1, 1, 1]
            for k in range(3):
                fr.origin[k] = float(lo[k])
                fr.inv_extent[k] = float(1.0 / ext[k])
            shifts = torch.zeros((1, 3), dtype=torch.int64)
            shift_vec = torch.zeros((1, 3), dtype=torch.float64)
            vol = float(np.prod(ext))
        for d in range(3):
            fr.nc[d], fr.rng[d], fr.nb[d] = nc[d], rng[d], nb[d]
        fr.ns = int(shifts.shape[0])
        w = [2 * nb[d] + 1 for d in range(3)]
        tab = np.zeros(w[0] * w[1] * w[2], dtype=np.int32)
        for p, s in enumerate(shifts.round().to(torch.int64).tolist()):
            tab[((s[0] + nb[0]) * w[1] + (s[1] + nb[1])) * w[2] + (s[2] + nb[2])] = p
        assert tab.size == fr.ns
        fr.img_off = img_off
        fr.cell_off = cell_off
        img_off += fr.ns
        cell_off += nc[0] * nc[1] * nc[2]
        img_tabs.append(tab)
        shift_tabs.append(shift_vec.numpy())
        int_tabs.append(shifts.round().to(torch.int64).numpy().astype(np.int32))
        frames.append(fr)
        if fr.nloc:
            dens = max(dens, fr.nloc / max(vol, 1e-12))
    arr = (Frame * nframes)(*frames)
    raw = bytes(ctypes.string_at(ctypes.addressof(arr), ctypes.sizeof(arr)))
    frames_dev = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(device)
    img_pos = torch.from_numpy(np.concatenate(img_tabs)).to(device)
    shift_vec = torch.from_numpy(np.concatenate(shift_tabs).reshape(-1, 3).copy()).to(device)
    img_shift = torch.from_numpy(np.concatenate(int_tabs).reshape(-1, 3).copy()).to(device)
    frame_of = torch.from_numpy(np.repeat(np.arange(nframes, dtype=np.int32),
                                          np.asarray(nlocs))).to(device)
    est = dens * 4.0 / 3.0 * math.pi * rc ** 3
    return FrameSet(frames_dev=frames_dev, frames_host=frames, img_pos=img_pos,
                    shift_vec=shift_vec, img_shift=img_shift, frame_of=frame_of,
                    atom_off=atom_off,
                    ncells=cell_off, nframes=nframes, natoms=natoms,
                    max_cand=_cap(int(est * 1.15) + 16), density_max=dens)

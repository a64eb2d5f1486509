"""``DeepPot``-style inference front end (deepmd/infer/deep_pot.py:91-175).

    dp = DeepPot("model.yaml" | "model.pt" | model_dict, precision="f64")
    e, f, v = dp.eval(coords, cells, atom_types)                 # [nf,1], [nf,N,3], [nf,9]
    e, f, v, ae, av = dp.eval(coords, cells, atom_types, atomic=True)

``coords`` [nf, natoms, 3] (or [nf, natoms*3]), ``cells`` [nf, 9] or None,
``atom_types`` [natoms] (or [nf, natoms] with ``mixed_type=True``).  Inputs
are copied host -> device, evaluated by the libdpk kernels and copied back.

An MD engine that owns the neighbour list (local + ghost atoms and a full
list built with rcut + skin) calls, as it would DeepPotPT::compute
(source/api_cc/src/DeepPotPT.cc:171-378):

    e, f, v = dp.compute_lower(coord, atype, box, nghost, InputNlist(...), ago)
"""

from __future__ import annotations

import numpy as np
import torch

from .model import SeAModel


def load_model_dict(path: str, head=None) -> dict:
    """Model dict from a serialized model (.yaml / .json, the dpmodel
    serialize() layout) or a PyTorch training checkpoint (.pt, state dict
    read with weights_only=True; ``head`` picks a multi-task branch)."""
    if path.endswith((".yaml", ".yml")):
        import yaml
        with open(path) as fh:
            return yaml.safe_load(fh)
    if path.endswith(".json"):
        import json
        with open(path) as fh:
            return json.load(fh)
    if path.endswith(".pt"):
        from .checkpoint import load_checkpoint
        return load_checkpoint(path, head)
    raise ValueError(f"unsupported model file {path!r} (serialized .yaml / .json, "
                     f"checkpoint .pt)")


class InputNlist:
    """Host neighbour list in the layout of deepmd::InputNlist
    (source/api_cc/include/common.h; an MD engine's full list): ``ilist`` the
    centre atoms, ``numneigh`` the row lengths and ``firstneigh`` the rows
    (indices into the caller's local + ghost atoms).  ``mask`` is and-ed into
    every index, as the engine's neighbour-bit mask is."""

    def __init__(self, ilist, numneigh, firstneigh, mask: int = -1):
        self.ilist = np.asarray(ilist, dtype=np.int64)
        self.numneigh = np.asarray(numneigh, dtype=np.int64)
        self.firstneigh = [np.asarray(r, dtype=np.int64) for r in firstneigh]
        self.mask = int(mask)
        if not (len(self.ilist) == len(self.numneigh) == len(self.firstneigh)):
            raise ValueError("ilist, numneigh and firstneigh differ in length")

    @classmethod
    def from_rows(cls, rows, mask: int = -1):
        return cls(np.arange(len(rows)), [len(r) for r in rows], rows, mask)

    @property
    def inum(self) -> int:
        return len(self.ilist)


def _select_real_atoms(atype, nghost: int, ntypes: int):
    """select_by_type (source/api_cc/src/common.cc:75-108): atoms with a type
    in [0, ntypes) are kept in order; fwd_map[old] = new or -1, bkw_map[new] =
    old, and the count of kept ghosts."""
    atype = np.asarray(atype)
    nall = atype.size
    keep = (atype >= 0) & (atype < ntypes)
    bkw = np.nonzero(keep)[0]
    fwd = np.full(nall, -1, dtype=np.int64)
    fwd[bkw] = np.arange(bkw.size)
    nghost_real = int(np.count_nonzero(keep[nall - nghost:])) if nghost else 0
    return fwd, bkw, nghost_real


def _shuffled_rows(nlist: InputNlist, nloc: int, fwd):
    """NeighborListData::copy_from_nlist + shuffle_exclude_empty + padding
    (common.cc:297-379): rows of the first ``nloc`` list entries, indices
    masked and remapped through fwd_map, rows and entries of removed atoms
    dropped, then padded with -1 to the longest row."""
    ilist = nlist.ilist[:nloc] & nlist.mask
    rows = []
    for ii in range(nloc):
        i = int(ilist[ii])
        i = int(fwd[i]) if 0 <= i < fwd.size else i
        if i < 0:
            continue
        r = np.asarray(nlist.firstneigh[ii][:int(nlist.numneigh[ii])], dtype=np.int64) & nlist.mask
        ok = (r >= 0) & (r < fwd.size)
        r = np.where(ok, fwd[np.where(ok, r, 0)], r)
        rows.append(r[r >= 0])
    width = max((len(r) for r in rows), default=0)
    out = np.full((len(rows), width), -1, dtype=np.int64)
    for k, r in enumerate(rows):
        out[k, :len(r)] = r
    return out


class DeepPot:
    def __init__(self, model, precision: str = "f64", device: str = "cuda", head=None):
        d = load_model_dict(model, head) if isinstance(model, str) else model
        self.model = SeAModel(d, precision=precision, device=device)
        self._cache = None

    def get_type_map(self):
        return self.model.type_map

    def get_rcut(self):
        return self.model.rcut

    def get_ntypes(self):
        return self.model.ntypes

    def get_sel(self):
        return list(self.model.sel)

    def get_dim_fparam(self):
        return self.model.numb_fparam

    def get_dim_aparam(self):
        return self.model.numb_aparam

    def has_default_fparam(self):
        return self.model.has_default_fparam()

    def _frame_atom_params(self, fparam, aparam, nf, natoms):
        """The reference's input shapes (deepmd/infer/deep_eval.py:951-972):
        fparam [nf, dim] or [dim] (tiled over frames); aparam [nf, natoms, dim],
        [natoms, dim] or [dim] (tiled)."""
        m = self.model
        if fparam is not None:
            fparam = np.asarray(fparam, dtype=np.float64)
            fdim = m.numb_fparam
            if fdim and fparam.size == nf * fdim:
                fparam = fparam.reshape(nf, fdim)
            elif fdim and fparam.size == fdim:
                fparam = np.tile(fparam.reshape(-1), (nf, 1))
            else:
                raise RuntimeError(f"got wrong size of frame param, should be either {nf} x "
                                   f"{fdim} or {fdim}")
        if aparam is not None:
            aparam = np.asarray(aparam, dtype=np.float64)
            fdim = m.numb_aparam
            if fdim and aparam.size == nf * natoms * fdim:
                aparam = aparam.reshape(nf, natoms * fdim)
            elif fdim and aparam.size == natoms * fdim:
                aparam = np.tile(aparam.reshape(-1), (nf, 1))
            elif fdim and aparam.size == fdim:
                aparam = np.tile(aparam.reshape(-1), (nf, natoms))
            else:
                raise RuntimeError(f"got wrong size of frame param, should be either {nf} x "
                                   f"{natoms} x {fdim} or {natoms} x {fdim} or {fdim}")
        return fparam, aparam

    def eval(self, coords, cells, atom_types, atomic: bool = False, fparam=None, aparam=None,
             mixed_type: bool = False, **kwargs):
        coords = np.asarray(coords, dtype=np.float64)
        nf = coords.shape[0]
        coords = coords.reshape(nf, -1, 3)
        natoms = coords.shape[1]
        at = np.asarray(atom_types, dtype=np.int32)
        at = at.reshape(nf, natoms) if mixed_type else np.tile(at.reshape(1, natoms), (nf, 1))
        boxes = None if cells is None else np.asarray(cells, dtype=np.float64).reshape(nf, 9)
        # the cell grid of open-boundary frames depends on the coordinates
        key = (coords.shape, at.tobytes(),
               coords.tobytes() if boxes is None else boxes.tobytes())
        if self._cache is None or self._cache[0] != key:
            s = self.model.system(coords, at, boxes)
            self._cache = (key, s)
        else:
            s = self._cache[1]
            s.coord.copy_(torch.from_numpy(coords.reshape(-1, 3)))
        if self.model.nx:
            fparam, aparam = self._frame_atom_params(fparam, aparam, nf, natoms)
            self.model.set_params(s, fparam, aparam)
        elif fparam is not None or aparam is not None:
            raise ValueError("the model takes no frame / atomic parameters")
        out = self.model.evaluate(s, atomic_virial=atomic)
        e = out["energy"].cpu().numpy().reshape(nf, 1)
        f = out["force"].cpu().numpy().reshape(nf, natoms, 3)
        v = out["virial"].cpu().numpy().reshape(nf, 9)
        if not atomic:
            return e, f, v
        ae = out["atom_energy"].cpu().numpy().reshape(nf, natoms, 1)
        av = out["atom_virial"].cpu().numpy().reshape(nf, natoms, 9)
        return e, f, v, ae, av

    def compute_lower(self, coord, atype, box, nghost: int, nlist: InputNlist | None, ago: int = 0,
                       atomic: bool = False, fparam=None, aparam=None):
        """DeepPotPT::compute with an engine neighbour list
        (source/api_cc/src/DeepPotPT.cc:171-378): ``coord`` [nall*3] local then
        ghost atoms, ``atype`` [nall] (types outside [0, ntypes), the engine's
        NULL type, are removed before the model and get zero outputs), ``nghost``
        ghost count, ``nlist`` the engine's list (stored when ``ago == 0`` and
        reused otherwise).  ``box`` is not needed by the model (the region is
        already extended).  Returns energy [1], force [nall*3] (extended: the
        engine adds ghost forces back to their owners), virial [9] and, with
        ``atomic``, atom energy [nall] and atom virial [nall*9].  ``fparam``
        [dim_fparam] and ``aparam`` [nloc, dim_aparam] (local atoms) feed the
        fitting's frame / atomic parameters."""
        del box
        coord = np.asarray(coord, dtype=np.float64).reshape(-1, 3)
        atype = np.asarray(atype, dtype=np.int64).reshape(-1)
        nall = atype.size
        if coord.shape[0] != nall:
            raise ValueError(f"coord holds {coord.shape[0]} atoms, atype {nall}")
        fwd, bkw, nghost_real = _select_real_atoms(atype, nghost, self.model.ntypes)
        nall_real = bkw.size
        nloc_real = nall_real - nghost_real
        if ago == 0:
            if nlist is None:
                raise ValueError("ago == 0 needs a neighbour list")
            self._lower_rows = _shuffled_rows(nlist, nall - nghost, fwd)
        elif getattr(self, "_lower_rows", None) is None:
            raise ValueError("ago > 0 before any neighbour list was given")
        rows = self._lower_rows
        if rows.shape[0] != nloc_real:
            raise ValueError(f"the neighbour list has {rows.shape[0]} rows for {nloc_real} local atoms")
        if aparam is not None:  # rows of the local atoms that stay (select_by_type)
            ap = np.asarray(aparam, dtype=np.float64).reshape(nall - nghost, -1)
            aparam = ap[bkw[:nloc_real]]
        sysl = getattr(self, "_lower_sys", None)
        # between re-neighbourings (ago > 0) the stored list, types and slot
        # order are reused without host synchronisation
        reuse = (ago > 0 and sysl is not None and fparam is None and aparam is None
                 and sysl.fs.nall == nall_real and sysl.fs.natoms == nloc_real)
        out = self.model.forward_lower(torch.from_numpy(coord[bkw]).view(1, nall_real, 3),
                                       torch.from_numpy(atype[bkw]).view(1, nall_real),
                                       None if reuse else torch.from_numpy(rows).view(1, nloc_real, -1),
                                       do_atomic_virial=atomic, system=sysl,
                                       fparam=fparam, aparam=aparam, ago=1 if reuse else 0)
        self._lower_sys = self.model._last_lower_system
        e = out["energy"].cpu().numpy().reshape(1)
        fr = out["extended_force"].cpu().numpy().reshape(nall_real, 3)
        f = np.zeros((nall, 3))
        f[bkw] = fr
        v = out["virial"].cpu().numpy().reshape(9)
        if not atomic:
            return e, f.reshape(-1), v
        ae = np.zeros(nall)
        ae[bkw[:nloc_real]] = out["atom_energy"].cpu().numpy().reshape(-1)
        av = np.zeros((nall, 9))
        av[bkw] = out["extended_virial"].cpu().numpy().reshape(nall_real, 9)
        return e, f.reshape(-1), v, ae, av.reshape(-1)

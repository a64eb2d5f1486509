"""``DeepPot``-style inference front end (deepmd/infer/deep_pot.py:91-175).

    dp = DeepPot("model.yaml" | model_dict, precision="f64")
    e, f, v = dp.eval(coords, cells, atom_types)                 # [nf,1], [nf,N,3], [nf,9]
    e, f, v, ae, av = dp.eval(coords, cells, atom_types, atomic=True)

``coords`` [nf, natoms, 3] (or [nf, natoms*3]), ``cells`` [nf, 9] or None,
``atom_types`` [natoms] (or [nf, natoms] with ``mixed_type=True``).  Inputs
are copied host -> device, evaluated by the libdpk kernels and copied back.
"""

from __future__ import annotations

import numpy as np
import torch

from .model import SeAModel


def load_model_dict(path: str) -> dict:
    if path.endswith((".yaml", ".yml")):
        import yaml
        with open(path) as fh:
            return yaml.safe_load(fh)
    if path.endswith(".json"):
        import json
        with open(path) as fh:
            return json.load(fh)
    raise ValueError(f"unsupported model file {path!r} (serialized .yaml / .json)")


class DeepPot:
    def __init__(self, model, precision: str = "f64", device: str = "cuda"):
        d = load_model_dict(model) if isinstance(model, str) else model
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

    def eval(self, coords, cells, atom_types, atomic: bool = False, fparam=None, aparam=None,
             mixed_type: bool = False, **kwargs):
        if fparam is not None or aparam is not None:
            raise NotImplementedError("frame / atomic parameters are not part of se_e2_a energy")
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
        out = self.model.evaluate(s, atomic_virial=atomic)
        e = out["energy"].cpu().numpy().reshape(nf, 1)
        f = out["force"].cpu().numpy().reshape(nf, natoms, 3)
        v = out["virial"].cpu().numpy().reshape(nf, 9)
        if not atomic:
            return e, f, v
        ae = out["atom_energy"].cpu().numpy().reshape(nf, natoms, 1)
        av = out["atom_virial"].cpu().numpy().reshape(nf, natoms, 9)
        return e, f, v, ae, av

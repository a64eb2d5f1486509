"""Oracle se_e2_a energy model: env matrix -> embedding -> descriptor -> fitting
-> energy, with forces and virials by torch autograd (CPU, FP64).

Test infrastructure only (see ``oracle/__init__.py``).

Restates, op for op:

* ``deepmd/pt/model/descriptor/env_mat.py:18-119`` and
  ``deepmd/pt/utils/preprocess.py:9-18``  (smooth env matrix)
* ``deepmd/pt/model/network/mlp.py:188-219``  (MLPLayer.forward)
* ``deepmd/pt/model/descriptor/se_a.py:742-858``  (DescrptBlockSeA.forward)
* ``deepmd/pt/model/task/fitting.py:734-899``  (energy fitting, non-mixed types)
* ``deepmd/pt/model/atomic_model/base_atomic_model.py:318-399,573-594``
* ``deepmd/pt/model/model/transform_output.py:20-209,367-433``  (E sum in FP64,
  F = -dE/dr_ext, virial F (x) r, atomic-virial correction, ghost -> local)
* ``deepmd/pt/model/model/make_model.py:136-220,385-466``

Model parameters (:class:`SeAModel`) are read from the reference's serialized
dict form (``source/tests/infer/deeppot_sea.yaml``); seeded random models are
produced in that same dict form by the product's model builder, so the oracle
and the HIP path always consume identical weights.
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np
import torch

from . import geometry


# ------------------------------------------------------------- parameters --

@dataclass
class Layer:
    w: torch.Tensor            # [in, out]
    b: torch.Tensor | None     # [out]
    idt: torch.Tensor | None   # [out]
    activation: str            # name of utils.py:174-212 ("none": identity)
    resnet: bool


@dataclass
class SeAModel:
    type_map: list
    rcut: float
    rcut_smth: float
    sel: list
    neuron: list
    axis_neuron: int
    type_one_side: bool
    davg: torch.Tensor         # [ntypes, nnei, 4]
    dstd: torch.Tensor         # [ntypes, nnei, 4]
    embeddings: list           # list[list[Layer]]  (NetworkCollection order)
    fitting: list              # list[list[Layer]]  one net per type
    bias_atom_e: torch.Tensor  # [ntypes, 1]
    out_bias: torch.Tensor     # [ntypes, 1]
    env_protection: float = 0.0
    exclude_types: list = field(default_factory=list)
    atom_exclude_types: list = field(default_factory=list)   # model level
    pair_exclude_types: list = field(default_factory=list)   # model level
    # frame / atomic parameters of the fitting (fitting.py:749-830)
    numb_fparam: int = 0
    numb_aparam: int = 0
    fparam_avg: torch.Tensor | None = None
    fparam_inv_std: torch.Tensor | None = None
    aparam_avg: torch.Tensor | None = None
    aparam_inv_std: torch.Tensor | None = None
    default_fparam: torch.Tensor | None = None
    use_aparam_as_mask: bool = False

    @property
    def ntypes(self):
        return len(self.sel)

    @property
    def nnei(self):
        return int(sum(self.sel))


def _arr(v):
    if v is None:
        return None
    if isinstance(v, dict):
        v = v["value"]
    if v is None:
        return None
    return torch.tensor(np.asarray(v, dtype=np.float64))


def _layers_from_dict(net: dict) -> list:
    out = []
    for ly in net["layers"]:
        var = ly["@variables"]
        act = ly.get("activation_function", "tanh")
        out.append(Layer(w=_arr(var["w"]), b=_arr(var.get("b")), idt=_arr(var.get("idt")),
                         activation="none" if act in (None, "none", "linear") else act,
                         resnet=bool(ly.get("resnet", False))))
    return out


def model_from_serialized(d: dict) -> SeAModel:
    """Read the reference's serialized model dict (``deeppot_sea.yaml`` 'model')."""
    m = d["model"] if "model" in d else d
    de, fi = m["descriptor"], m["fitting"]
    ntypes = len(de["sel"])
    nnei = sum(de["sel"])
    davg = _arr(de["@variables"]["davg"]).reshape(ntypes, nnei, 4)
    dstd = _arr(de["@variables"]["dstd"]).reshape(ntypes, nnei, 4)
    out_bias = _arr(m["@variables"]["out_bias"]).reshape(-1, ntypes, 1)[0]
    fv = fi["@variables"]
    nfp, nap = int(fi.get("numb_fparam", 0) or 0), int(fi.get("numb_aparam", 0) or 0)
    dfp = fi.get("default_fparam")
    return SeAModel(
        type_map=list(m.get("type_map", [])), rcut=float(de["rcut"]),
        rcut_smth=float(de["rcut_smth"]), sel=list(de["sel"]), neuron=list(de["neuron"]),
        axis_neuron=int(de["axis_neuron"]), type_one_side=bool(de["type_one_side"]),
        davg=davg, dstd=dstd,
        embeddings=[_layers_from_dict(n) for n in de["embeddings"]["networks"]],
        fitting=[_layers_from_dict(n) for n in fi["nets"]["networks"]],
        bias_atom_e=_arr(fi["@variables"]["bias_atom_e"]).reshape(ntypes, 1),
        out_bias=out_bias, env_protection=float(de.get("env_protection", 0.0)),
        exclude_types=list(de.get("exclude_types", [])),
        atom_exclude_types=sorted({int(t) for t in (m.get("atom_exclude_types") or [])}
                                  | {int(t) for t in (fi.get("exclude_types") or [])}),
        pair_exclude_types=[tuple(p) for p in (m.get("pair_exclude_types") or [])],
        numb_fparam=nfp, numb_aparam=nap,
        fparam_avg=_arr(fv.get("fparam_avg")) if nfp else None,
        fparam_inv_std=_arr(fv.get("fparam_inv_std")) if nfp else None,
        aparam_avg=_arr(fv.get("aparam_avg")) if nap else None,
        aparam_inv_std=_arr(fv.get("aparam_inv_std")) if nap else None,
        default_fparam=None if dfp is None else torch.tensor(dfp, dtype=torch.float64),
        use_aparam_as_mask=bool(fi.get("use_aparam_as_mask", False)),
    )


# ------------------------------------------------------------ primitives ----

def compute_smooth_weight(distance, rmin, rmax):
    """``preprocess.py:9-18``."""
    distance = torch.clamp(distance, min=rmin, max=rmax)
    uu = (distance - rmin) / (rmax - rmin)
    uu2 = uu * uu
    return uu2 * uu * (-6 * uu2 + 15 * uu - 10) + 1


def make_env_mat(nlist, coord, rcut, rcut_smth, protection=0.0, centers=None):
    """``env_mat.py:18-55``: (env_mat [nf,nloc,nnei,4], diff, sw).

    ``centers`` (optional int tensor [nloc]) selects which extended rows are the
    centres; the reference uses rows ``0..nloc-1``.
    """
    bsz, natoms, nnei = nlist.shape
    coord = coord.view(bsz, -1, 3)
    nall = coord.shape[1]
    mask = nlist >= 0
    nlist = torch.where(mask, nlist, nall)
    if centers is None:
        coord_l = coord[:, :natoms].view(bsz, -1, 1, 3)
    else:
        coord_l = coord[:, centers].view(bsz, -1, 1, 3)
    index = nlist.view(bsz, -1).unsqueeze(-1).expand(-1, -1, 3)
    coord_pad = torch.concat([coord, coord[:, -1:, :] + rcut], dim=1)
    coord_r = torch.gather(coord_pad, 1, index).view(bsz, natoms, nnei, 3)
    diff = coord_r - coord_l
    length = torch.linalg.norm(diff, dim=-1, keepdim=True)
    length = length + ~mask.unsqueeze(-1)
    t0 = 1 / (length + protection)
    t1 = diff / (length + protection) ** 2
    weight = compute_smooth_weight(length, rcut_smth, rcut) * mask.unsqueeze(-1)
    env_mat = torch.cat([t0, t1], dim=-1) * weight
    return env_mat, diff * mask.unsqueeze(-1), weight


def prod_env_mat(ext_coord, nlist, atype, mean, stddev, rcut, rcut_smth,
                 protection=0.0, centers=None):
    """``env_mat.py:58-119``: normalised env matrix (R - avg[t_i]) / std[t_i]."""
    em, diff, sw = make_env_mat(nlist, ext_coord, rcut, rcut_smth, protection, centers)
    return (em - mean[atype]) / stddev[atype], diff, sw


def activation(name, y):
    """``deepmd/pt/utils/utils.py:174-212`` (ActivationFn.forward) with
    ``SiLUT`` of ``:150-171`` for silut / custom_silu[:threshold]."""
    n = "linear" if name is None else str(name).lower()
    if n == "relu":
        return torch.nn.functional.relu(y)
    if n in ("gelu", "gelu_tf"):
        return torch.nn.functional.gelu(y, approximate="tanh")
    if n == "tanh":
        return torch.tanh(y)
    if n == "relu6":
        return torch.nn.functional.relu6(y)
    if n == "softplus":
        return torch.nn.functional.softplus(y)
    if n == "sigmoid":
        return torch.sigmoid(y)
    if n == "silu":
        return torch.nn.functional.silu(y)
    if n.startswith("silut") or n.startswith("custom_silu"):
        thr = float(n.split(":")[-1]) if ":" in n else 3.0
        sig = 1.0 / (1.0 + np.exp(-thr))
        slope = float(sig + thr * sig * (1.0 - sig))   # silu'(thr)
        const = float(thr * sig)                        # silu(thr)
        sig_y = torch.sigmoid(y)
        return torch.where(y >= thr, torch.tanh(slope * (y - thr)) + const, y * sig_y)
    if n in ("linear", "none"):
        return y
    raise RuntimeError(f"activation function {name} not supported")


def mlp_layer(x, ly: Layer):
    """``mlp.py:188-219``: act(x W + b) [* idt] [+ x | + (x, x)]."""
    y = torch.nn.functional.linear(x, ly.w.t(), ly.b)
    y = activation(ly.activation, y)
    if ly.idt is not None:
        y = y * ly.idt
    if ly.resnet:
        if x.shape[-1] == y.shape[-1]:
            y = y + x
        elif 2 * x.shape[-1] == y.shape[-1]:
            y = y + torch.concat([x, x], dim=-1)
    return y


def run_net(x, layers):
    for ly in layers:
        x = mlp_layer(x, ly)
    return x


def run_net_until_last(x, layers):
    for ly in layers[:-1]:
        x = mlp_layer(x, ly)
    return x


def exclude_mask(model: SeAModel, nlist, ext_atype, pairs=None, centre_types=None):
    """``deepmd/pt/utils/exclude_mask.py:64-140`` (PairExcludeMask): 0 where the
    (centre type, neighbour type) pair is excluded; padding slots (-1) count as
    a virtual type ``ntypes`` that is never excluded, so they keep mask 1, and
    a virtual centre (type -1) reads row ``ntypes`` (the negative index of
    ``type_i * (ntypes + 1) + type_j`` wraps there).  ``pairs`` defaults to
    the descriptor's exclude_types."""
    nf, nloc, nnei = nlist.shape
    pairs = model.exclude_types if pairs is None else pairs
    if not pairs:
        return torch.ones_like(nlist, dtype=torch.int32)
    nt = model.ntypes
    tab = torch.ones((nt + 1, nt + 1), dtype=torch.int32)
    for a, b in pairs:
        tab[a, b] = 0
        tab[b, a] = 0
    ti = ext_atype[:, :nloc].clone() if centre_types is None else centre_types.clone()
    ti = torch.where(ti >= 0, ti, nt)
    tj = torch.gather(ext_atype, 1, torch.where(nlist >= 0, nlist, 0).view(nf, -1)).view(nf, nloc, nnei)
    tj = torch.where(nlist >= 0, tj, nt)
    return tab[ti[:, :, None].expand(-1, -1, nnei), tj]


# ----------------------------------------------------------- descriptor -----

def descriptor_se_a(model: SeAModel, ext_coord, ext_atype, nlist, centers=None):
    """``se_a.py:742-858``: returns (descriptor [nf,nloc,ng*axis], rot_mat, sw)."""
    nf, nloc, nnei = nlist.shape
    atype = ext_atype[:, :nloc] if centers is None else ext_atype[:, centers]
    dmatrix, diff, sw = prod_env_mat(ext_coord, nlist, atype, model.davg, model.dstd,
                                     model.rcut, model.rcut_smth, model.env_protection,
                                     centers)
    dmatrix = dmatrix.view(-1, nnei, 4)
    emask = exclude_mask(model, nlist, ext_atype, centre_types=atype).view(dmatrix.shape[0], nnei)
    result, xyz = descriptor_from_env(model, dmatrix, atype, emask)
    x1 = xyz.permute(0, 2, 1)
    rot_mat = x1[:, :, 1:4]
    ng = model.neuron[-1]
    return result.view(nf, nloc, ng * model.axis_neuron), rot_mat.view(nf, nloc, ng, 3), sw


def descriptor_from_env(model: SeAModel, dmatrix, atype, emask):
    """``se_a.py:784-850`` from the normalised env matrix [nfnl, nnei, 4]:
    returns (descriptor [nfnl, ng, axis], xyz [nfnl, 4, ng] after /nnei)."""
    nfnl, nnei, _ = dmatrix.shape
    ng = model.neuron[-1]
    sec = [0] + list(np.cumsum(model.sel))
    xyz = torch.zeros([nfnl, 4, ng], dtype=dmatrix.dtype)
    for eidx, net in enumerate(model.embeddings):
        if model.type_one_side:
            ii, ti_mask = eidx, None
        else:
            ii, ti = eidx // model.ntypes, eidx % model.ntypes
            ti_mask = atype.ravel().eq(ti)
        if ti_mask is not None:
            mm = emask[ti_mask, sec[ii]:sec[ii + 1]]
            rr = dmatrix[ti_mask, sec[ii]:sec[ii + 1], :]
        else:
            mm = emask[:, sec[ii]:sec[ii + 1]]
            rr = dmatrix[:, sec[ii]:sec[ii + 1], :]
        rr = rr * mm[:, :, None]
        ss = rr[:, :, :1]
        gg = run_net(ss, net)
        gr = torch.matmul(rr.permute(0, 2, 1), gg)
        if ti_mask is not None:
            xyz[ti_mask] += gr
        else:
            xyz += gr
    xyz = xyz / nnei
    x1 = xyz.permute(0, 2, 1)
    x2 = xyz[:, :, 0:model.axis_neuron]
    return torch.matmul(x1, x2), xyz


# -------------------------------------------------------------- fitting -----

def fitting_input(model: SeAModel, descriptor, fparam=None, aparam=None):
    """``fitting.py:749-830``: the descriptor with the normalised frame
    parameters (tiled over atoms) and atomic parameters appended, or the
    default frame parameters when none are given."""
    nf, nloc, _ = descriptor.shape
    xx = descriptor
    if model.numb_fparam > 0:
        if fparam is None:
            if model.default_fparam is None:
                raise ValueError("fparam is required: the model has no default_fparam")
            fparam = torch.tile(model.default_fparam.view(1, -1), [nf, 1])
        fparam = torch.as_tensor(fparam, dtype=descriptor.dtype)
        if fparam.numel() != nf * model.numb_fparam:
            raise ValueError(f"input fparam: cannot reshape {list(fparam.shape)} "
                             f"into ({nf}, {model.numb_fparam}).")
        fparam = fparam.view(nf, model.numb_fparam)
        fparam = (fparam - model.fparam_avg.view(1, -1)) * model.fparam_inv_std.view(1, -1)
        xx = torch.cat([xx, torch.tile(fparam.view(nf, 1, -1), [1, nloc, 1])], dim=-1)
    if model.numb_aparam > 0 and not model.use_aparam_as_mask:
        if aparam is None:
            raise ValueError("aparam is required by the model")
        aparam = torch.as_tensor(aparam, dtype=descriptor.dtype)
        if aparam.numel() % (nf * model.numb_aparam) != 0:
            raise ValueError(f"input aparam: cannot reshape {list(aparam.shape)} "
                             f"into ({nf}, nloc, {model.numb_aparam}).")
        aparam = aparam.view(nf, -1, model.numb_aparam)
        aparam = (aparam - model.aparam_avg.view(1, 1, -1)) * model.aparam_inv_std.view(1, 1, -1)
        xx = torch.cat([xx, aparam], dim=-1)
    return xx


def fitting_energy(model: SeAModel, descriptor, atype, return_last_hidden=False, fparam=None,
                   aparam=None):
    """``fitting.py:734-899`` (non-mixed types): every type's net on every atom,
    masked by type, plus ``bias_atom_e``."""
    nf, nloc, _ = descriptor.shape
    descriptor = fitting_input(model, descriptor, fparam, aparam)
    outs = torch.zeros((nf, nloc, 1), dtype=descriptor.dtype)
    feat = None
    for ti, net in enumerate(model.fitting):
        mask = (atype == ti).unsqueeze(-1)
        prop = run_net(descriptor, net) + model.bias_atom_e[ti]
        outs = outs + torch.where(mask, prop, 0.0)
        if return_last_hidden:
            h = run_net_until_last(descriptor, net)
            feat = torch.where(mask, h, 0.0) if feat is None else feat + torch.where(mask, h, 0.0)
    return (outs, feat) if return_last_hidden else outs


def atomic_energy(model: SeAModel, ext_coord, ext_atype, nlist, centers=None, fparam=None,
                  aparam=None):
    """``base_atomic_model.py:318-399`` + ``dp_atomic_model.py:200-290``:
    per-atom energy [nf, nloc, 1] including ``out_bias`` and virtual masking.
    ``aparam`` [nf, nloc, numb_aparam] belongs to the rows of ``nlist``."""
    nf, nloc, _ = nlist.shape
    atype_l = ext_atype[:, :nloc] if centers is None else ext_atype[:, centers]
    if model.pair_exclude_types:  # base_atomic_model.py:380-383
        pm = exclude_mask(model, nlist, ext_atype, model.pair_exclude_types, atype_l)
        nlist = torch.where(pm == 1, nlist, -1)
    ext_mask = ext_atype >= 0
    eat = torch.where(ext_mask, ext_atype, 0)
    desc, _, _ = descriptor_se_a(model, ext_coord, eat, nlist, centers)
    atl = torch.where(atype_l >= 0, atype_l, 0)
    e = fitting_energy(model, desc, atl, fparam=fparam, aparam=aparam)
    e = e + model.out_bias[atl]
    amask = (atype_l >= 0).to(e.dtype)
    for t in model.atom_exclude_types:  # base_atomic_model.py:391-393 (AtomExcludeMask)
        amask = amask * (atype_l != t).to(e.dtype)
    return e * amask[:, :, None]


# ----------------------------------------------------------- full model -----

def atomic_virial_corr(ext_coord, atom_energy):
    """``transform_output.py:20-65``."""
    nall = ext_coord.shape[1]
    nloc = atom_energy.shape[1]
    coord = ext_coord[:, :nloc].detach()
    ce = coord * atom_energy
    s = torch.sum(ce, dim=1)
    corr = []
    for d in range(3):
        g = torch.autograd.grad([s[:, d:d + 1]], [ext_coord],
                                grad_outputs=[torch.ones_like(s[:, d:d + 1])],
                                create_graph=False, retain_graph=True)[0]
        corr.append(g.unsqueeze(-1))
    return torch.concat(corr, dim=-1)


def forward_lower(model: SeAModel, ext_coord, ext_atype, nlist, mapping=None,
                  do_atomic_virial=False, create_graph=False, extra_nlist_sort=False,
                  fparam=None, aparam=None):
    """``make_model.py:385-466`` + ``transform_output.py:155-209``.

    Returns dict with atom_energy [nf,nloc,1], energy [nf,1], extended_force
    [nf,nall,3], extended_virial [nf,nall,9] and virial [nf,9]."""
    nf = ext_atype.shape[0]
    ext_coord = ext_coord.view(nf, -1, 3)
    nlist = geometry.format_nlist(ext_coord, ext_atype, nlist, model.rcut, model.sel,
                                  extra_nlist_sort=extra_nlist_sort)
    cc = ext_coord.detach().clone().requires_grad_(True)
    ae = atomic_energy(model, cc, ext_atype, nlist, fparam=fparam, aparam=aparam)
    energy = torch.sum(ae.to(torch.float64), dim=1)
    g = torch.autograd.grad([energy], [cc], grad_outputs=[torch.ones_like(energy)],
                            create_graph=create_graph, retain_graph=True)[0]
    force = -g
    vir = torch.einsum("...ik,...ij->...ikj", force, cc)
    if do_atomic_virial:
        vir = vir + atomic_virial_corr(cc, ae)
    vir = vir.reshape(nf, -1, 9)
    return {"atom_energy": ae, "energy": energy, "extended_force": force,
            "extended_virial": vir, "virial": torch.sum(vir, dim=1), "nlist": nlist}


def communicate(ext_val, mapping, nloc):
    """``transform_output.py:367-433``: scatter-sum extended rows onto owners."""
    nf, nall, k = ext_val.shape
    out = torch.zeros((nf, nloc, k), dtype=ext_val.dtype)
    return torch.scatter_reduce(out, 1, mapping.view(nf, nall, 1).expand(-1, -1, k),
                                ext_val, reduce="sum")


def forward(model: SeAModel, coord, atype, box=None, do_atomic_virial=False, fparam=None,
            aparam=None):
    """``make_model.py:136-220``: DeepPot-style E/F/V for local atoms.

    ``coord`` [nf, nloc*3] FP64, ``atype`` [nf, nloc] int64, ``box`` [nf, 9] or
    None.  Returns energy [nf,1], atom_energy [nf,nloc,1], force [nf,nloc,3],
    virial [nf,9] and (if requested) atom_virial [nf,nloc,9]."""
    nf, nloc = atype.shape
    ext, ext_atype, mapping, nlist = geometry.extend_input_and_build_neighbor_list(
        coord.view(nf, nloc, 3), atype, model.rcut, model.sel, box=box, mixed_types=True)
    low = forward_lower(model, ext, ext_atype, nlist, mapping, do_atomic_virial,
                        fparam=fparam, aparam=aparam)
    out = {"energy": low["energy"], "atom_energy": low["atom_energy"],
           "force": communicate(low["extended_force"], mapping, nloc),
           "nlist": low["nlist"], "mapping": mapping, "extended_coord": ext,
           "extended_atype": ext_atype}
    av = communicate(low["extended_virial"], mapping, nloc)
    out["virial"] = torch.sum(av, dim=1)
    if do_atomic_virial:
        out["atom_virial"] = av
    return out


def forward_chunked(model: SeAModel, coord, atype, box, chunk=2048, nlist=None, fparam=None,
                    aparam=None):
    """Large-frame E/F/V with the same math as :func:`forward`, evaluating the
    centre atoms in chunks so per-slot activations stay bounded.

    E = Σ_chunks E_c and F_ext = -Σ_chunks dE_c/dr_ext (linearity of the
    gradient); the virial is Σ_ext F_ext (x) r_ext.  ``nlist`` may be supplied
    (already type-partitioned, extended indices) to skip the neighbour search."""
    nf, nloc = atype.shape
    assert nf == 1
    if nlist is None:
        ext, ext_atype, mapping, nl = geometry.extend_input_and_build_neighbor_list(
            coord.view(nf, nloc, 3), atype, model.rcut, model.sel, box=box, mixed_types=True)
        nlist = geometry.format_nlist(ext, ext_atype, nl, model.rcut, model.sel)
    else:
        nf_, nloc_ = atype.shape
        cn = geometry.normalize_coord(coord.view(nf, nloc, 3), box.view(nf, 3, 3))
        ext, ext_atype, mapping = geometry.extend_coord_with_ghosts(cn, atype, box, model.rcut)
        ext = ext.view(nf, -1, 3)
    cc = ext.detach().clone().requires_grad_(True)
    grad = torch.zeros_like(cc)
    energy = torch.zeros((nf, 1), dtype=torch.float64)
    ae_all = torch.zeros((nf, nloc, 1), dtype=torch.float64)
    for a0 in range(0, nloc, chunk):
        a1 = min(nloc, a0 + chunk)
        centers = torch.arange(a0, a1)
        ap = None if aparam is None else torch.as_tensor(aparam).view(nf, nloc, -1)[:, a0:a1]
        ae = atomic_energy(model, cc, ext_atype, nlist[:, a0:a1], centers=centers, fparam=fparam,
                           aparam=ap)
        e = torch.sum(ae, dim=1)
        g = torch.autograd.grad([e], [cc], grad_outputs=[torch.ones_like(e)])[0]
        grad += g
        energy += e.detach()
        ae_all[:, a0:a1] = ae.detach()
    force_ext = -grad
    vir_ext = torch.einsum("...ik,...ij->...ikj", force_ext, ext.detach()).reshape(nf, -1, 9)
    return {"energy": energy, "atom_energy": ae_all,
            "force": communicate(force_ext, mapping, nloc),
            "virial": torch.sum(vir_ext, dim=1), "nlist": nlist}

"""se_e2_a energy models from PyTorch training checkpoints (``.pt``).

The reference loads a ``.pt`` file as a state dict (``torch.load(...,
weights_only=True)``, the ``"model"`` entry when present), reads the model
parameters from ``state_dict["_extra_state"]["model_params"]``, picks a head
of a multi-task model and loads the weights into the model it builds
(deepmd/pt/infer/deep_eval.py:160-225).  Here the weights are read into the
serialized layout SeAModel takes (the dpmodel ``serialize()`` dict,
``source/tests/infer/deeppot_sea.yaml``), so nothing of the checkpoint is
executed: tensors and plain Python values only.

State-dict keys follow the module tree of a standard energy model:
ModelWrapper.model[head] (deepmd/pt/train/wrapper.py:43-53) -> atomic_model
(make_model.py:80-83) -> descriptor / fitting_net and the buffers out_bias,
out_std (dp_atomic_model.py:62-65, base_atomic_model.py:144-145);
DescrptSeA.sea (descriptor/se_a.py:123) with buffers mean, stddev and the
NetworkCollection filter_layers (se_a.py:537-556); EnergyFittingNet with
buffer bias_atom_e and filter_layers (task/fitting.py:473,528); a
NetworkCollection registers its ModuleList as both ``networks`` and
``_networks`` (network/mlp.py:381), each MLPLayer holds ``matrix``, ``bias``
and ``idt`` (mlp.py:99-110).  Missing model parameters take the reference's
argcheck defaults (deepmd/utils/argcheck.py:403-445, fitting_ener).
"""

from __future__ import annotations

import re

import numpy as np

from .model import _layer_dict, _nd

_NET = re.compile(r"^(?P<coll>.*\.filter_layers)\._?networks\.(?P<k>\d+)\.layers\.(?P<l>\d+)\."
                  r"(?P<v>matrix|bias|idt)$")


def _np(t) -> np.ndarray:
    return t.detach().to("cpu").double().numpy() if hasattr(t, "detach") else np.asarray(t)


def _head_prefix(state: dict, params: dict, head) -> tuple[str, dict]:
    """Key prefix of the chosen head and its model parameters
    (deep_eval.py:170-200: a multi-task model needs a head; an int picks the
    first; names match case-insensitively)."""
    if "model_dict" not in params:
        return "model.Default.", params
    heads = list(params["model_dict"])
    if head is None:
        if "Default" in heads:
            head = "Default"
        else:
            raise ValueError(f"multi-task checkpoint: choose a head among {heads}")
    if isinstance(head, int):
        head = heads[0]
    if head not in heads:
        low = {h.lower(): h for h in heads}
        if str(head).lower() not in low:
            raise ValueError(f"no head {head!r} in the checkpoint (heads: {heads})")
        head = low[str(head).lower()]
    return f"model.{head}.", params["model_dict"][head]


def _networks(state: dict, coll: str) -> list[list[dict]]:
    """networks[k][l] = {"matrix": .., "bias": .., "idt": ..} of one collection."""
    nets: dict = {}
    for key, val in state.items():
        m = _NET.match(key)
        if m is None or m.group("coll") != coll:
            continue
        k, l = int(m.group("k")), int(m.group("l"))
        nets.setdefault(k, {}).setdefault(l, {})[m.group("v")] = _np(val)
    if not nets:
        raise KeyError(f"no networks under {coll}")
    out = []
    for k in range(max(nets) + 1):
        layers = nets.get(k)
        if layers is None:
            raise KeyError(f"{coll}: network {k} missing")
        out.append([layers[l] for l in range(max(layers) + 1)])
    return out


def model_dict_from_state_dict(state: dict, head=None) -> dict:
    """Serialized-layout model dict (the input of SeAModel / DeepPot) of the
    se_e2_a energy model held in a PyTorch checkpoint's state dict."""
    if "model" in state and isinstance(state["model"], dict):
        state = state["model"]
    if "_extra_state" not in state or "model_params" not in state["_extra_state"]:
        raise ValueError("not a deepmd-kit PyTorch checkpoint (no _extra_state.model_params)")
    prefix, mp = _head_prefix(state, state["_extra_state"]["model_params"], head)
    de = dict(mp.get("descriptor", {}))
    fi = dict(mp.get("fitting_net", {}))
    if de.get("type", "se_e2_a") not in ("se_e2_a", "se_a"):
        raise NotImplementedError(f"descriptor {de.get('type')!r} (se_e2_a only)")
    if fi.get("type", "ener") != "ener":
        raise NotImplementedError(f"fitting {fi.get('type')!r} (energy only)")
    sel = de.get("sel")
    if not isinstance(sel, (list, tuple)):
        raise ValueError(f"descriptor sel {sel!r}: a trained checkpoint stores the list")
    sel = [int(x) for x in sel]
    ntypes, nnei = len(sel), int(sum(sel))
    type_map = list(mp.get("type_map", [str(i) for i in range(ntypes)]))
    am = prefix + "atomic_model."
    sea = am + "descriptor.sea."
    fit = am + "fitting_net."
    act = de.get("activation_function", "tanh")
    fact = fi.get("activation_function", "tanh")

    emb = []
    for layers in _networks(state, sea + "filter_layers"):
        emb.append({"layers": [_layer_dict(ly["matrix"], ly.get("bias"), ly.get("idt"), act, True)
                               for ly in layers]})
    fits = []
    for layers in _networks(state, fit + "filter_layers"):
        n = len(layers)
        fits.append({"layers": [_layer_dict(ly["matrix"], ly.get("bias"), ly.get("idt"),
                                            fact if i < n - 1 else "none", i < n - 1)
                                for i, ly in enumerate(layers)]})
    davg = _np(state[sea + "mean"]).reshape(ntypes, nnei, 4)
    dstd = _np(state[sea + "stddev"]).reshape(ntypes, nnei, 4)
    bias_atom_e = _np(state[fit + "bias_atom_e"]).reshape(ntypes, -1)
    out_bias = _np(state[am + "out_bias"]) if am + "out_bias" in state else np.zeros((1, ntypes, 1))
    out_std = _np(state[am + "out_std"]) if am + "out_std" in state else np.ones((1, ntypes, 1))
    type_one_side = bool(de.get("type_one_side", False))
    if len(emb) != (ntypes if type_one_side else ntypes * ntypes):
        raise ValueError(f"{len(emb)} embedding nets for {ntypes} types "
                         f"(type_one_side={type_one_side})")
    neuron = [int(x) for x in de.get("neuron", [10, 20, 40])]
    axis = int(de.get("axis_neuron", de.get("n_axis_neuron", 4)))
    fneuron = [int(x) for x in fi.get("neuron", fi.get("n_neuron", [120, 120, 120]))]
    # frame / atomic parameters: buffers of the fitting (fitting.py:473-519)
    nfp, nap = int(fi.get("numb_fparam", 0) or 0), int(fi.get("numb_aparam", 0) or 0)
    fvars = {"bias_atom_e": _nd(bias_atom_e)}
    for name, n in (("fparam", nfp), ("aparam", nap)):
        if n:
            for suf in ("avg", "inv_std"):
                key = f"{fit}{name}_{suf}"
                fvars[f"{name}_{suf}"] = _nd(_np(state[key]).reshape(n) if key in state
                                             else (np.zeros(n) if suf == "avg" else np.ones(n)))
    dfp = fi.get("default_fparam")
    if dfp is None and fit + "default_fparam_tensor" in state:
        dfp = _np(state[fit + "default_fparam_tensor"]).reshape(-1).tolist()
    return {"model": {
        "@class": "Model", "type": "standard", "type_map": type_map,
        "atom_exclude_types": [int(t) for t in mp.get("atom_exclude_types", [])],
        "pair_exclude_types": [[int(a), int(b)] for a, b in mp.get("pair_exclude_types", [])],
        "@variables": {"out_bias": _nd(out_bias), "out_std": _nd(out_std)},
        "descriptor": {
            "@class": "Descriptor", "type": "se_e2_a", "rcut": float(de.get("rcut", 6.0)),
            "rcut_smth": float(de.get("rcut_smth", 0.5)), "sel": sel, "neuron": neuron,
            "axis_neuron": axis, "type_one_side": type_one_side,
            "resnet_dt": bool(de.get("resnet_dt", False)), "activation_function": act,
            "env_protection": float(de.get("env_protection", 0.0)),
            "exclude_types": [list(p) for p in de.get("exclude_types", [])],
            "type_map": type_map,
            "@variables": {"davg": _nd(davg), "dstd": _nd(dstd)},
            "embeddings": {"@class": "NetworkCollection", "ndim": 1 if type_one_side else 2,
                           "ntypes": ntypes, "networks": emb},
        },
        "fitting": {
            "@class": "Fitting", "type": "ener", "neuron": fneuron,
            "resnet_dt": bool(fi.get("resnet_dt", True)), "dim_descrpt": neuron[-1] * axis,
            "mixed_types": False, "numb_fparam": nfp, "numb_aparam": nap,
            "default_fparam": None if dfp is None else [float(x) for x in dfp],
            "use_aparam_as_mask": bool(fi.get("use_aparam_as_mask", False)),
            "@variables": fvars,
            "nets": {"@class": "NetworkCollection", "ndim": 1, "ntypes": ntypes,
                     "networks": fits},
        },
    }}


def load_checkpoint(path: str, head=None) -> dict:
    """Model dict of a ``.pt`` training checkpoint (tensors only:
    ``torch.load(weights_only=True)``, as deep_eval.py:161-163)."""
    import torch
    state = torch.load(path, map_location="cpu", weights_only=True)
    return model_dict_from_state_dict(state, head)


def model_params_from_model_dict(model_dict: dict) -> dict:
    """``model_params`` (the input.json ``model`` section) describing a
    serialized se_e2_a energy model, as a checkpoint's ``_extra_state``
    holds it (deepmd/pt/train/wrapper.py:37-41,256-264)."""
    m = model_dict["model"] if "model" in model_dict else model_dict
    de, fi = m["descriptor"], m["fitting"]
    return {
        "type_map": list(m["type_map"]),
        "descriptor": {"type": "se_e2_a", "sel": [int(x) for x in de["sel"]],
                       "rcut": float(de["rcut"]), "rcut_smth": float(de["rcut_smth"]),
                       "neuron": [int(x) for x in de["neuron"]],
                       "axis_neuron": int(de["axis_neuron"]),
                       "type_one_side": bool(de["type_one_side"]),
                       "resnet_dt": bool(de.get("resnet_dt", False)),
                       "activation_function": de.get("activation_function", "tanh"),
                       "exclude_types": [list(p) for p in de.get("exclude_types", [])],
                       "env_protection": float(de.get("env_protection", 0.0)),
                       "precision": "float64"},
        "fitting_net": {"type": "ener", "neuron": [int(x) for x in fi["neuron"]],
                        "resnet_dt": bool(fi.get("resnet_dt", True)),
                        "activation_function": "tanh",
                        "numb_fparam": int(fi.get("numb_fparam", 0) or 0),
                        "numb_aparam": int(fi.get("numb_aparam", 0) or 0),
                        "default_fparam": fi.get("default_fparam"),
                        "use_aparam_as_mask": bool(fi.get("use_aparam_as_mask", False)),
                        "precision": "float64"},
    }


def state_dict_from_model_dict(model_dict: dict, head: str = "Default") -> dict:
    """State dict of the module tree a standard se_e2_a energy model has in
    the reference's trainer (the key names listed at the top of this
    module), FP64 tensors; every network is stored under both ``networks``
    and ``_networks`` as NetworkCollection registers it (mlp.py:381)."""
    import torch

    def t(v):
        return torch.tensor(np.asarray(v["value"] if isinstance(v, dict) else v,
                                       dtype=np.float64))

    m = model_dict["model"] if "model" in model_dict else model_dict
    de, fi = m["descriptor"], m["fitting"]
    am = f"model.{head}.atomic_model."
    sd = {am + "descriptor.sea.mean": t(de["@variables"]["davg"]),
          am + "descriptor.sea.stddev": t(de["@variables"]["dstd"]),
          am + "fitting_net.bias_atom_e": t(fi["@variables"]["bias_atom_e"]),
          am + "out_bias": t(m["@variables"]["out_bias"]),
          am + "out_std": t(m["@variables"]["out_std"])}
    for name in ("fparam_avg", "fparam_inv_std", "aparam_avg", "aparam_inv_std"):
        v = fi["@variables"].get(name)
        if v is not None and not (isinstance(v, dict) and v.get("value") is None):
            sd[am + "fitting_net." + name] = t(v)
    if fi.get("default_fparam") is not None:
        sd[am + "fitting_net.default_fparam_tensor"] = t(fi["default_fparam"])
    for coll, nets in ((am + "descriptor.sea.filter_layers", de["embeddings"]["networks"]),
                       (am + "fitting_net.filter_layers", fi["nets"]["networks"])):
        for k, net in enumerate(nets):
            for l, ly in enumerate(net["layers"]):
                var = ly["@variables"]
                for name, key in (("matrix", "w"), ("bias", "b"), ("idt", "idt")):
                    v = var.get(key)
                    if v is None or (isinstance(v, dict) and v.get("value") is None):
                        continue
                    for reg in ("networks", "_networks"):
                        sd[f"{coll}.{reg}.{k}.layers.{l}.{name}"] = t(v)
    return sd


def save_checkpoint(path: str, model_dict: dict, lr: float = 0.0, step: int = 0,
                    optimizer_state: dict | None = None) -> None:
    """Write ``{"model": state dict, "optimizer": ...}`` with
    ``_extra_state = {"model_params", "train_infos": {"lr", "step"}}``, the
    layout of the reference trainer's ``save_model`` (training.py:2027-2051,
    wrapper.py:256-264); ``load_checkpoint`` reads it back."""
    import torch
    sd = state_dict_from_model_dict(model_dict)
    sd["_extra_state"] = {"model_params": model_params_from_model_dict(model_dict),
                          "train_infos": {"lr": float(lr), "step": int(step)}}
    out = {"model": sd}
    if optimizer_state is not None:
        out["optimizer"] = optimizer_state
    torch.save(out, path)

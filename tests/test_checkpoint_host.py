"""PyTorch training checkpoints (.pt) on the host.

A checkpoint of the reference's deeppot_sea.yaml model is written the way the
reference's trainer stores one -- {"model": ModelWrapper state dict,
"optimizer": ...}, state-dict keys of the module tree (wrapper.py:43-53,
make_model.py:80-83, dp_atomic_model.py:62-65, base_atomic_model.py:144-145,
descriptor/se_a.py:123,537-556, task/fitting.py:473,528, network/mlp.py:
99-110,381: every network under both ``networks`` and ``_networks``) and
``_extra_state.model_params`` -- then loaded back with weights_only=True and
evaluated by the oracle against the reference's deeppot-testcase.yaml
goldens.  No se_e2_a .pt ships with the reference, so the key names are
pinned by this reading of its module tree, not by a checkpoint it wrote.
"""

import numpy as np
import pytest
import torch

from deepmd_amd.checkpoint import load_checkpoint, model_dict_from_state_dict
from deepmd_amd.deep_pot import load_model_dict
from oracle import sea


def _t(v):
    return torch.tensor(np.asarray(v["value"] if isinstance(v, dict) else v, dtype=np.float64))


def _model_params(m):
    de, fi = m["descriptor"], m["fitting"]
    return {
        "type_map": list(m["type_map"]),
        "descriptor": {"type": "se_e2_a", "sel": list(de["sel"]), "rcut": de["rcut"],
                       "rcut_smth": de["rcut_smth"], "neuron": list(de["neuron"]),
                       "axis_neuron": de["axis_neuron"], "type_one_side": de["type_one_side"],
                       "resnet_dt": de["resnet_dt"], "activation_function": "tanh",
                       "exclude_types": [], "precision": "float64", "seed": 1},
        "fitting_net": {"type": "ener", "neuron": list(fi["neuron"]),
                        "resnet_dt": fi["resnet_dt"], "activation_function": "tanh",
                        "numb_fparam": 0, "numb_aparam": 0, "precision": "float64"},
    }


def _state_dict(m, head="Default"):
    de, fi = m["descriptor"], m["fitting"]
    am = f"model.{head}.atomic_model."
    sd = {am + "descriptor.sea.mean": _t(de["@variables"]["davg"]),
          am + "descriptor.sea.stddev": _t(de["@variables"]["dstd"]),
          am + "fitting_net.bias_atom_e": _t(fi["@variables"]["bias_atom_e"]),
          am + "out_bias": _t(m["@variables"]["out_bias"]),
          am + "out_std": _t(m["@variables"]["out_std"]),
          # buffers the reader ignores
          am + "fitting_net.case_embd": torch.zeros(0),
          am + "descriptor.sea.compress_info.0": torch.zeros(6),
          am + "descriptor.sea.compress_data.0": torch.zeros(0)}
    for coll, nets in ((am + "descriptor.sea.filter_layers", de["embeddings"]["networks"]),
                       (am + "fitting_net.filter_layers", fi["nets"]["networks"])):
        for k, net in enumerate(nets):
            for l, ly in enumerate(net["layers"]):
                v = ly["@variables"]
                for name, src in (("matrix", "w"), ("bias", "b"), ("idt", "idt")):
                    if v.get(src) is not None:
                        for reg in ("networks", "_networks"):
                            sd[f"{coll}.{reg}.{k}.layers.{l}.{name}"] = _t(v[src])
    return sd


def _block_inputs(r):
    c = torch.tensor(r["coord"], dtype=torch.float64).view(1, -1)
    at = torch.tensor(r["atype"]).view(1, -1)
    box = None if r["box"] is None else torch.tensor(r["box"], dtype=torch.float64).view(1, 9)
    return c, at, box


def _rel(a, b):
    b = np.asarray(b, dtype=np.float64)
    return np.max(np.abs(np.asarray(a).ravel() - b.ravel())) / max(np.max(np.abs(b)), 1e-300)


def test_checkpoint_matches_testcase_goldens(sea_yaml, sea_testcase, tmp_path):
    m = sea_yaml["model"] if "model" in sea_yaml else sea_yaml
    sd = _state_dict(m)
    sd["_extra_state"] = {"model_params": _model_params(m), "train_infos": {"lr": 1e-3, "step": 7}}
    path = str(tmp_path / "model.pt")
    torch.save({"model": sd, "optimizer": {"state": {}, "param_groups": [{"lr": 1e-3}]}}, path)
    d = load_model_dict(path)
    om = sea.model_from_serialized(d)
    for r in sea_testcase["results"]:
        c, at, box = _block_inputs(r)
        out = sea.forward(om, c, at, box, do_atomic_virial=True)
        assert _rel(out["atom_energy"].detach().numpy(), r["atomic_energy"]) < 1e-10
        assert _rel(out["force"].detach().numpy(), r["force"]) < 1e-10
        assert _rel(out["atom_virial"].detach().numpy(), r["atomic_virial"]) < 1e-10


def test_checkpoint_multitask_heads(sea_yaml, tmp_path):
    m = sea_yaml["model"] if "model" in sea_yaml else sea_yaml
    mp = _model_params(m)
    sd = _state_dict(m, "water_1")
    sd.update(_state_dict(m, "water_2"))
    sd["model.water_2.atomic_model.out_bias"] = sd["model.water_2.atomic_model.out_bias"] + 0.5
    sd["_extra_state"] = {"model_params": {"model_dict": {"water_1": mp, "water_2": mp}},
                          "train_infos": {}}
    path = str(tmp_path / "multi.pt")
    torch.save(sd, path)
    with pytest.raises(ValueError, match="choose a head"):
        load_checkpoint(path)
    d1, d2 = load_checkpoint(path, head="water_1"), load_checkpoint(path, head="WATER_2")
    b1 = np.asarray(d1["model"]["@variables"]["out_bias"]["value"])
    b2 = np.asarray(d2["model"]["@variables"]["out_bias"]["value"])
    np.testing.assert_allclose(b2 - b1, 0.5)
    assert load_checkpoint(path, head=0)["model"]["type_map"] == list(m["type_map"])


def test_checkpoint_rejects_other_models(sea_yaml):
    m = sea_yaml["model"] if "model" in sea_yaml else sea_yaml
    with pytest.raises(ValueError, match="_extra_state"):
        model_dict_from_state_dict(_state_dict(m))
    sd = _state_dict(m)
    mp = _model_params(m)
    mp["descriptor"]["type"] = "dpa1"
    sd["_extra_state"] = {"model_params": mp}
    with pytest.raises(NotImplementedError, match="se_e2_a only"):
        model_dict_from_state_dict(sd)


def test_checkpoint_keeps_frame_atomic_parameters(tmp_path):
    # fitting buffers fparam_avg / _inv_std, aparam_avg / _inv_std and
    # default_fparam_tensor (deepmd/pt/model/task/fitting.py:473-519) survive a
    # state-dict round trip of the reference's fparam/aparam se_e2_a model
    import yaml
    from conftest import golden_path
    from deepmd_amd import checkpoint
    with open(golden_path("fparam_aparam_default.yaml")) as fh:
        d = yaml.safe_load(fh)
    path = str(tmp_path / "fp.pt")
    checkpoint.save_checkpoint(path, d)
    back = checkpoint.load_checkpoint(path)
    fa, fb = d["model"]["fitting"], back["model"]["fitting"]
    assert fb["numb_fparam"] == 1 and fb["numb_aparam"] == 1
    assert fb["default_fparam"] == fa["default_fparam"]
    for k in ("fparam_avg", "fparam_inv_std", "aparam_avg", "aparam_inv_std"):
        np.testing.assert_array_equal(np.asarray(fb["@variables"][k]["value"]).ravel(),
                                      np.asarray(fa["@variables"][k]["value"]).ravel())

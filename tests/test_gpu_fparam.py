"""Frame / atomic parameters of the energy fitting on the GPU
(deepmd/pt/model/task/fitting.py:749-830: normalised fparam / aparam appended
to the descriptor before the first fitting layer).

Pins: the reference's se_e2_a fparam/aparam case
(source/tests/infer/fparam_aparam-testcase.yaml on fparam_aparam_default.yaml,
values from its TensorFlow model, compared at 7 decimal places as the
reference's test_models.py:23-31 does) and the oracle at 1e-10 relative on
per-atom parameters, several frames with different parameters and the
forward_lower entry.
"""

import numpy as np
import pytest
import torch
import yaml

from conftest import golden_path

pytestmark = pytest.mark.gpu

ATOL_CASE = 5e-8   # assertAlmostEqual(places=7)


@pytest.fixture(scope="module", autouse=True)
def _hip():
    if not torch.cuda.is_available():
        pytest.fail("a HIP device is required for the gpu tests")


def _model_dict(default=True):
    with open(golden_path("fparam_aparam_default.yaml")) as fh:
        d = yaml.safe_load(fh)
    if not default:
        d["model"]["fitting"]["default_fparam"] = None
    return d


def _case():
    with open(golden_path("fparam_aparam-testcase.yaml")) as fh:
        return yaml.safe_load(fh)["results"][0]


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _oracle(d, x, at, box, fp, ap):
    from oracle import sea
    om = sea.model_from_serialized(d)
    nf = x.shape[0]
    out = sea.forward(om, torch.tensor(x).view(nf, -1), torch.tensor(at).view(nf, -1).long(),
                      None if box is None else torch.tensor(box).view(nf, 9),
                      do_atomic_virial=True,
                      fparam=None if fp is None else torch.tensor(np.asarray(fp, dtype=np.float64)).view(nf, -1),
                      aparam=None if ap is None else torch.tensor(np.asarray(ap, dtype=np.float64)).view(nf, at.shape[-1], -1))
    return {k: v.detach().numpy() for k, v in out.items() if torch.is_tensor(v)}


@pytest.mark.parametrize("mode", ["explicit", "default"])
def test_fparam_aparam_golden_case(mode):
    from deepmd_amd.deep_pot import DeepPot
    r = _case()
    dp = DeepPot(_model_dict(default=True), precision="f64")
    assert dp.get_dim_fparam() == 1 and dp.get_dim_aparam() == 1 and dp.has_default_fparam()
    x = np.array(r["coord"], dtype=np.float64).reshape(1, -1, 3)
    fp = [r["fparam"]] if mode == "explicit" else None
    e, f, v, ae, av = dp.eval(x, np.array(r["box"]).reshape(1, 9), np.array(r["atype"]),
                              atomic=True, fparam=fp, aparam=np.array(r["aparam"]))
    for got, ref in ((ae, r["atomic_energy"]), (f, r["force"]), (av, r["atomic_virial"])):
        assert np.max(np.abs(np.ravel(got) - np.ravel(ref))) < ATOL_CASE
    o = _oracle(_model_dict(), x, np.array(r["atype"]).reshape(1, -1),
                np.array(r["box"]).reshape(1, 9), [r["fparam"]], np.array(r["aparam"]))
    assert _rel(ae, o["atom_energy"]) < 1e-10
    assert _rel(f, o["force"]) < 1e-10
    assert _rel(av, o["atom_virial"]) < 1e-10
    assert _rel(v, o["virial"]) < 1e-10


def test_fparam_required_without_default():
    from deepmd_amd.deep_pot import DeepPot
    r = _case()
    dp = DeepPot(_model_dict(default=False), precision="f64")
    x = np.array(r["coord"], dtype=np.float64).reshape(1, -1, 3)
    with pytest.raises(ValueError, match="fparam is required"):
        dp.eval(x, np.array(r["box"]).reshape(1, 9), np.array(r["atype"]),
                aparam=np.array(r["aparam"]))
    with pytest.raises(RuntimeError, match="wrong size of frame param"):
        dp.eval(x, np.array(r["box"]).reshape(1, 9), np.array(r["atype"]), fparam=[1.0, 2.0],
                aparam=np.array(r["aparam"]))


@pytest.mark.parametrize("precision,tol", [("f64", 1e-10), ("f32", 1e-4), ("bf16", 3e-2)])
def test_frames_with_different_parameters(precision, tol):
    # two frames, different fparam each, per-atom random aparam: every value
    # reaches the first fitting layer of its own atom
    from deepmd_amd.deep_pot import DeepPot
    r = _case()
    rng = np.random.default_rng(5)
    x0 = np.array(r["coord"], dtype=np.float64).reshape(1, -1, 3)
    x = np.concatenate([x0, x0 + rng.normal(0, 0.05, size=x0.shape)])
    box = np.tile(np.array(r["box"]).reshape(1, 9), (2, 1))
    at = np.array(r["atype"])
    fp = np.array([[0.1], [0.7]])
    ap = rng.uniform(0.0, 1.0, size=(2, at.size, 1))
    d = _model_dict(default=False)
    dp = DeepPot(d, precision=precision)
    e, f, v, ae, av = dp.eval(x, box, at, atomic=True, fparam=fp, aparam=ap)
    o = _oracle(d, x, np.tile(at, (2, 1)), box, fp, ap)
    assert _rel(ae, o["atom_energy"]) < tol
    assert _rel(f, o["force"]) < tol
    assert _rel(v, o["virial"]) < tol
    # the same evaluation with the parameters swapped between frames differs
    e2 = dp.eval(x, box, at, fparam=fp[::-1].copy(), aparam=ap)[0]
    assert np.max(np.abs(e2 - e)) > 1e-6


def test_forward_lower_with_parameters():
    from oracle import geometry, sea
    from deepmd_amd.model import SeAModel
    r = _case()
    d = _model_dict(default=False)
    om = sea.model_from_serialized(d)
    x = torch.tensor(r["coord"], dtype=torch.float64).view(1, -1, 3)
    at = torch.tensor(r["atype"]).view(1, -1).long()
    box = torch.tensor(r["box"], dtype=torch.float64).view(1, 9)
    ext, eat, mp, nl = geometry.extend_input_and_build_neighbor_list(x, at, om.rcut, om.sel, box)
    rng = np.random.default_rng(7)
    ap = torch.tensor(rng.uniform(0, 1, size=(1, at.shape[1], 1)))
    fp = torch.tensor([[0.4]], dtype=torch.float64)
    ref = sea.forward_lower(om, ext, eat, nl, fparam=fp, aparam=ap, do_atomic_virial=True)
    m = SeAModel(d, "f64", "cuda")
    out = m.forward_lower(ext, eat, nl, do_atomic_virial=True, fparam=fp, aparam=ap)
    assert _rel(out["atom_energy"].cpu(), ref["atom_energy"].detach()) < 1e-10
    assert _rel(out["extended_force"].cpu(), ref["extended_force"].detach()) < 1e-10
    assert _rel(out["virial"].cpu(), ref["virial"].detach()) < 1e-10

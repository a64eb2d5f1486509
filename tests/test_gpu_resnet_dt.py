"""Embedding nets with time steps (resnet_dt, deepmd/pt/model/network/mlp.py:
188-219: act(x W + b) * idt + resnet term) on the HIP path, against the oracle
(its forces from torch autograd): one type (one net: the FP64 kernels hold
the weights in LDS) and two types with type_one_side False (four nets, read
through the caches), in FP64, FP32 and split FP16, with tanh and (FP64 /
FP32, whose kernels evaluate other activations) gelu.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda"


@pytest.fixture(scope="module", autouse=True)
def _hip():
    if not torch.cuda.is_available():
        pytest.fail("a HIP device is required for the gpu tests")


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("act", ["tanh", "gelu"])
@pytest.mark.parametrize("ntypes", [1, 2])
def test_resnet_dt_embedding_vs_oracle(ntypes, act):
    from deepmd_amd import systems
    from deepmd_amd.model import SeAModel, random_model_dict
    from oracle import sea
    sel = [46, 92] if ntypes == 2 else [138]
    d = random_model_dict(3, 6.0, 0.5, sel, [25, 50, 100], 16, [32, 32],
                          type_one_side=ntypes == 1, resnet_dt_emb=True, activation=act)
    x, at, box = systems.water(3, 9.33, 5)
    if ntypes == 1:
        at = np.zeros_like(at)
    om = sea.model_from_serialized(d)
    ref = sea.forward(om, torch.tensor(x).view(1, -1), torch.tensor(at).view(1, -1).long(),
                      torch.tensor(box.ravel()).view(1, 9))
    print(f"ntypes {ntypes} {act}: oracle energy {ref['energy'].item():.6f}", flush=True)
    precs = [("f64", 1e-10), ("f32", 2e-4)] + ([("f16x2", 2e-3)] if act == "tanh" else [])
    for prec, tol in precs:
        m = SeAModel(d, prec, DEV)
        out = m.evaluate(m.system(x[None], at[None], box.reshape(1, 9)))
        torch.cuda.synchronize()
        ee = _rel(out["energy"].cpu(), ref["energy"].detach())
        fe = _rel(out["force"].cpu(), ref["force"].detach().view(-1, 3))
        ve = _rel(out["virial"].cpu(), ref["virial"].detach())
        print(f"  {prec}: E {ee:.1e} F {fe:.1e} V {ve:.1e}", flush=True)
        assert ee < tol and fe < tol and ve < tol, (prec, ee, fe, ve)

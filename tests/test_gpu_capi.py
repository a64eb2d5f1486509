"""The C++ MD-engine runner (libdpk_deeppot.so, include/dpk_deeppot.h) --
DeepPotPT::compute with an InputNlist (source/api_cc/src/DeepPotPT.cc:171-378)
driven through libdpk's C ABI only -- against SeAModel.forward_lower on the
same extended region and list (split FP16): the same kernels, so E / F / V
agree to the rounding of the force kernel's FP64 atomics.  Then ago > 0
steps on moved coordinates (the list kept), and the runner's errors."""

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


def _csr(rows, perm):
    """InputNlist form of dense rows (-1 padded): row ii is local atom perm[ii]."""
    ilist = np.asarray(perm, dtype=np.int32)
    nbrs = [rows[i][rows[i] >= 0] for i in perm]
    return ilist, np.array([len(r) for r in nbrs], dtype=np.int32), nbrs


def _case(kind):
    from deepmd_amd.model import random_model_dict
    from oracle import geometry
    from deepmd_amd import systems
    if kind == "water":
        d = random_model_dict(5, 6.0, 0.5, [46, 92], [25, 50, 100], 16, [240, 240, 240])
        x, at, box = systems.water(3, 12.44 * 3 / 4, 4)
        rc, sel = 6.5, [46, 92]
    else:
        d = random_model_dict(6, 8.0, 2.0, [200], [32, 64, 128], 16, [240, 240, 240])
        x, at, box = systems.fcc(4, 3.615, 0.05, 7)
        rc, sel = 8.5, [200]
    ec, ea, mp, nl = geometry.extend_input_and_build_neighbor_list(
        torch.tensor(x).view(1, -1), torch.tensor(at).view(1, -1).long(), rc, sel,
        box=torch.tensor(box.ravel()).view(1, 9))
    return d, ec.view(-1, 3).numpy(), ea.view(-1).numpy().astype(np.int32), nl[0].numpy(), len(at)


@pytest.mark.parametrize("kind", ["water", "cu"])
def test_cpp_runner_matches_forward_lower(kind, tmp_path):
    from deepmd_amd.export import CppDeepPot, export_lower_model
    from deepmd_amd.model import SeAModel
    d, ec, ea, rows, nloc = _case(kind)
    nall = ea.size
    m = SeAModel(d, "f16x2", DEV)
    path = str(tmp_path / "model.dpkm")
    export_lower_model(m, path)
    dp = CppDeepPot(path)
    rng = np.random.default_rng(3)
    ilist, numneigh, nbrs = _csr(rows, rng.permutation(nloc))
    keys = ("energy", "extended_force", "virial")

    def ref(x):
        out = m.forward_lower(torch.from_numpy(x)[None], torch.from_numpy(ea)[None],
                              torch.from_numpy(rows)[None])
        return [out[k].double().cpu().numpy() for k in keys]

    e, f, v = dp.compute(ec, ea, nall - nloc, ilist, numneigh, nbrs, ago=0)
    re, rf, rv = ref(ec)
    print(f"{kind}: nloc {nloc}, nall {nall}, E {e[0]:.6f} vs {re[0]:.6f}", flush=True)
    assert _rel(e, re) <= 1e-12 and _rel(f, rf) <= 1e-12 and _rel(v, rv) <= 1e-12
    assert np.abs(f[nloc:]).max() > 0  # ghost atoms carry forces for the engine
    # MD steps between re-neighbourings: the list and types are kept
    for k in range(3):
        x = ec + 0.02 * rng.standard_normal(ec.shape)
        e, f, v = dp.compute(x, ea, nall - nloc, ago=k + 1)
        re, rf, rv = ref(x)
        assert _rel(e, re) <= 1e-12 and _rel(f, rf) <= 1e-12 and _rel(v, rv) <= 1e-12, k
    # errors: a neighbour outside the region, a type outside the model
    bad = [r.copy() for r in nbrs]
    bad[0][0] = nall + 5
    with pytest.raises(RuntimeError, match="outside"):
        dp.compute(ec, ea, nall - nloc, ilist, numneigh, bad, ago=0)
    ea2 = ea.copy()
    ea2[0] = 9
    with pytest.raises(RuntimeError, match="type"):
        dp.compute(ec, ea2, nall - nloc, ilist, numneigh, nbrs, ago=0)
    with pytest.raises(RuntimeError, match="ago > 0"):
        dp.compute(ec[:-1], ea[:-1], nall - nloc - 1, ago=1)


def test_cpp_runner_rejects_unexported_models(tmp_path):
    from deepmd_amd.export import CppDeepPot, export_lower_model
    from deepmd_amd.model import SeAModel, random_model_dict
    d = random_model_dict(5, 6.0, 0.5, [46, 92], [25, 50, 100], 16, [240, 240, 240])
    with pytest.raises(ValueError, match="split-FP16"):
        export_lower_model(SeAModel(d, "f64", DEV), str(tmp_path / "m"))
    p = tmp_path / "junk"
    p.write_bytes(b"not a model\n")
    with pytest.raises(RuntimeError, match="not an exported model"):
        CppDeepPot(str(p))

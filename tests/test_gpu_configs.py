"""Each benchmark configuration's own model and system on the HIP path,
against the CPU oracle (bench.configs(): C1-C4).

* C4 -- the Cu model (rcut 8, rcut_smth 2, sel [200], embedding [32,64,128],
  axis 16, fitting [240,240,240]) at a physical force scale (input
  statistics, forces scaled to 1 eV/A RMS: bench.physical_model) on an
  oracle-sized 8x8x8-cell box (2048 atoms): neighbour list bit-exact, FP64
  E/F/V <= 1e-10 relative, split FP16 (f16x2, the benchmark's precision)
  force RMSE <= 1e-3 eV/A; BF16 and plain FP16 with their measured bounds
  (both above the gate).  Then the full 131,072-atom frame: capacity status,
  finite outputs and f16x2 vs the FP64 GPU path at the same gate.
* C3 -- the ragged 4-species model (type_one_side False) on 16 of the 256
  frames (every 16th by size, 24..768 atoms) packed in one system: per frame
  neighbour list bit-exact, FP64 <= 1e-10, FP32 (its precision) <= 1e-4 of
  the force RMS.
* C2 -- the training step at 8 x 192-atom water with fitting [240]^3: loss
  and weight gradients vs the oracle's autograd (FP64 <= 1e-8; BF16, its
  precision, within a stated relative bound).
* C1 -- the 12,288-atom water model compressed to FP32 quintic tables vs the
  uncompressed FP64 oracle: force RMSE <= 1e-4 of the RMS (the FP32 bar of
  source/tests/pt/test_model_compression_se_a.py:22-24 is 4 decimal places).
"""

import numpy as np
import pytest
import torch

import bench

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


def _oracle(d, x, at, box, chunk=1024):
    from oracle import sea
    om = sea.model_from_serialized(d)
    out = sea.forward_chunked(om, torch.tensor(x).view(1, -1),
                              torch.tensor(at).view(1, -1).long(),
                              torch.tensor(box.ravel()).view(1, 9), chunk=chunk)
    return {k: v.detach().numpy() for k, v in out.items()}


def _eval(d, prec, x, at, box):
    from deepmd_amd.model import SeAModel
    m = SeAModel(d, prec, DEV)
    s = m.system(x[None], at[None], box.reshape(1, 9))
    out = m.evaluate(s)
    torch.cuda.synchronize()
    res = {k: out[k].double().cpu().numpy().copy() for k in ("energy", "force", "virial")}
    res["nlist"] = s.nlist.cpu().numpy().copy()
    res["status"] = s.status.cpu().numpy().copy()
    res["max_cand"] = s.fs.max_cand
    return res


def _canon_rows(nl, x, at, box, rcut, sel):
    from test_gpu_parity import _canon
    return _canon(nl, x[None], at[None], box.reshape(1, 9), rcut, sel)


@pytest.fixture(scope="module")
def c4_physical():
    cfg = bench.configs()["c4"]
    d, info = bench.physical_model(cfg)
    return cfg, d, info


def test_c4_model_on_oracle_box(c4_physical):
    from deepmd_amd import systems
    cfg, d, info = c4_physical
    x, at, box = systems.fcc(8, cfg["a"], cfg["sigma"], cfg["geom_seed"])
    print(f"C4 model, 2048-atom box, output scale {info['output_scale']:.3f}", flush=True)
    ref = _oracle(d, x, at, box)
    n = len(at)
    r64 = _eval(d, "f64", x, at, box)
    # neighbour list: extended indices in the oracle's image order, canonical ties
    got = r64["nlist"].reshape(1, n, -1)
    assert np.array_equal(got, _canon_rows(got, x, at, box, cfg["rcut"], cfg["sel"]))
    assert np.array_equal(got, _canon_rows(ref["nlist"], x, at, box, cfg["rcut"], cfg["sel"]))
    assert _rel(r64["energy"], ref["energy"]) < 1e-10
    assert _rel(r64["force"], ref["force"]) < 1e-10
    assert _rel(r64["virial"], ref["virial"]) < 1e-10
    f0 = ref["force"].reshape(-1, 3)
    rms = float(np.sqrt(np.mean(f0 ** 2)))
    assert 0.5 < rms < 2.0, rms  # the physical scale held on a second box
    # mixed precisions: force RMSE (eV/A) against the oracle
    bounds = {"f16x2": 1e-3, "f16": 5e-3, "bf16": 4e-2}
    for prec, bound in bounds.items():
        r = _eval(d, prec, x, at, box)
        rmse = float(np.sqrt(np.mean((r["force"] - f0) ** 2)))
        de = abs(r["energy"][0] - ref["energy"].ravel()[0]) / n
        print(f"C4 {prec}: force RMSE {rmse:.3e} eV/A (RMS {rms:.3f}), energy "
              f"{de:.2e} eV/atom", flush=True)
        assert np.all(np.isfinite(r["force"]))
        assert rmse <= bound, (prec, rmse)
    # the gate separates the precisions: only the split operands pass it
    r = _eval(d, "f16", x, at, box)
    assert float(np.sqrt(np.mean((r["force"] - f0) ** 2))) > 1e-3


def test_c4_full_frame_split_fp16(c4_physical):
    cfg, d, _ = c4_physical
    x, at, box = bench.make_system(cfg)[0]
    assert len(at) == 131072
    r64 = _eval(d, "f64", x, at, box)
    r = _eval(d, "f16x2", x, at, box)
    for res in (r64, r):
        assert int(res["status"][0]) <= res["max_cand"]  # no candidate overflow
        assert np.all(np.isfinite(res["force"])) and np.all(np.isfinite(res["virial"]))
    # identical neighbour lists (geometry is FP64 in every precision)
    assert np.array_equal(r64["nlist"], r["nlist"])
    rmse = float(np.sqrt(np.mean((r["force"] - r64["force"]) ** 2)))
    rms = float(np.sqrt(np.mean(r64["force"] ** 2)))
    de = abs(r["energy"][0] - r64["energy"][0]) / len(at)
    print(f"C4 131072 atoms f16x2: force RMSE {rmse:.3e} eV/A (RMS {rms:.3f}), energy "
          f"{de:.2e} eV/atom", flush=True)
    assert rmse <= 1e-3
    assert de <= 1e-4 * abs(r64["energy"][0]) / len(at)


def test_c3_ragged_frames_vs_oracle():
    from deepmd_amd.model import SeAModel
    cfg = bench.configs()["c3"]
    frames = sorted(bench.make_system(cfg), key=lambda f: len(f[1]))[::16]
    assert len(frames) == 16 and len(frames[0][1]) < 100 and len(frames[-1][1]) > 700
    d = bench.model_dict(cfg)
    refs = [_oracle(d, *f) for f in frames]
    nlocs = [len(f[1]) for f in frames]
    for prec, tol in (("f64", 1e-10), ("f32", 1e-4)):
        m = SeAModel(d, prec, DEV)
        s = m.system([f[0] for f in frames], [f[1] for f in frames],
                     [f[2].reshape(9) for f in frames], nlocs=nlocs)
        out = m.evaluate(s)
        torch.cuda.synchronize()
        e = out["energy"].cpu().numpy()
        f = out["force"].cpu().numpy()
        v = out["virial"].cpu().numpy()
        nl = s.nlist.cpu().numpy()
        off = 0
        for k, (fr, ref) in enumerate(zip(frames, refs)):
            n = nlocs[k]
            fk = f[off:off + n]
            if prec == "f64":
                got = nl[off:off + n].reshape(1, n, -1)
                # extended indices of a packed frame are its own (image * n + j)
                assert np.array_equal(got, _canon_rows(ref["nlist"], fr[0], fr[1], fr[2],
                                                       cfg["rcut"], cfg["sel"])), k
                assert _rel(e[k], ref["energy"]) < tol, k
                assert _rel(fk, ref["force"]) < tol, k
                assert _rel(v[k], ref["virial"]) < tol, k
            else:
                f0 = ref["force"].reshape(-1, 3)
                rms = float(np.sqrt(np.mean(f0 ** 2)))
                rmse = float(np.sqrt(np.mean((fk - f0) ** 2)))
                assert rmse <= tol * rms, (k, rmse, rms)
                assert _rel(e[k], ref["energy"]) < 1e-5, k
            off += n
        print(f"C3 16 frames ({sum(nlocs)} atoms) {prec}: ok", flush=True)


def test_c2_training_gradients_vs_oracle():
    from test_gpu_train import _oracle_loss_grads
    from deepmd_amd import systems
    from deepmd_amd.model import SeAModel, random_model_dict
    from deepmd_amd.train import SeATrainer
    cfg = bench.configs()["c2"]
    fr = [systems.water(cfg["nside"], cfg["box"], sd) for sd in cfg["geom_seeds"]]
    coords = np.stack([f[0] for f in fr])
    atypes = np.stack([f[1] for f in fr])
    boxes = np.stack([f[2].reshape(9) for f in fr])
    assert coords.shape[:2] == (8, 192)
    args = (cfg["rcut"], cfg["rcut_smth"], cfg["sel"], cfg["neuron"], cfg["axis"], cfg["fit"])
    d, d2 = random_model_dict(cfg["weight_seed"], *args), random_model_dict(cfg["label_seed"], *args)
    lm = SeAModel(d2, "f64", DEV)
    out = lm.evaluate(lm.system(coords, atypes, boxes))
    nf, n = coords.shape[:2]
    label = {"energy": out["energy"].clone(), "force": out["force"].view(nf, n, 3).clone(),
             "virial": out["virial"].clone()}
    rloss, rgrads = _oracle_loss_grads(d, coords, atypes, boxes, label)
    r = np.concatenate([g.ravel() for g in rgrads])
    for prec, tol in (("f64", 1e-8), ("bf16", 2e-3)):
        tr = SeATrainer(d, prec, DEV)
        loss, grads, _ = tr.loss_and_grads(tr.batch(coords, atypes, boxes), label)
        a = np.concatenate([g.detach().double().cpu().numpy().ravel() for g in grads])
        rel = float(np.linalg.norm(a - r) / np.linalg.norm(r))
        lrel = abs(loss.item() - rloss) / abs(rloss)
        print(f"C2 {prec}: loss rel err {lrel:.2e}, gradient rel err {rel:.2e}", flush=True)
        if prec == "f64":
            assert lrel <= 1e-10
            for k, (g, rg) in enumerate(zip(grads, rgrads)):
                assert _rel(g.detach().cpu().numpy(), rg) < tol, k
        else:
            # the BF16 trainer runs the embedding, its backward and the force
            # path's second derivative on the hand-written training kernels
            assert tr.kernels is not None
            assert lrel <= 1e-3 and rel <= tol
            nemb = sum(len(n.params()) for n in tr.emb)
            for k in range(nemb):
                ek = float(np.linalg.norm(grads[k].detach().double().cpu().numpy() - rgrads[k])
                           / np.linalg.norm(rgrads[k]))
                assert ek <= 1e-2, (k, ek)


def test_c1_compressed_water_vs_uncompressed_oracle():
    cfg = bench.configs()["c1"]
    x, at, box = bench.make_system(cfg)[0]
    assert len(at) == 12288
    d = bench.model_dict(cfg)
    from deepmd_amd.model import SeAModel
    m = SeAModel(d, "f32", DEV)
    m.enable_compression(**cfg["compress"])
    s = m.system(x[None], at[None], box.reshape(1, 9))
    out = m.evaluate(s)
    torch.cuda.synchronize()
    ref = _oracle(d, x, at, box)
    f0 = ref["force"].reshape(-1, 3)
    f = out["force"].cpu().numpy()
    rms = float(np.sqrt(np.mean(f0 ** 2)))
    rmse = float(np.sqrt(np.mean((f - f0) ** 2)))
    e = _rel(out["energy"].cpu().numpy(), ref["energy"])
    v = _rel(out["virial"].cpu().numpy(), ref["virial"])
    print(f"C1 12288 atoms compressed f32: force RMSE {rmse:.3e} (RMS {rms:.3e}), E {e:.2e}, "
          f"V {v:.2e}", flush=True)
    assert rmse <= 1e-4 * rms
    assert e < 1e-6 and v < 1e-4

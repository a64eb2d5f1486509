"""Host-side pieces of the training loop: learning-rate schedules, loss
prefactor annealing, the set.*/*.npy data reader and batch sampler, the
statistics helpers, the model initialisation and the checkpoint writer.

Pins: the reference's own learning-rate tests (source/tests/pt/test_lr.py:
decay_rate_pt's identities, the cosine curve's three values), closed forms
(a staircase r^floor(t/s) that must land on stop_lr at the last step, a
linear warmup, least squares that recovers exactly planted per-type
energies), the reference's water system (tests/golden/water_data_0, copied
by make_water_data.py) and a checkpoint round trip through the independent
reader ``model_dict_from_state_dict``.
"""

import math
import os
import types

import numpy as np
import pytest
import torch

from deepmd_amd import training as T
from deepmd_amd.checkpoint import (model_dict_from_state_dict, save_checkpoint,
                                   state_dict_from_model_dict, load_checkpoint)
from deepmd_amd.train import SeATrainer

WATER = os.path.join(os.path.dirname(__file__), "golden", "water_data_0")
START, STOP = 1e-3, 3.51e-8


@pytest.mark.parametrize("decay_steps,num_steps", [(400, 500), (400, 1000), (400, 1500),
                                                   (500, 500), (500, 1000), (500, 1500)])
def test_lr_exp_matches_reference_identities(decay_steps, num_steps):
    lr = T.LearningRateExp(start_lr=START, stop_lr=STOP, decay_steps=decay_steps,
                           num_steps=num_steps)
    ds = lr.decay_steps
    if decay_steps >= num_steps:  # auto-adjusted: T // 10 > 100 ? 100 : T // 100 + 1
        assert ds == (100 if num_steps // 10 > 100 else num_steps // 100 + 1)
    else:
        assert ds == decay_steps
    # staircase: constant inside a period, ratio r between periods
    vals = np.array([lr.value(t) for t in range(num_steps)])
    assert vals[0] == START
    assert np.all(vals[1:ds] == START)
    r = vals[ds] / vals[0]
    np.testing.assert_allclose(r ** (num_steps / ds), STOP / START, rtol=1e-10)
    # test_lr.py decay_rate_pt: an explicit decay rate equal to the computed one
    # reproduces the schedule; a larger stop_lr truncates it
    rate = math.exp(math.log(STOP / START) / (num_steps / ds))
    same = T.LearningRateExp(start_lr=START, stop_lr=1e-10, decay_steps=ds,
                             num_steps=num_steps, decay_rate=rate)
    trunc = T.LearningRateExp(start_lr=START, stop_lr=1e-5, decay_steps=ds,
                              num_steps=num_steps, decay_rate=rate)
    keep = [t for t in range(num_steps) if t % ds != 0]
    np.testing.assert_allclose([same.value(t) for t in keep], vals[keep], rtol=1e-12)
    np.testing.assert_allclose([trunc.value(t) for t in keep],
                               np.clip(vals[keep], 1e-5, None), rtol=1e-12)


def test_lr_exp_smooth_and_warmup():
    lr = T.LearningRateExp(start_lr=1.0, stop_lr=0.01, decay_steps=10, num_steps=110,
                           warmup_steps=10, warmup_start_factor=0.1, smooth=True)
    np.testing.assert_allclose([lr.value(t) for t in range(10)],
                               0.1 + 0.9 * np.arange(10) / 10, rtol=1e-15)
    assert lr.value(10) == 1.0
    # smooth: lr(t) = r^((t - warmup) / s), reaching stop_lr after the decay steps
    np.testing.assert_allclose(lr.value(110), 0.01, rtol=1e-12)
    np.testing.assert_allclose(lr.value(15) ** 2, lr.value(20), rtol=1e-12)
    with pytest.raises(ValueError, match="mutually exclusive"):
        T.LearningRateExp(start_lr=1.0, stop_lr=0.1, stop_lr_ratio=0.1, num_steps=10)
    with pytest.raises(ValueError, match="smaller than num_steps"):
        T.LearningRateExp(start_lr=1.0, stop_lr=0.1, num_steps=10, warmup_steps=10)


def test_lr_cosine_reference_curve():
    lr = T.LearningRateCosine(start_lr=1.0, stop_lr=0.1, num_steps=10)
    assert np.isclose(lr.value(0), 1.0)
    assert np.isclose(lr.value(10), 0.1)
    assert np.isclose(lr.value(15), 0.1)
    assert np.isclose(lr.value(5), 0.1 + 0.9 * 0.5)


@pytest.mark.parametrize("kind", ["inverse_linear", "cosine", "linear"])
def test_lr_wsd_reference_curves(kind):
    """test_lr.py TestLearningRateWSD: stable until step 8 of 10, the decay
    midpoint at step 9, stop_lr from step 10."""
    lr = T.LearningRateWSD(start_lr=1.0, stop_lr=0.1, num_steps=10, decay_phase_ratio=0.2,
                           decay_type=kind)
    mid = {"inverse_linear": 1.0 / (0.5 / 0.1 + 0.5 / 1.0), "cosine": 0.1 + 0.9 * 0.5,
           "linear": 1.0 + (0.1 - 1.0) * 0.5}[kind]
    assert np.isclose(lr.value(0), 1.0) and np.isclose(lr.value(8), 1.0)
    assert np.isclose(lr.value(9), mid)
    assert np.isclose(lr.value(10), 0.1) and np.isclose(lr.value(15), 0.1)


def test_make_learning_rate_defaults():
    lr = T.make_learning_rate({"type": "exp", "start_lr": 1e-3, "decay_steps": 10}, 100)
    assert lr.stop_lr == 5e-8
    assert isinstance(T.make_learning_rate({"type": "wsd", "start_lr": 1e-3}, 100),
                      T.LearningRateWSD)
    with pytest.raises(NotImplementedError):
        T.make_learning_rate({"type": "poly", "start_lr": 1e-3}, 100)


def _fake_trainer(pref, limit, start_lr):
    ns = types.SimpleNamespace(pref=pref, limit_pref=limit, start_lr=start_lr,
                               device=torch.device("cpu"),
                               has=tuple(a != 0 and b != 0 for a, b in zip(pref, limit)))
    ns.prefactors = lambda lr=None: SeATrainer.prefactors(ns, lr)
    return ns


def test_prefactor_annealing_and_loss_terms():
    tr = _fake_trainer((0.02, 1000.0, 0.0), (1.0, 1.0, 0.0), 1e-3)
    assert SeATrainer.prefactors(tr, 1e-3) == pytest.approx((0.02, 1000.0, 0.0))
    assert SeATrainer.prefactors(tr, 0.0) == pytest.approx((1.0, 1.0, 0.0))
    pe, pf, _ = SeATrainer.prefactors(tr, 5e-4)
    assert pe == pytest.approx(0.51) and pf == pytest.approx(500.5)
    nloc = 4
    pred = {"energy": torch.tensor([1.0, 2.0], dtype=torch.float64),
            "force": torch.zeros(2, nloc, 3, dtype=torch.float64),
            "virial": torch.zeros(2, 9, dtype=torch.float64)}
    lab = {"energy": torch.tensor([1.5, 1.0], dtype=torch.float64),
           "force": torch.full((2, nloc, 3), 0.5, dtype=torch.float64),
           "virial": torch.ones(2, 9, dtype=torch.float64)}
    loss, more = SeATrainer.loss_terms(tr, pred, lab, nloc, 0.0)
    l2e, l2f = (0.25 + 1.0) / 2, 0.25
    assert float(loss) == pytest.approx(l2e / nloc + l2f)
    assert float(more["rmse_e"]) == pytest.approx(math.sqrt(l2e) / nloc)
    assert float(more["rmse_f"]) == pytest.approx(0.5)
    assert "rmse_v" not in more  # both virial prefactors zero: term off
    assert float(more["rmse"]) == pytest.approx(math.sqrt(float(loss)))
    lab["find_force"] = 0.0
    loss2, more2 = SeATrainer.loss_terms(tr, pred, lab, nloc, 0.0)
    assert float(loss2) == pytest.approx(l2e / nloc) and math.isnan(float(more2["rmse_f"]))
    # start nonzero but limit zero: the term is off altogether (ener.py:160)
    tr0 = _fake_trainer((1.0, 1.0, 1.0), (1.0, 0.0, 1.0), 1e-3)
    _, more3 = SeATrainer.loss_terms(tr0, pred, lab, nloc, 1e-3)
    assert "rmse_f" not in more3 and "rmse_v" in more3


def test_water_system_reader_and_sampler():
    d = T.DeepmdData(WATER)
    assert d.natoms == 192 and d.nframes == 6 and d.pbc
    assert (np.bincount(d.atom_type) == [64, 128]).all()
    fr = d.frames([0, 3])
    assert fr["coord"].shape == (2, 192, 3) and fr["box"].shape == (2, 9)
    assert fr["force"].shape == (2, 192, 3) and fr["energy"].shape == (2, 1)
    assert fr["find_energy"] == 1.0 and fr["find_virial"] == 0.0
    raw = np.load(os.path.join(WATER, "set.000", "coord.npy"))
    np.testing.assert_array_equal(fr["coord"][1].reshape(-1), raw[3].astype(np.float64))
    # remap to a type map that lists H first (data.py:101-124)
    d2 = T.DeepmdData(WATER, type_map=["H", "O"])
    np.testing.assert_array_equal(d2.atom_type, 1 - d.atom_type)
    with pytest.raises(ValueError, match="not present"):
        T.DeepmdData(WATER, type_map=["H", "C"])
    # auto batch sizes: smallest b with b * 192 >= rule
    assert T.DataSystem([WATER], "auto").batch_size == [1]
    assert T.DataSystem([WATER], "auto:500").batch_size == [3]
    # two systems: probability follows batch counts (6 // 2 = 3 and 6 // 4 = 1)
    ds = T.DataSystem([WATER, WATER], [2, 4], seed=1)
    np.testing.assert_allclose(ds.probs, [0.75, 0.25])
    # one epoch of system 0 visits every frame once
    seen = []
    for _ in range(3):
        _, b = ds.get_batch(0)
        seen += [int(np.where((d.coord == c).all(axis=(1, 2)))[0][0]) for c in b["coord"]]
    assert sorted(seen) == list(range(6))
    st = ds.stat_frames(2)
    assert [s["coord"].shape[0] for s in st] == [4, 6]


def test_energy_bias_and_std():
    rng = np.random.default_rng(0)
    bias = np.array([-93.5, -187.25])
    frames = []
    for _ in range(3):
        at = rng.integers(0, 2, size=(4, 10))
        e = np.array([bias[a].sum() for a in at]).reshape(-1, 1)
        frames.append({"atype": at, "energy": e, "find_energy": 1.0})
    frames.append({"atype": at, "energy": e + 100.0, "find_energy": 0.0})  # no labels: ignored
    np.testing.assert_allclose(T.energy_bias(frames, 2), bias, rtol=1e-10)
    assert T._std(0, 0, 0) == 0.1
    assert T._std(10, 10.0, 10.0) == 0.01  # zero spread -> protection
    assert T._std(2, 0.0, 2.0) == pytest.approx(1.0)


def test_init_model_dict_layout_and_distribution():
    mp = {"type_map": ["O", "H"],
          "descriptor": {"type": "se_e2_a", "sel": [46, 92], "rcut": 6.0, "rcut_smth": 0.5,
                         "neuron": [25, 50, 100], "axis_neuron": 16, "resnet_dt": False},
          "fitting_net": {"neuron": [240, 240, 240], "resnet_dt": True}}
    md = T.init_model_dict(mp, seed=3)
    m = md["model"]
    emb = m["descriptor"]["embeddings"]
    assert emb["ndim"] == 2 and len(emb["networks"]) == 4
    fit = m["fitting"]["nets"]["networks"]
    assert len(fit) == 2 and len(fit[0]["layers"]) == 4
    w = np.asarray(fit[0]["layers"][1]["@variables"]["w"]["value"])
    assert w.shape == (240, 240)
    assert abs(w.std() - 1 / math.sqrt(480)) < 0.03 / math.sqrt(480)
    b = np.asarray(fit[0]["layers"][1]["@variables"]["b"]["value"])
    assert abs(b.std() - 1.0) < 0.2
    # time steps only on residual layers: fitting 240->240 yes, first and output layers no
    assert fit[0]["layers"][0]["@variables"]["idt"] is None
    idt = np.asarray(fit[0]["layers"][1]["@variables"]["idt"]["value"])
    assert abs(idt.mean() - 0.1) < 1e-3
    assert fit[0]["layers"][3]["@variables"]["idt"] is None
    assert fit[0]["layers"][3]["activation_function"] == "none"
    assert all(ly["@variables"]["idt"] is None for ly in emb["networks"][0]["layers"])
    with pytest.raises(NotImplementedError):
        T.init_model_dict(dict(mp, descriptor=dict(mp["descriptor"], type="dpa1")))


def test_checkpoint_writer_round_trip(sea_yaml, tmp_path):
    m = sea_yaml["model"] if "model" in sea_yaml else sea_yaml
    md = {"model": m}
    path = str(tmp_path / "model.ckpt-7.pt")
    save_checkpoint(path, md, lr=2.5e-4, step=7, optimizer_state={"state": {}})
    raw = torch.load(path, map_location="cpu", weights_only=True)
    assert raw["model"]["_extra_state"]["train_infos"] == {"lr": 2.5e-4, "step": 7}
    back = load_checkpoint(path)["model"]

    def nets(x, k1, k2):
        return [[(np.asarray(ly["@variables"]["w"]["value"]),
                  np.asarray(ly["@variables"]["b"]["value"]))
                 for ly in n["layers"]] for n in x[k1][k2]["networks"]]

    for k1, k2 in (("descriptor", "embeddings"), ("fitting", "nets")):
        for n0, n1 in zip(nets(m, k1, k2), nets(back, k1, k2)):
            for (w0, b0), (w1, b1) in zip(n0, n1):
                np.testing.assert_array_equal(w0, w1)
                np.testing.assert_array_equal(b0, b1)
    for key in ("davg", "dstd"):
        np.testing.assert_array_equal(
            np.asarray(m["descriptor"]["@variables"][key]["value"]).reshape(-1),
            np.asarray(back["descriptor"]["@variables"][key]["value"]).reshape(-1))
    sd = state_dict_from_model_dict(md)
    assert not any(k.startswith("model.Default.atomic_model.descriptor.sea.filter_layers.n")
                   and k.replace("networks", "_networks", 1) not in sd
                   for k in sd if ".networks." in k)
    with pytest.raises(ValueError):
        model_dict_from_state_dict(sd)  # no _extra_state


def test_evaluation_metrics_and_merge():
    from deepmd_amd import evaluate as E
    # MAE / RMSE of a known difference, per-atom scaling
    mae, rmse, w = E.error_stat(np.array([1.0, -3.0]), np.zeros(2), scale=0.5)
    assert (mae, rmse, w) == (1.0, math.sqrt(5.0) * 0.5, 2.0)
    n = 4
    pred = {"energy": np.array([[1.0], [2.0]]), "force": np.zeros((2, n, 3)),
            "virial": np.ones((2, 9))}
    lab = {"energy": np.array([[0.0], [0.0]]), "force": np.ones((2, n, 3)),
           "virial": np.zeros((2, 9)), "find_energy": 1.0, "find_force": 1.0,
           "find_virial": 0.0}
    m = E.system_metrics(pred, lab, n, pbc=True)
    assert set(m) == {"energy", "energy_per_atom", "force"}  # no virial labels
    assert m["energy"][:2] == (1.5, math.sqrt(2.5))
    assert m["energy_per_atom"][:2] == (1.5 / n, math.sqrt(2.5) / n)
    lab["find_virial"] = 1.0
    assert "virial" not in E.system_metrics(pred, lab, n, pbc=False)  # open boundaries
    # merging two systems equals the metrics of their concatenation
    a, b = np.array([1.0, 2.0, -1.0]), np.array([0.5, -4.0])
    both = E.merge([{"force": E.error_stat(a, 0 * a)}, {"force": E.error_stat(b, 0 * b)}])
    cat = np.concatenate([a, b])
    assert both["mae_f"] == pytest.approx(np.mean(np.abs(cat)))
    assert both["rmse_f"] == pytest.approx(np.sqrt(np.mean(cat * cat)))


def _ddp_worker(rank, world, port, q):
    import torch.distributed as dist
    dist.init_process_group("gloo", init_method=f"tcp://127.0.0.1:{port}", rank=rank,
                            world_size=world)
    try:
        ps = [torch.zeros(3, dtype=torch.float32, requires_grad=True),
              torch.zeros(2, 2, dtype=torch.float64, requires_grad=True),
              torch.zeros(1, dtype=torch.float64, requires_grad=True)]
        ps[0].grad = torch.full((3,), float(rank + 1), dtype=torch.float32)
        ps[1].grad = torch.arange(4, dtype=torch.float64).view(2, 2) * (rank + 1)
        if rank == 1:  # rank 0 has no gradient for the last parameter: zeros
            ps[2].grad = torch.tensor([6.0], dtype=torch.float64)
        T.allreduce_gradients(ps)
        q.put((rank, [p.grad.tolist() for p in ps], [p.grad.dtype for p in ps]))
    finally:
        dist.destroy_process_group()


def test_allreduce_gradients_gloo_two_ranks():
    """Data-parallel gradient averaging over a 2-rank gloo group (CPU)."""
    import socket
    import torch.multiprocessing as mp
    if torch.cuda.device_count() > 0:
        pytest.skip("multi-process CPU test: run where no GPU is visible")
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        port = s.getsockname()[1]
    ctx = mp.get_context("spawn")
    q = ctx.Queue()
    procs = [ctx.Process(target=_ddp_worker, args=(r, 2, port, q)) for r in range(2)]
    for p in procs:
        p.start()
    out = dict((r, (g, d)) for r, g, d in (q.get(timeout=120) for _ in procs))
    for p in procs:
        p.join(timeout=60)
        assert p.exitcode == 0
    for r in range(2):
        g, d = out[r]
        assert g[0] == [1.5, 1.5, 1.5]
        assert g[1] == [[0.0, 1.5], [3.0, 4.5]]
        assert g[2] == [3.0]
        assert d == [torch.float32, torch.float64, torch.float64]
    T.allreduce_gradients([torch.zeros(1, requires_grad=True)])  # no group: no-op


@pytest.mark.parametrize("neuron,type_one_side", [([25, 50, 100], True), ([25, 50, 100], False),
                                                  ([32, 64, 128], True), ([16, 48, 96], True)])
def test_training_kernel_gather_map_equals_pack_embedding(neuron, type_one_side):
    """The training kernels' packed parameters are a gather of the master
    weights; the gather map must reproduce the inference packing
    (model.pack_embedding: channel plan, transposes, padding, idt = 1)."""
    from deepmd_amd.model import _layers, pack_embedding, random_model_dict
    from deepmd_amd.train_embed import packing_index
    d = random_model_dict(3, 6.0, 0.5, [46, 92], neuron, 16, [32],
                          type_one_side=type_one_side)
    nets = [_layers(n) for n in d["model"]["descriptor"]["embeddings"]["networks"]]
    packed, _, _, _ = pack_embedding(nets, neuron)
    flat = np.concatenate([np.concatenate([ly.w.ravel(), ly.b.ravel()])
                           for net in nets for ly in net] + [np.array([0.0, 1.0])])
    idx = packing_index(neuron, len(nets))
    assert idx.shape == packed.shape
    assert np.array_equal(flat[idx], packed)

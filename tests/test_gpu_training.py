"""Training loop on the GPU: input statistics, Adam steps on the reference's
water system, learning curve, checkpoints.

* Input statistics (davg / dstd of the raw environment matrix per centre
  type) from the device path are compared with the same sums over the
  oracle's neighbour list and environment matrix (pinned to the reference's
  goldens elsewhere) on frames of tests/golden/water_data_0.
* A trainer built from the reference's se_e2_a.json (source/tests/pt/water/
  se_e2_a.json: sel [46, 92], neuron [25, 50, 100], axis 16, fitting
  [240, 240, 240], start_pref_e 0.02 -> 1, start_pref_f 1000 -> 1) takes 150
  Adam steps; the force error must fall, the learning curve follows the
  reference's format, and a saved checkpoint loads back into DeepPot with
  the trained weights (E / F equal to the trainer's own prediction).
"""

import math
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

WATER = os.path.join(os.path.dirname(__file__), "golden", "water_data_0")
DEV = "cuda"


@pytest.fixture(scope="module", autouse=True)
def _hip():
    if not torch.cuda.is_available():
        pytest.fail("a HIP device is required for the gpu tests")


def _config(steps, disp_freq=1):
    return {
        "model": {"type_map": ["O", "H"],
                  "descriptor": {"type": "se_e2_a", "sel": [46, 92], "rcut_smth": 0.5,
                                 "rcut": 6.0, "neuron": [25, 50, 100], "resnet_dt": False,
                                 "axis_neuron": 16, "seed": 1},
                  "fitting_net": {"neuron": [240, 240, 240], "resnet_dt": True, "seed": 1},
                  "data_stat_nbatch": 3},
        "learning_rate": {"type": "exp", "decay_steps": 50, "start_lr": 1e-3,
                          "stop_lr": 3e-4},
        "loss": {"type": "ener", "start_pref_e": 0.02, "limit_pref_e": 1,
                 "start_pref_f": 1000, "limit_pref_f": 1},
        "optimizer": {"type": "Adam", "adam_beta1": 0.9, "adam_beta2": 0.999},
        "training": {"training_data": {"systems": [WATER], "batch_size": 1},
                     "validation_data": {"systems": [WATER], "batch_size": 1, "numb_btch": 2},
                     "numb_steps": steps, "seed": 10, "disp_file": "lcurve.out",
                     "disp_freq": disp_freq, "save_freq": steps,
                     "gradient_max_norm": 50.0},
    }


def test_input_stats_match_oracle():
    from deepmd_amd import training as T
    from oracle import geometry
    from oracle import sea as osea
    cfg = _config(1)
    md = T.init_model_dict(cfg["model"], seed=1)
    data = T.DataSystem([WATER], 1)
    frames = data.stat_frames(3)
    got = T.env_mat_sums(md, frames, DEV)
    want = np.zeros_like(got)
    sel = [46, 92]
    for fr in frames:
        c = torch.tensor(fr["coord"], dtype=torch.float64)
        nf, nloc = c.shape[:2]
        at = torch.tensor(fr["atype"], dtype=torch.int64)
        box = torch.tensor(fr["box"], dtype=torch.float64)
        ext, eat, _, nl = geometry.extend_input_and_build_neighbor_list(
            c, at, 6.0, sel, box=box, mixed_types=True)
        nl = geometry.format_nlist(ext.view(nf, -1, 3), eat, nl, 6.0, sel)
        em, _, _ = osea.make_env_mat(nl, ext, 6.0, 0.5)
        em = em.reshape(nf * nloc, -1, 4).numpy()
        a = at.reshape(-1).numpy()
        for t in range(2):
            rows = em[a == t]
            r, ang = rows[..., :1], rows[..., 1:]
            want[t] += [r.size, r.sum(), (r * r).sum(), ang.size, ang.sum(), (ang * ang).sum()]
    np.testing.assert_array_equal(got[:, [0, 3]], want[:, [0, 3]])
    np.testing.assert_allclose(got, want, rtol=1e-12)
    md = T.compute_input_stats(md, frames, DEV)
    davg = np.asarray(md["model"]["descriptor"]["@variables"]["davg"]["value"])
    dstd = np.asarray(md["model"]["descriptor"]["@variables"]["dstd"]["value"])
    for t in range(2):
        n, s, s2 = want[t, :3]
        assert np.allclose(davg[t, :, 0], s / n, rtol=1e-12) and not davg[t, :, 1:].any()
        assert np.allclose(dstd[t, :, 0], math.sqrt(s2 / n - (s / n) ** 2), rtol=1e-10)
    # O and H energies of the water frames: bias O + 2 H ~ frame energy / 64
    b = np.asarray(md["model"]["@variables"]["out_bias"]["value"]).reshape(-1)
    e = np.concatenate([fr["energy"].reshape(-1) for fr in frames])
    np.testing.assert_allclose(64 * b[0] + 128 * b[1], e.mean(), rtol=1e-10)


def test_trainer_steps_curve_and_checkpoint(tmp_path):
    from deepmd_amd import training as T
    from deepmd_amd.deep_pot import DeepPot
    steps, freq = 150, 10
    tr = T.Trainer(_config(steps, freq), precision="f64", device=DEV, output_dir=str(tmp_path))
    print(f"trainer: {steps} steps, water 192 atoms, f64", flush=True)
    rec = tr.run()
    assert [s for s, _ in rec] == [1] + list(range(freq, steps + 1, freq))
    for _, r in rec:
        assert all(math.isfinite(r[k]) for k in ("rmse", "rmse_e", "rmse_f"))
    assert rec[0][1]["lr"] == 1e-3 and rec[-1][1]["lr"] < rec[0][1]["lr"]
    f0 = np.mean([r["rmse_f"] for _, r in rec[:2]])
    f1 = np.mean([r["rmse_f"] for _, r in rec[-3:]])
    print("rmse_f", [round(r["rmse_f"], 4) for _, r in rec], flush=True)
    assert f1 < 0.9 * f0
    lines = open(tmp_path / "lcurve.out").read().splitlines()
    assert lines[0].split() == ["#", "step", "rmse_val", "rmse_trn", "rmse_e_val", "rmse_e_trn",
                                "rmse_f_val", "rmse_f_trn", "lr"]
    assert all(math.isfinite(r["rmse_f_val"]) for _, r in rec)
    assert float(lines[-1].split()[5]) == pytest.approx(rec[-1][1]["rmse_f_val"], rel=1e-2)
    assert len(lines) == 2 + len(rec) and int(lines[-1].split()[0]) == steps
    ck = tmp_path / f"model.ckpt-{steps}.pt"
    assert ck.exists() and (tmp_path / "checkpoint").read_text() == ck.name
    # the checkpoint holds the trained weights: DeepPot(f64) = trainer prediction
    data = T.DeepmdData(WATER)
    fr = data.frames([0])
    pred = tr.model.predict(tr.model.batch(fr["coord"], fr["atype"], fr["box"]),
                            create_graph=False)
    dp = DeepPot(str(ck), precision="f64", device=DEV)
    e, f, _ = dp.eval(fr["coord"].reshape(1, -1), fr["box"], data.atom_type)
    np.testing.assert_allclose(np.asarray(e).reshape(-1),
                               pred["energy"].detach().cpu().numpy(), rtol=1e-11)
    fp = pred["force"].detach().cpu().numpy().reshape(-1)
    assert np.max(np.abs(np.asarray(f).reshape(-1) - fp)) < 1e-9 * np.max(np.abs(fp))
    # dp-test metrics of the checkpoint (fused DeepPot path) = errors of the
    # trainer's autograd prediction on the same frame
    from deepmd_amd import evaluate as E
    res = E.test_model(dp, [WATER], numb_test=1)["average"]
    fl = fr["force"].reshape(-1)
    assert res["rmse_f"] == pytest.approx(float(np.sqrt(np.mean((fp - fl) ** 2))), rel=1e-8)
    de = abs(float(pred["energy"].item()) - float(fr["energy"].reshape(-1)[0]))
    assert res["rmse_e"] == pytest.approx(de, rel=1e-8)
    assert res["mae_ea"] == pytest.approx(de / 192, rel=1e-8)
    assert "rmse_v" not in res  # the water system has no virial labels
    # resuming from the checkpoint continues from its weights
    tr2 = T.Trainer(_config(steps), init_model=str(ck), precision="f64", device=DEV,
                    output_dir=str(tmp_path))
    p2 = tr2.model.predict(tr2.model.batch(fr["coord"], fr["atype"], fr["box"]),
                           create_graph=False)
    np.testing.assert_allclose(p2["energy"].detach().cpu().numpy(),
                               pred["energy"].detach().cpu().numpy(), rtol=1e-12)



def test_trainer_batch_with_boxes_of_different_size(tmp_path):
    """One batch of two frames in a 12.4 A and a 5.85 A box (one and two
    image shells at rcut 6): the step's display terms equal the root mean
    square of the terms of each frame evaluated on its own beforehand."""
    import shutil
    from deepmd_amd import training as T
    d = T.DeepmdData(WATER)
    sysd = tmp_path / "sys"
    (sysd / "set.000").mkdir(parents=True)
    for f in ("type.raw", "type_map.raw"):
        shutil.copyfile(os.path.join(WATER, f), sysd / f)
    sc = 0.47
    np.save(sysd / "set.000" / "coord.npy", np.stack([d.coord[0], d.coord[1] * sc]).reshape(2, -1))
    np.save(sysd / "set.000" / "box.npy", np.stack([d.box[0], d.box[1] * sc]))
    np.save(sysd / "set.000" / "energy.npy", d.labels["energy"][:2].reshape(-1))
    np.save(sysd / "set.000" / "force.npy", d.labels["force"][:2].reshape(2, -1))
    cfg = _config(2)
    cfg["training"]["training_data"] = {"systems": [str(sysd)], "batch_size": 2}
    tr = T.Trainer(cfg, precision="f64", device=DEV, output_dir=str(tmp_path))
    m, lr = tr.model, tr.lr.value(0)
    sq = {}
    for k in range(2):
        fr = tr.data.systems[0].frames([k])
        b = m.batch(fr["coord"], fr["atype"], fr["box"])
        _, more = m.loss_terms(m.predict(b, create_graph=False), tr._label(fr), b.nloc, lr)
        for key, v in more.items():
            sq[key] = sq.get(key, 0.0) + 0.5 * float(v) ** 2
    got = tr.step()
    for key in ("rmse", "rmse_e", "rmse_f"):
        assert got[key] == pytest.approx(math.sqrt(sq[key]), rel=1e-10), key


def test_final_checkpoint_step_and_restart(tmp_path):
    """A run whose length is no multiple of save_freq still ends with a
    checkpoint; train_infos.step is the 0-based index of the last step
    (pt/train/training.py:1744-1760), and a restart continues from it
    (training.py:781-785) with the saved optimiser state."""
    from deepmd_amd import training as T
    cfg = _config(5)
    cfg["training"]["save_freq"] = 3
    cfg["training"].pop("validation_data")
    tr = T.Trainer(cfg, precision="f64", device=DEV, output_dir=str(tmp_path))
    tr.run()
    names = sorted(p.name for p in tmp_path.glob("model.ckpt-*.pt"))
    assert names == ["model.ckpt-3.pt", "model.ckpt-5.pt"]
    raw = torch.load(tmp_path / "model.ckpt-5.pt", map_location="cpu", weights_only=True)
    assert raw["model"]["_extra_state"]["train_infos"]["step"] == 4
    assert raw["model"]["_extra_state"]["train_infos"]["lr"] == pytest.approx(tr.lr.value(4))
    tr2 = T.Trainer(cfg, init_model=str(tmp_path / "model.ckpt-3.pt"), precision="f64",
                    device=DEV, output_dir=str(tmp_path / "r"), restart=True)
    assert tr2.step_id == 2
    st = tr2.model.opt.state_dict()["state"]
    assert st and all(int(v["step"]) == 3 for v in st.values())
    os.makedirs(tmp_path / "r", exist_ok=True)
    tr2.run()
    assert tr2.step_id == 5
    # the restart step itself writes nothing; the last step does
    assert sorted(p.name for p in (tmp_path / "r").glob("model.ckpt-*.pt")) == ["model.ckpt-5.pt"]

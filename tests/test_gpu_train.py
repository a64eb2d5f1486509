"""Training step (energy + force + virial loss) on the GPU.

The force-loss gradient with respect to the weights runs through the HIP
second-order kernels (prod_force_se_a_grad / prod_virial_se_a_grad).  Checked
against: the fused evaluation (E/F/V), central finite differences of the
energy in the positions (h = 1e-5 A, FP64, 32 random atoms), the oracle's
autograd gradients of the same loss (create_graph through the torch CPU
restatement), and central finite differences of the loss in the weights.
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


def _frames(nf=2, nside=3, seed0=20):
    from deepmd_amd import systems
    L = 12.44 * nside / 4
    out = [systems.water(nside, L, seed0 + k) for k in range(nf)]
    coords = np.stack([o[0] for o in out])
    atypes = np.stack([o[1] for o in out])
    boxes = np.stack([o[2].reshape(9) for o in out])
    return coords, atypes, boxes


def _dicts(type_one_side=True):
    from deepmd_amd.model import random_model_dict
    args = (6.0, 0.5, [46, 92], [25, 50, 100], 16, [32, 32])
    return (random_model_dict(1, *args, type_one_side=type_one_side),
            random_model_dict(2, *args, type_one_side=type_one_side))


def _labels(d2, coords, atypes, boxes):
    from deepmd_amd.model import SeAModel
    m = SeAModel(d2, "f64", DEV)
    s = m.system(coords, atypes, boxes)
    out = m.evaluate(s)
    nf, n = coords.shape[:2]
    return {"energy": out["energy"].clone(), "force": out["force"].view(nf, n, 3).clone(),
            "virial": out["virial"].clone()}


@pytest.mark.parametrize("type_one_side", [True, False])
def test_predict_matches_fused_eval(type_one_side):
    from deepmd_amd.model import SeAModel
    from deepmd_amd.train import SeATrainer
    d, _ = _dicts(type_one_side)
    coords, atypes, boxes = _frames()
    tr = SeATrainer(d, "f64", DEV)
    b = tr.batch(coords, atypes, boxes)
    p = tr.predict(b, create_graph=False)
    m = SeAModel(d, "f64", DEV)
    s = m.system(coords, atypes, boxes)
    out = m.evaluate(s)
    torch.cuda.synchronize()
    assert _rel(p["energy"].detach().cpu(), out["energy"].cpu()) < 1e-10
    assert _rel(p["force"].detach().cpu().reshape(-1), out["force"].cpu().reshape(-1)) < 1e-10
    assert _rel(p["virial"].detach().cpu(), out["virial"].cpu()) < 1e-10


def test_force_central_fd_32_atoms():
    from deepmd_amd.train import SeATrainer
    d, _ = _dicts()
    coords, atypes, boxes = _frames(nf=1)
    tr = SeATrainer(d, "f64", DEV)
    f = tr.predict(tr.batch(coords, atypes, boxes), create_graph=False)["force"]
    f = f.detach().cpu().numpy()[0]
    rng = np.random.default_rng(5)
    atoms = rng.choice(coords.shape[1], size=32, replace=False)
    h = 1e-5
    for i in atoms:
        for k in range(3):
            e = []
            for sgn in (1, -1):
                c = coords.copy()
                c[0, i, k] += sgn * h
                e.append(tr.predict(tr.batch(c, atypes, boxes), create_graph=False)["energy"]
                         .item())
            fd = -(e[0] - e[1]) / (2 * h)
            assert abs(fd - f[i, k]) < 1e-6 * max(1.0, abs(f[i, k])), (i, k, fd, f[i, k])


def _oracle_loss_grads(d, coords, atypes, boxes, label, pref=(1.0, 1.0, 1.0)):
    from oracle import geometry, sea
    om = sea.model_from_serialized(d)
    params = []
    for net in om.embeddings + om.fitting:
        for ly in net:
            for name in ("w", "b", "idt"):
                t = getattr(ly, name)
                if t is not None:
                    t.requires_grad_(True)
                    params.append(t)
    nf, n = coords.shape[:2]
    c = torch.tensor(coords).view(nf, n, 3)
    a = torch.tensor(atypes).view(nf, n).long()
    bx = torch.tensor(boxes).view(nf, 9)
    ext, eat, mp, nl = geometry.extend_input_and_build_neighbor_list(c, a, om.rcut, om.sel, bx)
    low = sea.forward_lower(om, ext, eat, nl, mp, create_graph=True)
    force = sea.communicate(low["extended_force"], mp, n)
    energy = low["energy"].view(-1)
    virial = low["virial"]
    pe, pf, pv = pref
    lab = {k: v.detach().cpu() for k, v in label.items()}
    loss = (pe / n * torch.mean((energy - lab["energy"]) ** 2)
            + pf * torch.mean((lab["force"] - force) ** 2)
            + pv / n * torch.mean((lab["virial"] - virial) ** 2))
    grads = torch.autograd.grad(loss, params)
    return loss.item(), [g.numpy() for g in grads]


@pytest.mark.parametrize("type_one_side", [True, False])
def test_loss_gradients_vs_oracle_autograd(type_one_side):
    from deepmd_amd.train import SeATrainer
    d, d2 = _dicts(type_one_side)
    coords, atypes, boxes = _frames()
    label = _labels(d2, coords, atypes, boxes)
    tr = SeATrainer(d, "f64", DEV)
    loss, grads, _ = tr.loss_and_grads(tr.batch(coords, atypes, boxes), label)
    rloss, rgrads = _oracle_loss_grads(d, coords, atypes, boxes, label)
    print(f"loss {loss.item():.12e} oracle {rloss:.12e}", flush=True)
    assert abs(loss.item() - rloss) <= 1e-10 * abs(rloss)
    assert len(grads) == len(rgrads)
    for k, (g, r) in enumerate(zip(grads, rgrads)):
        assert g.shape == r.shape, k
        assert _rel(g.detach().cpu().numpy(), r) < 1e-8, k


def test_loss_gradient_weight_fd():
    from deepmd_amd.train import SeATrainer
    d, d2 = _dicts()
    coords, atypes, boxes = _frames(nf=1)
    label = _labels(d2, coords, atypes, boxes)
    tr = SeATrainer(d, "f64", DEV)
    b = tr.batch(coords, atypes, boxes)
    _, grads, _ = tr.loss_and_grads(b, label)
    params = tr.parameters()
    rng = np.random.default_rng(7)
    h = 1e-6
    for _ in range(8):
        k = int(rng.integers(len(params)))
        p = params[k]
        j = int(rng.integers(p.numel()))
        with torch.no_grad():
            old = p.view(-1)[j].item()
            p.view(-1)[j] = old + h
        lp = tr.loss(tr.predict(b, create_graph=True), label, b.nloc).item()
        with torch.no_grad():
            p.view(-1)[j] = old - h
        lm = tr.loss(tr.predict(b, create_graph=True), label, b.nloc).item()
        with torch.no_grad():
            p.view(-1)[j] = old
        fd = (lp - lm) / (2 * h)
        g = grads[k].view(-1)[j].item()
        assert abs(fd - g) <= 1e-6 * max(abs(g), 1e-3), (k, j, fd, g)


@pytest.mark.parametrize("prec", ["bf16", "f32"])
def test_mixed_precision_gradients_and_steps(prec):
    from deepmd_amd.train import SeATrainer
    d, d2 = _dicts()
    coords, atypes, boxes = _frames()
    label = _labels(d2, coords, atypes, boxes)
    ref = SeATrainer(d, "f64", DEV)
    _, g64, _ = ref.loss_and_grads(ref.batch(coords, atypes, boxes), label)
    tr = SeATrainer(d, prec, DEV, lr=1e-3)
    b = tr.batch(coords, atypes, boxes)
    _, g, _ = tr.loss_and_grads(b, label)
    a = torch.cat([x.detach().double().reshape(-1) for x in g])
    r = torch.cat([x.detach().reshape(-1) for x in g64])
    cos = float(torch.dot(a, r) / (a.norm() * r.norm()))
    print(f"{prec}: gradient cosine vs f64 {cos:.6f}", flush=True)
    assert cos > (0.999 if prec == "bf16" else 0.999999)
    losses = [tr.step(b, label).item() for _ in range(5)]
    print(f"{prec}: losses {losses}", flush=True)
    assert losses[-1] < losses[0]


def test_exclude_types_predict_and_gradients():
    """Type-pair exclusion in the training path: predictions equal the fused
    evaluation, loss gradients equal the oracle's autograd."""
    from deepmd_amd.model import SeAModel, _nd
    from deepmd_amd.train import SeATrainer
    d, d2 = _dicts(type_one_side=False)
    de = d["model"]["descriptor"]
    de["exclude_types"] = [[0, 1]]
    rng = np.random.default_rng(4)
    de["@variables"]["davg"] = _nd(rng.normal(0.0, 0.05, size=(2, 138, 4)))
    de["@variables"]["dstd"] = _nd(rng.uniform(0.5, 1.5, size=(2, 138, 4)))
    coords, atypes, boxes = _frames()
    label = _labels(d2, coords, atypes, boxes)
    tr = SeATrainer(d, "f64", DEV)
    b = tr.batch(coords, atypes, boxes)
    loss, grads, p = tr.loss_and_grads(b, label)
    m = SeAModel(d, "f64", DEV)
    out = m.evaluate(m.system(coords, atypes, boxes))
    torch.cuda.synchronize()
    assert _rel(p["energy"].detach().cpu(), out["energy"].cpu()) < 1e-10
    assert _rel(p["force"].detach().cpu().reshape(-1), out["force"].cpu().reshape(-1)) < 1e-10
    assert _rel(p["virial"].detach().cpu(), out["virial"].cpu()) < 1e-10
    rloss, rgrads = _oracle_loss_grads(d, coords, atypes, boxes, label)
    assert abs(loss.item() - rloss) <= 1e-10 * abs(rloss)
    for k, (g, r) in enumerate(zip(grads, rgrads)):
        assert _rel(g.detach().cpu().numpy(), r) < 1e-8, k


@pytest.mark.parametrize("type_one_side", [True, False])
def test_bf16_training_kernels_vs_f64(type_one_side):
    """BF16 training path: embedding forward, backward and the force path's
    second derivative as hand-written kernels (dpk_embed_fwd / _bwd /
    dpk_embed_train_vjp), the fitting in torch.  Against the FP64 trainer
    (torch autograd with create_graph): every embedding weight gradient
    within 1e-2 relative (BF16 operands; measured <= 3.2e-3), the whole
    gradient within 1e-3 (measured 7.5e-5), the loss within 1e-3."""
    from deepmd_amd.train import SeATrainer
    d, d2 = _dicts(type_one_side)
    coords, atypes, boxes = _frames()
    label = _labels(d2, coords, atypes, boxes)
    ref = SeATrainer(d, "f64", DEV)
    l64, g64, _ = ref.loss_and_grads(ref.batch(coords, atypes, boxes), label)
    tr = SeATrainer(d, "bf16", DEV)
    assert tr.kernels is not None
    l, g, _ = tr.loss_and_grads(tr.batch(coords, atypes, boxes), label)
    nemb = sum(len(n.params()) for n in tr.emb)
    errs = [float((a.double() - r).norm() / r.norm()) for a, r in zip(g, g64)]
    print("embedding gradient errors", " ".join(f"{e:.1e}" for e in errs[:nemb]), flush=True)
    assert max(errs[:nemb]) < 1e-2
    a = torch.cat([x.detach().double().reshape(-1) for x in g])
    r = torch.cat([x.detach().reshape(-1) for x in g64])
    assert float((a - r).norm() / r.norm()) < 1e-3
    assert abs(l.item() - l64.item()) / abs(l64.item()) < 1e-3


@pytest.mark.parametrize("prec", ["bf16", "f64"])
def test_captured_training_step_replays_match_eager(prec):
    """A training step captured in a HIP graph (forward, double backward,
    Adam): each replay, with eager work and host synchronisation between
    replays, gives the loss of the same step taken eagerly, and the weights
    after the replays equal the eager ones."""
    from deepmd_amd.train import SeATrainer
    d, d2 = _dicts()
    coords, atypes, boxes = _frames()
    label = _labels(d2, coords, atypes, boxes)
    ea = SeATrainer(d, prec, DEV, capturable=True)
    ba = ea.batch(coords, atypes, boxes)
    eager = [ea.step(ba, label).item() for _ in range(6)]
    gr = SeATrainer(d, prec, DEV, capturable=True)
    bg = gr.batch(coords, atypes, boxes)
    graph = gr.capture_step(bg, label, warmup=2)
    got = []
    for _ in range(4):
        got.append(graph.replay().item())
        junk = torch.randn(4096, device=DEV)
        assert torch.isfinite(junk.sum()).item()
    tol = 1e-9 if prec == "f64" else 1e-5
    print(f"{prec}: eager {eager[2:]} graph {got}", flush=True)
    for e, g in zip(eager[2:], got):
        assert abs(e - g) <= tol * abs(e), (e, g)
    for pa, pg in zip(ea.parameters(), gr.parameters()):
        assert float((pa - pg).detach().norm()) <= tol * max(float(pa.detach().norm()), 1e-30)

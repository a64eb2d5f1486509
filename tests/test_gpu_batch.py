"""Ragged multi-frame batches and the HIP-graph captured step on the GPU.

A packed batch of frames with uneven atom counts, four species and small
periodic boxes runs through one launch per kernel; every frame must equal the
oracle's single-frame result (neighbour lists bit-exact in canonical order,
FP64 energy / force / virial <= 1e-10 relative).  A captured step replayed
several times, with eager work and synchronisation in between and with the
coordinates moved in place, must equal the eager evaluation every time.
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


def _frames(n=5, nmin=24, nmax=150, seed=31):
    from deepmd_amd import systems
    return systems.ragged_metal(n, nmin, nmax, 3.6, 0.08, 4, seed)


def _model(prec, type_one_side=False, rcut=6.0):
    from deepmd_amd.model import SeAModel, random_model_dict
    d = random_model_dict(5, rcut, 0.5, [60, 60, 60, 60], [16, 32, 64], 16, [32, 32],
                          type_one_side=type_one_side)
    return d, SeAModel(d, prec, DEV)


def _system(m, frames):
    return m.system([f[0] for f in frames], [f[1] for f in frames],
                    [f[2].reshape(9) for f in frames], nlocs=[len(f[1]) for f in frames])


def _canon_rows(nl, ext, sel):
    out = nl.copy()
    sec = np.concatenate([[0], np.cumsum(sel)])
    for i in range(nl.shape[0]):
        for t in range(len(sel)):
            row = nl[i, sec[t]:sec[t + 1]]
            k = row[row >= 0]
            r = torch.linalg.norm(torch.tensor(ext[k] - ext[i]), dim=-1).numpy()
            out[i, sec[t]:sec[t] + len(k)] = k[np.lexsort((k, r))]
    return out


@pytest.mark.parametrize("type_one_side", [False, True])
def test_ragged_batch_f64_vs_oracle(type_one_side):
    from oracle import geometry, sea
    frames = _frames()
    d, m = _model("f64", type_one_side)
    om = sea.model_from_serialized(d)
    s = _system(m, frames)
    print(f"ragged batch: {[len(f[1]) for f in frames]} atoms, type_one_side={type_one_side}",
          flush=True)
    out = m.evaluate(s, atomic_virial=True)
    torch.cuda.synchronize()
    nl_all = s.nlist.cpu().numpy()
    off = 0
    for f, (x, at, box) in enumerate(frames):
        n = len(at)
        c = torch.tensor(x).view(1, n, 3)
        a = torch.tensor(at).view(1, n).long()
        b = torch.tensor(box.ravel()).view(1, 9)
        ext, eat, mp, nl = geometry.extend_input_and_build_neighbor_list(c, a, m.rcut, m.sel, b)
        nl = geometry.format_nlist(ext, eat, nl, m.rcut, m.sel).numpy()[0]
        e = ext.numpy()[0]
        mine = nl_all[off:off + n]
        assert np.array_equal(_canon_rows(mine, e, m.sel), _canon_rows(nl, e, m.sel)), f"frame {f}"
        ref = sea.forward(om, torch.tensor(x).view(1, -1), a, b, do_atomic_virial=True)
        assert _rel(out["energy"][f].item(), ref["energy"].item()) < 1e-10
        assert _rel(out["force"][off:off + n].cpu().numpy(), ref["force"].detach().numpy()) < 1e-10
        assert _rel(out["virial"][f].cpu().numpy(), ref["virial"].detach().numpy()) < 1e-10
        assert _rel(out["atom_virial"][off:off + n].cpu().numpy(),
                    ref["atom_virial"].detach().numpy()) < 1e-10
        off += n


@pytest.mark.parametrize("prec", ["f32", "bf16"])
def test_mixed_ragged_batch_error(prec):
    frames = _frames(seed=32)
    _, m64 = _model("f64")
    _, mm = _model(prec)
    ref = m64.evaluate(_system(m64, frames))
    out = mm.evaluate(_system(mm, frames))
    torch.cuda.synchronize()
    df = (out["force"] - ref["force"]).cpu().numpy()
    rms = float(np.sqrt(np.mean(ref["force"].cpu().numpy() ** 2)))
    rmse = float(np.sqrt(np.mean(df ** 2)))
    print(f"{prec}: force rmse {rmse:.3e} eV/A (rms {rms:.3e})", flush=True)
    assert rmse <= 1e-3
    assert rmse <= (2e-5 if prec == "f32" else 5e-2) * rms


@pytest.mark.parametrize("prec", ["f32", "bf16"])
def test_graph_replay_matches_eager(prec):
    frames = _frames(n=6, seed=33)
    _, m = _model(prec)
    s = _system(m, frames)
    g = m.capture(s)
    rng = np.random.default_rng(0)
    for rep in range(3):
        print(f"graph replay {rep} ({prec})", flush=True)
        eager = {k: v.clone() for k, v in m.evaluate(s).items()}
        torch.cuda.synchronize()
        out = g.replay()
        torch.cuda.synchronize()
        for k in ("energy", "force", "virial", "atom_energy"):
            assert _rel(out[k].cpu().numpy(), eager[k].cpu().numpy()) < 1e-10, (rep, k)
        # move the atoms in place, as an MD step would, and replay again
        dx = torch.tensor(rng.normal(0.0, 0.02, size=tuple(s.coord.shape)), dtype=torch.float64,
                          device=DEV)
        s.coord.add_(dx)
    g.check()

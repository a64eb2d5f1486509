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


@pytest.mark.parametrize("case", ["water", "cu_dense_nc2", "small_box"])
def test_force_cells_equals_per_atom_kernel(case):
    """The cell-blocked force kernel (LDS table + fallback atomics) equals the
    per-atom kernel (direct atomics) on the same net_deriv: a water box, Cu
    with 2 cells per axis and > 64 atoms per cell (table overflow, images
    of one atom in two relative cells), and a box smaller than rcut."""
    from deepmd_amd import _lib, systems
    from deepmd_amd.model import SeAModel, random_model_dict
    if case == "water":
        x, at, box = systems.water(4, 12.44, 3)
        d = random_model_dict(1, 6.0, 0.5, [46, 92], [8, 16, 32], 4, [16, 16])
    elif case == "cu_dense_nc2":
        x, at, box = systems.fcc(6, 3.615, 0.05, 7)
        d = random_model_dict(1, 10.0, 2.0, [400], [8, 16, 32], 4, [16, 16])
    else:
        x, at, box = systems.water(1, 5.0, 4)
        d = random_model_dict(1, 6.0, 0.5, [120, 240], [8, 16, 32], 4, [16, 16])
    m = SeAModel(d, "f64", DEV)
    s = m.system(x[None], at[None], box.reshape(1, 9))
    m.evaluate(s, atomic_virial=True)
    torch.cuda.synchronize()
    fs = s.fs
    print(f"force kernels: {case}, {fs.natoms} atoms, {fs.ncells} cells, "
          f"rng {list(fs.frames_host[0].rng)}", flush=True)
    f_cells, av_cells = s.force.clone(), s.atom_virial.clone()
    v_cells = s.vrow.clone()
    force = torch.zeros_like(s.force)
    av = torch.zeros_like(s.atom_virial)
    _lib.call("dpk_force_virial_fused", _lib.ptr(force), _lib.ptr(av), _lib.ptr(s.nd), m.prec,
              _lib.ptr(s.wrapped), _lib.ptr(s.atype), _lib.ptr(fs.frame_of),
              _lib.ptr(fs.frames_dev), _lib.ptr(s.nlist), _lib.ptr(fs.shift_vec),
              _lib.ptr(m.dstd_dev), fs.natoms, m.nnei, m.rcut, m.rcut_smth, m.env_protection,
              _lib.ptr(s.vrow),
              s.vrow.numel(), None, m.ntypes, _lib.stream_ptr())
    torch.cuda.synchronize()
    assert _rel(f_cells.cpu().numpy(), force.cpu().numpy()) < 1e-12
    assert _rel(av_cells.cpu().numpy(), av.cpu().numpy()) < 1e-12
    assert torch.equal(v_cells, s.vrow)


@pytest.mark.parametrize("prec", ["f64", "bf16"])
def test_every_tunable_value_gives_the_same_result(prec):
    from deepmd_amd import autotune
    frames = _frames(n=4, seed=34)
    _, m = _model(prec, type_one_side=True)
    s = _system(m, frames)
    defaults = autotune.get(m)
    ref = {k: v.clone() for k, v in m.evaluate(s, atomic_virial=True).items()}
    nl_ref = s.nlist.clone()
    try:
        for name, (_, values) in autotune.TUNABLES.items():
            for v in values:
                autotune.set_values(dict(defaults, **{name: v}), m)
                out = m.evaluate(s, atomic_virial=True)
                torch.cuda.synchronize()
                print(f"tunable {name}={v} ({prec})", flush=True)
                assert torch.equal(s.nlist, nl_ref), (name, v)
                for k in ("energy", "force", "virial", "atom_virial"):
                    assert _rel(out[k].cpu().numpy(), ref[k].cpu().numpy()) < 1e-12, (name, v, k)
    finally:
        autotune.set_values(defaults, m)
    res = autotune.autotune(m, s, reps=1)
    assert set(res) == set(autotune.TUNABLES) | {"ms"}
    assert autotune.get(m) == {k: v for k, v in res.items() if k != "ms"}


def test_models_keep_their_own_tunables():
    """Two models with different launch tunables, evaluated alternately:
    each runs with its own set (the process-wide defaults untouched), and
    both give their reference results."""
    from deepmd_amd import autotune
    frames = _frames(n=3, seed=35)
    _, ma = _model("f64", type_one_side=True)
    _, mb = _model("f64", type_one_side=True)
    defaults = autotune.get()
    sa, sb = _system(ma, frames), _system(mb, frames)
    ref = {k: v.clone() for k, v in ma.evaluate(sa).items()}
    autotune.set_values({"nlist_mode": 1, "force_mode": 1, "nlist_sort": 1}, ma)
    assert autotune.get(ma)["nlist_mode"] == 1 and autotune.get(mb)["nlist_mode"] == 0
    for _ in range(2):
        oa = {k: v.clone() for k, v in ma.evaluate(sa).items()}
        ob = {k: v.clone() for k, v in mb.evaluate(sb).items()}
        torch.cuda.synchronize()
        for k in ("energy", "force", "virial"):
            assert _rel(oa[k].cpu().numpy(), ref[k].cpu().numpy()) < 1e-12, k
            assert _rel(ob[k].cpu().numpy(), ref[k].cpu().numpy()) < 1e-12, k
    assert autotune.get() == defaults


def _with_exclusion(d, pairs, seed=3):
    """Exclude type pairs and give the model nonzero env-matrix statistics, so
    padding rows (-avg/std) and excluded rows (zero) differ."""
    from deepmd_amd.model import _nd
    de = d["model"]["descriptor"]
    de["exclude_types"] = [list(p) for p in pairs]
    nt, nnei = len(de["sel"]), int(sum(de["sel"]))
    rng = np.random.default_rng(seed)
    de["@variables"]["davg"] = _nd(rng.normal(0.0, 0.05, size=(nt, nnei, 4)))
    de["@variables"]["dstd"] = _nd(rng.uniform(0.5, 1.5, size=(nt, nnei, 4)))
    return d


@pytest.mark.parametrize("type_one_side", [False, True])
def test_exclude_types_f64_vs_oracle(type_one_side):
    from deepmd_amd.model import SeAModel, random_model_dict
    from oracle import sea
    frames = _frames(n=3, seed=35)
    d = random_model_dict(5, 6.0, 0.5, [60, 60, 60, 60], [16, 32, 64], 16, [32, 32],
                          type_one_side=type_one_side)
    d = _with_exclusion(d, [(0, 1), (2, 2)])
    m = SeAModel(d, "f64", DEV)
    om = sea.model_from_serialized(d)
    assert om.exclude_types
    s = _system(m, frames)
    print(f"exclude_types {m.exclude_types}, type_one_side={type_one_side}", flush=True)
    out = m.evaluate(s, atomic_virial=True)
    torch.cuda.synchronize()
    d0 = _with_exclusion(random_model_dict(5, 6.0, 0.5, [60, 60, 60, 60], [16, 32, 64], 16,
                                           [32, 32], type_one_side=type_one_side), [])
    m0 = SeAModel(d0, "f64", DEV)
    plain = m0.evaluate(_system(m0, frames))
    torch.cuda.synchronize()
    assert _rel(plain["force"].cpu().numpy(), out["force"].cpu().numpy()) > 1e-3
    off = 0
    for f, (x, at, box) in enumerate(frames):
        n = len(at)
        a = torch.tensor(at).view(1, n).long()
        b = torch.tensor(box.ravel()).view(1, 9)
        ref = sea.forward(om, torch.tensor(x).view(1, -1), a, b, do_atomic_virial=True)
        assert _rel(out["energy"][f].item(), ref["energy"].item()) < 1e-10
        assert _rel(out["force"][off:off + n].cpu().numpy(), ref["force"].detach().numpy()) < 1e-10
        assert _rel(out["virial"][f].cpu().numpy(), ref["virial"].detach().numpy()) < 1e-10
        assert _rel(out["atom_virial"][off:off + n].cpu().numpy(),
                    ref["atom_virial"].detach().numpy()) < 1e-10
        off += n


def test_exclude_types_compressed_and_bf16():
    """Exclusion through the tabulated (FP64) and BF16 embedding paths."""
    from deepmd_amd.model import SeAModel, random_model_dict
    frames = _frames(n=3, seed=36)
    d = _with_exclusion(random_model_dict(5, 6.0, 0.5, [60, 60, 60, 60], [16, 32, 64], 16,
                                          [32, 32]), [(1, 3)])
    m = SeAModel(d, "f64", DEV)
    ref = {k: v.clone() for k, v in m.evaluate(_system(m, frames)).items()}
    mc = SeAModel(d, "f64", DEV)
    mc.enable_compression(0.5)
    outc = mc.evaluate(_system(mc, frames))
    mb = SeAModel(d, "bf16", DEV)
    outb = mb.evaluate(_system(mb, frames))
    torch.cuda.synchronize()
    rc = float(torch.sqrt(torch.mean((outc["force"] - ref["force"]) ** 2)))
    rb = float(torch.sqrt(torch.mean((outb["force"] - ref["force"]) ** 2)))
    print(f"exclusion: compressed f64 force rmse {rc:.3e}, bf16 {rb:.3e}", flush=True)
    assert rc <= 1e-5
    assert rb <= 1e-3


def test_edge_frames_vs_oracle():
    """One ragged batch holding a lone atom (no neighbour inside rcut), an
    open-boundary pair, an empty frame and a frame with a virtual atom (type
    -1): each frame equals the oracle (FP64, 1e-10); the empty frame gives
    zero energy and virial."""
    from deepmd_amd.model import SeAModel, random_model_dict
    from oracle import sea
    d = random_model_dict(5, 6.0, 0.5, [60, 60], [16, 32, 64], 16, [32, 32])
    m = SeAModel(d, "f64", DEV)
    om = sea.model_from_serialized(d)
    rng = np.random.default_rng(0)
    x4 = rng.uniform(0, 9, (40, 3))
    at4 = rng.integers(0, 2, 40)
    at4[7] = -1
    frames = [(np.array([[1.0, 2.0, 3.0]]), np.array([0]), np.eye(3).ravel() * 10),
              (np.array([[0.0, 0.0, 0.0], [1.5, 0.3, 0.2]]), np.array([0, 1]), None),
              (np.zeros((0, 3)), np.zeros(0, dtype=int), np.eye(3).ravel() * 8),
              (x4, at4, np.eye(3).ravel() * 9)]
    s = m.system([f[0] for f in frames], [f[1] for f in frames], [f[2] for f in frames],
                 nlocs=[len(f[1]) for f in frames])
    print("edge frames:", [len(f[1]) for f in frames], flush=True)
    out = m.evaluate(s, atomic_virial=True)
    torch.cuda.synchronize()
    off = 0
    for f, (x, at, box) in enumerate(frames):
        n = len(at)
        if n == 0:
            assert out["energy"][f].item() == 0.0
            assert torch.all(out["virial"][f] == 0)
            continue
        a = torch.tensor(at).view(1, n).long()
        b = None if box is None else torch.tensor(box).view(1, 9)
        ref = sea.forward(om, torch.tensor(x).view(1, -1), a, b, do_atomic_virial=True)
        assert _rel(out["energy"][f].item(), ref["energy"].item()) < 1e-10, f
        assert _rel(out["force"][off:off + n].cpu().numpy(),
                    ref["force"].detach().numpy()) < 1e-10, f
        assert _rel(out["virial"][f].cpu().numpy(), ref["virial"].detach().numpy()) < 1e-10, f
        off += n


@pytest.mark.parametrize("prec,type_one_side", [("bf16", False), ("bf16", True),
                                                 ("f32", False)])
def test_padding_tail_table_vs_full_blocks(prec, type_one_side, monkeypatch):
    """The BF16 and FP32 embeddings stop each type block at its last real
    neighbour and take the padding rows (-avg/std, nonzero here) from a
    per-(centre type, slot) table.  Against the same model computing every
    slot, on a ragged batch with excluded pairs (zero rows inside the blocks)
    and virtual atoms (zero rows, no table), and against FP64."""
    from deepmd_amd import model as M
    frames = _frames(n=4, seed=7)
    frames[1][1][:3] = -1  # virtual atoms
    d, _ = _model("f64", type_one_side=type_one_side)
    d = _with_exclusion(d, [(0, 2), (1, 1)])
    mt = M.SeAModel(d, prec, DEV)
    assert mt.emb_pad is not None
    with monkeypatch.context() as mp:
        mp.setattr(M.SeAModel, "SKIP_PADDING", False)
        mf = M.SeAModel(d, prec, DEV)
        assert mf.emb_pad is None
        st, sf = _system(mt, frames), _system(mf, frames)
        out_t = {k: v.clone() for k, v in mt.evaluate(st, atomic_virial=True).items()}
        out_f = mf.evaluate(sf, atomic_virial=True)
    m64 = M.SeAModel(d, "f64", DEV)
    ref = m64.evaluate(_system(m64, frames), atomic_virial=True)
    torch.cuda.synchronize()
    real = (st.nlist >= 0).view(st.nlist.shape[0], -1, 1)
    assert torch.all(st.nd.view(real.shape[0], -1, 4)[~real.expand(-1, -1, 4)] == 0)
    for k in ("energy", "force", "virial", "atom_virial"):
        e = _rel(out_t[k].cpu().numpy(), out_f[k].cpu().numpy())
        print(f"padding table vs full blocks ({prec}, {type_one_side}) {k}: {e:.2e}", flush=True)
        assert e < (2e-2 if prec == "bf16" else 1e-5), k
    rms = float(torch.sqrt(torch.mean(ref["force"] ** 2)))
    et = float(torch.sqrt(torch.mean((out_t["force"] - ref["force"]) ** 2)))
    ef = float(torch.sqrt(torch.mean((out_f["force"] - ref["force"]) ** 2)))
    print(f"force rmse vs f64: table {et:.3e}, full {ef:.3e} (rms {rms:.3e})", flush=True)
    assert et <= 1.5 * ef + 1e-3 * rms


@pytest.mark.parametrize("prec", ["f64", "bf16"])
def test_model_level_exclusions_vs_oracle(prec):
    """Model-level atom_exclude_types / pair_exclude_types
    (base_atomic_model.py:380-399): excluded pairs leave the formatted list
    (padding rows, -avg/std, with nonzero davg), excluded centre types get no
    energy; every frame equals the oracle (FP64 1e-10; BF16 to its force
    bound), and the exclusions change the result."""
    from deepmd_amd.model import SeAModel, random_model_dict
    from oracle import sea
    frames = _frames(n=3, seed=37)
    d = _with_exclusion(random_model_dict(5, 6.0, 0.5, [60, 60, 60, 60], [16, 32, 64], 16,
                                          [32, 32]), [])
    d["model"]["atom_exclude_types"] = [1]
    d["model"]["pair_exclude_types"] = [[0, 2], [3, 3]]
    om = sea.model_from_serialized(d)
    assert om.atom_exclude_types == [1] and om.pair_exclude_types
    m = SeAModel(d, prec, DEV)
    s = _system(m, frames)
    print(f"model-level exclusions, {prec}", flush=True)
    out = m.evaluate(s, atomic_virial=True)
    torch.cuda.synchronize()
    nl = s.nlist.cpu().numpy()
    m0 = SeAModel(_with_exclusion(random_model_dict(5, 6.0, 0.5, [60, 60, 60, 60],
                                                     [16, 32, 64], 16, [32, 32]), []), prec, DEV)
    plain = m0.evaluate(_system(m0, frames))
    torch.cuda.synchronize()
    assert _rel(plain["force"].cpu().numpy(), out["force"].cpu().numpy()) > 1e-3
    off = 0
    for f, (x, at, box) in enumerate(frames):
        n = len(at)
        a = torch.tensor(at).view(1, n).long()
        b = torch.tensor(box.ravel()).view(1, 9)
        ref = sea.forward(om, torch.tensor(x).view(1, -1), a, b, do_atomic_virial=True)
        ae = out["atom_energy"][off:off + n].cpu().numpy()
        assert not ae[at == 1].any()
        # excluded pairs are gone from the list: no (0, 2), (2, 0) or (3, 3) entry
        ext_t = np.tile(at, 27)
        rows = nl[off:off + n]
        tj = np.where(rows >= 0, ext_t[np.maximum(rows, 0) % n], -1)
        ti = at[:, None]
        assert not np.any((rows >= 0) & (((ti == 0) & (tj == 2)) | ((ti == 2) & (tj == 0))
                                         | ((ti == 3) & (tj == 3))))
        if prec == "f64":
            assert _rel(out["energy"][f].item(), ref["energy"].item()) < 1e-10
            assert _rel(out["force"][off:off + n].cpu().numpy(),
                        ref["force"].detach().numpy()) < 1e-10
            assert _rel(out["virial"][f].cpu().numpy(), ref["virial"].detach().numpy()) < 1e-10
            assert _rel(out["atom_virial"][off:off + n].cpu().numpy(),
                        ref["atom_virial"].detach().numpy()) < 1e-10
        else:
            fr = ref["force"].detach().numpy().reshape(-1, 3)
            fg = out["force"][off:off + n].cpu().numpy()
            assert np.sqrt(np.mean((fg - fr) ** 2)) < 0.05 * np.sqrt(np.mean(fr ** 2))
        off += n

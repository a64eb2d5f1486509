"""CPU tests of the product's host side: the C ABI library loads and exports
every entry point of include/dpk.h, frame set-up reproduces the reference's
image order / shift vectors, embedding packing round-trips, and seeded model
dicts load in the oracle."""

import ctypes
import os
import re

import numpy as np
import pytest
import torch

from conftest import ROOT

from deepmd_amd import _lib, frames, model
from oracle import geometry, sea


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "dpk.h")).read()
    return sorted(set(re.findall(r"\b(dpk_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    path = _lib.lib_path()
    if not os.path.exists(path):
        pytest.skip("libdpk.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(path)
    syms = _header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # every exported entry point has a ctypes signature in the binding
    assert set(syms) <= set(_lib.EXPORTS)
    lib = _lib.lib()
    assert lib.dpk_abi_version() == 8
    assert lib.dpk_embed_net_elems(32, 64, 128) == 2 * 32 + 64 * 32 + 2 * 64 + 128 * 64 + 2 * 128


def test_frame_struct_layout():
    # dpk_frame: 24 doubles + 16 ints
    assert ctypes.sizeof(_lib.Frame) == 24 * 8 + 16 * 4


@pytest.mark.parametrize("nb", [(1, 1, 1), (2, 1, 3)])
def test_shift_order_matches_oracle(nb):
    assert torch.equal(frames.shift_list(list(nb)), geometry.image_shifts(list(nb)))


def test_frames_tables():
    box = np.array([[10.0, 0, 0], [1.0, 11.0, 0], [0.5, 0.3, 12.0]])
    fs = frames.make_frames([5, 7], [box, np.diag([6.0, 7.0, 8.0])], 6.0, "cpu")
    f0, f1 = fs.frames_host
    assert f0.pbc == 1 and f0.atom_off == 0 and f1.atom_off == 5
    # shared image count = max over frames (nlist.py:479)
    assert list(f0.nb) == list(f1.nb) == [1, 1, 1]
    sv = fs.shift_vec.numpy()
    bt = torch.tensor(box)
    ref = torch.einsum("sd,fdk->fsk", geometry.image_shifts([1, 1, 1]).to(torch.float64), bt[None])[0]
    assert np.array_equal(sv[:27], ref.numpy())
    # img_pos is the inverse of the shift list
    sh = geometry.image_shifts([1, 1, 1]).numpy().astype(int)
    tab = fs.img_pos.numpy()[:27]
    for p, s in enumerate(sh):
        assert tab[((s[0] + 1) * 3 + s[1] + 1) * 3 + s[2] + 1] == p
    # cell width > rcut and cells cover the frame
    face = frames.to_face_distance(bt).numpy()
    for d in range(3):
        assert face[d] / f0.nc[d] > 6.0
    assert fs.ncells == f0.nc[0] * f0.nc[1] * f0.nc[2] + f1.nc[0] * f1.nc[1] * f1.nc[2]


def _net(rng, dims):
    layers = []
    for i in range(3):
        layers.append(model.Layer(rng.normal(size=(dims[i], dims[i + 1])), rng.normal(size=dims[i + 1]),
                                  None, "tanh", True))
    return layers


@pytest.mark.parametrize("neuron", [[25, 50, 100], [32, 64, 128], [3, 6, 12]])
def test_embedding_pack_roundtrip(neuron):
    rng = np.random.default_rng(0)
    nets = [_net(rng, [1] + neuron) for _ in range(2)]
    packed, ch_out, r2, r3 = model.pack_embedding(nets, neuron)
    assert (r2, r3) == (2, 2)
    p1, p2, p3, _, _ = model.channel_plans(neuron)
    P1, P2, P3 = 32, 64, 128
    for k, n in enumerate(nets):
        v = packed[k]
        w2t = v[2 * P1: 2 * P1 + P2 * P1].reshape(P2, P1)
        np.testing.assert_array_equal(w2t[np.ix_(p2, p1)], n[1].w.T)
        o3 = 2 * P1 + P2 * P1 + 2 * P2
        w3t = v[o3:o3 + P3 * P2].reshape(P3, P2)
        np.testing.assert_array_equal(w3t[np.ix_(p3, p2)], n[2].w.T)
        # padded rows/cols are zero
        assert np.count_nonzero(w3t) == np.count_nonzero(n[2].w)
    # resnet copies stay lane-local: internal channel of (x, x)[q] == plan_in[q % d_in] mod P_in
    assert np.all(p2 % P1 == p1[np.arange(neuron[1]) % neuron[0]])
    assert np.all(p3 % P2 == p2[np.arange(neuron[2]) % neuron[1]])
    assert sorted(ch_out[ch_out >= 0].tolist()) == list(range(neuron[2]))
    # internal MLP with zero padding reproduces the true net
    s = rng.normal(size=7)
    def internal(s, v):
        y1 = np.tanh(np.outer(s, v[:P1]) + v[P1:2 * P1])
        y2 = np.tanh(y1 @ w2t.T + v[2 * P1 + P2 * P1: 2 * P1 + P2 * P1 + P2]) + np.tile(y1, 2)
        G = np.tanh(y2 @ w3t.T + v[o3 + P3 * P2: o3 + P3 * P2 + P3]) + np.tile(y2, 2)
        return G
    G = internal(s, packed[-1])
    x = s[:, None]
    for ly in nets[-1]:
        y = np.tanh(x @ ly.w + ly.b)
        if y.shape[1] == 2 * x.shape[1]:
            y = y + np.tile(x, 2)
        x = y
    np.testing.assert_allclose(G[:, p3], x, rtol=1e-12, atol=1e-12)


def test_random_model_dict_loads_in_oracle():
    d = model.random_model_dict(1, 6.0, 0.5, [46, 92], [25, 50, 100], 16, [240, 240, 240])
    m = sea.model_from_serialized(d)
    assert m.sel == [46, 92] and m.neuron == [25, 50, 100]
    assert len(m.embeddings) == 2 and len(m.fitting) == 2
    assert m.fitting[0][1].idt is not None and m.fitting[0][0].idt is None


def test_tuning_registry():
    # host-only: defaults, set/get round trip, range checks (no kernel launched)
    path = _lib.lib_path()
    if not os.path.exists(path):
        pytest.skip("libdpk.so not built")
    from deepmd_amd import autotune
    L = _lib.lib()
    before = autotune.get()
    assert before == {"nlist_mode": 0, "nlist_stage_pct": 125, "force_mode": 0,
                      "embed_grid_pct": 100, "nlist_waves": 0, "force_waves": 0,
                      "nlist_sort": 0}
    for name, (key, values) in autotune.TUNABLES.items():
        for v in values:
            assert L.dpk_tune_set(key, v) == 0
            assert L.dpk_tune_get(key) == v
    assert L.dpk_tune_set(0, 7) != 0 and b"outside" in L.dpk_last_error()
    assert L.dpk_tune_set(99, 0) != 0
    assert L.dpk_tune_get(99) == -1
    autotune.set_values(before)
    assert autotune.get() == before


def test_tuning_sets_are_per_owner_and_thread_local():
    """dpk_tuning: a set starts from the defaults, holds its own values
    (range-checked), is current only on the thread that made it current, and
    dpk_tuning_use returns the previous set for restoring.  Host-only calls."""
    import threading
    from deepmd_amd import _lib
    L = _lib.lib()
    a, b = _lib.Tuning(), _lib.Tuning()
    d0 = L.dpk_tune_get(1)
    assert a.get(1) == d0 and b.get(1) == d0
    a.set(1, 200)
    assert a.get(1) == 200 and b.get(1) == d0 and L.dpk_tune_get(1) == d0
    with pytest.raises(_lib.DpkError):
        a.set(1, 5)  # below the range
    assert b"outside" in L.dpk_last_error()
    assert L.dpk_tuning_set(None, 1, 100) != 0
    assert L.dpk_tuning_get(None, 1) == d0  # NULL: the default
    seen = {}
    with a.active():
        prev = L.dpk_tuning_use(b.handle)  # nested: b current, a returned
        assert prev == a.handle
        assert L.dpk_tuning_use(prev) == b.handle  # restore a

        def other():
            seen["other"] = L.dpk_tuning_use(None)  # nothing current on a new thread

        t = threading.Thread(target=other)
        t.start()
        t.join()
    assert seen["other"] is None
    assert L.dpk_tuning_use(None) is None  # the scope restored "no set"


def test_runner_library_exports_its_header_symbols():
    """libdpk_deeppot.so (the C++ MD-engine runner) exports every function of
    include/dpk_deeppot.h; loading a file that is not an exported model fails
    with a message, not a crash (before any device work on a machine with a
    GPU; here the device query itself fails first)."""
    from deepmd_amd.build import DEEPPOT_LIB
    if not os.path.exists(DEEPPOT_LIB):
        pytest.skip("libdpk_deeppot.so not built (run __graft_entry__.build())")
    _lib.lib()
    L = ctypes.CDLL(DEEPPOT_LIB)
    txt = open(os.path.join(ROOT, "include", "dpk_deeppot.h")).read()
    syms = sorted(set(re.findall(r"\b(dpk_deeppot_[a-z_]+)\s*\(", txt)))
    assert len(syms) == 6, syms
    assert not [s for s in syms if not hasattr(L, s)]
    L.dpk_deeppot_load.restype = ctypes.c_void_p
    L.dpk_deeppot_load.argtypes = [ctypes.c_char_p, ctypes.c_int]
    L.dpk_deeppot_last_error.restype = ctypes.c_char_p
    assert L.dpk_deeppot_load(b"/nonexistent/model.dpkm", 0) is None
    assert L.dpk_deeppot_last_error()

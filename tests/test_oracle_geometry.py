"""Pins the oracle's geometry (images, neighbour lists, formatting) against the
reference's gtest goldens, brute force and the torch rounding contract."""

from fractions import Fraction

import numpy as np
import pytest
import torch

from oracle import geometry, refcapi, sea


def _fma_norm(v):
    x, y, z = (Fraction(float(t)) for t in v)
    acc = float(x * x)
    acc = float(Fraction(acc) + y * y)
    acc = float(Fraction(acc) + z * z)
    return np.sqrt(acc)


def test_norm_rounding_contract():
    # torch's CPU 3-vector norm == sqrt(fma(z,z,fma(y,y,x*x))), the rounding the
    # HIP distance kernels reproduce; a plain ((x*x+y*y)+z*z) fails this check.
    rng = np.random.default_rng(0)
    d = rng.uniform(-8, 8, size=(3000, 3))
    t = torch.linalg.norm(torch.from_numpy(d.reshape(10, 300, 3)), dim=-1).numpy().ravel()
    f = np.array([_fma_norm(v) for v in d])
    assert np.array_equal(t, f)
    plain = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    assert not np.array_equal(t, plain)


def test_image_shift_order():
    s = geometry.image_shifts([1, 1, 1]).numpy()
    assert s.shape == (27, 3)
    assert np.all(s[0] == 0)
    n = np.linalg.norm(s, axis=1)
    assert np.all(np.diff(n) >= 0)
    assert len({tuple(x) for x in s.astype(int)}) == 27


def _water_box(nmol=64, L=12.44, seed=0):
    rng = np.random.default_rng(seed)
    g = int(round(nmol ** (1 / 3)))
    idx = np.stack(np.meshgrid(*[np.arange(g)] * 3, indexing="ij"), -1).reshape(-1, 3)
    o = (idx + 0.5) * L / g + rng.uniform(-0.3, 0.3, size=(nmol, 3))
    h1 = o + rng.normal(size=(nmol, 3)) * 0.55
    h2 = o + rng.normal(size=(nmol, 3)) * 0.55
    return np.stack([o, h1, h2], 1).reshape(-1, 3), np.tile([0, 1, 1], nmol), L


def test_dense_equals_cell_list():
    xyz, at, L = _water_box()
    c = torch.tensor(xyz).view(1, -1, 3)
    a = torch.tensor(at).view(1, -1)
    box = torch.tensor(np.diag([L] * 3).ravel()).view(1, 9)
    cn = geometry.normalize_coord(c, box.view(1, 3, 3))
    ext, eat, mp = geometry.extend_coord_with_ghosts(cn, a, box, 6.0)
    d = geometry.build_neighbor_list_dense(ext, eat, len(at), 6.0, [46, 92])
    cl = geometry.build_neighbor_list_cells(ext, eat, len(at), 6.0, [46, 92])
    assert torch.equal(d, cl)


def test_nlist_brute_force():
    # every kept index is within rcut, every within-rcut atom is kept (sel is
    # not exceeded here), ordering is ascending distance inside type blocks
    xyz, at, L = _water_box(nmol=27, L=9.3, seed=4)
    c = torch.tensor(xyz).view(1, -1, 3)
    a = torch.tensor(at).view(1, -1)
    box = torch.tensor(np.diag([L] * 3).ravel()).view(1, 9)
    ext, eat, mp, nl = geometry.extend_input_and_build_neighbor_list(c, a, 6.0, [46, 92], box)
    nl = geometry.format_nlist(ext, eat, nl, 6.0, [46, 92]).numpy()[0]
    e = ext.numpy()[0]
    nloc = len(at)
    for i in range(nloc):
        r = np.linalg.norm(e - e[i], axis=1)
        r[i] = np.inf
        for t, (s0, s1) in enumerate([(0, 46), (46, 138)]):
            want = np.nonzero((r <= 6.0) & (eat.numpy()[0] == t))[0]
            want = want[np.lexsort((want, r[want]))]
            got = nl[i, s0:s1]
            got = got[got >= 0]
            assert np.array_equal(got, want)


def test_cc_fmt_nlist_golden(cc_goldens, cc_geometry):
    # The gtest's expected formatted lists (C++ extended layout, FP32 keys) equal
    # the oracle's per-type partition of the distance-sorted list over the same
    # extended coordinates (no distance ties in this fixture).
    ext = torch.tensor(cc_geometry["posi_cpy"], dtype=torch.float64).view(1, -1, 3)
    eat = torch.tensor(cc_geometry["atype_cpy"]).view(1, -1)
    for cls, sel in (("TestFormatNlist", [10, 10]), ("TestFormatNlistShortSel", [2, 2])):
        full = geometry.build_neighbor_list_dense(ext.view(1, -1), eat, 6, 6.0, [200],
                                                  distinguish_types=False)
        got = geometry.nlist_distinguish_types(full, eat, sel).numpy()[0].ravel()
        want = np.array(cc_goldens["test_fmt_nlist.cc"][cls]["expect_nlist_cpy"])
        assert np.array_equal(got, want), cls


def test_cc_env_golden_pt_layout(cc_goldens):
    # expected_env of test_env_mat_a.cc (6-atom water, 13 A box, rc 6, rc_smth
    # 0.8, sel [10, 10]) reproduced end-to-end by the oracle's own images,
    # neighbour list and env matrix (PT layout), to the golden's 5 decimals.
    g = cc_goldens["test_env_mat_a.cc"]["TestEnvMatA"]
    c = torch.tensor(g["posi"], dtype=torch.float64).view(1, -1, 3)
    a = torch.tensor(g["atype"]).view(1, -1)
    box = torch.tensor(np.diag([13.0] * 3).ravel()).view(1, 9)
    ext, eat, mp, nl = geometry.extend_input_and_build_neighbor_list(c, a, 6.0, [10, 10], box)
    nl = geometry.format_nlist(ext, eat, nl, 6.0, [10, 10])
    em, _, _ = sea.make_env_mat(nl, ext, 6.0, 0.8)
    np.testing.assert_allclose(em.numpy().ravel(), g["expected_env"], atol=1e-5)


@pytest.mark.skipif(not refcapi.available(), reason="oracle/_ref not built")
def test_neighbor_sets_vs_compiled_reference():
    # same within-rcut neighbour sets (as coordinates) as the reference's
    # copy_coord_cpu + build_nlist_cpu on a 192-atom water box
    xyz, at, L = _water_box()
    box = np.diag([L] * 3).ravel()
    ext_c, ext_t, mapping = refcapi.copy_coord(xyz, at, 6.0, box)
    raw, nn = refcapi.build_nlist(ext_c, len(at), 6.0)
    c = torch.tensor(xyz).view(1, -1, 3)
    a = torch.tensor(at).view(1, -1)
    ext, eat, mp, nl = geometry.extend_input_and_build_neighbor_list(
        c, a, 6.0, [200], torch.tensor(box).view(1, 9), mixed_types=True)
    e = ext.numpy()[0]
    nl = nl.numpy()[0]
    for i in range(0, len(at), 7):
        ours = {tuple(np.round(e[j] - e[i], 8)) for j in nl[i] if j >= 0}
        ref = {tuple(np.round(ext_c[j] - ext_c[i], 8)) for j in raw[i, :nn[i]]}
        assert ours == ref

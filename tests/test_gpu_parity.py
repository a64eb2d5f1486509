"""GPU parity of the libdpk kernels against the CPU oracle.

Bars: neighbour lists bit-exact; FP64 E / F / V <= 1e-10 relative (the
golden deeppot-testcase values and the oracle); op-level FP64 1e-12;
tabulate 1e-10 (FP64) / 1e-5 (FP32) as the reference's op test; mixed
precision (BF16 / FP32 GEMMs) force RMSE <= 1e-3 eV/A vs the FP64 oracle.
"""

import json

import numpy as np
import pytest
import torch

from conftest import golden_path

pytestmark = pytest.mark.gpu

DEV = "cuda"


@pytest.fixture(scope="module", autouse=True)
def _hip():
    if not torch.cuda.is_available():
        pytest.fail("a HIP device is required for the gpu tests")
    import deepmd_amd.ops  # noqa: F401  registers torch.ops.deepmd


def _oracle():
    from oracle import cops, geometry, sea
    return cops, geometry, sea


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


# ----------------------------------------------------------------- nlist ----

def _canon(nl, coords, atypes, boxes, rcut, sel):
    """Order each type block by (FP64 distance, extended index): the canonical
    tie-break.  torch.topk leaves the order of exactly equal distances
    unspecified (water's two O-H bonds are both 0.9572 A), so parity is
    checked on this canonical form plus the kept sets themselves."""
    _, geometry, _ = _oracle()
    nf = coords.shape[0]
    c = torch.tensor(coords).view(nf, -1, 3)
    a = torch.tensor(atypes).view(nf, -1).to(torch.int64)
    b = None if boxes is None else torch.tensor(boxes).view(nf, 9)
    ext, _, _, _ = geometry.extend_input_and_build_neighbor_list(c, a, rcut, sel, b)
    e = ext.numpy()
    out = nl.copy()
    sec = np.concatenate([[0], np.cumsum(sel)])
    for f in range(nf):
        for i in range(nl.shape[1]):
            for t in range(len(sel)):
                row = nl[f, i, sec[t]:sec[t + 1]]
                k = row[row >= 0]
                r = torch.linalg.norm(torch.tensor(e[f, k] - e[f, i]), dim=-1).numpy()
                out[f, i, sec[t]:sec[t] + len(k)] = k[np.lexsort((k, r))]
    return out


def _oracle_nlist(coords, atypes, boxes, rcut, sel):
    _, geometry, _ = _oracle()
    nf = coords.shape[0]
    c = torch.tensor(coords).view(nf, -1, 3)
    a = torch.tensor(atypes).view(nf, -1).to(torch.int64)
    b = None if boxes is None else torch.tensor(boxes).view(nf, 9)
    ext, eat, mp, nl = geometry.extend_input_and_build_neighbor_list(c, a, rcut, sel, b)
    return geometry.format_nlist(ext, eat, nl, rcut, sel).numpy()


def _gpu_nlist(coords, atypes, boxes, rcut, sel, neuron=(3, 6, 12)):
    from deepmd_amd import model
    d = model.random_model_dict(0, rcut, min(0.5, rcut / 2), sel, list(neuron), 2, [8, 8])
    m = model.SeAModel(d, "f64", DEV)
    s = m.system(coords, atypes, boxes)
    m.build_nlist(s)
    torch.cuda.synchronize()
    return s.nlist.cpu().numpy().reshape(coords.shape[0], coords.shape[1], -1), s, m


def _water(seed, g, L):
    from deepmd_amd import systems
    x, at, box = systems.water(g, L, seed)
    return x[None], at[None], box.reshape(1, 9)


@pytest.mark.parametrize("seed", [0, 1])
def test_nlist_bitexact_water(seed):
    x, at, box = _water(seed, 4, 12.44)
    ref = _oracle_nlist(x, at, box, 6.0, [46, 92])
    got, _, _ = _gpu_nlist(x, at, box, 6.0, [46, 92])
    # GPU order is already canonical; the oracle's (topk) order is canonicalised
    assert np.array_equal(got, _canon(got, x, at, box, 6.0, [46, 92]))
    assert np.array_equal(got, _canon(ref, x, at, box, 6.0, [46, 92]))


def test_nlist_bitexact_testcase_blocks(sea_testcase):
    for r in sea_testcase["results"]:
        x = np.array(r["coord"], dtype=np.float64).reshape(1, -1, 3)
        at = np.array(r["atype"]).reshape(1, -1)
        box = None if r["box"] is None else np.array(r["box"]).reshape(1, 9)
        ref = _oracle_nlist(x, at, box, 6.0, [46, 92])
        got, _, _ = _gpu_nlist(x, at, box, 6.0, [46, 92])
        assert np.array_equal(got, ref)


def test_nlist_bitexact_triclinic_multiframe_truncation():
    rng = np.random.default_rng(5)
    nf, n = 3, 60
    boxes, xs = [], []
    for f in range(nf):
        b = np.array([[9.0, 0, 0], [1.3, 9.5, 0], [-0.7, 0.9, 10.1]]) * (1 + 0.02 * f)
        xs.append(rng.uniform(0, 1, size=(n, 3)) @ b + rng.normal(0, 2.0, size=(n, 3)))
        boxes.append(b.ravel())
    x = np.stack(xs)
    at = rng.integers(0, 3, size=(nf, n))
    at[:, -3:] = -1  # virtual atoms: never neighbours, empty rows
    sel = [6, 9, 4]  # small sel: exercises the nnei-nearest truncation
    ref = _oracle_nlist(x, at, np.stack(boxes), 4.5, sel)
    got, _, _ = _gpu_nlist(x, at, np.stack(boxes), 4.5, sel)
    assert np.array_equal(got, ref)


def test_nlist_small_box_two_images():
    # face distance < rcut: nbuff = 2 images per direction
    rng = np.random.default_rng(2)
    L = 5.0
    x = rng.uniform(0, L, size=(1, 12, 3))
    at = rng.integers(0, 2, size=(1, 12))
    box = np.diag([L, L, L]).reshape(1, 9)
    ref = _oracle_nlist(x, at, box, 6.0, [60, 60])
    got, _, _ = _gpu_nlist(x, at, box, 6.0, [60, 60])
    # a centre's own images at +-L are exactly equidistant: compare canonically
    assert np.array_equal(got, _canon(got, x, at, box, 6.0, [60, 60]))
    assert np.array_equal(got, _canon(ref, x, at, box, 6.0, [60, 60]))


@pytest.mark.parametrize("sort_mode", [0, 1])
def test_nlist_near_equal_and_equal_distances(sort_mode):
    """The default search sorts one 64-bit key per candidate (the distance's
    bits with the low ones replaced by the extended index) and hands rows
    with neighbours of equal distance part to the full-key sort.  Case 1:
    two neighbours 16 ulp apart in distance (3.0 and 3.0 + 16 ulp, inside
    the 9 index bits of a 500-atom open frame), the farther one with the
    smaller index, so the packed order alone would be wrong.  Case 2: a
    perfect FCC lattice (exact ties in every shell).  Both sort modes must
    give the canonical (distance, index) order and the oracle's sets."""
    from deepmd_amd import autotune, systems
    defaults = autotune.get()
    rng = np.random.default_rng(9)
    n = 500
    x = rng.uniform(8.0, 24.0, size=(n, 3))
    x[0] = (0.0, 0.0, 0.0)
    x[1] = (0.0, 3.0 + 16 * np.spacing(3.0), 0.0)  # farther, smaller index
    x[2] = (3.0, 0.0, 0.0)
    at = np.zeros(n, dtype=np.int32)
    xf, af, bf = systems.fcc(4, 3.615, 0.0, 0)
    cases = [(x[None], at[None], None, [120]),
             (xf[None], af[None], bf.reshape(1, 9), [90])]
    try:
        autotune.set_values(dict(defaults, nlist_sort=sort_mode))
        for ci, (cx, ca, cb, sel) in enumerate(cases):
            print(f"nlist sort mode {sort_mode}, case {ci}", flush=True)
            ref = _oracle_nlist(cx, ca, cb, 6.0, sel)
            got, _, _ = _gpu_nlist(cx, ca, cb, 6.0, sel)
            assert np.array_equal(got, _canon(got, cx, ca, cb, 6.0, sel))
            assert np.array_equal(got, _canon(ref, cx, ca, cb, 6.0, sel))
        got, _, _ = _gpu_nlist(cases[0][0], cases[0][1], None, 6.0, [120])
        # the centre's only neighbours: 3.0 before 3.0 + 16 ulp
        assert list(got[0, 0, :3]) == [2, 1, -1], got[0, 0, :3]
    finally:
        autotune.set_values(defaults)


def test_nlist_wrapped_coords_match_oracle():
    _, geometry, _ = _oracle()
    x, at, box = _water(3, 4, 12.44)
    x = x + np.array([13.0, -7.5, 30.2])  # far outside the cell
    got, s, _ = _gpu_nlist(x, at, box, 6.0, [46, 92])
    cn = geometry.normalize_coord(torch.tensor(x).view(1, -1, 3), torch.tensor(box).view(1, 3, 3))
    assert np.array_equal(s.wrapped.cpu().numpy(), cn.numpy()[0])


# ------------------------------------------------------------- op forms -----

def _op_system(seed=7, nloc=40, L=9.0, sel=(12, 24), rc=4.5):
    _, geometry, sea = _oracle()
    rng = np.random.default_rng(seed)
    c = torch.tensor(rng.uniform(0, L, size=(nloc, 3))).view(1, -1, 3)
    a = torch.tensor(rng.integers(0, 2, size=nloc)).view(1, -1)
    box = torch.tensor(np.diag([L] * 3).ravel()).view(1, 9)
    ext, eat, mp, nl = geometry.extend_input_and_build_neighbor_list(c, a, rc, list(sel), box)
    nl = geometry.format_nlist(ext, eat, nl, rc, list(sel))
    nnei = sum(sel)
    avg = rng.normal(0, 0.1, size=(2, nnei * 4))
    std = rng.uniform(0.5, 1.5, size=(2, nnei * 4))
    return ext.numpy(), eat.numpy(), nl.numpy(), avg, std, rc, 1.0, rng


def test_prod_env_mat_a_op():
    cops, _, _ = _oracle()
    ext, eat, nl, avg, std, rc, rcs, _ = _op_system()
    em, ed, rij = torch.ops.deepmd.prod_env_mat_a(
        torch.tensor(ext, device=DEV).view(1, -1), torch.tensor(eat, device=DEV),
        torch.tensor(nl, device=DEV), torch.tensor(avg, device=DEV), torch.tensor(std, device=DEV),
        rc, rcs)
    rem, red, rrij = cops.prod_env_mat_a(ext, eat, nl, avg, std, rc, rcs)
    np.testing.assert_allclose(em.cpu().numpy(), rem, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(ed.cpu().numpy(), red, rtol=1e-12, atol=1e-13)
    np.testing.assert_array_equal(rij.cpu().numpy(), rrij)


def test_prod_force_virial_ops_and_grads():
    cops, _, _ = _oracle()
    ext, eat, nl, avg, std, rc, rcs, rng = _op_system(8)
    nloc, nall = nl.shape[1], ext.shape[1]
    em, ed, rij = cops.prod_env_mat_a(ext, eat, nl, avg, std, rc, rcs)
    g = rng.normal(size=em.shape)
    td = lambda a: torch.tensor(a, device=DEV)  # noqa: E731
    f = torch.ops.deepmd.prod_force_se_a(td(g), td(ed), td(nl), nloc, nall)
    np.testing.assert_allclose(f.cpu().numpy().reshape(1, nall, 3), cops.prod_force_a(g, ed, nl, nall),
                               rtol=1e-12, atol=1e-12)
    v, av = torch.ops.deepmd.prod_virial_se_a(td(g), td(ed), td(rij), td(nl), nloc, nall)
    rv, rav = cops.prod_virial_a(g, ed, rij, nl, nall)
    np.testing.assert_allclose(v.cpu().numpy(), rv, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(av.cpu().numpy().reshape(rav.shape), rav, rtol=1e-11, atol=1e-11)
    gr = rng.normal(size=(1, nloc * 3))
    gn = torch.ops.deepmd.prod_force_se_a_grad(td(gr), td(g), td(ed), td(nl), nloc)
    np.testing.assert_allclose(gn.cpu().numpy().reshape(1, nloc, -1),
                               cops.prod_force_grad_a(gr.reshape(1, nloc, 3), ed, nl),
                               rtol=1e-12, atol=1e-12)
    gv = rng.normal(size=(1, 9))
    gn = torch.ops.deepmd.prod_virial_se_a_grad(td(gv), td(g), td(ed), td(rij), td(nl), nloc)
    np.testing.assert_allclose(gn.cpu().numpy().reshape(1, nloc, -1),
                               cops.prod_virial_grad_a(gv, ed, rij, nl), rtol=1e-12, atol=1e-12)
    # autograd: d <w, force> / d net_deriv is the exact adjoint (extended rows)
    gt = td(g).requires_grad_(True)
    w = rng.normal(size=(1, nall * 3))
    (torch.ops.deepmd.prod_force_se_a(gt, td(ed), td(nl), nloc, nall) * td(w)).sum().backward()
    lhs = np.sum(w.reshape(1, nall, 3) * cops.prod_force_a(g, ed, nl, nall))
    np.testing.assert_allclose(np.sum(gt.grad.cpu().numpy() * g), lhs, rtol=1e-10)


def test_prod_env_mat_a_autograd_is_negative_force():
    cops, _, _ = _oracle()
    ext, eat, nl, avg, std, rc, rcs, rng = _op_system(9)
    nall = ext.shape[1]
    c = torch.tensor(ext, device=DEV).view(1, -1).requires_grad_(True)
    em, ed, rij = torch.ops.deepmd.prod_env_mat_a(
        c, torch.tensor(eat, device=DEV), torch.tensor(nl, device=DEV),
        torch.tensor(avg, device=DEV), torch.tensor(std, device=DEV), rc, rcs)
    g = rng.normal(size=em.shape)
    (em * torch.tensor(g, device=DEV)).sum().backward()
    _, red, _ = cops.prod_env_mat_a(ext, eat, nl, avg, std, rc, rcs)
    f = cops.prod_force_a(g, red, nl, nall)
    np.testing.assert_allclose(c.grad.cpu().numpy().reshape(f.shape), -f, rtol=1e-11, atol=1e-12)


# -------------------------------------------------------------- tabulate ----

@pytest.mark.parametrize("dtype,prec", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_tabulate_op_reference_golden(dtype, prec):
    with open(golden_path("pt_tabulate_op.json")) as fh:
        p = json.load(fh)
    table = torch.tensor(p["table_tensor"]["value"], dtype=dtype, device=DEV).view(-1, 48)
    info = torch.tensor(p["table_info_tensor"]["value"], dtype=dtype)
    emx = torch.tensor(p["em_x_tensor"]["value"], dtype=dtype, device=DEV).view(16, 1).requires_grad_(True)
    em = torch.tensor(p["em_tensor"]["value"], dtype=dtype, device=DEV).view(4, 4, 4).requires_grad_(True)
    out = torch.ops.deepmd.tabulate_fusion_se_a(table, info, emx, em, 8)[0]
    exp = torch.tensor(p["expected_descriptor_tensor"]["value"], dtype=dtype).view(out.shape)
    torch.testing.assert_close(out.detach().cpu(), exp, atol=prec, rtol=prec)
    out.sum().backward()
    torch.testing.assert_close(emx.grad.cpu().view(-1),
                               torch.tensor(p["expected_dy_dem_x"]["value"], dtype=dtype).view(-1),
                               atol=prec, rtol=prec)
    torch.testing.assert_close(em.grad.cpu().view(-1),
                               torch.tensor(p["expected_dy_dem"]["value"], dtype=dtype).view(-1),
                               atol=prec, rtol=prec)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_tabulate_ops_vs_oracle_with_padding(dtype):
    cops, _, _ = _oracle()
    rng = np.random.default_rng(3)
    nloc, nnei, L = 6, 20, 32
    table = (rng.normal(size=(40, 6 * L)) * 0.1).astype(dtype)
    info = np.array([0.0, 0.2, 2.0, 0.01, 0.1, -1], dtype=dtype)
    emx = rng.uniform(-0.1, 2.3, size=(nloc, nnei)).astype(dtype)
    em = rng.normal(size=(nloc, nnei, 4)).astype(dtype)
    for i in range(nloc):
        k = 3 * i
        if k:
            emx[i, -k:] = dtype(-0.05)
            em[i, -k:, 1:] = 0
    tol = 1e-10 if dtype == np.float64 else 2e-5
    td = lambda a: torch.tensor(a, device=DEV)  # noqa: E731
    out = torch.ops.deepmd.tabulate_fusion_se_a(td(table), torch.tensor(info), td(emx), td(em), L)[0]
    np.testing.assert_allclose(out.cpu().numpy(), cops.tabulate_fusion_se_a(table, info, emx, em, L),
                               rtol=tol, atol=tol)
    dy = rng.normal(size=out.shape).astype(dtype)
    dx, dem = torch.ops.deepmd.tabulate_fusion_se_a_grad(td(table), torch.tensor(info), td(emx),
                                                         td(em), td(dy), out)
    rdx, rdem = cops.tabulate_fusion_se_a_grad(table, info, emx, em, dy)
    np.testing.assert_allclose(dx.cpu().numpy(), rdx, rtol=tol, atol=tol)
    np.testing.assert_allclose(dem.cpu().numpy(), rdem, rtol=tol, atol=tol)
    zx = rng.normal(size=emx.shape).astype(dtype)
    zem = rng.normal(size=em.shape).astype(dtype)
    gg = torch.ops.deepmd.tabulate_fusion_se_a_grad_grad(td(table), torch.tensor(info), td(emx),
                                                         td(em), td(zx), td(zem), out)
    np.testing.assert_allclose(gg.cpu().numpy(),
                               cops.tabulate_fusion_se_a_grad_grad(table, info, emx, em, zx, zem),
                               rtol=tol, atol=tol)


def test_tabulate_double_backward_fd():
    # second-order: d/d em of <dy, grad> matches central differences (FP64)
    rng = np.random.default_rng(4)
    nloc, nnei, L = 3, 8, 16
    table = torch.tensor(rng.normal(size=(40, 6 * L)) * 0.1, device=DEV)
    info = torch.tensor([0.0, 0.2, 2.0, 0.01, 0.1, -1], dtype=torch.float64)
    emx = torch.tensor(rng.uniform(0.05, 1.5, size=(nloc, nnei)), device=DEV).requires_grad_(True)
    em = torch.tensor(rng.normal(size=(nloc, nnei, 4)), device=DEV).requires_grad_(True)
    w = torch.tensor(rng.normal(size=(nloc, 4, L)), device=DEV)
    out = torch.ops.deepmd.tabulate_fusion_se_a(table, info, emx, em, L)[0]
    gx, gem = torch.autograd.grad((out * w).sum(), (emx, em), create_graph=True)
    u = torch.tensor(rng.normal(size=(nloc, nnei, 4)), device=DEV)
    s = (gem * u).sum()
    (d2,) = torch.autograd.grad(s, em)
    h = 1e-6

    def fgem(e):
        e = e.detach().requires_grad_(True)
        o = torch.ops.deepmd.tabulate_fusion_se_a(table, info, emx.detach(), e, L)[0]
        return torch.autograd.grad((o * w).sum(), e)[0]
    v = torch.tensor(rng.normal(size=(nloc, nnei, 4)), device=DEV)
    fd = ((fgem(em.detach() + h * v) - fgem(em.detach() - h * v)) * u).sum() / (2 * h)
    np.testing.assert_allclose((d2 * v).sum().item(), fd.item(), rtol=1e-5, atol=1e-9)


# ----------------------------------------------------------------- model ----

@pytest.mark.parametrize("block", [0, 1, 2])
def test_model_f64_golden_testcase(sea_yaml, sea_testcase, block):
    from deepmd_amd.deep_pot import DeepPot
    dp = DeepPot(sea_yaml, precision="f64")
    r = sea_testcase["results"][block]
    x = np.array(r["coord"], dtype=np.float64).reshape(1, -1, 3)
    box = None if r["box"] is None else np.array(r["box"]).reshape(1, 9)
    e, f, v, ae, av = dp.eval(x, box, np.array(r["atype"]), atomic=True)
    assert _rel(ae, r["atomic_energy"]) < 1e-10
    assert _rel(f, r["force"]) < 1e-10
    assert _rel(av, r["atomic_virial"]) < 1e-10
    assert _rel(e, [np.sum(r["atomic_energy"])]) < 1e-12
    assert _rel(v, np.array(r["atomic_virial"]).reshape(-1, 9).sum(0)) < 1e-10


def test_model_f64_from_pt_checkpoint(sea_yaml, sea_testcase, tmp_path):
    # the reference model written as a PyTorch training checkpoint, loaded by
    # DeepPot("...pt") and evaluated on the GPU against the goldens
    from test_checkpoint_host import _model_params, _state_dict
    from deepmd_amd.deep_pot import DeepPot
    m = sea_yaml["model"] if "model" in sea_yaml else sea_yaml
    sd = _state_dict(m)
    sd["_extra_state"] = {"model_params": _model_params(m), "train_infos": {}}
    path = str(tmp_path / "model.pt")
    torch.save({"model": sd}, path)
    dp = DeepPot(path, precision="f64")
    for r in sea_testcase["results"]:
        x = np.array(r["coord"], dtype=np.float64).reshape(1, -1, 3)
        box = None if r["box"] is None else np.array(r["box"]).reshape(1, 9)
        e, f, v, ae, av = dp.eval(x, box, np.array(r["atype"]), atomic=True)
        assert _rel(ae, r["atomic_energy"]) < 1e-10
        assert _rel(f, r["force"]) < 1e-10
        assert _rel(av, r["atomic_virial"]) < 1e-10


def _c0_model_and_system(seed_geom=0):
    from deepmd_amd import model, systems
    d = model.random_model_dict(1, 6.0, 0.5, [46, 92], [25, 50, 100], 16, [240, 240, 240])
    x, at, box = systems.water(4, 12.44, seed_geom)
    return d, x, at, box


def _oracle_efv(d, x, at, box):
    _, _, sea = _oracle()
    m = sea.model_from_serialized(d)
    out = sea.forward(m, torch.tensor(x).view(1, -1), torch.tensor(at).view(1, -1).long(),
                      torch.tensor(box.ravel()).view(1, 9))
    return (out["energy"].detach().numpy().ravel(), out["force"].detach().numpy().reshape(-1, 3),
            out["virial"].detach().numpy().ravel(), out["atom_energy"].detach().numpy().ravel())


def test_model_f64_c0_vs_oracle():
    from deepmd_amd.model import SeAModel
    d, x, at, box = _c0_model_and_system()
    e0, f0, v0, ae0 = _oracle_efv(d, x, at, box)
    m = SeAModel(d, "f64", DEV)
    s = m.system(x[None], at[None], box.reshape(1, 9))
    out = m.evaluate(s)
    assert _rel(out["energy"].cpu().numpy(), e0) < 1e-10
    assert _rel(out["atom_energy"].cpu().numpy(), ae0) < 1e-10
    assert _rel(out["force"].cpu().numpy(), f0) < 1e-10
    assert _rel(out["virial"].cpu().numpy(), v0) < 1e-10


@pytest.mark.parametrize("prec", ["f32", "bf16"])
def test_model_mixed_c0_force_rmse(prec):
    from deepmd_amd.model import SeAModel
    d, x, at, box = _c0_model_and_system(1)
    e0, f0, v0, _ = _oracle_efv(d, x, at, box)
    m = SeAModel(d, prec, DEV)
    s = m.system(x[None], at[None], box.reshape(1, 9))
    out = m.evaluate(s)
    f = out["force"].cpu().numpy()
    rmse = float(np.sqrt(np.mean((f - f0) ** 2)))
    print(f"{prec}: force rmse {rmse:.3e} eV/A (|F| rms {np.sqrt(np.mean(f0 ** 2)):.3e}), "
          f"dE/atom {abs(out['energy'].item() - e0[0]) / len(at):.3e}")
    assert rmse <= 1e-3
    rel = np.sqrt(np.mean((f - f0) ** 2) / np.mean(f0 ** 2))
    assert rel < (2e-5 if prec == "f32" else 5e-2)
    assert abs(out["energy"].item() - e0[0]) / len(at) < (1e-6 if prec == "f32" else 1e-3)


@pytest.mark.parametrize("prec,tol", [("f64", 1e-10), ("f32", 2e-5), ("bf16", 5e-2)])
def test_model_stages_vs_oracle(prec, tol):
    # every stage of the fused path against the oracle on a tie-free random
    # system: env matrix, xyz = R^T G / nnei, descriptor, dE/d em, forces
    _, geometry, sea = _oracle()
    from deepmd_amd.model import SeAModel, random_model_dict
    rng = np.random.default_rng(11)
    n, L, sel = 48, 10.0, [30, 60]
    x = rng.uniform(0, L, size=(n, 3))
    at = rng.integers(0, 2, size=n)
    box = np.diag([L] * 3)
    d = random_model_dict(3, 5.0, 1.0, sel, [25, 50, 100], 8, [32, 32])
    om = sea.model_from_serialized(d)
    c = torch.tensor(x).view(1, -1, 3)
    a = torch.tensor(at).view(1, -1)
    b = torch.tensor(box.ravel()).view(1, 9)
    ext, eat, mp, nl = geometry.extend_input_and_build_neighbor_list(c, a, 5.0, sel, b)
    nl = geometry.format_nlist(ext, eat, nl, 5.0, sel)
    dm, _, _ = sea.prod_env_mat(ext, nl, eat[:, :n], om.davg, om.dstd, 5.0, 1.0)
    dm = dm.view(n, -1, 4).detach().requires_grad_(True)
    desc, xyz = sea.descriptor_from_env(om, dm, eat[:, :n], torch.ones(n, sum(sel), dtype=torch.int32))
    e = sea.fitting_energy(om, desc.view(1, n, -1), a)
    ref_nd, = torch.autograd.grad(e.sum(), dm)
    m = SeAModel(d, prec, DEV)
    s = m.system(x[None], at[None], box.reshape(1, 9))
    out = m.evaluate(s)
    torch.cuda.synchronize()
    assert np.array_equal(s.nlist.cpu().numpy(), nl.numpy()[0])
    em_tol = 1e-13 if prec == "f64" else 1e-6
    assert _rel(s.em.cpu().numpy(), dm.detach().numpy()) < em_tol
    assert _rel(s.xyz.cpu().numpy(), xyz.detach().numpy()) < tol
    # the fitting sees the descriptor in slot order (atoms grouped by type)
    order = s.order.long().cpu().numpy()
    assert sorted(order.tolist()) == list(range(n))
    assert np.all(np.diff(at[order]) >= 0)
    d_slots = s.fit["D"].double().cpu().numpy()
    assert _rel(d_slots, desc.detach().numpy().reshape(n, -1)[order]) < tol
    # dE/d em on the real neighbours; the BF16 and FP32 paths stop each type
    # block at its last real neighbour and leave the padding slots (which
    # move no atom) at zero, the FP64 path evaluates them as the reference
    real = s.nlist.cpu().numpy().reshape(n, -1) >= 0
    nd = s.nd.cpu().numpy().reshape(n, -1, 4)
    rnd = ref_nd.numpy().reshape(n, -1, 4)
    assert _rel(nd[real], rnd[real]) < tol * 10
    if prec in ("bf16", "f32"):
        assert not np.any(nd[~real])
    else:
        assert _rel(nd, rnd) < tol * 10
    ref = sea.forward(om, torch.tensor(x).view(1, -1), a.long(), b)
    assert _rel(out["force"].cpu().numpy(), ref["force"].detach().numpy()) < tol * 10

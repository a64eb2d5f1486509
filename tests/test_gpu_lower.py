"""forward_lower on a caller's extended region and neighbour list (GPU).

The MD-engine entry (DeepPotPT::compute with an InputNlist,
source/api_cc/src/DeepPotPT.cc:171-378) hands the model extended coordinates
and a neighbour list built with rcut + skin in any order.  Checked here:
dpk_nlist_format against the oracle's format_nlist (pad branch: exact, input
order kept; sort branch and extra_nlist_sort: canonical ties), forward_lower
against the oracle's forward_lower (FP64, <= 1e-10 relative, extended force
and extended atomic virial), the over-cut / reversed list cases of
source/tests/pt/model/test_format_nlist_overcut.py, agreement with the
DeepPot path after the ghost -> local reduction, and the InputNlist front end
with NULL-type atoms.
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


def _water_model(type_one_side=True, seed=1):
    from deepmd_amd.model import random_model_dict
    return random_model_dict(seed, 6.0, 0.5, [46, 92], [25, 50, 100], 16, [32, 32],
                             type_one_side=type_one_side)


def _extended(rcut_build, sel, seed=3, nside=3):
    """Extended region + a type-mixed nlist of width sum(sel) built with
    ``rcut_build`` by the oracle (as an MD engine would, cutoff + skin)."""
    from deepmd_amd import systems
    from oracle import geometry
    x, at, box = systems.water(nside, 12.44 * nside / 4, seed)
    n = len(at)
    ec, ea, mp, nl = geometry.extend_input_and_build_neighbor_list(
        torch.tensor(x).view(1, -1), torch.tensor(at).view(1, -1).long(), rcut_build, sel,
        box=torch.tensor(box.ravel()).view(1, 9))
    return ec, ea, mp, nl, n


def _canon(nl, ext, sel):
    """Ties (equal distance) in index order, per type block."""
    out = nl.copy()
    sec = np.concatenate([[0], np.cumsum(sel)])
    for i in range(nl.shape[0]):
        for t in range(len(sel)):
            row = nl[i, sec[t]:sec[t + 1]]
            k = row[row >= 0]
            r = torch.linalg.norm(torch.tensor(ext[k] - ext[i]), dim=-1).numpy()
            out[i, sec[t]:sec[t] + len(k)] = k[np.lexsort((k, r))]
    return out


def _format_gpu(m, ec, ea, nl, extra_sort):
    from deepmd_amd import _lib
    nf, nloc, nnei_in = nl.shape
    nall = ea.shape[1]
    out = torch.empty((nloc, m.nnei), dtype=torch.int32, device=DEV)
    status = torch.zeros(4, dtype=torch.int32, device=DEV)
    nl_d = nl.to(torch.int32).contiguous().to(DEV)
    ec_d = ec.reshape(-1).contiguous().to(DEV)
    ea_d = ea.reshape(-1).to(torch.int32).contiguous().to(DEV)
    _lib.call("dpk_nlist_format", _lib.ptr(out), _lib.ptr(status), _lib.ptr(nl_d), nnei_in,
              _lib.ptr(ec_d), _lib.ptr(ea_d), 1, nloc, nall, m.rcut, m.sel_host[1], m.ntypes,
              int(extra_sort), _lib.stream_ptr())
    torch.cuda.synchronize()
    assert int(status[1]) == 0
    return out.cpu().numpy()


@pytest.mark.parametrize("case", ["pad_overcut", "pad_reversed", "sort_wide", "extra_sort",
                                  "sort_wide_400", "sort_wide_900"])
def test_nlist_format_vs_oracle(case):
    from deepmd_amd.model import SeAModel
    from oracle import geometry
    m = SeAModel(_water_model(), "f64", DEV)
    ec, ea, mp, nl, n = _extended(m.rcut + 1.0, m.sel)
    if case == "pad_reversed":
        nl = torch.flip(nl, dims=[-1])
    if case.startswith("sort_wide"):
        # shuffled rows padded with -1 to 145 (256-entry kernel), 400 (512,
        # register sort) and 900 (1024, LDS sort) entries
        width = 145 if case == "sort_wide" else int(case.rsplit("_", 1)[1])
        perm = torch.randperm(nl.shape[-1], generator=torch.Generator().manual_seed(0))
        pad = -torch.ones(nl.shape[:-1] + (width - nl.shape[-1],), dtype=nl.dtype)
        nl = torch.cat([nl[..., perm], pad], dim=-1)
    extra = case == "extra_sort"
    print(f"format {case}: {n} centres, rows {nl.shape[-1]} wide", flush=True)
    mine = _format_gpu(m, ec, ea, nl, extra)
    ref = geometry.format_nlist(ec, ea, nl, m.rcut, m.sel, extra_nlist_sort=extra).numpy()[0]
    if case.startswith("pad"):
        assert np.array_equal(mine, ref)      # input order kept: no ties to resolve
    else:
        e = ec.view(-1, 3).numpy()
        assert np.array_equal(_canon(mine, e, m.sel), _canon(ref, e, m.sel))


@pytest.mark.parametrize("type_one_side", [True, False])
def test_forward_lower_f64_vs_oracle(type_one_side):
    from deepmd_amd.model import SeAModel
    from oracle import sea
    d = _water_model(type_one_side)
    m = SeAModel(d, "f64", DEV)
    om = sea.model_from_serialized(d)
    ec, ea, mp, nl, n = _extended(m.rcut + 1.0, m.sel, seed=4)
    nl = torch.flip(nl, dims=[-1])
    out = m.forward_lower(ec, ea, nl, mp, do_atomic_virial=True)
    torch.cuda.synchronize()
    ref = sea.forward_lower(om, ec, ea, nl, mp, do_atomic_virial=True)
    assert _rel(out["energy"].cpu(), ref["energy"].detach()) < 1e-10
    assert _rel(out["atom_energy"].cpu(), ref["atom_energy"].detach()) < 1e-10
    assert _rel(out["extended_force"].cpu(), ref["extended_force"].detach()) < 1e-10
    assert _rel(out["virial"].cpu(), ref["virial"].detach()) < 1e-10
    assert _rel(out["extended_virial"].cpu(), ref["extended_virial"].detach()) < 1e-10


@pytest.mark.parametrize("reverse", [False, True])
def test_forward_lower_overcut_matches_canonical(reverse):
    """test_format_nlist_overcut.py on the HIP path: an rcut + 2 list, as it is
    and reversed, gives the rcut-bounded energy and force (1e-10)."""
    from deepmd_amd.model import SeAModel, random_model_dict
    from oracle import geometry, sea
    d = random_model_dict(1, 4.0, 0.5, [46, 92, 4], [25, 50, 100], 16, [24, 24, 24])
    m = SeAModel(d, "f64", DEV)
    g = torch.Generator().manual_seed(20)
    coord = 5.5 * torch.rand([6, 3], dtype=torch.float64, generator=g)
    atype = torch.tensor([0, 0, 1, 1, 2, 2])
    cell = 6.0 * torch.eye(3, dtype=torch.float64)

    def ef(rb, rev):
        ec, ea, mp, nl = geometry.extend_input_and_build_neighbor_list(
            coord.view(1, -1), atype.view(1, -1), rb, m.sel, box=cell.view(1, 9))
        if rev:
            nl = torch.flip(nl, dims=[-1])
        out = m.forward_lower(ec, ea, nl, mp)
        torch.cuda.synchronize()
        f = sea.communicate(out["extended_force"].cpu().view(1, -1, 3), mp, 6)
        return out["energy"].item(), f.numpy()

    e0, f0 = ef(m.rcut, False)
    e1, f1 = ef(m.rcut + 2.0, reverse)
    assert abs(e1 - e0) <= 1e-10 * max(1.0, abs(e0))
    np.testing.assert_allclose(f1, f0, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("prec", ["f64", "bf16"])
def test_forward_lower_matches_deeppot_path(prec):
    """Ghost -> local reduction of forward_lower's extended force equals the
    DeepPot path (implicit images) on the same frame."""
    from deepmd_amd.model import SeAModel
    from oracle import sea
    m = SeAModel(_water_model(), prec, DEV)
    ec, ea, mp, nl, n = _extended(m.rcut + 0.5, m.sel, seed=5)
    low = m.forward_lower(ec, ea, nl, mp, do_atomic_virial=True)
    x = ec.view(-1, 3)[:n].numpy()
    from deepmd_amd import systems
    _, at, box = systems.water(3, 12.44 * 3 / 4, 5)
    s = m.system(x[None], at[None], box.reshape(1, 9))
    out = m.evaluate(s, atomic_virial=True)
    torch.cuda.synchronize()
    f = sea.communicate(low["extended_force"].cpu().view(1, -1, 3), mp, n).numpy()[0]
    av = sea.communicate(low["extended_virial"].cpu().view(1, -1, 9), mp, n).numpy()[0]
    tol = 1e-10 if prec == "f64" else 1e-5
    assert _rel(low["energy"].cpu(), out["energy"].cpu()) < tol
    assert _rel(f, out["force"].cpu()) < tol
    assert _rel(low["virial"].cpu(), out["virial"].cpu()) < tol
    assert _rel(av, out["atom_virial"].cpu()) < tol


def test_forward_lower_compressed_extra_sort():
    from deepmd_amd.model import SeAModel
    d = _water_model()
    m = SeAModel(d, "f64", DEV)
    mc = SeAModel(d, "f64", DEV)
    mc.enable_compression(0.5)
    ec, ea, mp, nl, n = _extended(m.rcut + 1.0, m.sel, seed=6)
    ref = m.forward_lower(ec, ea, nl, mp)
    out = mc.forward_lower(ec, ea, torch.flip(nl, dims=[-1]), mp, extra_nlist_sort=True)
    torch.cuda.synchronize()
    assert _rel(out["energy"].cpu(), ref["energy"].cpu()) < 1e-9
    assert _rel(out["extended_force"].cpu(), ref["extended_force"].cpu()) < 1e-7


def test_compute_with_input_nlist_and_null_atoms():
    """DeepPot.compute_lower (DeepPotPT::compute with an InputNlist): CSR
    neighbour rows, NULL-type atoms (type < 0) removed before the model and
    zero in the outputs, extended forces mapped back to the caller's order."""
    from deepmd_amd.deep_pot import DeepPot, InputNlist
    d = _water_model()
    dp = DeepPot(d, precision="f64")
    m = dp.model
    ec, ea, mp, nl, n = _extended(m.rcut + 1.0, m.sel, seed=7)
    nall = ea.shape[1]
    x = ec.view(-1, 3).numpy()
    at = ea.view(-1).numpy().astype(np.int32)
    rows = [[int(j) for j in r if j >= 0] for r in nl[0].numpy()]
    # the caller's arrays carry one NULL atom among the locals and one among
    # the ghosts; neighbour indices refer to the caller's numbering
    ins_loc, ins_gh = 5, n + 11
    old_to_new = np.array([i + (i >= ins_loc) + (i >= ins_gh - 1) for i in range(nall)])
    xc = np.insert(x, [ins_loc, ins_gh - 1], [[0.1, 0.2, 0.3], [0.4, 0.5, 0.6]], axis=0)
    atc = np.insert(at, [ins_loc, ins_gh - 1], [-1, -1])
    assert atc[ins_loc] == -1 and atc[ins_gh] == -1
    new_rows = [[int(old_to_new[j]) for j in r] for r in rows]
    new_rows.insert(ins_loc, [0, 1])       # the NULL local's own row is dropped
    nloc_c = n + 1
    nlist = InputNlist.from_rows(new_rows)
    e, f, v, ae, av = dp.compute_lower(xc.reshape(-1), atc, None, nghost=len(atc) - nloc_c,
                                       nlist=nlist, ago=0, atomic=True)
    ref = m.forward_lower(ec, ea, nl, mp, do_atomic_virial=True)
    torch.cuda.synchronize()
    fr = ref["extended_force"].cpu().numpy()[0]
    f = np.asarray(f).reshape(-1, 3)
    assert np.all(f[[ins_loc, ins_gh]] == 0)
    assert _rel(np.delete(f, [ins_loc, ins_gh], axis=0), fr) < 1e-12
    assert _rel(e, ref["energy"].cpu()) < 1e-12
    assert _rel(v, ref["virial"].cpu()) < 1e-12
    ae = np.asarray(ae).reshape(-1)
    assert ae[ins_loc] == 0
    # ago > 0 reuses the stored list with new coordinates
    xc2 = xc.copy()
    xc2[:nloc_c] += 0.01
    e2 = dp.compute_lower(xc2.reshape(-1), atc, None, nghost=len(atc) - nloc_c, nlist=None,
                          ago=1)[0]
    assert np.isfinite(e2).all() and abs(float(np.ravel(e2)[0]) - float(np.ravel(e)[0])) > 0


@pytest.mark.parametrize("prec", ["f64", "f16x2"])
def test_forward_lower_ago_reuse_and_graph_replays(prec):
    """MD-engine stepping (DeepPotPT::compute's ``ago``): after one call with
    the list (ago == 0), calls with ago > 0 reuse the list and slot order with
    no host synchronisation and give the same E/F/V as a fresh ago == 0 call
    on the moved coordinates (to the rounding of the force kernel's FP64
    atomics, 1e-12); the ago > 0 call captured in a HIP graph and replayed on
    new coordinates (with eager work and synchronisations between replays)
    matches the eager result at every replay."""
    from deepmd_amd.model import SeAModel, random_model_dict
    d = random_model_dict(1, 6.0, 0.5, [46, 92], [25, 50, 100], 16, [240, 240, 240])
    m = SeAModel(d, prec, DEV)
    ec, ea, mp, nl, n = _extended(m.rcut + 0.5, m.sel, seed=7)
    g = torch.Generator().manual_seed(11)
    first = m.forward_lower(ec, ea, nl, mp)
    s = m._last_lower_system
    keys = ("energy", "extended_force", "virial")

    def moved(k):
        return (ec + 0.01 * k * torch.randn(ec.shape, generator=g, dtype=ec.dtype)).to(DEV)

    def fresh(x):  # a new system, ago == 0
        out = m.forward_lower(x, ea, nl, mp)
        return {k: out[k].double().cpu() for k in keys}

    x1 = moved(1)
    r1 = m.forward_lower(x1, ea, None, system=s, ago=1)
    r1 = {k: r1[k].double().cpu() for k in keys}
    f1 = fresh(x1)
    # the same kernels on the same list; forces gather with FP64 atomics, whose
    # order varies between runs
    for k in keys:
        assert _rel(r1[k], f1[k]) <= 1e-12, k
    assert not torch.equal(r1["extended_force"], first["extended_force"].double().cpu())
    # graph capture of the ago > 0 step on a static input buffer
    xin = x1.clone()
    s2 = m._last_lower_system
    m.forward_lower(xin, ea, None, system=s2, ago=1, copy_outputs=False)  # warm-up
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        gout = m.forward_lower(xin, ea, None, system=s2, ago=1, copy_outputs=False)
    for rep in range(3):
        xr = moved(rep + 2)
        xin.copy_(xr)
        graph.replay()
        torch.cuda.synchronize()
        got = {k: gout[k].double().cpu().clone() for k in keys}
        ref = fresh(xr)  # eager work between replays
        torch.cuda.synchronize()
        for k in keys:
            rel = _rel(got[k], ref[k])
            assert rel <= 1e-12, (rep, k, rel)
    # errors: ago > 0 without an earlier call, or on another extended region
    m2 = SeAModel(d, prec, DEV)
    with pytest.raises(ValueError, match="ago > 0"):
        m2.forward_lower(x1, ea, None, ago=1)
    with pytest.raises(ValueError, match="ago > 0"):
        m.forward_lower(x1.view(-1, 3)[:-1], ea[:, :-1], None, system=s2, ago=1)


def test_forward_lower_bucket_overflow_vs_oracle():
    """The MD path's neighbour-force gather bins slots into buckets of
    capacity ~nnei; an atom in more lists than that spills to the overflow
    list (atomics).  A centre with 100 atoms on a unit shell around it, sel
    [40]: it is a neighbour of all 100 (capacity 64), forces and atomic
    virials vs the oracle (1e-10)."""
    from deepmd_amd.model import SeAModel, random_model_dict
    from oracle import geometry, sea
    d = random_model_dict(3, 6.0, 0.5, [40], [8, 16, 32], 16, [16])
    m = SeAModel(d, "f64", DEV)
    om = sea.model_from_serialized(d)
    rng = np.random.default_rng(11)
    v = rng.normal(size=(100, 3))
    shell = v / np.linalg.norm(v, axis=1, keepdims=True)
    coord = np.concatenate([np.zeros((1, 3)), shell]) + 10.0
    n = coord.shape[0]
    c = torch.tensor(coord).view(1, -1)
    a = torch.zeros((1, n), dtype=torch.long)
    ec, ea, mp, nl = geometry.extend_input_and_build_neighbor_list(c, a, m.rcut, m.sel, box=None)
    assert int((nl[0] == 0).sum()) > 64  # the centre sits in more lists than a bucket holds
    out = m.forward_lower(ec, ea, nl, mp, do_atomic_virial=True)
    torch.cuda.synchronize()
    ref = sea.forward_lower(om, ec, ea, nl, mp, do_atomic_virial=True)
    assert _rel(out["energy"].cpu(), ref["energy"].detach()) < 1e-10
    assert _rel(out["extended_force"].cpu(), ref["extended_force"].detach()) < 1e-10
    assert _rel(out["virial"].cpu(), ref["virial"].detach()) < 1e-10
    assert _rel(out["extended_virial"].cpu(), ref["extended_virial"].detach()) < 1e-10

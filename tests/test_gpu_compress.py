"""Compressed (tabulated) embedding inside the fused model, on the GPU.

Kernel parity: the fused forward / backward over type blocks equal the
oracle's tabulate_fusion_se_a / _grad restatement (padding fold included)
applied block by block, as se_a.py:795-843 does, with the same tables.
Model level: the compressed FP64 model equals the uncompressed FP64 oracle to
the interpolation error of the tables, and FP32 tables stay within the mixed
precision bar.
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


def _setup(prec, type_one_side, nside=3, L=12.44 * 3 / 4, seed=4):
    from deepmd_amd import systems
    from deepmd_amd.model import SeAModel, random_model_dict
    d = random_model_dict(2, 6.0, 0.5, [46, 92], [25, 50, 100], 16, [32, 32],
                          type_one_side=type_one_side)
    x, at, box = systems.water(nside, L, seed)
    m = SeAModel(d, prec, DEV)
    m.enable_compression(0.5)
    s = m.system(x[None], at[None], box.reshape(1, 9))
    return d, m, s, x, at, box


def _oracle_blocks(m, em, atype, fn):
    """Apply fn(table, info, em_x, em_block, rows) per (centre type, neighbour
    block) like se_a.py:795-843; returns a list of (rows, block, result)."""
    nt = m.ntypes
    sec = np.concatenate([[0], np.cumsum(m.sel)])
    out = []
    for ti in range(nt):
        rows = np.nonzero(atype == ti)[0]
        if rows.size == 0:
            continue
        for tb in range(nt):
            net = tb if m.type_one_side else tb * nt + ti
            table = m.tab_raw[net].cpu().numpy().astype(em.dtype)
            info = m.tab_info_np[net].astype(em.dtype)
            blk = em[rows, sec[tb]:sec[tb + 1], :]
            out.append((rows, (sec[tb], sec[tb + 1]), fn(table, info, blk[:, :, 0].copy(), blk)))
    return out


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-5)])
@pytest.mark.parametrize("type_one_side", [True, False])
def test_compressed_kernels_vs_oracle_op(prec, tol, type_one_side):
    from oracle import cops
    d, m, s, x, at, box = _setup(prec, type_one_side)
    print(f"compressed kernels {prec} type_one_side={type_one_side}", flush=True)
    m.evaluate(s)
    torch.cuda.synchronize()
    em = s.em.cpu().numpy()
    L = m.d3
    nnei = m.nnei
    xyz_ref = np.zeros((len(at), 4, L))
    for rows, _, r in _oracle_blocks(m, em, at, lambda tb, inf, ex, e:
                                     cops.tabulate_fusion_se_a(tb, inf, ex, e, L)):
        xyz_ref[rows] += r
    xyz_ref /= nnei
    assert _rel(s.xyz.cpu().numpy(), xyz_ref) < tol
    dy = s.gxyz.cpu().numpy().astype(em.dtype) / nnei
    nd_ref = np.zeros_like(em, dtype=np.float64)
    nt = m.ntypes
    sec = np.concatenate([[0], np.cumsum(m.sel)])
    for ti in range(nt):
        rows = np.nonzero(at == ti)[0]
        for tb in range(nt):
            net = tb if m.type_one_side else tb * nt + ti
            table = m.tab_raw[net].cpu().numpy().astype(em.dtype)
            info = m.tab_info_np[net].astype(em.dtype)
            blk = em[rows, sec[tb]:sec[tb + 1], :]
            gx, gm = cops.tabulate_fusion_se_a_grad(table, info, blk[:, :, 0].copy(), blk, dy[rows])
            gm = gm.astype(np.float64)
            gm[:, :, 0] += gx
            nd_ref[rows, sec[tb]:sec[tb + 1], :] = gm
    assert _rel(s.nd.cpu().numpy(), nd_ref) < tol * 10
    # the tables, not the nets, produced xyz: the uncompressed path differs
    from deepmd_amd.model import SeAModel
    m2 = SeAModel(d, prec, DEV)
    s2 = m2.system(x[None], at[None], box.reshape(1, 9))
    m2.evaluate(s2)
    torch.cuda.synchronize()
    assert m.compressed and not m2.compressed
    assert not torch.equal(s.xyz, s2.xyz)


def test_compressed_f64_model_vs_uncompressed_oracle():
    from oracle import sea
    for type_one_side in (True, False):
        d, m, s, x, at, box = _setup("f64", type_one_side)
        out = m.evaluate(s, atomic_virial=True)
        torch.cuda.synchronize()
        om = sea.model_from_serialized(d)
        ref = sea.forward(om, torch.tensor(x).view(1, -1), torch.tensor(at).view(1, -1).long(),
                          torch.tensor(box.ravel()).view(1, 9), do_atomic_virial=True)
        e = _rel(out["energy"].cpu().numpy(), ref["energy"].detach().numpy())
        f = _rel(out["force"].cpu().numpy(), ref["force"].detach().numpy())
        v = _rel(out["virial"].cpu().numpy(), ref["virial"].detach().numpy())
        print(f"compressed f64 vs uncompressed (type_one_side={type_one_side}): "
              f"E {e:.2e} F {f:.2e} V {v:.2e}", flush=True)
        # interpolation error of stride-0.01 quintic tables
        assert e < 1e-9 and f < 1e-7 and v < 1e-7


def test_compressed_f32_force_error():
    from oracle import sea
    d, m, s, x, at, box = _setup("f32", True)
    out = m.evaluate(s)
    torch.cuda.synchronize()
    om = sea.model_from_serialized(d)
    ref = sea.forward(om, torch.tensor(x).view(1, -1), torch.tensor(at).view(1, -1).long(),
                      torch.tensor(box.ravel()).view(1, 9))
    df = out["force"].cpu().numpy() - ref["force"].detach().numpy().reshape(-1, 3)
    rmse = float(np.sqrt(np.mean(df ** 2)))
    rms = float(np.sqrt(np.mean(ref["force"].detach().numpy() ** 2)))
    print(f"compressed f32: force rmse {rmse:.3e} (rms {rms:.3e})", flush=True)
    assert rmse <= 1e-3 and rmse <= 1e-4 * rms

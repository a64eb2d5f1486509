"""Fitting-net paths on the GPU.

Uniform hidden widths above 128 are padded (240 -> 256) with zero weights,
biases and output weights, which leaves every real activation and gradient
unchanged: the padded FP64 model must equal the unpadded one to rounding of
the GEMM summation order, and the BF16 one must stay within its error budget.
A two-type model with a doubling resnet ([120, 240, 240], y += (x, x)) runs
unpadded through the per-type segments.
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


def _cu(prec, pad: bool, monkeypatch):
    from deepmd_amd import model as M
    from deepmd_amd import systems
    d = M.random_model_dict(1, 6.0, 1.0, [120], [32, 64, 128], 16, [240, 240, 240])
    with monkeypatch.context() as mp:
        if not pad:
            mp.setattr(M.Fitting, "PAD", 10 ** 9)
        m = M.SeAModel(d, prec, DEV)
    assert (m.fitting.widths == [256] * 3) == pad
    x, at, box = systems.fcc(5, 3.615, 0.05, 11)
    return m, m.system(x[None], at[None], box.reshape(1, 9))


@pytest.mark.parametrize("prec", ["f64", "bf16"])
def test_padded_fitting_equals_unpadded(prec, monkeypatch):
    mp, sp = _cu(prec, True, monkeypatch)
    mu, su = _cu(prec, False, monkeypatch)
    out = {k: v.clone() for k, v in mp.evaluate(sp, atomic_virial=True).items()}
    ref = mu.evaluate(su, atomic_virial=True)
    torch.cuda.synchronize()
    tol = 1e-12 if prec == "f64" else 2e-2
    for k in ("energy", "atom_energy", "force", "virial", "atom_virial"):
        e = _rel(out[k].cpu(), ref[k].cpu())
        print(f"{prec} padded vs unpadded {k}: {e:.2e}", flush=True)
        assert e < tol, k


def test_fitting_two_types_resnet_doubling_bf16():
    from deepmd_amd import systems
    from deepmd_amd.model import SeAModel, random_model_dict
    d = random_model_dict(3, 6.0, 0.5, [46, 92], [25, 50, 100], 16, [120, 240, 240])
    m = SeAModel(d, "bf16", DEV)
    assert [ly["res"] for ly in m.fitting.nets[0]["hidden"]] == [0, 2, 1]
    m64 = SeAModel(d, "f64", DEV)
    x, at, box = systems.water(3, 12.44 * 3 / 4, 9)
    out = m.evaluate(m.system(x[None], at[None], box.reshape(1, 9)))
    ref = m64.evaluate(m64.system(x[None], at[None], box.reshape(1, 9)))
    torch.cuda.synchronize()
    rms = float(torch.sqrt(torch.mean(ref["force"] ** 2)))
    rmse = float(torch.sqrt(torch.mean((out["force"] - ref["force"]) ** 2)))
    de = float((out["energy"] - ref["energy"]).abs().max()) / len(at)
    print(f"two types: force rmse {rmse:.3e} (rms {rms:.3e}), energy/atom err {de:.3e}", flush=True)
    assert rmse <= 0.05 * rms
    assert de <= 1e-3


@pytest.mark.parametrize("fit,ntypes", [([240, 240, 240], 1), ([256, 256], 2),
                                        ([240, 240, 240], 2)])
def test_fused_fitting_tail_matches_per_layer_path(fit, ntypes, monkeypatch):
    """BF16 nets of 2-3 hidden layers of width 256 (240 padded) run layers
    2..L, the energy and the backward in one kernel; the per-layer path
    (GEMM + epilogue kernels) is the reference.  Partial 64-slot tiles: 500
    atoms, and per-type segments for two types."""
    from deepmd_amd import model as M
    from deepmd_amd import systems
    d = M.random_model_dict(ntypes - 1, 6.0, 1.0, [120] * ntypes if ntypes == 1 else [60, 60],
                            [32, 64, 128], 16, fit)
    mf = M.SeAModel(d, "bf16", DEV)
    assert mf.fitting.tail is not None
    with monkeypatch.context() as mp:
        mp.setattr(M.Fitting, "FUSED_TAIL", False)
        mu = M.SeAModel(d, "bf16", DEV)
    assert mu.fitting.tail is None
    m64 = M.SeAModel(d, "f64", DEV)
    x, at, box = systems.fcc(5, 3.615, 0.05, 11, ntypes=ntypes)
    outs = []
    for m in (mf, mu, m64):
        s = m.system(x[None], at[None], box.reshape(1, 9))
        o = m.evaluate(s, atomic_virial=True)
        outs.append({k: v.clone() for k, v in o.items()})
    torch.cuda.synchronize()
    fu, un, ref = outs
    for k in ("energy", "atom_energy", "force", "virial", "atom_virial"):
        e = _rel(fu[k].cpu(), un[k].cpu())
        print(f"fit {fit} ntypes {ntypes}: fused vs per-layer {k}: {e:.2e}", flush=True)
        assert e < 2e-2, k
    rms = float(torch.sqrt(torch.mean(ref["force"] ** 2)))
    ef = float(torch.sqrt(torch.mean((fu["force"] - ref["force"]) ** 2)))
    eu = float(torch.sqrt(torch.mean((un["force"] - ref["force"]) ** 2)))
    print(f"force rmse vs f64: fused {ef:.3e}, per-layer {eu:.3e} (rms {rms:.3e})", flush=True)
    assert ef <= 1.5 * eu + 1e-3 * rms
    de = float((fu["energy"] - ref["energy"]).abs().max()) / len(at)
    assert de <= 1e-3


@pytest.mark.parametrize("kind", ["water", "cu"])
def test_fused_layer1_split_fp16_vs_gemm_path(kind):
    """dpk_fit_l1_fwd / _bwd (descriptor and layer 1 fused, split FP16) equal
    the path that writes [D_hi | D_lo] and dE/dD around two FP16 GEMMs with
    FP32 output: the same ~22-bit operands, another summation order."""
    from deepmd_amd import model as M, systems
    if kind == "water":
        d = M.random_model_dict(3, 6.0, 0.5, [46, 92], [25, 50, 100], 16, [240, 240, 240])
        x, at, box = systems.water(4, 12.44, 1)
    else:
        d = M.random_model_dict(3, 8.0, 2.0, [200], [32, 64, 128], 16, [240, 240, 240])
        x, at, box = systems.fcc(6, 3.615, 0.05, 2)
    outs = {}
    for fused in (True, False):
        old = M.Fitting.FUSED_L1
        M.Fitting.FUSED_L1 = fused
        try:
            m = M.SeAModel(d, "f16x2", DEV)
            assert (m.fitting.l1 is not None) == fused
            s = m.system(x[None], at[None], box.reshape(1, 9))
            out = m.evaluate(s)
            torch.cuda.synchronize()
            outs[fused] = {k: out[k].double().cpu().numpy().copy() for k in ("energy", "force")}
            outs[fused]["gxyz"] = s.gxyz.double().cpu().numpy().copy()
        finally:
            M.Fitting.FUSED_L1 = old
    a, b = outs[True], outs[False]
    for k in ("energy", "force", "gxyz"):
        rel = np.max(np.abs(a[k] - b[k])) / np.max(np.abs(b[k]))
        print(f"{kind} {k}: fused vs GEMM path {rel:.2e}", flush=True)
        assert rel < 1e-5, (k, rel)


@pytest.mark.parametrize("d3,n", [(16, 1), (18, 70), (33, 129), (100, 200), (128, 333)])
def test_fused_layer1_kernels_vs_fp64(d3, n):
    """dpk_fit_l1_fwd / _bwd on their own, through the C-ABI, against FP64
    math with the same FP16 weights: z1 = D W1 and
    dE/dxyz[c][p] = sum_m GD[16 p + m] xyz[c][m] + [p < 16] sum_k GD[16 k + p] xyz[c][k],
    GD = gz1 W1^T (se_a.py:844-850 and its autograd).  Covers odd d3 and d3
    not a multiple of 4 (partial pipeline stages and output runs), a partial
    last workgroup, a permuted atom order, and that rows of atoms outside the
    segment are left untouched."""
    from deepmd_amd._lib import DPK_PREC_F16X2, call, lib, ptr, stream_ptr
    g = torch.Generator().manual_seed(d3 * 1000 + n)
    natoms = n + 7
    h = 256
    xyz = torch.randn(natoms, 4, d3, generator=g, dtype=torch.float64).float().double()
    w1 = torch.randn(16 * d3, h, generator=g, dtype=torch.float64) / (16 * d3) ** 0.5
    order = torch.randperm(natoms, generator=g)[:n].to(torch.int32)
    gz = torch.randn(n, h, generator=g, dtype=torch.float64)
    w1h = w1.float().half().double()          # the kernels' FP16 weights
    gzh = gz.float().half()
    gzl = (gz.float() - gzh.float()).half()   # split operand: hi | lo
    gzs = gzh.double() + gzl.double()
    # FP64 reference
    xa = xyz[order.long()]                    # [n, 4, d3]
    D = torch.einsum("nck,ncm->nkm", xa, xa[:, :, :16]).reshape(n, 16 * d3)
    z_ref = D @ w1h
    GD = (gzs @ w1h.T).reshape(n, d3, 16)
    gx_ref = torch.einsum("nkm,ncm->nck", GD, xa[:, :, :16])
    gx_ref[:, :, :16] += torch.einsum("nkp,nck->ncp", GD, xa)
    # device
    nb = lib().dpk_fit_l1_image_bytes(d3, h)
    assert nb > 0
    img = torch.empty(nb, dtype=torch.uint8, device=DEV)
    st = stream_ptr()
    call("dpk_fit_l1_pack", ptr(img), DPK_PREC_F16X2, ptr(w1.float().to(DEV)), d3, 16, h, st)
    xd = xyz.float().to(DEV).contiguous()
    od = order.to(DEV)
    z1 = torch.full((n, h), float("nan"), dtype=torch.float32, device=DEV)
    call("dpk_fit_l1_fwd", ptr(z1), DPK_PREC_F16X2, ptr(xd), ptr(od), n, d3, 16, ptr(img), h, st)
    gzd = torch.cat([gzh, gzl], 1).to(DEV).contiguous()
    gx = torch.full((natoms, 4, d3), 7.0, dtype=torch.float32, device=DEV)
    call("dpk_fit_l1_bwd", ptr(gx), DPK_PREC_F16X2, ptr(gzd), ptr(xd), ptr(od), n, d3, 16,
         ptr(img), h, st)
    torch.cuda.synchronize()
    z1 = z1.double().cpu()
    gx = gx.double().cpu()
    rz = float((z1 - z_ref).abs().max() / z_ref.abs().max())
    rg = float((gx[order.long()] - gx_ref).abs().max() / gx_ref.abs().max())
    print(f"d3={d3} n={n}: z1 rel {rz:.2e}, dE/dxyz rel {rg:.2e}", flush=True)
    # xyz enters D as an FP32 product split into ~22-bit operands
    assert rz < 2e-5 and rg < 2e-5, (rz, rg)
    rest = torch.ones(natoms, dtype=torch.bool)
    rest[order.long()] = False
    assert torch.all(gx[rest] == 7.0)

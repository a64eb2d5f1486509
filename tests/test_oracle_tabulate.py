"""Pins the oracle's tabulate_fusion_se_a (forward / grad / grad_grad) and the
table builder against the reference goldens
(``source/lib/tests/test_tabulate_se_a.cc``,
``source/tests/pt/test_tabulate_fusion_se_a.py``), the compiled reference,
adjoint identities and finite differences."""

import json

import numpy as np
import pytest
import torch

from conftest import golden_path
from oracle import cops, refcapi, sea, tabulate


def _cc(cc_goldens):
    g = cc_goldens["test_tabulate_se_a.cc"]["TestTabulateSeA"]
    table = np.array(g["table"]).reshape(-1, 6 * 8)
    return g, table, np.array(g["info"]), np.array(g["em_x"]).reshape(4, 4), np.array(g["em"]).reshape(4, 4, 4)


def test_cc_golden_forward_grad(cc_goldens):
    g, table, info, emx, em = _cc(cc_goldens)
    out = cops.tabulate_fusion_se_a(table, info, emx, em, 8)
    np.testing.assert_allclose(out.ravel(), g["expected_xyz_scatter"], atol=1e-5)
    dx, dem = cops.tabulate_fusion_se_a_grad(table, info, emx, em, np.ones((4, 4, 8)))
    np.testing.assert_allclose(dx.ravel(), g["expected_dy_dem_x"], atol=1e-5)
    np.testing.assert_allclose(dem.ravel(), g["expected_dy_dem"], atol=1e-5)


@pytest.mark.parametrize("dtype,prec", [(np.float64, 1e-10), (np.float32, 1e-5)])
def test_pt_op_golden(dtype, prec):
    with open(golden_path("pt_tabulate_op.json")) as fh:
        p = json.load(fh)
    table = np.array(p["table_tensor"]["value"], dtype=dtype).reshape(-1, 48)
    info = np.array(p["table_info_tensor"]["value"], dtype=dtype)
    emx = np.array(p["em_x_tensor"]["value"], dtype=dtype).reshape(4, 4)
    em = np.array(p["em_tensor"]["value"], dtype=dtype).reshape(4, 4, 4)
    out = cops.tabulate_fusion_se_a(table, info, emx, em, 8)
    np.testing.assert_allclose(out.ravel(), np.array(p["expected_descriptor_tensor"]["value"]).ravel(),
                               atol=prec, rtol=prec)
    # backward of loss = out.sum()  ->  dy = ones
    dx, dem = cops.tabulate_fusion_se_a_grad(table, info, emx, em, np.ones((4, 4, 8), dtype))
    np.testing.assert_allclose(dx.ravel(), np.array(p["expected_dy_dem_x"]["value"]).ravel(),
                               atol=prec, rtol=prec)
    np.testing.assert_allclose(dem.ravel(), np.array(p["expected_dy_dem"]["value"]).ravel(),
                               atol=prec, rtol=prec)


def _padded_inputs(rng, nloc=5, nnei=12, L=8, dtype=np.float64, nspline=40):
    """Sorted rows with trailing padding (em_x at the padding value, em[1:4]=0)."""
    table = rng.normal(size=(nspline, 6 * L)).astype(dtype) * 0.1
    info = np.array([0.0, 0.2, 2.0, 0.01, 0.1, -1], dtype=dtype)
    emx = rng.uniform(-0.1, 2.3, size=(nloc, nnei)).astype(dtype)
    em = rng.normal(size=(nloc, nnei, 4)).astype(dtype)
    pad_val = dtype(-0.05)
    for i in range(nloc):
        npad = i * 2
        if npad:
            emx[i, -npad:] = pad_val
            em[i, -npad:, 1:] = 0
            em[i, -npad:, 0] = rng.normal()
    em[:, :, 0] = np.where(emx == pad_val, em[:, :, 0], emx)
    return table, info, emx, em


@pytest.mark.skipif(not refcapi.available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_vs_compiled_reference_with_padding_fold(dtype):
    rng = np.random.default_rng(3)
    table, info, emx, em = _padded_inputs(rng, dtype=dtype)
    tol = 1e-12 if dtype == np.float64 else 2e-5
    out = cops.tabulate_fusion_se_a(table, info, emx, em, 8)
    np.testing.assert_allclose(out, refcapi.tabulate_fusion_se_a(table, info, emx, em, 8), rtol=tol, atol=tol)
    dy = rng.normal(size=out.shape).astype(dtype)
    dx, dem = cops.tabulate_fusion_se_a_grad(table, info, emx, em, dy)
    rdx, rdem = refcapi.tabulate_fusion_se_a_grad(table, info, emx, em, dy)
    np.testing.assert_allclose(dx, rdx, rtol=tol, atol=tol)
    np.testing.assert_allclose(dem, rdem, rtol=tol, atol=tol)
    zx = rng.normal(size=emx.shape).astype(dtype)
    zem = rng.normal(size=em.shape).astype(dtype)
    gg = cops.tabulate_fusion_se_a_grad_grad(table, info, emx, em, zx, zem)
    np.testing.assert_allclose(gg, refcapi.tabulate_fusion_se_a_grad_grad(table, info, emx, em, zx, zem, 8),
                               rtol=tol, atol=tol)


def test_grad_is_adjoint_of_forward_fd():
    # <dy, d out> = <grad(dy), d(em_x, em)> along a random direction (central FD)
    rng = np.random.default_rng(4)
    table, info, emx, em = _padded_inputs(rng, nloc=3, nnei=6)
    emx[:] = rng.uniform(0.05, 1.5, size=emx.shape)   # away from pieces' edges, no padding
    em[:, :, 1:] = rng.normal(size=em[:, :, 1:].shape)
    dy = rng.normal(size=(3, 4, 8))
    dxx = rng.normal(size=emx.shape)
    dee = rng.normal(size=em.shape)
    h = 1e-7
    fp = cops.tabulate_fusion_se_a(table, info, emx + h * dxx, em + h * dee, 8)
    fm = cops.tabulate_fusion_se_a(table, info, emx - h * dxx, em - h * dee, 8)
    lhs = np.sum(dy * (fp - fm) / (2 * h))
    gx, gem = cops.tabulate_fusion_se_a_grad(table, info, emx, em, dy)
    rhs = np.sum(gx * dxx) + np.sum(gem * dee)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-6)


def test_grad_grad_is_adjoint_of_grad():
    # grad is linear in dy, so <grad_grad(zx, zem), dy> = <zx, gx(dy)> + <zem, gem(dy)>
    rng = np.random.default_rng(5)
    table, info, emx, em = _padded_inputs(rng)
    dy = rng.normal(size=(5, 4, 8))
    zx = rng.normal(size=emx.shape)
    zem = rng.normal(size=em.shape)
    gx, gem = cops.tabulate_fusion_se_a_grad(table, info, emx, em, dy)
    gg = cops.tabulate_fusion_se_a_grad_grad(table, info, emx, em, zx, zem)
    np.testing.assert_allclose(np.sum(gg * dy), np.sum(zx * gx) + np.sum(zem * gem), rtol=1e-12)


def test_builder_matches_net(sea_yaml):
    # every piece reproduces y, y', y'' of the embedding net at both knots, and
    # the table tracks the net between knots
    m = sea.model_from_serialized(sea_yaml)
    net = m.embeddings[1]
    table, xx = tabulate.build_table(net, -1.0, 3.0, 0.01, 0.1, 5)
    nspline = table.shape[0]
    L = m.neuron[-1]
    c = table.reshape(nspline, L, 6)
    y, d1, d2 = tabulate.net_derivs(net, xx)
    h = np.diff(xx)[:nspline, None]
    p = sum(c[:, :, k] * h ** k for k in range(6))
    dp = sum(k * c[:, :, k] * h ** (k - 1) for k in range(1, 6))
    ddp = sum(k * (k - 1) * c[:, :, k] * h ** (k - 2) for k in range(2, 6))
    np.testing.assert_allclose(c[:, :, 0], y[:nspline], rtol=0, atol=0)
    np.testing.assert_allclose(p, y[1:nspline + 1], atol=1e-10)
    np.testing.assert_allclose(dp, d1[1:nspline + 1], atol=1e-8)
    np.testing.assert_allclose(ddp, d2[1:nspline + 1], atol=1e-6)
    info = tabulate.table_info(-1.0, 3.0, 5, 0.01, 0.1)
    xs = np.random.default_rng(0).uniform(-0.9, 9.9, size=(1, 64))
    em = np.concatenate([xs[..., None], np.zeros((1, 64, 3))], -1)
    em[0, :, 0] = 1.0
    out = cops.tabulate_fusion_se_a(table, info, xs, em, L, is_sorted=False)  # [1, 4, L]
    ref = sea.run_net(torch.tensor(xs.reshape(-1, 1)), net).numpy().sum(0)
    np.testing.assert_allclose(out[0, 0], ref, rtol=1e-6, atol=1e-7)

"""Pins the oracle's C++-op semantics (env_mat_a with derivative, prod_force_a,
prod_virial_a and their grads) against the reference gtest goldens
(``source/lib/tests/test_{env_mat_a,prod_force_a,prod_virial_a,
prod_force_grad_a,prod_virial_grad_a}.cc``), the compiled reference, finite
differences and torch autograd of the PT env matrix."""

import numpy as np
import pytest
import torch

from oracle import cops, refcapi, sea


def _geo(cc_geometry, key):
    ext = np.array(cc_geometry["posi_cpy"]).reshape(1, -1, 3)
    at = np.array(cc_geometry["atype_cpy"]).reshape(1, -1)
    fmt = np.array(cc_geometry["fmt"][key]).reshape(1, 6, -1)
    return ext, at, fmt


def _envd(cc_geometry, key):
    ext, at, fmt = _geo(cc_geometry, key)
    nnei = fmt.shape[-1]
    avg = np.zeros((2, nnei * 4))
    std = np.ones((2, nnei * 4))
    em, ed, rij = cops.prod_env_mat_a(ext, at, fmt, avg, std, 6.0, 0.8, float32_cutoffs=True)
    return ext, at, fmt, em, ed, rij


def test_env_mat_golden(cc_goldens, cc_geometry):
    for cls, key in (("TestEnvMatA", "10_10"), ("TestEnvMatAShortSel", "2_2")):
        _, _, _, em, _, _ = _envd(cc_geometry, key)
        np.testing.assert_allclose(em.ravel(), cc_goldens["test_env_mat_a.cc"][cls]["expected_env"],
                                   atol=1e-5)


def test_prod_force_golden(cc_goldens, cc_geometry):
    ext, at, fmt, em, ed, rij = _envd(cc_geometry, "5_5")
    nall = ext.shape[1]
    nd = em.size
    net = (10 - np.arange(nd) * 0.01).reshape(em.shape)
    # the gtest doubles everything into 2 identical frames
    f = cops.prod_force_a(np.concatenate([net, net]), np.concatenate([ed, ed]),
                          np.concatenate([fmt, fmt]), nall)
    want = cc_goldens["test_prod_force_a.cc"]["TestProdForceA"]["expected_force"] * 2
    np.testing.assert_allclose(f.ravel(), want, atol=1e-5)


def test_prod_virial_golden(cc_goldens, cc_geometry):
    ext, at, fmt, em, ed, rij = _envd(cc_geometry, "5_5")
    net = (10 - np.arange(em.size) * 0.01).reshape(em.shape)
    v, av = cops.prod_virial_a(net, ed, rij, fmt, ext.shape[1])
    g = cc_goldens["test_prod_virial_a.cc"]["TestProdVirialA"]
    np.testing.assert_allclose(v.ravel(), g["expected_virial"], atol=1e-5)
    np.testing.assert_allclose(av.ravel(), g["expected_atom_virial"], atol=1e-5)


def test_prod_force_grad_golden(cc_goldens, cc_geometry):
    ext, at, fmt, em, ed, rij = _envd(cc_geometry, "5_5")
    grad = (10 - np.arange(18) * 0.1).reshape(1, 6, 3)
    gn = cops.prod_force_grad_a(np.concatenate([grad, grad]), np.concatenate([ed, ed]),
                                np.concatenate([fmt, fmt]))
    want = cc_goldens["test_prod_force_grad_a.cc"]["TestProdForceGradA"]["expected_grad_net"] * 2
    np.testing.assert_allclose(gn.ravel(), want, atol=1e-5)


def test_prod_virial_grad_golden(cc_goldens, cc_geometry):
    ext, at, fmt, em, ed, rij = _envd(cc_geometry, "5_5")
    grad = (10 - np.arange(9) * 1.0).reshape(1, 9)
    gn = cops.prod_virial_grad_a(grad, ed, rij, fmt)
    want = cc_goldens["test_prod_virial_grad_a.cc"]["TestProdVirialGradA"]["expected_grad_net"]
    np.testing.assert_allclose(gn.ravel(), want, atol=1e-5)


def _random_system(seed, nloc=40, L=9.0, sel=(12, 24), rc=4.5, rcs=1.0):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0, L, size=(nloc, 3))
    at = rng.integers(0, 2, size=nloc)
    c = torch.tensor(xyz).view(1, -1, 3)
    a = torch.tensor(at).view(1, -1)
    box = torch.tensor(np.diag([L] * 3).ravel()).view(1, 9)
    ext, eat, mp, nl = sea.geometry.extend_input_and_build_neighbor_list(c, a, rc, list(sel), box)
    nl = sea.geometry.format_nlist(ext, eat, nl, rc, list(sel))
    nnei = sum(sel)
    avg = rng.normal(0, 0.1, size=(2, nnei * 4))
    std = rng.uniform(0.5, 1.5, size=(2, nnei * 4))
    return ext.numpy(), eat.numpy(), nl.numpy(), avg, std, rc, rcs, rng


@pytest.mark.skipif(not refcapi.available(), reason="oracle/_ref not built")
def test_vs_compiled_reference_random():
    ext, eat, nl, avg, std, rc, rcs, rng = _random_system(11)
    nloc, nnei = nl.shape[1:]
    nall = ext.shape[1]
    em, ed, rij = cops.prod_env_mat_a(ext, eat, nl, avg, std, rc, rcs, float32_cutoffs=True)
    rem, red, rrij = refcapi.env_mat_a(ext[0], eat[0], nloc, nl[0], [12, 24], rcs, rc)
    tc = eat[0, :nloc]
    np.testing.assert_allclose(em[0], (rem - avg[tc]) / std[tc], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(ed[0], red / np.repeat(std[tc], 3, axis=1), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(rij[0], rrij, rtol=0, atol=0)
    net = rng.normal(size=em.shape)
    np.testing.assert_allclose(cops.prod_force_a(net, ed, nl, nall)[0],
                               refcapi.prod_force_a(net, ed, nl, nloc, nall)[0], rtol=1e-12, atol=1e-12)
    v, av = cops.prod_virial_a(net, ed, rij, nl, nall)
    rv, rav = refcapi.prod_virial_a(net, ed, rij, nl, nloc, nall)
    np.testing.assert_allclose(v[0], rv, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(av[0], rav, rtol=1e-11, atol=1e-11)
    g = rng.normal(size=(1, nloc, 3))
    np.testing.assert_allclose(cops.prod_force_grad_a(g, ed, nl),
                               refcapi.prod_force_grad_a(g, ed, nl, nloc), rtol=1e-12, atol=1e-12)
    gv = rng.normal(size=(1, 9))
    np.testing.assert_allclose(cops.prod_virial_grad_a(gv, ed, rij, nl)[0],
                               refcapi.prod_virial_grad_a(gv, ed, rij, nl, nloc), rtol=1e-12, atol=1e-12)


def test_env_deriv_central_fd():
    # em_deriv is d em / d r_centre (prod_env_mat.cc convention)
    ext, eat, nl, avg, std, rc, rcs, rng = _random_system(12, nloc=20)
    em, ed, _ = cops.prod_env_mat_a(ext, eat, nl, avg, std, rc, rcs)
    nloc, nnei = nl.shape[1:]
    ed = ed.reshape(nloc, nnei, 4, 3)
    h = 1e-6
    for i in (0, 5, 13):
        for d in range(3):
            p = ext.copy(); p[0, i, d] += h
            m = ext.copy(); m[0, i, d] -= h
            # only the centre moves (its images are separate extended atoms)
            ep, _, _ = cops.prod_env_mat_a(p, eat, nl, avg, std, rc, rcs)
            emm, _, _ = cops.prod_env_mat_a(m, eat, nl, avg, std, rc, rcs)
            fd = ((ep - emm) / (2 * h)).reshape(1, nloc, nnei, 4)[0, i]
            np.testing.assert_allclose(ed[i, :, :, d], fd, atol=1e-6, rtol=1e-5)


def test_prod_force_virial_equal_pt_autograd():
    # the C++ adjoint form equals torch autograd through the PT env matrix:
    # for E = <g, em>, F_ext = -dE/dr_ext and virial = sum F (x) r
    ext, eat, nl, avg, std, rc, rcs, rng = _random_system(13)
    nloc, nnei = nl.shape[1:]
    nall = ext.shape[1]
    em, ed, rij = cops.prod_env_mat_a(ext, eat, nl, avg, std, rc, rcs)
    g = rng.normal(size=em.shape)
    cc = torch.tensor(ext).requires_grad_(True)
    tc = torch.tensor(eat[:, :nloc])
    tem, _, _ = sea.prod_env_mat(cc, torch.tensor(nl), tc,
                                 torch.tensor(avg).view(2, nnei, 4), torch.tensor(std).view(2, nnei, 4),
                                 rc, rcs)
    np.testing.assert_allclose(tem.detach().numpy().reshape(em.shape), em, rtol=1e-12, atol=1e-13)
    E = (tem.reshape(em.shape) * torch.tensor(g)).sum()
    gr, = torch.autograd.grad(E, cc)
    f_pt = -gr.numpy()
    f = cops.prod_force_a(g, ed, nl, nall)
    np.testing.assert_allclose(f, f_pt, rtol=1e-10, atol=1e-11)
    v, _ = cops.prod_virial_a(g, ed, rij, nl, nall)
    v_pt = np.einsum("fik,fij->fkj", f_pt, ext).reshape(1, 9)
    np.testing.assert_allclose(v, v_pt, rtol=1e-9, atol=1e-10)

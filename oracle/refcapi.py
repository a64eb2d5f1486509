"""ctypes bindings for ``oracle/_ref/libdeepmd_ref.so`` (the reference's own
C++ CPU kernels compiled by ``oracle/Makefile``).

Test infrastructure only.  :func:`load` returns None when the library has not
been built (e.g. on a machine without ``/root/reference``); callers skip.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

_LIB = None
_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libdeepmd_ref.so")


def available() -> bool:
    """True when the library has been built (checked without loading it, so
    that test collection maps nothing)."""
    return os.path.exists(_PATH)


def load():
    global _LIB
    if _LIB is None and os.path.exists(_PATH):
        _LIB = ctypes.CDLL(_PATH)
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def copy_coord(coord, atype, rcut, box, mem_factor=40):
    """Reference periodic images: (ext_coord [nall,3], ext_atype, mapping)."""
    lib = load()
    coord = _c(coord, np.float64).reshape(-1)
    atype = _c(atype, np.int32)
    nloc = atype.size
    mem = nloc * mem_factor
    oc = np.zeros(mem * 3)
    ot = np.zeros(mem, np.int32)
    mp = np.zeros(mem, np.int32)
    lib.ref_copy_coord.restype = ctypes.c_int
    nall = lib.ref_copy_coord(_p(oc), _p(ot), _p(mp), _p(coord), _p(atype), nloc, mem,
                              ctypes.c_double(rcut), _p(_c(box, np.float64)))
    assert nall > 0
    return oc[: nall * 3].reshape(nall, 3), ot[:nall], mp[:nall]


def build_nlist(ext_coord, nloc, rcut, max_nnei=1024):
    lib = load()
    c = _c(ext_coord, np.float64).reshape(-1)
    nall = c.size // 3
    nl = np.full((nloc, max_nnei), -1, np.int32)
    nn = np.zeros(nloc, np.int32)
    ret = lib.ref_build_nlist(_p(nl), _p(nn), max_nnei, _p(c), nloc, nall, ctypes.c_double(rcut))
    assert ret == 0
    return nl, nn


def format_nlist_i(posi, atype, i_idx, nei, rcut, sel):
    lib = load()
    sec = np.concatenate([[0], np.cumsum(sel)]).astype(np.int32)
    out = np.full(int(sec[-1]), -1, np.int32)
    p = _c(posi, np.float64).reshape(-1)
    t = _c(atype, np.int32)
    n = _c(nei, np.int32)
    lib.ref_format_nlist_i(_p(out), _p(p), _p(t), t.size, i_idx, _p(n), n.size,
                           ctypes.c_double(rcut), _p(sec), sec.size)
    return out


def env_mat_a(posi, atype, nloc, fmt_nlist, sel, rcut_smth, rcut):
    lib = load()
    sec = np.concatenate([[0], np.cumsum(sel)]).astype(np.int32)
    nnei = int(sec[-1])
    p = _c(posi, np.float64).reshape(-1)
    t = _c(atype, np.int32)
    f = _c(fmt_nlist, np.int32).reshape(-1)
    em = np.zeros(nloc * nnei * 4)
    ed = np.zeros(nloc * nnei * 12)
    rij = np.zeros(nloc * nnei * 3)
    lib.ref_env_mat_a(_p(em), _p(ed), _p(rij), _p(p), _p(t), nloc, t.size, _p(f), _p(sec),
                      sec.size, ctypes.c_double(rcut_smth), ctypes.c_double(rcut))
    return em.reshape(nloc, -1), ed.reshape(nloc, -1), rij.reshape(nloc, -1)


def prod_env_mat_a(coord, atype, raw_nlist, raw_numneigh, avg, std, nloc, rcut, rcut_smth, sel):
    lib = load()
    sec = np.concatenate([[0], np.cumsum(sel)]).astype(np.int32)
    nnei = int(sec[-1])
    c = _c(coord, np.float64).reshape(-1)
    t = _c(atype, np.int32)
    raw = _c(raw_nlist, np.int32)
    nn = _c(raw_numneigh, np.int32)
    em = np.zeros(nloc * nnei * 4)
    ed = np.zeros(nloc * nnei * 12)
    rij = np.zeros(nloc * nnei * 3)
    nl = np.zeros(nloc * nnei, np.int32)
    lib.ref_prod_env_mat_a(_p(em), _p(ed), _p(rij), _p(nl), _p(c), _p(t), _p(raw), _p(nn),
                           raw.shape[1], _p(_c(avg, np.float64)), _p(_c(std, np.float64)),
                           nloc, t.size, ctypes.c_double(rcut), ctypes.c_double(rcut_smth),
                           _p(sec), sec.size)
    return em.reshape(nloc, -1), ed.reshape(nloc, -1), rij.reshape(nloc, -1), nl.reshape(nloc, nnei)


def prod_force_a(net_deriv, in_deriv, nlist, nloc, nall, nframes=1):
    lib = load()
    nnei = nlist.shape[-1]
    f = np.zeros(nframes * nall * 3)
    lib.ref_prod_force_a(_p(f), _p(_c(net_deriv, np.float64)), _p(_c(in_deriv, np.float64)),
                         _p(_c(nlist, np.int32)), nloc, nall, nnei, nframes)
    return f.reshape(nframes, nall, 3)


def prod_virial_a(net_deriv, in_deriv, rij, nlist, nloc, nall):
    lib = load()
    nnei = nlist.shape[-1]
    v = np.zeros(9)
    av = np.zeros(nall * 9)
    lib.ref_prod_virial_a(_p(v), _p(av), _p(_c(net_deriv, np.float64)),
                          _p(_c(in_deriv, np.float64)), _p(_c(rij, np.float64)),
                          _p(_c(nlist, np.int32)), nloc, nall, nnei)
    return v, av.reshape(nall, 9)


def prod_force_grad_a(grad, env_deriv, nlist, nloc, nframes=1):
    lib = load()
    nnei = nlist.shape[-1]
    out = np.zeros(nframes * nloc * nnei * 4)
    lib.ref_prod_force_grad_a(_p(out), _p(_c(grad, np.float64)), _p(_c(env_deriv, np.float64)),
                              _p(_c(nlist, np.int32)), nloc, nnei, nframes)
    return out.reshape(nframes, nloc, nnei * 4)


def prod_virial_grad_a(grad, env_deriv, rij, nlist, nloc):
    lib = load()
    nnei = nlist.shape[-1]
    out = np.zeros(nloc * nnei * 4)
    lib.ref_prod_virial_grad_a(_p(out), _p(_c(grad, np.float64)), _p(_c(env_deriv, np.float64)),
                               _p(_c(rij, np.float64)), _p(_c(nlist, np.int32)), nloc, nnei)
    return out.reshape(nloc, nnei * 4)


def _tab_dt(table):
    return (np.float64, "f64") if table.dtype == np.float64 else (np.float32, "f32")


def tabulate_fusion_se_a(table, info, em_x, em, L):
    lib = load()
    dt, sfx = _tab_dt(table)
    nloc, nnei, nd = em.shape
    out = np.zeros(nloc * nd * L, dt)
    getattr(lib, "ref_tabulate_fusion_se_a_" + sfx)(
        _p(out), _p(_c(table, dt)), _p(_c(info, dt)), _p(_c(em_x, dt)), _p(_c(em, dt)),
        nloc, nnei, L, nd)
    return out.reshape(nloc, nd, L)


def tabulate_fusion_se_a_grad(table, info, em_x, em, dy):
    lib = load()
    dt, sfx = _tab_dt(table)
    nloc, nnei, nd = em.shape
    L = dy.shape[-1]
    dx = np.zeros(nloc * nnei, dt)
    dem = np.zeros(nloc * nnei * nd, dt)
    getattr(lib, "ref_tabulate_fusion_se_a_grad_" + sfx)(
        _p(dx), _p(dem), _p(_c(table, dt)), _p(_c(info, dt)), _p(_c(em_x, dt)),
        _p(_c(em, dt)), _p(_c(dy, dt)), nloc, nnei, L, nd)
    return dx.reshape(nloc, nnei), dem.reshape(nloc, nnei, nd)


def tabulate_fusion_se_a_grad_grad(table, info, em_x, em, dzx, dzem, L):
    lib = load()
    dt, sfx = _tab_dt(table)
    nloc, nnei, nd = em.shape
    out = np.zeros(nloc * nd * L, dt)
    getattr(lib, "ref_tabulate_fusion_se_a_grad_grad_" + sfx)(
        _p(out), _p(_c(table, dt)), _p(_c(info, dt)), _p(_c(em_x, dt)), _p(_c(em, dt)),
        _p(_c(dzx, dt)), _p(_c(dzem, dt)), nloc, nnei, L, nd)
    return out.reshape(nloc, nd, L)

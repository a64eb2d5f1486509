"""Export a model for the C++ MD-engine runner (``libdpk_deeppot.so``).

The runner (``deepmd-kit_amd/cpp/deep_pot_lower.cpp``, C API in
``include/dpk_deeppot.h``) is the C++ counterpart of ``forward_lower`` for
an engine that hands host arrays -- the role of deepmd::DeepPotPT::compute
with an InputNlist (source/api_cc/src/DeepPotPT.cc:171-378).  It calls
libdpk's C ABI only, so it takes the model in the device-ready forms the
Python model builds: the packed embedding image, the fused layer-1 and
fitting-tail images, davg / dstd, the type-pair mask.  This module writes
them to one file:

    DPKM 1
    <key> <values...>          scalars and small integer lists, one per line
    blob <name> <nbytes>       followed by nbytes raw bytes and a newline
    ...
    end

Only the split-FP16 precision is exported: its fitting is made of libdpk
kernels alone (the other precisions use hipBLASLt GEMMs for layer 1).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .model import _P1, _P2, _P3, SeAModel

MAGIC = "DPKM 1"


def _bytes(t: torch.Tensor) -> bytes:
    return t.detach().contiguous().cpu().view(torch.uint8).numpy().tobytes()


def export_lower_model(m: SeAModel, path: str) -> None:
    """Write ``m`` (split FP16, fused fitting) for the C++ runner."""
    fit = m.fitting
    if m.precision != "f16x2" or fit.l1 is None or fit.tail is None:
        raise ValueError("the C++ runner takes split-FP16 models with the fused fitting "
                         "(SeAModel(d, 'f16x2') with widths 256 and 2-3 hidden layers)")
    if m.compressed:
        raise ValueError("compressed models are not exported")
    if m.nx:
        raise ValueError("models with frame / atomic parameters are not exported")
    if m.model_pair_mask is not None:
        raise ValueError("models with pair_exclude_types are not exported")
    keep = [t for t in range(m.ntypes) if t not in m.atom_exclude_types]
    hid = [net["hidden"] for net in fit.nets]
    lines = [MAGIC,
             f"ntypes {m.ntypes}", f"nnei {m.nnei}", f"rcut {m.rcut!r}",
             f"rcut_smth {m.rcut_smth!r}",
             f"env_protection {m.env_protection!r}", f"d3 {m.d3}", f"m2 {m.axis_neuron}",
             f"prec {m.prec}", f"fit_prec {fit.prec_code}", f"nnets {m.nnets}",
             f"P {_P1} {_P2} {_P3}", f"res {m.res2} {m.res3}", f"has_idt {m.emb_has_idt}",
             f"skip_padding {int(bool(m.skip_padding))}", f"h1 {hid[0][0]['h']}",
             f"layers {len(hid[0])}",
             "sel " + " ".join(str(int(x)) for x in m.sel),
             "sec " + " ".join(str(int(x)) for x in m.sec),
             "keep " + " ".join(str(t) for t in keep),
             "tail_res " + " ".join(str(int(r)) for r in fit.tail_res),
             "ebias " + " ".join(repr(float(net["bout"]) + float(fit.ebias[t]))
                                 for t, net in enumerate(fit.nets))]
    blobs = [("davg", _bytes(m.davg_dev)), ("dstd", _bytes(m.dstd_dev)),
             ("emb_params", _bytes(m.emb_params)), ("emb_image", _bytes(m.emb_image)),
             ("net_of", _bytes(m.net_of)), ("ch_out", _bytes(m.ch_out))]
    if m.pair_mask is not None:
        blobs.append(("pair_mask", _bytes(m.pair_mask)))
    if m.emb_pad is not None:
        blobs.append(("emb_pad", _bytes(m.emb_pad)))
    for t in range(m.ntypes):
        blobs.append((f"l1_{t}", _bytes(fit.l1[t])))
        blobs.append((f"tail_{t}", _bytes(fit.tail[t])))
    with open(path, "wb") as f:
        f.write(("\n".join(lines) + "\n").encode())
        for name, b in blobs:
            f.write(f"blob {name} {len(b)}\n".encode())
            f.write(b)
            f.write(b"\n")
        f.write(b"end\n")


class CppDeepPot:
    """ctypes handle on the C++ runner (for tests and Python callers that want
    its host-array interface)."""

    def __init__(self, path: str, device: int = 0):
        import ctypes
        from .build import DEEPPOT_LIB
        _lib.lib()  # libdpk first (the runner links it)
        L = ctypes.CDLL(DEEPPOT_LIB)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.dpk_deeppot_load.restype = P
        L.dpk_deeppot_load.argtypes = [ctypes.c_char_p, I]
        L.dpk_deeppot_free.restype = None
        L.dpk_deeppot_free.argtypes = [P]
        L.dpk_deeppot_compute.restype = I
        L.dpk_deeppot_compute.argtypes = [P, P, P, P, P, P, I, I, I, P, P, P, I]
        L.dpk_deeppot_last_error.restype = ctypes.c_char_p
        L.dpk_deeppot_last_error.argtypes = []
        L.dpk_deeppot_rcut.restype = D
        L.dpk_deeppot_rcut.argtypes = [P]
        L.dpk_deeppot_ntypes.restype = I
        L.dpk_deeppot_ntypes.argtypes = [P]
        self._L = L
        self.h = L.dpk_deeppot_load(path.encode(), device)
        if not self.h:
            raise RuntimeError(L.dpk_deeppot_last_error().decode())

    def compute(self, coord, atype, nghost: int, ilist=None, numneigh=None, neighbors=None,
                ago: int = 0):
        """coord [nall, 3] float64 and atype [nall] (host arrays, local atoms
        first), the engine's list as ilist [inum], numneigh [inum] and
        neighbors (one int array per row) when ago == 0 -> energy [1], force
        [nall, 3], virial [9] (float64)."""
        import ctypes
        coord = np.ascontiguousarray(coord, dtype=np.float64).reshape(-1, 3)
        atype = np.ascontiguousarray(atype, dtype=np.int32).reshape(-1)
        nall = atype.size
        keep = []
        inum, il, nn, fn = 0, None, None, None
        if ilist is not None:
            il = np.ascontiguousarray(ilist, dtype=np.int32)
            nn = np.ascontiguousarray(numneigh, dtype=np.int32)
            rows = [np.ascontiguousarray(r, dtype=np.int32) for r in neighbors]
            keep = [il, nn, rows]
            inum = il.size
            fn = (ctypes.c_void_p * max(inum, 1))(*[r.ctypes.data for r in rows])
            keep.append(fn)
        e = np.zeros(1)
        f = np.zeros((nall, 3))
        v = np.zeros(9)
        rc = self._L.dpk_deeppot_compute(
            self.h, e.ctypes.data, f.ctypes.data, v.ctypes.data, coord.ctypes.data,
            atype.ctypes.data, nall, nghost, inum, None if il is None else il.ctypes.data,
            None if nn is None else nn.ctypes.data,
            None if fn is None else ctypes.addressof(fn), ago)
        del keep
        if rc != 0:
            raise RuntimeError(self._L.dpk_deeppot_last_error().decode())
        return e, f, v

    def __del__(self):
        h, self.h = getattr(self, "h", None), None
        if h:
            self._L.dpk_deeppot_free(h)

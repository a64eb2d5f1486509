"""ctypes binding of libdpk.so (the C ABI declared in include/dpk.h).

The library is required: there is no CPU fallback.  ``lib()`` raises
``DpkUnavailable`` when the shared object has not been built, and every
entry point's status code is checked (``DpkError`` carries
``dpk_last_error()``).  torch is imported first so that libdpk binds to the
HIP runtime torch already loaded (one runtime, one set of streams).
"""

from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (must precede loading libdpk; see module docstring)

from . import build as _build

_LIB = None

c_int, c_double, c_size_t, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p
P = c_void_p

DPK_PREC_F64, DPK_PREC_F32, DPK_PREC_BF16, DPK_PREC_F16, DPK_PREC_F16X2 = 0, 1, 2, 3, 4
# activation codes (DPK_ACT_*)
ACTIVATIONS = {"tanh": 0, "linear": 1, "none": 1, "relu": 2, "relu6": 3, "softplus": 4,
               "sigmoid": 5, "silu": 6, "gelu": 7, "gelu_tf": 7, "silut": 8, "custom_silu": 8}


class DpkUnavailable(RuntimeError):
    pass


class DpkError(RuntimeError):
    pass


class Frame(ctypes.Structure):
    """Mirror of ``dpk_frame`` (include/dpk.h)."""
    _fields_ = [
        ("box", c_double * 9), ("rec", c_double * 9), ("origin", c_double * 3),
        ("inv_extent", c_double * 3), ("atom_off", c_int), ("nloc", c_int),
        ("cell_off", c_int), ("nc", c_int * 3), ("rng", c_int * 3), ("pbc", c_int),
        ("img_off", c_int), ("nb", c_int * 3), ("ns", c_int), ("out_off", c_int),
    ]


_SIGS = {
    "dpk_last_error": (ctypes.c_char_p, []),
    "dpk_abi_version": (c_int, []),
    "dpk_tune_set": (c_int, [c_int, c_int]),
    "dpk_tune_get": (c_int, [c_int]),
    "dpk_tuning_create": (P, []),
    "dpk_tuning_destroy": (None, [P]),
    "dpk_tuning_set": (c_int, [P, c_int, c_int]),
    "dpk_tuning_get": (c_int, [P, c_int]),
    "dpk_tuning_use": (P, [P]),
    "dpk_nlist_workspace_bytes": (c_size_t, [c_int, c_int]),
    "dpk_nlist_build": (c_int, [P, P, P, P, P, P, P, c_int, c_int, c_int, P, P, c_double, P,
                                c_int, c_int, P, c_size_t, P]),
    "dpk_nlist_build_noalloc": (c_int, [P, P, P, P, P, P, P, c_int, c_int, c_int, P, P, c_double,
                                        P, c_int, c_int, P]),
    "dpk_nlist_format": (c_int, [P, P, P, c_int, P, P, c_int, c_int, c_int, c_double, P, c_int,
                                 c_int, P]),
    "dpk_nlist_pair_exclude": (c_int, [P, P, P, P, c_int, c_int, P, c_int, P]),
    "dpk_prod_env_mat_a_f64": (c_int, [P, P, P, P, P, P, P, P, c_int, c_int, c_int, c_int,
                                       c_int, c_double, c_double, c_double, P]),
    "dpk_prod_env_mat_a_f32": (c_int, [P, P, P, P, P, P, P, P, c_int, c_int, c_int, c_int,
                                       c_int, c_double, c_double, c_double, P]),
    "dpk_prod_force_a_f64": (c_int, [P, P, P, P, c_int, c_int, c_int, c_int, P]),
    "dpk_prod_virial_workspace_bytes": (c_size_t, [c_int, c_int]),
    "dpk_prod_virial_a_f64": (c_int, [P, P, P, P, P, P, c_int, c_int, c_int, c_int, P,
                                      c_size_t, P]),
    "dpk_prod_force_grad_a_f64": (c_int, [P, P, P, P, c_int, c_int, c_int, c_int, P]),
    "dpk_prod_virial_grad_a_f64": (c_int, [P, P, P, P, P, c_int, c_int, c_int, P]),
    "dpk_tabulate_fusion_se_a_f64": (c_int, [P, P, c_int, P, P, P, c_int, c_int, c_int, c_int,
                                             P]),
    "dpk_tabulate_fusion_se_a_f32": (c_int, [P, P, c_int, P, P, P, c_int, c_int, c_int, c_int,
                                             P]),
    "dpk_tabulate_fusion_se_a_grad_f64": (c_int, [P, P, P, c_int, P, P, P, P, c_int, c_int,
                                                  c_int, c_int, P]),
    "dpk_tabulate_fusion_se_a_grad_f32": (c_int, [P, P, P, c_int, P, P, P, P, c_int, c_int,
                                                  c_int, c_int, P]),
    "dpk_tabulate_fusion_se_a_grad_grad_f64": (c_int, [P, P, c_int, P, P, P, P, P, c_int,
                                                       c_int, c_int, c_int, P]),
    "dpk_tabulate_fusion_se_a_grad_grad_f32": (c_int, [P, P, c_int, P, P, P, P, P, c_int,
                                                       c_int, c_int, c_int, P]),
    "dpk_env_mat_fused": (c_int, [P, c_int, P, P, P, P, P, P, P, P, c_int, c_int, c_double,
                                  c_double, c_double, P, c_int, P]),
    "dpk_embed_net_elems": (c_size_t, [c_int, c_int, c_int]),
    "dpk_embed_image_bytes": (c_size_t, [c_int, c_int, c_int]),
    "dpk_embed_pack_bf16": (c_int, [P, P, c_int, c_int, c_int, c_int, P]),
    "dpk_embed_pack_f16": (c_int, [P, P, c_int, c_int, c_int, c_int, P]),
    "dpk_embed_fwd": (c_int, [P, c_int, P, P, P, P, P, P, c_int, c_int, c_int, c_int, c_int,
                              c_int, c_int, c_int, c_int, c_double, c_int, P, c_int, c_int, P,
                              P, P]),
    "dpk_embed_bwd": (c_int, [P, c_int, P, P, P, P, P, P, P, c_int, c_int, c_int, c_int,
                              c_int, c_int, c_int, c_int, c_int, c_double, c_int, P, c_int,
                              c_int, P, P]),
    "dpk_embed_train_bimage_bytes": (c_size_t, []),
    "dpk_embed_train_pack_bwd": (c_int, [P, P, c_int, P]),
    "dpk_embed_train_parts_elems": (c_size_t, [c_int, c_int]),
    "dpk_embed_train_vjp": (c_int, [P, P, P, c_int, P, P, P, P, P, P, P, P, P, P, P, c_int,
                                    c_int, c_int, c_int, c_int, c_int, c_int, c_int, P]),
    "dpk_tab_embed_fwd": (c_int, [P, c_int, P, P, P, P, P, P, P, c_int, c_int, P, c_int, c_int,
                                  c_int, P]),
    "dpk_tab_embed_bwd": (c_int, [P, c_int, P, P, P, P, P, P, P, P, c_int, c_int, P, c_int,
                                  c_int, c_int, P]),
    "dpk_descriptor_fwd": (c_int, [P, c_int, P, c_int, c_int, c_int, c_int, P]),
    "dpk_descriptor_bwd": (c_int, [P, c_int, P, c_int, P, c_int, c_int, c_int, P]),
    "dpk_fit_descriptor_fwd": (c_int, [P, c_int, P, P, c_int, c_int, c_int, P]),
    "dpk_fit_descriptor_bwd": (c_int, [P, c_int, P, P, P, c_int, c_int, c_int, P]),
    "dpk_fit_layer_fwd": (c_int, [P, P, P, c_int, P, P, P, P, c_int, c_int, c_int, c_int, c_int,
                                  c_double, P]),
    "dpk_fit_energy": (c_int, [P, P, c_int, P, P, P, c_double, P, c_int, c_int, P]),
    "dpk_fit_layer_bwd": (c_int, [P, P, c_int, P, P, P, P, c_int, c_int, c_int, c_int, P]),
    "dpk_fit_tail_bytes": (c_size_t, [c_int]),
    "dpk_fit_tail_bf16": (c_int, [P, P, P, P, c_int, c_int, c_double, P, c_int, P]),
    "dpk_fit_tail_f16": (c_int, [P, P, P, P, c_int, c_int, c_double, P, c_int, P]),
    "dpk_fit_tail_f16x2": (c_int, [P, P, P, P, c_int, c_int, c_double, P, c_int, P]),
    "dpk_fit_l1_image_bytes": (c_size_t, [c_int, c_int]),
    "dpk_fit_l1_pack": (c_int, [P, c_int, P, c_int, c_int, c_int, P]),
    "dpk_fit_l1_fwd": (c_int, [P, c_int, P, P, c_int, c_int, c_int, P, c_int, P]),
    "dpk_fit_l1_bwd": (c_int, [P, c_int, P, P, P, c_int, c_int, c_int, P, c_int, P]),
    "dpk_force_virial_workspace_bytes": (c_size_t, [c_int]),
    "dpk_force_virial_fused": (c_int, [P, P, P, c_int, P, P, P, P, P, P, P, c_int, c_int,
                                       c_double, c_double, c_double, P, c_size_t, P, c_int, P]),
    "dpk_force_virial_cells": (c_int, [P, P, P, c_int, P, P, P, P, P, P, P, P, c_int, c_int,
                                       c_int, c_double, c_double, c_double, P, c_size_t, P,
                                       c_size_t, P, c_int, P]),
    "dpk_force_virial_gather_workspace_bytes": (c_size_t, [c_int, c_int, c_int]),
    "dpk_force_virial_gather": (c_int, [P, P, P, P, c_int, P, P, P, P, c_int, c_int, c_int,
                                        c_double, c_double, c_double, P, c_int, P, c_size_t, P]),
    "dpk_frame_sum_workspace_bytes": (c_size_t, [c_int, c_int, c_int]),
    "dpk_frame_sum": (c_int, [P, P, c_int, P, c_int, c_int, P, c_size_t, P]),
}

EXPORTS = tuple(_SIGS)


def lib_path() -> str:
    return os.environ.get("DPK_LIB", _build.LIB)


def lib():
    """The loaded libdpk (raises DpkUnavailable if it was never built)."""
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise DpkUnavailable(
                f"libdpk.so not found at {path}; build it with "
                "`python -m deepmd_amd.build` (deepmd-kit_amd/) -- there is no CPU fallback")
        L = ctypes.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().dpk_last_error().decode(errors="replace")
        raise DpkError(f"{what or 'libdpk'} failed (status {rc}): {msg}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)


def ptr(t) -> int | None:
    """Device (or host) address of a tensor, None for None."""
    if t is None:
        return None
    return t.data_ptr()


class Tuning:
    """One set of launch tunables (dpk_tuning), owned by a model and made the
    calling thread's current set around its launches (``with t.active():``)."""

    def __init__(self):
        self._lib = lib()
        self.handle = self._lib.dpk_tuning_create()
        if not self.handle:
            raise MemoryError(self._lib.dpk_last_error().decode())

    def set(self, key: int, value: int) -> None:
        check(self._lib.dpk_tuning_set(self.handle, key, value), "dpk_tuning_set")

    def get(self, key: int) -> int:
        return int(self._lib.dpk_tuning_get(self.handle, key))

    def active(self):
        return _TuningScope(self)

    def __del__(self):
        h, self.handle = getattr(self, "handle", None), None
        if h:
            self._lib.dpk_tuning_destroy(h)


class _TuningScope:
    def __init__(self, t: Tuning):
        self.t = t

    def __enter__(self):
        self.prev = self.t._lib.dpk_tuning_use(self.t.handle)
        return self.t

    def __exit__(self, *exc):
        self.t._lib.dpk_tuning_use(self.prev)
        return False


def stream_ptr(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def require_hip(*tensors) -> None:
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise RuntimeError("deepmd_amd kernels run on the MI355X (HIP) device only; "
                               f"got a tensor on {t.device}")

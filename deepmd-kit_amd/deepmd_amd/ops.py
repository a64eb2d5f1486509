"""The ``torch.ops.deepmd`` operator surface, backed by libdpk HIP kernels.

Importing this module loads ``lib/libdeepmd_op_pt.so`` (built from
``csrc/torch_op/deepmd_op_pt.cpp``) with ``torch.ops.load_library``, the way
the reference loads its own op library (deepmd/pt/cxx_op.py:42-56).  The
library registers, in the ``deepmd`` namespace:

* ``enable_mpi()`` -- the sentinel of source/op/pt/print_summary.cc:5-14
  (a [1] bool tensor, False: no MPI in this build);
* ``tabulate_fusion_se_a(table, table_info, em_x, em, last_layer_size) -> Tensor[]``
  -- the reference's schema and autograd chain (forward -> grad -> grad_grad,
  is_sorted padding fold; tabulate_multi_device.cc:731-962,1630-1684), plus
  ``tabulate_fusion_se_a_grad`` / ``_grad_grad`` as ops of their own;
* ``prod_env_mat_a``, ``prod_force_se_a``, ``prod_virial_se_a``,
  ``prod_force_se_a_grad``, ``prod_virial_se_a_grad`` -- torch forms of the
  TensorFlow ops of source/op/tf/prod_env_mat_multi_device.cc:13-30,
  prod_force_multi_device.cc:6-14, prod_virial_multi_device.cc:5-15,
  prod_force_grad_multi_device.cc:5-14, prod_virial_grad_multi_device.cc:5-15,
  with autograd (env matrix -> coordinates, force / virial -> net_deriv).

Only HIP tensors are computed (a CPU tensor raises after validation).  The
ops read no device data on the host and never synchronise, so a step built
from them can be captured in a HIP graph.  A C++ process loads the same file
(``dlopen`` / linking) to make the ops visible to TorchScript models.
"""

from __future__ import annotations

import os

import torch

from . import _lib
from .build import OP_LIB


def _namespace_owner() -> str | None:
    """Path of the library that already registered torch.ops.deepmd, "" when
    the namespace exists but came from elsewhere, None when it is free."""
    for p in torch.ops.loaded_libraries:
        if os.path.basename(p) == os.path.basename(OP_LIB):
            return p
    try:
        torch.ops.deepmd.enable_mpi  # noqa: B018
        return ""
    except (AttributeError, RuntimeError):
        return None


def load() -> str:
    """Load the operator library once per process; returns its path."""
    owner = _namespace_owner()
    if owner:
        if os.path.realpath(owner) != os.path.realpath(OP_LIB):
            raise RuntimeError(f"torch.ops.deepmd is already provided by {owner}")
        return owner
    if owner == "":
        raise RuntimeError("the torch.ops.deepmd namespace is already registered in this process "
                           "(load either the reference op library or deepmd_amd, not both)")
    if not os.path.exists(OP_LIB):
        raise _lib.DpkUnavailable(
            f"{OP_LIB} not found; build it with `python -m deepmd_amd.build` -- there is no "
            "CPU fallback")
    _lib.lib()  # libdpk first: one HIP runtime (torch's), one copy of the library
    torch.ops.load_library(OP_LIB)
    return OP_LIB


load()

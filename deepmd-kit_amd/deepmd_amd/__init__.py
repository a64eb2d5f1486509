"""deepmd_amd -- MI355X-native Deep Potential se_e2_a energy / force / virial.

Hand-written gfx950 HIP kernels (libdpk.so, C ABI in include/dpk.h) behind:

* ``deepmd_amd.ops``      -- registers the ``torch.ops.deepmd`` operators
  (``tabulate_fusion_se_a`` and its grads, ``prod_env_mat_a``,
  ``prod_force_se_a``, ``prod_virial_se_a`` and their grads);
* ``deepmd_amd.model``    -- the fused E/F/V step (``SeAModel``);
* ``deepmd_amd.deep_pot`` -- a ``DeepPot``-style ``eval`` front end;
* ``deepmd_amd.build``    -- compiles libdpk.so in-tree with hipcc.
"""

__version__ = "0.1.0"

"""CPU oracle for the se_e2_a energy / force / virial path.

TEST INFRASTRUCTURE ONLY.  This package is a plain CPU restatement (torch-CPU
FP64 and numpy) of the reference algorithm, written to check the MI355X
product path.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
as the timed CPU baseline -- never as part of the product: the package under
``deepmd-kit_amd/`` never imports anything from here, and its GPU path fails
loudly when the HIP library is missing instead of falling back to this code.

Modules (each function cites the reference file:line it restates):

* ``geometry``  -- PBC wrap, periodic images, dense and cell-list neighbour
  lists, per-type formatting (``deepmd/pt/utils/{region,nlist}.py``,
  ``deepmd/pt/model/model/make_model.py``).
* ``sea``       -- env matrix, embedding net, se_e2_a descriptor, energy
  fitting, forces / virials by autograd, model loading from the reference's
  serialized dict (``deeppot_sea.yaml``) and seeded initialisation.
* ``cops``      -- the C++ library's op semantics: ``prod_env_mat_a`` with
  ``em_deriv`` / ``rij``, ``prod_force_a``, ``prod_virial_a`` and their
  ``_grad`` forms, ``tabulate_fusion_se_a`` forward / grad / grad_grad.
* ``tabulate``  -- the compression table builder (quintic Hermite pieces).
* ``_ref``      -- (built, not committed) the reference's own C++ CPU kernels
  compiled from ``/root/reference/source/lib/src`` by ``oracle/Makefile``;
  ``oracle.refcapi`` loads it with ctypes.

Parity status of each function is recorded in its docstring: every function is
pinned by at least one golden vector, compiled-reference comparison, finite
difference or closed-form check in ``tests/test_oracle_*.py``.
"""

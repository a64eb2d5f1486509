"""The operator library libdeepmd_op_pt.so on its own: loaded in a fresh
process with ``torch.ops.load_library`` only (as deepmd/pt/cxx_op.py:42-56
loads the reference's), without importing deepmd_amd.

CPU: the enable_mpi sentinel, the op schemas (tabulate_fusion_se_a inferred
from its C++ signature, as source/op/pt/tabulate_multi_device.cc:1682-1684
registers it), the host-side validation the reference performs
(tabulate_multi_device.cc:21-50,72-83; test_tabulate_device_validation.py)
plus dtype / size checks of every raw-pointer operand, and TorchScript
compiling a call to the op.
GPU: the reference's tabulate golden (test_tabulate_fusion_se_a.py) forward
and backward, prod_force / prod_virial against the oracle, and a scripted,
saved and reloaded TorchScript module calling the op.
"""

import json
import os
import subprocess
import sys
import textwrap

import pytest

from conftest import GOLDEN, PKG, ROOT


def _op_lib():
    from deepmd_amd.build import OP_LIB
    return OP_LIB


def _run(code: str, tmp, timeout=300):
    """Run ``code`` in a fresh interpreter (from a file: TorchScript reads
    the source of what it compiles)."""
    env = dict(os.environ)
    env["PYTHONPATH"] = os.pathsep.join([ROOT, PKG, env.get("PYTHONPATH", "")])
    script = os.path.join(str(tmp), "child.py")
    with open(script, "w") as fh:
        fh.write(textwrap.dedent(code))
    r = subprocess.run([sys.executable, script], capture_output=True,
                       text=True, timeout=timeout, env=env, cwd=os.path.dirname(__file__))
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_op_library_loads_alone_and_validates(tmp_path):
    lib = _op_lib()
    if not os.path.exists(lib):
        pytest.skip("libdeepmd_op_pt.so not built (run __graft_entry__.build())")
    out = _run(f"""
        import sys, torch
        torch.ops.load_library({lib!r})
        assert "deepmd_amd" not in sys.modules
        assert torch.ops.deepmd.enable_mpi().tolist() == [False]
        s = str(torch.ops.deepmd.tabulate_fusion_se_a.default._schema)
        assert s == ("deepmd::tabulate_fusion_se_a(Tensor _0, Tensor _1, Tensor _2, Tensor _3, "
                     "int _4) -> Tensor[] _0"), s
        for name in ("prod_env_mat_a", "prod_force_se_a", "prod_virial_se_a",
                     "prod_force_se_a_grad", "prod_virial_se_a_grad",
                     "tabulate_fusion_se_a_grad", "tabulate_fusion_se_a_grad_grad"):
            getattr(torch.ops.deepmd, name).default
        f64, f32 = torch.float64, torch.float32
        t = torch.zeros((2, 48), dtype=f64)
        info = torch.zeros(6, dtype=f64)
        def err(args, kind, msg):
            try:
                torch.ops.deepmd.tabulate_fusion_se_a(*args)
            except Exception as e:
                assert type(e).__name__ == kind and msg in str(e), (type(e), str(e))
                return
            raise AssertionError("no error for " + msg)
        z = lambda *s, dt=f64: torch.zeros(s, dtype=dt)
        err((t, info, z(3, 4), z(3, 4, 4), 8), "RuntimeError", "the input tensor is on the CPU")
        err((t, info, z(3, 4, dt=f32), z(3, 4, 4), 8), "RuntimeError",
            "em_x must have the dtype of table")
        err((t, info, z(3, 3), z(3, 4, 4), 8), "RuntimeError", "em_x holds 9 values")
        err((t, info, z(3, 4), z(3, 4, 5), 8), "RuntimeError",
            "The environment basis dimension must be 4, 9, 16, or 25, got 5")
        err((t, info, z(12), z(3, 4, 4), 8), "ValueError", "Dim of input should be 2")
        err((z(96), info, z(3, 4), z(3, 4, 4), 8), "ValueError", "Dim of table should be 2")
        err((t, info, z(3, 4), z(3, 4, 4), 9), "RuntimeError", "cannot hold 9 channels")
        g = torch.zeros((3, 4, 8), dtype=f64)
        try:
            torch.ops.deepmd.tabulate_fusion_se_a_grad(t, info, z(3, 4), z(3, 4, 4),
                                                       z(3, 4, 7), g)
        except RuntimeError as e:
            assert "dy holds 84 values" in str(e)
        else:
            raise AssertionError("short dy accepted")
        for fn, args in ((torch.ops.deepmd.prod_force_se_a,
                          (z(1, 2 * 3 * 4), z(1, 2 * 3 * 12), torch.zeros((1, 2, 3), dtype=torch.int32), 2, 5)),):
            try:
                fn(*args)
            except RuntimeError as e:
                assert "MI355X (HIP) tensors only" in str(e)
        @torch.jit.script
        def scripted(table, tinfo, ex, em):
            return torch.ops.deepmd.tabulate_fusion_se_a(table, tinfo, ex, em, 8)[0]
        assert "deepmd::tabulate_fusion_se_a" in str(scripted.graph)
        print("ok")
    """, tmp_path)
    assert out.strip().endswith("ok")


def test_ops_module_loads_the_library(tmp_path):
    """deepmd_amd.ops is the loader: after it, the namespace comes from the
    .so (a second load is a no-op)."""
    lib = _op_lib()
    if not os.path.exists(lib):
        pytest.skip("libdeepmd_op_pt.so not built (run __graft_entry__.build())")
    out = _run(f"""
        import os, torch
        from deepmd_amd import ops
        assert any(os.path.basename(p) == "libdeepmd_op_pt.so" for p in torch.ops.loaded_libraries)
        assert ops.load() == ops.OP_LIB
        print("ok")
    """, tmp_path)
    assert out.strip().endswith("ok")


@pytest.mark.gpu
def test_op_library_gpu_golden_and_torchscript(tmp_path):
    lib = _op_lib()
    gold = os.path.join(GOLDEN, "pt_tabulate_op.json")
    ts = str(tmp_path / "scripted.pt")
    out = _run(f"""
        import json, sys, numpy as np, torch
        torch.ops.load_library({lib!r})
        assert "deepmd_amd" not in sys.modules
        DEV = "cuda"
        p = json.load(open({gold!r}))
        for dtype, prec in ((torch.float64, 1e-10), (torch.float32, 1e-5)):
            table = torch.tensor(p["table_tensor"]["value"], dtype=dtype, device=DEV).view(-1, 48)
            info = torch.tensor(p["table_info_tensor"]["value"], dtype=dtype)
            emx = torch.tensor(p["em_x_tensor"]["value"], dtype=dtype, device=DEV).view(16, 1)
            em = torch.tensor(p["em_tensor"]["value"], dtype=dtype, device=DEV).view(4, 4, 4)
            emx.requires_grad_(True); em.requires_grad_(True)
            out = torch.ops.deepmd.tabulate_fusion_se_a(table, info, emx, em, 8)[0]
            exp = torch.tensor(p["expected_descriptor_tensor"]["value"], dtype=dtype).view(out.shape)
            torch.testing.assert_close(out.detach().cpu(), exp, atol=prec, rtol=prec)
            out.sum().backward()
            torch.testing.assert_close(emx.grad.cpu().view(-1), torch.tensor(
                p["expected_dy_dem_x"]["value"], dtype=dtype), atol=prec, rtol=prec)
            torch.testing.assert_close(em.grad.cpu().view(-1), torch.tensor(
                p["expected_dy_dem"]["value"], dtype=dtype), atol=prec, rtol=prec)

        class M(torch.nn.Module):
            def forward(self, table, tinfo, ex, em):
                return torch.ops.deepmd.tabulate_fusion_se_a(table, tinfo, ex, em, 8)[0]
        table = torch.tensor(p["table_tensor"]["value"], dtype=torch.float64, device=DEV).view(-1, 48)
        info = torch.tensor(p["table_info_tensor"]["value"], dtype=torch.float64)
        emx = torch.tensor(p["em_x_tensor"]["value"], dtype=torch.float64, device=DEV).view(16, 1)
        em = torch.tensor(p["em_tensor"]["value"], dtype=torch.float64, device=DEV).view(4, 4, 4)
        sm = torch.jit.script(M())
        torch.jit.save(sm, {ts!r})
        lm = torch.jit.load({ts!r})
        got = lm(table, info, emx, em).cpu()
        exp = torch.tensor(p["expected_descriptor_tensor"]["value"], dtype=torch.float64).view(got.shape)
        torch.testing.assert_close(got, exp, atol=1e-10, rtol=1e-10)

        # prod_force / prod_virial with their autograd: adjoint identities
        g = torch.Generator().manual_seed(5)
        nf, nloc, nall, nnei = 2, 6, 14, 8
        nl = torch.randint(-1, nall, (nf, nloc, nnei), generator=g, dtype=torch.int32).to(DEV)
        nd = torch.randn(nf, nloc * nnei * 4, generator=g, dtype=torch.float64).to(DEV)
        ed = torch.randn(nf, nloc * nnei * 12, generator=g, dtype=torch.float64).to(DEV)
        rij = torch.randn(nf, nloc * nnei * 3, generator=g, dtype=torch.float64).to(DEV)
        nd.requires_grad_(True)
        f = torch.ops.deepmd.prod_force_se_a(nd, ed, nl, nloc, nall)
        w = torch.randn(f.shape, generator=g, dtype=torch.float64).to(DEV)
        (gnd,) = torch.autograd.grad((f * w).sum(), nd)
        # <w, F(nd)> = <F^T w, nd> for the linear map nd -> F
        lhs = float((f * w).sum()); rhs = float((gnd * nd).sum())
        assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1.0), (lhs, rhs)
        v, av = torch.ops.deepmd.prod_virial_se_a(nd, ed, rij, nl, nloc, nall)
        wv = torch.randn(v.shape, generator=g, dtype=torch.float64).to(DEV)
        (gv,) = torch.autograd.grad((v * wv).sum(), nd)
        lhs = float((v * wv).sum()); rhs = float((gv * nd).sum())
        assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1.0), (lhs, rhs)
        assert torch.allclose(av.view(nf, nall, 9).sum(1), v, rtol=1e-12, atol=1e-12)
        torch.cuda.synchronize()
        print("ok")
    """, tmp_path)
    assert out.strip().endswith("ok")

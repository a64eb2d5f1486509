"""The ``torch.ops.deepmd`` operator surface, backed by libdpk HIP kernels.

Importing this module registers, in the reference's ``deepmd`` namespace:

* ``enable_mpi()`` -- the library sentinel of source/op/pt/print_summary.cc:5-14
  (returns a [1] bool tensor, False: no MPI in this build);
* ``tabulate_fusion_se_a(table, table_info, em_x, em, last_layer_size) -> Tensor[]``
  -- same schema and autograd chain as source/op/pt/tabulate_multi_device.cc
  (forward -> tabulate_fusion_se_a_grad -> tabulate_fusion_se_a_grad_grad,
  is_sorted padding fold, lines 731-962, 1630-1684);
* ``prod_env_mat_a``, ``prod_force_se_a``, ``prod_virial_se_a``,
  ``prod_force_se_a_grad``, ``prod_virial_se_a_grad`` -- torch forms of the
  TensorFlow ops of source/op/tf/prod_env_mat_multi_device.cc:13-30,
  prod_force_multi_device.cc:6-14, prod_virial_multi_device.cc:5-15,
  prod_force_grad_multi_device.cc:5-14, prod_virial_grad_multi_device.cc:5-15.

Only HIP tensors are accepted (no CPU kernels: a CPU tensor raises).  Errors
mirror the reference's messages where its tests match on them
(source/tests/pt/test_tabulate_device_validation.py).
"""

from __future__ import annotations

import torch

from . import _lib
from ._lib import call, ptr, stream_ptr

_NS = "deepmd"


def _already_defined() -> bool:
    try:
        torch.ops.deepmd.enable_mpi  # noqa: B018
        return True
    except (AttributeError, RuntimeError):
        return False


if _already_defined():  # another deepmd op library owns the namespace
    raise RuntimeError("the torch.ops.deepmd namespace is already registered in this process "
                       "(load either the reference op library or deepmd_amd, not both)")

LIB = torch.library.Library(_NS, "DEF")

LIB.define("enable_mpi() -> Tensor")
LIB.define("tabulate_fusion_se_a(Tensor table, Tensor table_info, Tensor em_x, Tensor em, "
           "int last_layer_size) -> Tensor[]")
LIB.define("tabulate_fusion_se_a_grad(Tensor table, Tensor table_info, Tensor em_x, Tensor em, "
           "Tensor dy, Tensor descriptor) -> Tensor[]")
LIB.define("tabulate_fusion_se_a_grad_grad(Tensor table, Tensor table_info, Tensor em_x, "
           "Tensor em, Tensor dz_dy_dem_x, Tensor dz_dy_dem, Tensor descriptor) -> Tensor")
LIB.define("prod_env_mat_a(Tensor coord, Tensor atype, Tensor nlist, Tensor davg, Tensor dstd, "
           "float rcut, float rcut_smth) -> (Tensor, Tensor, Tensor)")
LIB.define("prod_force_se_a(Tensor net_deriv, Tensor in_deriv, Tensor nlist, int nloc, "
           "int nall) -> Tensor")
LIB.define("prod_virial_se_a(Tensor net_deriv, Tensor in_deriv, Tensor rij, Tensor nlist, "
           "int nloc, int nall) -> (Tensor, Tensor)")
LIB.define("prod_force_se_a_grad(Tensor grad, Tensor net_deriv, Tensor in_deriv, Tensor nlist, "
           "int nloc) -> Tensor")
LIB.define("prod_virial_se_a_grad(Tensor grad, Tensor net_deriv, Tensor in_deriv, Tensor rij, "
           "Tensor nlist, int nloc) -> Tensor")


def _enable_mpi():
    return torch.zeros(1, dtype=torch.bool)


LIB.impl("enable_mpi", _enable_mpi, "CompositeExplicitAutograd")


def _cpu_unsupported(name):
    def f(*args, **kwargs):
        raise RuntimeError(
            f"deepmd::{name}: the input tensor is on the CPU; deepmd_amd implements this "
            "operator for MI355X (HIP) tensors only")
    return f


for _op in ("tabulate_fusion_se_a", "tabulate_fusion_se_a_grad", "tabulate_fusion_se_a_grad_grad",
            "prod_env_mat_a", "prod_force_se_a", "prod_virial_se_a", "prod_force_se_a_grad",
            "prod_virial_se_a_grad"):
    LIB.impl(_op, _cpu_unsupported(_op), "CPU")


# --------------------------------------------------------------- tabulate --

def _check_tab_devices(table, table_info, named):
    if table_info.device.type != "cpu":
        raise RuntimeError("table_info must be on the CPU")
    for name, t in named:
        if t is not None and t.device != table.device:
            raise RuntimeError(f"{name} must be on the same device as table; table is on "
                               f"{table.device} but {name} is on {t.device}")


def _check_nd(nd):
    if nd not in (4, 9, 16, 25):
        raise RuntimeError(f"The environment basis dimension must be 4, 9, 16, or 25, got {nd}")


def _sfx(dtype):
    if dtype == torch.float64:
        return "f64"
    if dtype == torch.float32:
        return "f32"
    raise RuntimeError(f"tabulate: unsupported dtype {dtype}")


def _tab_fwd(table, table_info, em_x, em, last_layer_size):
    _check_tab_devices(table, table_info, [("em_x", em_x), ("em", em)])
    if table.dim() != 2:
        raise ValueError("Dim of table should be 2")
    if em_x.dim() != 2:
        raise ValueError("Dim of input should be 2")
    if em.dim() != 3:
        raise ValueError("Dim of input should be 3")
    nloc, nnei, nd = em.shape
    _check_nd(nd)
    sfx = _sfx(table.dtype)
    info = table_info.to(table.dtype).contiguous()
    out = torch.empty((nloc, nd, last_layer_size), dtype=table.dtype, device=table.device)
    tb, ex, e = table.contiguous(), em_x.contiguous(), em.contiguous()
    call(f"dpk_tabulate_fusion_se_a_{sfx}", ptr(out), ptr(tb), table.shape[0], ptr(info),
         ptr(ex), ptr(e), nloc, nnei, last_layer_size, nd, stream_ptr(table.device))
    return [out]


def _tab_grad(table, table_info, em_x, em, dy, descriptor):
    _check_tab_devices(table, table_info, [("em_x", em_x), ("em", em), ("dy", dy),
                                           ("descriptor", descriptor)])
    if dy.dim() != 3:
        raise ValueError("Dim of dy_tensor should be 3")
    nloc, nnei, nd = em.shape
    _check_nd(nd)
    L = descriptor.shape[2]
    sfx = _sfx(table.dtype)
    info = table_info.to(table.dtype).contiguous()
    dx = torch.zeros_like(em_x)
    dem = torch.zeros_like(em)
    tb, ex, e, dyc = table.contiguous(), em_x.contiguous(), em.contiguous(), dy.contiguous()
    call(f"dpk_tabulate_fusion_se_a_grad_{sfx}", ptr(dx), ptr(dem), ptr(tb), table.shape[0],
         ptr(info), ptr(ex), ptr(e), ptr(dyc), nloc, nnei, L, nd, stream_ptr(table.device))
    return [dx, dem]


def _tab_gg(table, table_info, em_x, em, dzx, dzem, descriptor):
    _check_tab_devices(table, table_info, [("em_x", em_x), ("em", em), ("dz_dy_dem_x", dzx),
                                           ("dz_dy_dem", dzem), ("descriptor", descriptor)])
    if dzx.dim() != 2:
        raise ValueError("Dim of dz_dy_dem_x should be 2")
    if dzem.dim() != 3:
        raise ValueError("Dim of dz_dy_dem should be 3")
    nloc, nnei, nd = em.shape
    _check_nd(nd)
    L = descriptor.shape[2]
    if L > 1024:
        raise RuntimeError("In the process of model compression, the size of the last layer "
                           "of embedding net must be less than 1024!")
    sfx = _sfx(table.dtype)
    info = table_info.to(table.dtype).contiguous()
    out = torch.empty_like(descriptor)
    tb, ex, e = table.contiguous(), em_x.contiguous(), em.contiguous()
    zx, ze = dzx.contiguous(), dzem.contiguous()
    call(f"dpk_tabulate_fusion_se_a_grad_grad_{sfx}", ptr(out), ptr(tb), table.shape[0],
         ptr(info), ptr(ex), ptr(e), ptr(zx), ptr(ze), nloc, nnei, L, nd,
         stream_ptr(table.device))
    return out


LIB.impl("tabulate_fusion_se_a", _tab_fwd, "CUDA")
LIB.impl("tabulate_fusion_se_a_grad", _tab_grad, "CUDA")
LIB.impl("tabulate_fusion_se_a_grad_grad", _tab_gg, "CUDA")


def _tab_setup(ctx, inputs, output):
    table, table_info, em_x, em, _ = inputs
    ctx.save_for_backward(table, table_info, em_x, em, output[0])


def _tab_backward(ctx, grads):
    table, table_info, em_x, em, desc = ctx.saved_tensors
    dy = grads[0].contiguous()
    dx, dem = torch.ops.deepmd.tabulate_fusion_se_a_grad(table, table_info, em_x, em, dy, desc)
    return None, None, dx, dem, None


def _tab_grad_setup(ctx, inputs, output):
    table, table_info, em_x, em, dy, desc = inputs
    ctx.save_for_backward(table, table_info, em_x, em, desc)


def _tab_grad_backward(ctx, grads):
    table, table_info, em_x, em, desc = ctx.saved_tensors
    gx = grads[0] if grads[0] is not None else torch.zeros_like(em_x)
    gm = grads[1] if grads[1] is not None else torch.zeros_like(em)
    dz = torch.ops.deepmd.tabulate_fusion_se_a_grad_grad(table, table_info, em_x, em,
                                                         gx.contiguous(), gm.contiguous(), desc)
    return None, None, None, None, dz, None


torch.library.register_autograd("deepmd::tabulate_fusion_se_a", _tab_backward,
                                setup_context=_tab_setup)
torch.library.register_autograd("deepmd::tabulate_fusion_se_a_grad", _tab_grad_backward,
                                setup_context=_tab_grad_setup)


# ---------------------------------------------------- env mat / force / virial --

def _f64(t):
    return t.to(torch.float64).contiguous()


def _check_nlist(nlist, nall, name):
    """Neighbour indices must address the nall extended rows (checked on the
    host before any kernel dereferences them)."""
    if nlist.numel() and int(nlist.max()) >= nall:
        raise RuntimeError(f"{name}: nlist holds index >= nall ({nall})")


def _i32(t):
    return t.to(torch.int32).contiguous()


def _env_fwd(coord, atype, nlist, davg, dstd, rcut, rcut_smth):
    _lib.require_hip(coord, atype, nlist, davg, dstd)
    nf, nloc, nnei = nlist.shape
    nall = atype.shape[1]
    _check_nlist(nlist, nall, "prod_env_mat_a")
    if coord.numel() != nf * nall * 3 or davg.numel() < nnei * 4 or dstd.numel() != davg.numel():
        raise RuntimeError("prod_env_mat_a: inconsistent coord / davg / dstd shapes")
    if atype.numel() and int(atype.max()) * nnei * 4 >= davg.numel():
        raise RuntimeError("prod_env_mat_a: atom type outside davg / dstd")
    dt = coord.dtype
    sfx = "f64" if dt == torch.float64 else "f32"
    if dt not in (torch.float64, torch.float32):
        raise RuntimeError(f"prod_env_mat_a: unsupported dtype {dt}")
    em = torch.empty((nf, nloc, nnei * 4), dtype=dt, device=coord.device)
    ed = torch.empty((nf, nloc, nnei * 12), dtype=dt, device=coord.device)
    rij = torch.empty((nf, nloc, nnei * 3), dtype=dt, device=coord.device)
    # converted inputs are bound to names so they outlive the launch
    c_, t_, n_ = coord.reshape(nf, nall * 3).contiguous(), _i32(atype), _i32(nlist)
    a_, s_ = davg.to(dt).contiguous(), dstd.to(dt).contiguous()
    call(f"dpk_prod_env_mat_a_{sfx}", ptr(em), ptr(ed), ptr(rij), ptr(c_), ptr(t_), ptr(n_),
         ptr(a_), ptr(s_), nf, nloc, nall, nnei, float(rcut), float(rcut_smth),
         stream_ptr(coord.device))
    return em, ed, rij


def _force_fwd(net_deriv, in_deriv, nlist, nloc, nall):
    _lib.require_hip(net_deriv, in_deriv, nlist)
    nf = net_deriv.shape[0]
    nnei = nlist.numel() // (nf * nloc) if nloc else 0
    _check_nlist(nlist, nall, "prod_force_se_a")
    force = torch.empty((nf, nall * 3), dtype=torch.float64, device=net_deriv.device)
    g_, d_, n_ = _f64(net_deriv), _f64(in_deriv), _i32(nlist)
    call("dpk_prod_force_a_f64", ptr(force), ptr(g_), ptr(d_), ptr(n_), nf, nloc, nall,
         max(nnei, 1), stream_ptr(net_deriv.device))
    return force.to(net_deriv.dtype)


def _virial_fwd(net_deriv, in_deriv, rij, nlist, nloc, nall):
    _lib.require_hip(net_deriv, in_deriv, rij, nlist)
    nf = net_deriv.shape[0]
    nnei = nlist.numel() // (nf * nloc) if nloc else 0
    _check_nlist(nlist, nall, "prod_virial_se_a")
    dev = net_deriv.device
    vir = torch.empty((nf, 9), dtype=torch.float64, device=dev)
    av = torch.empty((nf, nall * 9), dtype=torch.float64, device=dev)
    ws = torch.empty(int(_lib.lib().dpk_prod_virial_workspace_bytes(nf, nloc)) + 8,
                     dtype=torch.uint8, device=dev)
    g_, d_, r_, n_ = _f64(net_deriv), _f64(in_deriv), _f64(rij), _i32(nlist)
    call("dpk_prod_virial_a_f64", ptr(vir), ptr(av), ptr(g_), ptr(d_), ptr(r_), ptr(n_), nf,
         nloc, nall, max(nnei, 1), ptr(ws), ws.numel(), stream_ptr(dev))
    return vir.to(net_deriv.dtype), av.to(net_deriv.dtype)


def _force_grad_impl(grad, in_deriv, nlist, nloc, ngrad, out_dtype):
    nf = grad.shape[0]
    nnei = nlist.numel() // (nf * nloc) if nloc else 0
    gn = torch.empty((nf, nloc * max(nnei, 1) * 4), dtype=torch.float64, device=grad.device)
    g_, d_, n_ = _f64(grad), _f64(in_deriv), _i32(nlist)
    call("dpk_prod_force_grad_a_f64", ptr(gn), ptr(g_), ptr(d_), ptr(n_), nf, nloc,
         max(nnei, 1), ngrad, stream_ptr(grad.device))
    return gn.to(out_dtype)


def _force_grad_fwd(grad, net_deriv, in_deriv, nlist, nloc):
    _lib.require_hip(grad, net_deriv, in_deriv, nlist)
    return _force_grad_impl(grad, in_deriv, nlist, nloc, nloc, net_deriv.dtype)


def _virial_grad_fwd(grad, net_deriv, in_deriv, rij, nlist, nloc):
    _lib.require_hip(grad, net_deriv, in_deriv, rij, nlist)
    nf = grad.shape[0]
    nnei = nlist.numel() // (nf * nloc) if nloc else 0
    gn = torch.empty((nf, nloc * max(nnei, 1) * 4), dtype=torch.float64, device=grad.device)
    g_, d_, r_, n_ = _f64(grad), _f64(in_deriv), _f64(rij), _i32(nlist)
    call("dpk_prod_virial_grad_a_f64", ptr(gn), ptr(g_), ptr(d_), ptr(r_), ptr(n_), nf, nloc,
         max(nnei, 1), stream_ptr(grad.device))
    return gn.to(net_deriv.dtype)


LIB.impl("prod_env_mat_a", _env_fwd, "CUDA")
LIB.impl("prod_force_se_a", _force_fwd, "CUDA")
LIB.impl("prod_virial_se_a", _virial_fwd, "CUDA")
LIB.impl("prod_force_se_a_grad", _force_grad_fwd, "CUDA")
LIB.impl("prod_virial_se_a_grad", _virial_grad_fwd, "CUDA")


def _env_setup(ctx, inputs, output):
    coord, atype, nlist, davg, dstd, rcut, rcut_smth = inputs
    ctx.save_for_backward(output[1], nlist)
    ctx.coord_shape = coord.shape
    ctx.nall = atype.shape[1]


def _env_backward(ctx, g_em, g_ed, g_rij):
    # d E / d coord = -(force of net_deriv = g_em): centre -g.dR, neighbour +g.dR
    if g_ed is not None and bool(torch.any(g_ed != 0)):
        raise RuntimeError("prod_env_mat_a: em_deriv is not differentiable")
    ed, nlist = ctx.saved_tensors
    nf, nloc, _ = nlist.shape
    f = torch.ops.deepmd.prod_force_se_a(g_em.contiguous(), ed, nlist, nloc, ctx.nall)
    gcoord = -f.reshape(ctx.coord_shape)
    if g_rij is not None and bool(torch.any(g_rij != 0)):
        # rij = r_j - r_i: scatter +g to neighbours, -g to centres
        g = g_rij.reshape(nf, nloc, -1, 3).to(torch.float64)
        gc = torch.zeros((nf, ctx.nall, 3), dtype=torch.float64, device=g.device)
        m = nlist >= 0
        gm = torch.where(m[..., None], g, 0.0)
        gc[:, :nloc] -= gm.sum(2)
        idx = torch.where(m, nlist, 0).reshape(nf, -1, 1).expand(-1, -1, 3).long()
        gc.scatter_add_(1, idx, gm.reshape(nf, -1, 3))
        gcoord = gcoord + gc.reshape(ctx.coord_shape).to(gcoord.dtype)
    return gcoord, None, None, None, None, None, None


def _force_setup(ctx, inputs, output):
    net_deriv, in_deriv, nlist, nloc, nall = inputs
    ctx.save_for_backward(in_deriv, nlist)
    ctx.nloc, ctx.nall = nloc, nall
    ctx.nd_dtype = net_deriv.dtype
    ctx.nd_shape = net_deriv.shape


def _force_backward(ctx, g_force):
    in_deriv, nlist = ctx.saved_tensors
    gn = _force_grad_impl(g_force.contiguous(), in_deriv, nlist, ctx.nloc, ctx.nall, ctx.nd_dtype)
    return gn.view(ctx.nd_shape), None, None, None, None


def _virial_setup(ctx, inputs, output):
    net_deriv, in_deriv, rij, nlist, nloc, nall = inputs
    ctx.save_for_backward(net_deriv, in_deriv, rij, nlist)
    ctx.nloc = nloc


def _virial_backward(ctx, g_vir, g_av):
    net_deriv, in_deriv, rij, nlist = ctx.saved_tensors
    if g_av is not None and bool(torch.any(g_av != 0)):
        raise RuntimeError("prod_virial_se_a: atom_virial is not differentiable")
    gn = torch.ops.deepmd.prod_virial_se_a_grad(g_vir.contiguous(), net_deriv, in_deriv, rij,
                                                nlist, ctx.nloc)
    return gn.view(net_deriv.shape), None, None, None, None, None


torch.library.register_autograd("deepmd::prod_env_mat_a", _env_backward, setup_context=_env_setup)
torch.library.register_autograd("deepmd::prod_force_se_a", _force_backward,
                                setup_context=_force_setup)
torch.library.register_autograd("deepmd::prod_virial_se_a", _virial_backward,
                                setup_context=_virial_setup)

"""Hand-written BF16 kernels for the embedding net and the R^T G contraction
of the training step, behind two ``torch.autograd.Function``s.

A force-loss training step (deepmd/pt/model/model/make_model.py:461,
transform_output.py:74-86; deepmd/pt/loss/ener.py:216-330) needs the energy,
its derivative with respect to the environment matrix (the forces), and the
derivative of both with respect to the weights.  Through the descriptor
(deepmd/pt/model/descriptor/se_a.py:784-858) these are two functions of the
embedding weights:

    EmbedX:  xyz = (1/S) sum_j R_j^T G(s_j)                   [nfl, 4, d3]
    EmbedN:  nd  = dE/dem given gx = dE/dxyz                   [nfl, S, 4]

EmbedX.forward / EmbedN.forward are the BF16 inference kernels
(``dpk_embed_fwd`` / ``dpk_embed_bwd``); both backward passes are
``dpk_embed_train_vjp`` (csrc/embed_train.hip): the weight gradient of the
energy path (EmbedX) and of the force path (EmbedN, reverse mode through the
net and its tangent dG/ds, i.e. the second derivative), plus dL/dgx, which
carries the force loss into torch's double backward of the fitting net.
The env matrix is an input that does not depend on the weights: no
gradient with respect to it is formed (the trainer never asks for one).

The weights enter as the packed FP32 parameter matrix of include/dpk.h
(``dpk_embed_net_elems`` per net), built from the trainer's master weights by
one differentiable gather (``EmbedTrainer.packed``), so the gradients land on
the master weights through torch's own index backward.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import DPK_PREC_BF16, call, lib, ptr, stream_ptr
from .model import _P1, _P2, _P3, channel_plans

_ELEMS = 2 * _P1 + _P2 * _P1 + 2 * _P2 + _P3 * _P2 + 2 * _P3


def supported(geo) -> str | None:
    """None when the kernels cover the model's embedding nets, else why not."""
    if geo.emb_has_idt:
        return "embedding layers with a time step (resnet_dt)"
    if geo.emb_act_name != "tanh":
        return f"the embedding activation {geo.emb_act_name!r} (built for tanh)"
    if geo.nnets > 16:
        return f"{geo.nnets} embedding nets (at most 16)"
    return None


def packing_index(neuron, nnets) -> np.ndarray:
    """Gather map [nnets, ELEMS] from the flat vector [w1, b1, w2, b2, w3, b3
    of every net (reference layouts, w [in, out] row-major) ..., 0, 1] to the
    packed parameters of include/dpk.h (internal channel plan of
    ``model.channel_plans``; padding 0, absent time steps 1)."""
    d1, d2, d3 = neuron
    p1, p2, p3, _, _ = channel_plans(neuron)
    per_net = 2 * d1 + d1 * d2 + d2 + d2 * d3 + d3
    zero, one = nnets * per_net, nnets * per_net + 1
    idx = np.full((nnets, _ELEMS), zero, dtype=np.int64)
    o_w2, o_b2 = 2 * _P1, 2 * _P1 + _P2 * _P1
    o_i2, o_w3 = o_b2 + _P2, o_b2 + 2 * _P2
    o_b3 = o_w3 + _P3 * _P2
    o_i3 = o_b3 + _P3
    for k in range(nnets):
        f_w1 = k * per_net
        f_b1 = f_w1 + d1
        f_w2 = f_b1 + d1
        f_b2 = f_w2 + d1 * d2
        f_w3 = f_b2 + d2
        f_b3 = f_w3 + d2 * d3
        row = idx[k]
        row[p1] = f_w1 + np.arange(d1)
        row[_P1 + p1] = f_b1 + np.arange(d1)
        kk, oo = np.meshgrid(np.arange(d1), np.arange(d2), indexing="ij")
        row[o_w2 + p2[oo] * _P1 + p1[kk]] = f_w2 + kk * d2 + oo   # w2t[o][k] = w2[k][o]
        row[o_b2 + p2] = f_b2 + np.arange(d2)
        row[o_i2:o_i2 + _P2] = one
        kk, oo = np.meshgrid(np.arange(d2), np.arange(d3), indexing="ij")
        row[o_w3 + p3[oo] * _P2 + p2[kk]] = f_w3 + kk * d3 + oo
        row[o_b3 + p3] = f_b3 + np.arange(d3)
        row[o_i3:o_i3 + _P3] = one
    return idx


def _iarr(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ctypes.c_void_p)


class EmbedTrainer:
    """Device state of the kernel path for one trainer (one model)."""

    def __init__(self, geo, nets, device):
        why = supported(geo)
        if why is not None:
            raise NotImplementedError(f"BF16 training kernels do not support {why}")
        self.geo = geo
        self.device = torch.device(device)
        self.nnets = geo.nnets
        self.d3 = geo.d3
        self.nnei = geo.nnei
        _, _, _, r2, r3 = channel_plans(geo.neuron)
        self.res2, self.res3 = r2, r3
        self.ch_out = geo.ch_out
        idx = packing_index(geo.neuron, self.nnets).ravel()
        self.idx = torch.tensor(idx, device=self.device)
        # inverse map: every weight appears once in the packed matrix, so the
        # gather's backward is a gather too (the two constants map anywhere)
        inv = np.zeros(int(idx.max()) + 1, dtype=np.int64)  # flat length (the 1 is last)
        inv[idx] = np.arange(idx.size)
        self.inv = torch.tensor(inv, device=self.device)
        self.consts = torch.tensor([0.0, 1.0], dtype=torch.float32, device=self.device)
        self.nets = nets
        L = lib()
        self.img_bytes = L.dpk_embed_image_bytes(_P1, _P2, _P3)
        self.bimg_bytes = L.dpk_embed_train_bimage_bytes()
        props = torch.cuda.get_device_properties(self.device)
        self.ncu = int(props.multi_processor_count)

    # -- parameters -----------------------------------------------------------
    def packed(self) -> torch.Tensor:
        """[nnets, ELEMS] FP32 packed parameters (differentiable gather of the
        master weights)."""
        flat = [p.reshape(-1) for net in self.nets for p in net.params()]
        flat = torch.cat(flat + [self.consts])
        return _Gather.apply(flat, self.idx, self.inv).view(self.nnets, _ELEMS)

    def images(self, P: torch.Tensor):
        st = stream_ptr()
        img = torch.empty(self.nnets * self.img_bytes, dtype=torch.uint8, device=self.device)
        bimg = torch.empty(self.nnets * self.bimg_bytes, dtype=torch.uint8, device=self.device)
        Pc = P.detach().contiguous()
        call("dpk_embed_pack_bf16", ptr(img), ptr(Pc), self.nnets, _P1, _P2, _P3, st)
        call("dpk_embed_train_pack_bwd", ptr(bimg), ptr(Pc), self.nnets, st)
        return img, bimg

    # -- per batch --------------------------------------------------------------
    def bind(self, atype_flat: torch.Tensor) -> "Binding":
        """Launch data of one batch: centre types (int32, device) and each
        net's atom list and slot range (one host read of the types)."""
        return Binding(self, atype_flat)

    # -- kernels ----------------------------------------------------------------
    def fwd(self, bd: "Binding", em, img):
        g = self.geo
        xyz = torch.empty((bd.natoms, 4, self.d3), dtype=torch.float32, device=self.device)
        call("dpk_embed_fwd", ptr(xyz), DPK_PREC_BF16, ptr(em), ptr(bd.atype), None, ptr(img),
             ptr(g.net_of), ptr(self.ch_out), g.ntypes, self.nnets, _P1, _P2, _P3, self.res2,
             self.res3, 0, 0, 0.0, self.d3, g.sec_host[1], bd.natoms, self.nnei, None, None,
             stream_ptr())
        return xyz

    def bwd(self, bd: "Binding", em, gx, img):
        g = self.geo
        nd = torch.empty((bd.natoms, self.nnei, 4), dtype=torch.float32, device=self.device)
        call("dpk_embed_bwd", ptr(nd), DPK_PREC_BF16, ptr(em), ptr(gx), ptr(bd.atype), None,
             ptr(img), ptr(g.net_of), ptr(self.ch_out), g.ntypes, self.nnets, _P1, _P2, _P3,
             self.res2, self.res3, 0, 0, 0.0, self.d3, g.sec_host[1], bd.natoms, self.nnei,
             None, stream_ptr())
        return nd

    def vjp(self, bd: "Binding", em, gmat, img, bimg, vnd=None):
        """dL/dparams [nnets, ELEMS] (and dL/dgx when vnd is given)."""
        gP = torch.empty((self.nnets, _ELEMS), dtype=torch.float32, device=self.device)
        ggx = None
        if vnd is not None:
            ggx = torch.zeros((bd.nblk, bd.natoms, 4, self.d3), dtype=torch.float32,
                              device=self.device)
        call("dpk_embed_train_vjp", ptr(gP), ptr(ggx), ptr(bd.parts), int(vnd is not None),
             ptr(em), ptr(gmat), ptr(vnd), ptr(img), ptr(bimg), ptr(bd.atoms), bd.h_off[1],
             bd.h_b0[1], bd.h_b1[1], bd.h_blk[1], ptr(self.ch_out), self.nnets, bd.nx,
             self.res2, self.res3, 0, self.d3, bd.natoms, self.nnei, stream_ptr())
        if ggx is not None:  # per-block parts (elementwise adds, no reduction kernel)
            out = ggx[0]
            for b in range(1, bd.nblk):
                out = out + ggx[b]
            ggx = out
        return gP, ggx


class Binding:
    """Per-batch launch data of the training kernels."""

    def __init__(self, et: EmbedTrainer, atype_flat: torch.Tensor):
        g = et.geo
        nt = g.ntypes
        self.atype = atype_flat.to(device=et.device, dtype=torch.int32).contiguous()
        self.natoms = int(self.atype.numel())
        at_host = self.atype.cpu().numpy()
        lists, off, b0, b1, blk = [], [0], [], [], []
        for k in range(et.nnets):
            if g.type_one_side:
                tb, atoms = k, np.arange(len(at_host), dtype=np.int32)
            else:
                tb, ti = divmod(k, nt)
                atoms = np.nonzero(at_host == ti)[0].astype(np.int32)
            lists.append(atoms)
            off.append(off[-1] + len(atoms))
            b0.append(int(g.sec[tb]))
            b1.append(int(g.sec[tb + 1]))
            blk.append(tb)
        self.atoms = torch.tensor(np.concatenate(lists) if off[-1] else np.zeros(1, np.int32),
                                  dtype=torch.int32, device=et.device)
        self.h_off, self.h_b0, self.h_b1, self.h_blk = (_iarr(off), _iarr(b0), _iarr(b1),
                                                        _iarr(blk))
        nat_max = max(len(a) for a in lists)
        # one 4-wavefront workgroup per CU (the kernel's LDS), split over the nets
        self.nx = max(1, min(-(-max(nat_max, 1) // 4), -(-et.ncu // et.nnets)))
        n = lib().dpk_embed_train_parts_elems(et.nnets, self.nx)
        self.parts = torch.empty(n, dtype=torch.float32, device=et.device)
        self.nblk = nt


class _Gather(torch.autograd.Function):
    """out = flat[idx] where every non-constant position of flat appears once
    in idx: the backward is the gather grad[inv] (no sort, no atomics)."""

    @staticmethod
    def forward(ctx, flat, idx, inv):
        ctx.save_for_backward(inv)
        return flat[idx]

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, g):
        inv, = ctx.saved_tensors
        return g.reshape(-1)[inv], None, None


class EmbedX(torch.autograd.Function):
    """xyz = (1/S) sum_j R_j^T G_j (se_a.py:836-843) of the packed parameters."""

    @staticmethod
    def forward(ctx, P, em, et: EmbedTrainer, bd: Binding):
        img, bimg = et.images(P)
        ctx.et, ctx.bd, ctx.img, ctx.bimg = et, bd, img, bimg
        ctx.save_for_backward(em)
        return et.fwd(bd, em, img)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gxyz):
        em, = ctx.saved_tensors
        gP, _ = ctx.et.vjp(ctx.bd, em, gxyz.contiguous(), ctx.img, ctx.bimg)
        return gP, None, None, None


class EmbedN(torch.autograd.Function):
    """nd = dE/dem for gx = dE/dxyz (the forces' input), differentiable in the
    parameters and in gx."""

    @staticmethod
    def forward(ctx, P, em, gx, et: EmbedTrainer, bd: Binding):
        gx = gx.contiguous()
        img, bimg = et.images(P)
        ctx.et, ctx.bd, ctx.img, ctx.bimg = et, bd, img, bimg
        ctx.save_for_backward(em, gx)
        return et.bwd(bd, em, gx, img)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, v):
        em, gx = ctx.saved_tensors
        gP, ggx = ctx.et.vjp(ctx.bd, em, gx, ctx.img, ctx.bimg, vnd=v.contiguous())
        return gP, None, ggx, None, None

"""Energy / force / virial training step of the se_e2_a model on MI355X.

The step follows the reference's PyTorch training path (deepmd/pt/train/
training.py -> EnergyStdLoss.forward, deepmd/pt/loss/ener.py:216-330, on the
model of make_model.py:136-220, with the force from autograd and
create_graph=True, make_model.py:461, transform_output.py:74-86):

    em, em_deriv, rij = prod_env_mat_a(extended coords, ...)      (HIP, FP64)
    xyz = R^T G / S                      (embedding + contraction)
    E = fitting(descriptor(xyz))                                  (torch)
    dE/dem  (forces' input, differentiable in the weights)
    F = prod_force_se_a(dE/dem, em_deriv, nlist), folded over images  (HIP)
    V = prod_virial_se_a(dE/dem, em_deriv, rij, nlist)[0]             (HIP)
    loss = pref_e/N MSE(E) + pref_f MSE(F) + pref_v/N MSE(V)

``precision`` "bf16" runs the embedding net and R^T G contraction, dE/dem
and the backward of both -- the weight gradients of the energy path and the
force path's second derivative -- on hand-written BF16 MFMA kernels
(``train_embed``: EmbedX / EmbedN over dpk_embed_fwd, dpk_embed_bwd and
dpk_embed_train_vjp), FP32 master weights; the fitting net runs in torch
(FP32) and its double backward is torch's.  "f32" and "f64" run the
embedding as eager torch GEMMs with autograd (the FP64 trainer is the
reference the kernels are tested against).  The force-loss gradient passes
through the hand-written second-order kernels prod_force_se_a_grad /
prod_virial_se_a_grad (the ops' registered backward) in every mode.
Positions are not differentiated: em_deriv does not depend on the weights.
Geometry (neighbour list, images, env matrix, forces, virials, energy sums)
is FP64.  ``capture_step`` records a whole step (forward, double backward,
Adam) in a HIP graph.
"""

from __future__ import annotations

import contextlib
import copy
import math

import numpy as np
import torch

from . import ops  # noqa: F401  (registers torch.ops.deepmd)
from ._lib import call, lib, ptr, stream_ptr
from .model import SeAModel, _res_mode
from .train_embed import EmbedN, EmbedTrainer, EmbedX


def _t(a, dtype, device, grad=True):
    t = torch.tensor(np.asarray(a, dtype=np.float64), dtype=dtype, device=device)
    return t.requires_grad_(grad)


def _activation(name: str, y):
    """ActivationFn of deepmd/pt/utils/utils.py:174-212 ("none": identity)."""
    n = str(name).lower()
    base, _, arg = n.partition(":")
    if base == "tanh":
        return torch.tanh(y)
    if base == "relu":
        return torch.nn.functional.relu(y)
    if base == "relu6":
        return torch.nn.functional.relu6(y)
    if base == "softplus":
        return torch.nn.functional.softplus(y)
    if base == "sigmoid":
        return torch.sigmoid(y)
    if base == "silu":
        return torch.nn.functional.silu(y)
    if base in ("gelu", "gelu_tf"):
        return torch.nn.functional.gelu(y, approximate="tanh")
    if base in ("silut", "custom_silu"):
        thr = float(arg) if arg else 3.0
        s = 1.0 / (1.0 + math.exp(-thr))
        slope, cst = s + thr * s * (1.0 - s), thr * s
        return torch.where(y >= thr, torch.tanh(slope * (y - thr)) + cst,
                           torch.nn.functional.silu(y))
    return y


class _Net:
    """Trainable MLP in the reference's layer semantics (mlp.py:188-219)."""

    def __init__(self, layers, wdtype, device):
        self.layers = []
        for ly in layers:
            self.layers.append(dict(
                w=_t(ly.w, wdtype, device), b=_t(ly.b, wdtype, device),
                idt=None if ly.idt is None else _t(ly.idt, wdtype, device),
                act=ly.act, res=_res_mode(ly.w.shape[0], ly.w.shape[1], ly.resnet)))

    def params(self):
        out = []
        for ly in self.layers:
            out += [ly["w"], ly["b"]] + ([ly["idt"]] if ly["idt"] is not None else [])
        return out

    def __call__(self, x, gemm_dtype):
        for ly in self.layers:
            w = ly["w"]
            if gemm_dtype != w.dtype:  # BF16 operands, FP32 accumulation inside the GEMM
                y = torch.matmul(x.to(gemm_dtype), w.to(gemm_dtype)).to(w.dtype) + ly["b"]
            else:  # bias in the GEMM (one kernel, and fewer nodes in the double backward)
                y = torch.addmm(ly["b"], x, w) if x.dim() == 2 else torch.matmul(x, w) + ly["b"]
            y = _activation(ly["act"], y)
            if ly["idt"] is not None:
                y = y * ly["idt"]
            if ly["res"] == 1:
                y = y + x
            elif ly["res"] == 2:
                y = y + torch.cat([x, x], dim=-1)
            x = y
        return x


class _Virial(torch.autograd.Function):
    """virial [nf, 9] of prod_virial_se_a (the op's first output) without its
    per-atom virial, which the loss never reads (the op would add it with
    nine FP64 atomics per slot); backward prod_virial_se_a_grad."""

    @staticmethod
    def forward(ctx, nd, in_deriv, rij, nlist, nloc, nall):
        nf = nd.shape[0]
        nnei = nlist.shape[-1]
        vir = torch.empty((nf, 9), dtype=torch.float64, device=nd.device)
        ws = torch.empty(lib().dpk_prod_virial_workspace_bytes(nf, nloc), dtype=torch.uint8,
                         device=nd.device)
        nd = nd.contiguous()
        call("dpk_prod_virial_a_f64", ptr(vir), None, ptr(nd), ptr(in_deriv), ptr(rij),
             ptr(nlist), nf, nloc, nall, nnei, ptr(ws), ws.numel(), stream_ptr())
        ctx.save_for_backward(in_deriv, rij, nlist)
        ctx.nloc = nloc
        return vir

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gv):
        in_deriv, rij, nlist = ctx.saved_tensors
        nf, nnei = nlist.shape[0], nlist.shape[-1]
        gn = torch.empty((nf, ctx.nloc * nnei * 4), dtype=torch.float64, device=gv.device)
        gv = gv.contiguous()
        call("dpk_prod_virial_grad_a_f64", ptr(gn), ptr(gv), ptr(in_deriv), ptr(rij),
             ptr(nlist), nf, ctx.nloc, nnei, stream_ptr())
        return gn, None, None, None, None, None


class _rocblas:
    """The BF16 kernel path's torch GEMMs (fitting net, descriptor products)
    through rocBLAS: for these skinny FP32/FP64 shapes (e.g. the [240, 512] x
    [512, 240] weight gradients) hipBLASLt picks a 256 x 256 macro tile that
    runs on one workgroup (135 us a call on MI355X; the step takes 2.85 ms
    with it, 2.18 ms with rocBLAS).  torch's preferred BLAS is process-wide:
    it is set around each step and restored."""

    def __enter__(self):
        self.prev = torch.backends.cuda.preferred_blas_library()
        torch.backends.cuda.preferred_blas_library("cublas")  # rocBLAS on ROCm

    def __exit__(self, *exc):
        torch.backends.cuda.preferred_blas_library(self.prev)
        return False


class TrainBatch:
    """A uniform batch (nf frames of nloc atoms) with its FP64 geometry."""

    def __init__(self, model: "SeATrainer", coords, atypes, boxes):
        g = model.geo
        coords = np.asarray(coords, dtype=np.float64)
        nf, nloc = coords.shape[0], coords.shape[1]
        self.nf, self.nloc = nf, nloc
        self.sys = g.system(coords, atypes, boxes)
        g.build_nlist(self.sys)
        fs = self.sys.fs
        ns = fs.frames_host[0].ns
        if any(fr.ns != ns or fr.nloc != nloc for fr in fs.frames_host):
            raise ValueError("training batches need frames of equal size and image count")
        self.ns = ns
        self.nall = ns * nloc
        dev = g.device
        wrapped = self.sys.wrapped.view(nf, 1, nloc, 3)
        offs = torch.tensor([fr.img_off for fr in fs.frames_host], device=dev)
        idx = offs.view(nf, 1) + torch.arange(ns, device=dev).view(1, ns)
        shifts = fs.shift_vec[idx.view(-1)].view(nf, ns, 1, 3)
        self.ext = (wrapped + shifts).reshape(nf, self.nall * 3).contiguous()
        at = self.sys.atype.view(nf, 1, nloc).expand(nf, ns, nloc)
        self.ext_atype = at.reshape(nf, self.nall).contiguous()
        self.atype = self.sys.atype.view(nf, nloc).long()
        self.nlist = self.sys.nlist.view(nf, nloc, -1)
        em, ed, rij = torch.ops.deepmd.prod_env_mat_a(self.ext, self.ext_atype, self.nlist,
                                                      g.davg_dev, g.dstd_dev, g.rcut, g.rcut_smth,
                                                      g.env_protection)
        if g.pair_mask is not None:
            # excluded type pairs: zero env-matrix rows (se_a.py:811-819, rr * mask) and
            # zero derivative rows, so they add nothing to forces or virials
            nt = g.ntypes
            tab = g.pair_mask.view(nt + 1, nt + 1)
            nl = self.nlist.long()
            tj = torch.gather(self.ext_atype.long(), 1, nl.clamp(min=0).view(nf, -1))
            tj = torch.where(nl.view(nf, -1) >= 0, tj, nt).view(nf, nloc, -1)
            ti = self.sys.atype.view(nf, nloc, 1).long().expand_as(tj)
            mask = tab[ti, tj].to(em.dtype)
            em = (em.view(nf, nloc, -1, 4) * mask.unsqueeze(-1)).reshape(nf, -1)
            ed = (ed.view(nf, nloc, -1, 12) * mask.unsqueeze(-1)).reshape(nf, -1)
        self.em = em.view(nf, nloc, -1, 4)
        self.em_deriv, self.rij = ed, rij
        # rows of each (centre type) for the per-type nets
        flat = self.atype.reshape(-1)
        self.rows = [torch.nonzero(flat == t).view(-1) for t in range(g.ntypes)]
        self.em32 = self.em.detach().to(torch.float32).reshape(nf * nloc, -1, 4).contiguous()
        self._binding = None

    def kernel_binding(self, kernels):
        """Launch data of the training kernels for this batch (built once)."""
        if self._binding is None or self._binding[0] is not kernels:
            self._binding = (kernels, kernels.bind(self.atype.reshape(-1)))
        return self._binding[1]


class SeATrainer:
    """Trainable se_e2_a energy model and its loss (EnergyStdLoss, mse).

    ``pref_e/f/v`` are the start prefactors; with ``limit_pref_*`` and
    ``start_lr`` given, a term's prefactor at learning rate lr is
    limit + (start - limit) lr / start_lr (ener.py:245-251), otherwise it is
    constant."""

    def __init__(self, model_dict: dict, precision: str = "f64", device="cuda",
                 pref_e: float = 1.0, pref_f: float = 1.0, pref_v: float = 1.0,
                 lr: float = 1e-3, limit_pref_e: float | None = None,
                 limit_pref_f: float | None = None, limit_pref_v: float | None = None,
                 start_lr: float | None = None, capturable: bool = False):
        self.src = model_dict
        self.geo = SeAModel(model_dict, "f64", device)
        g = self.geo
        if g.nx:
            raise NotImplementedError("training with frame / atomic fitting parameters "
                                      "(numb_fparam / numb_aparam > 0) is not built; "
                                      "evaluation supports them")
        self.device = g.device
        self.precision = precision
        self.wdtype = torch.float64 if precision == "f64" else torch.float32
        self.gemm_dtype = {"f64": torch.float64, "f32": torch.float32,
                           "bf16": torch.bfloat16}[precision]
        self.emb = [_Net(n, self.wdtype, self.device) for n in g.emb_layers]
        self.fit = [_Net(n, self.wdtype, self.device) for n in g.fit_layers]
        # bf16: the embedding net and R^T G contraction, their backward and the
        # force path's second derivative run as hand-written BF16 MFMA kernels
        self.kernels = EmbedTrainer(g, self.emb, self.device) if precision == "bf16" else None
        self.bias = torch.tensor(g.bias_atom_e.reshape(-1) + g.out_bias.reshape(-1),
                                 dtype=torch.float64, device=self.device)
        self.pref = (pref_e, pref_f, pref_v)
        lim = (limit_pref_e, limit_pref_f, limit_pref_v)
        self.limit_pref = tuple(p if l is None else l for p, l in zip(self.pref, lim))
        self.start_lr = start_lr
        # a term is on when both of its prefactors are nonzero (ener.py:160-162)
        self.has = tuple(a != 0.0 and b != 0.0 for a, b in zip(self.pref, self.limit_pref))
        # capturable: Adam keeps its step count on the device, so that whole
        # training steps can be captured in a HIP graph (capture_step)
        # (with the fused multi-tensor Adam: a handful of launches per step)
        self.opt = torch.optim.Adam(self.parameters(), lr=lr, capturable=capturable,
                                    fused=True if capturable else None)

    def _blas(self):
        """rocBLAS around the kernel path's steps (see _rocblas); the eager
        f32 / f64 paths keep torch's default (their batched embedding
        products run slower on rocBLAS)."""
        return _rocblas() if self.kernels is not None else contextlib.nullcontext()

    def parameters(self):
        out = []
        for n in self.emb + self.fit:
            out += n.params()
        return out

    def batch(self, coords, atypes, boxes) -> TrainBatch:
        return TrainBatch(self, coords, atypes, boxes)

    # -- model ----------------------------------------------------------------
    def _xyz(self, b: TrainBatch, em):
        """R^T G / S [nf*nloc, 4, d3] from em by eager torch (se_a.py:784-843)."""
        g = self.geo
        nf, nloc, nnei = b.nf, b.nloc, g.nnei
        rd = self.wdtype
        x = em.to(rd).view(nf * nloc, nnei, 4)
        sec = np.concatenate([[0], np.cumsum(g.sel)])
        nt = g.ntypes
        xyz = torch.zeros((nf * nloc, 4, g.d3), dtype=rd, device=self.device)
        for tb in range(nt):
            blk = x[:, sec[tb]:sec[tb + 1], :]
            if g.type_one_side:
                gg = self.emb[tb](blk[:, :, :1], self.gemm_dtype)
                xyz = xyz + torch.matmul(blk.transpose(1, 2), gg)
            else:
                for ti in range(nt):
                    rows = b.rows[ti]
                    if rows.numel() == 0:
                        continue
                    bi = blk.index_select(0, rows)
                    gg = self.emb[tb * nt + ti](bi[:, :, :1], self.gemm_dtype)
                    xyz = xyz.index_add(0, rows, torch.matmul(bi.transpose(1, 2), gg))
        return xyz / nnei

    def _fit(self, b: TrainBatch, xyz):
        """Per-atom energies [nf, nloc] (FP64) from xyz (se_a.py:844-850,
        fitting.py:734-899)."""
        g = self.geo
        nf, nloc, nt = b.nf, b.nloc, g.ntypes
        xyz = xyz.to(self.wdtype)
        desc = torch.matmul(xyz.transpose(1, 2), xyz[:, :, :g.axis_neuron])
        desc = desc.reshape(nf * nloc, -1)
        ae = torch.zeros(nf * nloc, dtype=torch.float64, device=self.device)
        for t in range(nt):
            rows = b.rows[t]
            if rows.numel() == 0 or t in g.atom_exclude_types:
                continue
            y = self.fit[t](desc.index_select(0, rows), self.wdtype).reshape(-1)
            ae = ae.index_add(0, rows, y.to(torch.float64) + self.bias[t])
        return ae.view(nf, nloc)

    def _energy(self, b: TrainBatch, em):
        """Per-atom energies [nf, nloc] (FP64) from em (eager torch)."""
        return self._fit(b, self._xyz(b, em))

    def predict(self, b: TrainBatch, create_graph: bool = True):
        """energy [nf], atom_energy [nf, nloc], force [nf, nloc, 3], virial [nf, 9]."""
        if self.kernels is not None:
            # xyz and nd = dE/dem by the BF16 kernels; the fitting's first
            # derivative gx = dE/dxyz by torch, kept in the graph so that the
            # force loss reaches the fitting weights (and, through EmbedN's
            # backward, the embedding weights)
            bd = b.kernel_binding(self.kernels)
            em = b.em32
            P = self.kernels.packed()
            xyz = EmbedX.apply(P, em, self.kernels, bd)
            ae = self._fit(b, xyz)
            energy = ae.sum(dim=1)
            gx, = torch.autograd.grad(energy.sum(), xyz, create_graph=create_graph)
            nd = EmbedN.apply(P, em, gx, self.kernels, bd)
        else:
            em = b.em.detach().requires_grad_(True)
            ae = self._energy(b, em)
            energy = ae.sum(dim=1)
            nd, = torch.autograd.grad(energy.sum(), em, create_graph=create_graph)
        nd = nd.reshape(b.nf, -1).to(torch.float64)
        fe = torch.ops.deepmd.prod_force_se_a(nd, b.em_deriv, b.nlist, b.nloc, b.nall)
        force = fe.view(b.nf, b.ns, b.nloc, 3).sum(dim=1)
        vir = _Virial.apply(nd, b.em_deriv, b.rij, b.nlist, b.nloc, b.nall)
        return {"energy": energy, "atom_energy": ae, "force": force, "virial": vir}

    def loss(self, pred, label, natoms: int):
        """ener.py:216-330, mse, no mask, intensive_ener_virial False."""
        pe, pf, pv = self.pref
        norm = 1.0 / natoms
        out = torch.zeros((), dtype=torch.float64, device=self.device)
        if pe:
            out = out + norm * pe * torch.mean(torch.square(pred["energy"] - label["energy"]))
        if pf:
            out = out + pf * torch.mean(torch.square(label["force"] - pred["force"]).reshape(-1))
        if pv:
            out = out + norm * pv * torch.mean(torch.square(label["virial"].reshape(-1, 9)
                                                            - pred["virial"].reshape(-1, 9)))
        return out

    def prefactors(self, lr: float | None = None):
        """(pref_e, pref_f, pref_v) at learning rate ``lr`` (ener.py:245-251)."""
        if lr is None or self.start_lr is None:
            return self.pref
        coef = lr / self.start_lr
        return tuple(l + (s - l) * coef for s, l in zip(self.pref, self.limit_pref))

    def loss_terms(self, pred, label, natoms: int, lr: float | None = None):
        """Loss and display terms of EnergyStdLoss.forward (ener.py:245-330,
        461-476, 640-663, 835; mse, no atom mask, intensive_ener_virial
        False).  ``label`` may carry ``find_energy`` / ``find_force`` /
        ``find_virial`` (1.0 when the system has the property, default 1.0);
        a missing property adds nothing and displays nan.  Returns (loss,
        {"rmse", "rmse_e", "rmse_f"[, "rmse_v"]}) as device tensors."""
        pe, pf, pv = self.prefactors(lr)
        he, hf, hv = self.has
        norm = 1.0 / natoms
        out = torch.zeros((), dtype=torch.float64, device=self.device)
        more = {}
        nan = torch.tensor(float("nan"), dtype=torch.float64, device=self.device)
        if he and "energy" in label:
            find = float(label.get("find_energy", 1.0))
            l2 = torch.mean(torch.square(pred["energy"] - label["energy"].reshape(-1)))
            out = out + norm * (pe * find * l2)
            more["rmse_e"] = l2.detach().sqrt() * norm if find else nan
        if hf and "force" in label:
            find = float(label.get("find_force", 1.0))
            l2 = torch.mean(torch.square(label["force"].reshape(-1) - pred["force"].reshape(-1)))
            out = out + pf * find * l2
            more["rmse_f"] = l2.detach().sqrt() if find else nan
        if hv and "virial" in label:
            find = float(label.get("find_virial", 1.0))
            l2 = torch.mean(torch.square(label["virial"].reshape(-1, 9)
                                         - pred["virial"].reshape(-1, 9)))
            out = out + norm * (pv * find * l2)
            more["rmse_v"] = l2.detach().sqrt() * norm if find else nan
        more["rmse"] = out.detach().sqrt()
        return out, more

    def model_dict(self) -> dict:
        """The model in the serialized layout with the current weights (the
        input dict with each layer's w / b / idt replaced)."""
        md = copy.deepcopy(self.src)
        m = md["model"] if "model" in md else md
        colls = ((m["descriptor"]["embeddings"]["networks"], self.emb),
                 (m["fitting"]["nets"]["networks"], self.fit))
        for nets, trained in colls:
            for net, tn in zip(nets, trained):
                for ly, t in zip(net["layers"], tn.layers):
                    var = ly["@variables"]
                    for key, name in (("w", "w"), ("b", "b"), ("idt", "idt")):
                        if t[name] is not None:
                            val = t[name].detach().to("cpu", torch.float64).numpy()
                            if isinstance(var.get(key), dict):
                                var[key] = dict(var[key], value=val.tolist())
                            else:
                                var[key] = val.tolist()
        return md

    def loss_and_grads(self, b: TrainBatch, label):
        with self._blas():
            pred = self.predict(b, create_graph=True)
            loss = self.loss(pred, label, b.nloc)
            grads = torch.autograd.grad(loss, self.parameters())
        return loss, grads, pred

    def step(self, b: TrainBatch, label):
        """One optimiser step (Adam); returns the loss (device tensor)."""
        self.opt.zero_grad(set_to_none=True)
        with self._blas():
            pred = self.predict(b, create_graph=True)
            loss = self.loss(pred, label, b.nloc)
            loss.backward()
        self.opt.step()
        return loss.detach()

    def capture_step(self, b: TrainBatch, label, warmup: int = 2) -> "TrainGraph":
        """One training step on ``b`` (forward, force/virial, loss, double
        backward, Adam) captured in a HIP graph.  Needs ``capturable=True``.
        The step reads the batch's device tensors and ``label`` in place:
        copying new frames of the same shape and types into them trains on
        those.  ``warmup`` eager steps run first (on a side stream, as graph
        capture requires); they are real optimiser steps.  No autograd graph
        of this trainer (a loss or prediction with grad_fn) may be alive when
        it is called: the weights' gradient accumulators must be created on
        the capture stream."""
        if not self.opt.defaults.get("capturable", False):
            raise RuntimeError("capture_step needs a trainer built with capturable=True")
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            for _ in range(max(warmup, 1)):
                self.step(b, label)
        torch.cuda.current_stream(self.device).wait_stream(side)
        graph = torch.cuda.CUDAGraph()
        self.opt.zero_grad(set_to_none=True)
        with torch.cuda.graph(graph), self._blas():
            pred = self.predict(b, create_graph=True)
            loss = self.loss(pred, label, b.nloc)
            loss.backward()
            self.opt.step()
        return TrainGraph(graph, loss.detach())


class TrainGraph:
    """A captured training step: ``replay()`` runs it and returns the loss
    tensor it writes (valid until the next replay)."""

    def __init__(self, graph, loss):
        self.graph, self.loss = graph, loss

    def replay(self):
        self.graph.replay()
        return self.loss

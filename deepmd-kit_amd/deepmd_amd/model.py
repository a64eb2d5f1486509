"""se_e2_a energy model on MI355X: parameters, packing and the E/F/V step.

The step (``SeAModel.evaluate``) is the DeepPot.eval / forward path of the
reference (deepmd/pt/model/model/make_model.py:136-220 ->
forward_common_lower:385-466 -> se_a.py:742-858 -> fitting.py:734-899 ->
transform_output.py:68-209,367-433) with the autograd backward replaced by
hand-written adjoint kernels:

    nlist (cell list)      -> dpk_nlist_build
    env matrix             -> dpk_env_mat_fused
    embedding + R^T G      -> dpk_embed_fwd          (MFMA)
    G^T R R^T G            -> dpk_fit_descriptor_fwd (rows grouped by type)
    fitting net (+ grad)   -> hipBLASLt GEMMs + dpk_fit_layer_fwd / _bwd, dpk_fit_energy
    dE/dxyz                -> dpk_fit_descriptor_bwd
    dE/d em (recompute)    -> dpk_embed_bwd          (MFMA)
    forces, virial         -> dpk_force_virial_cells + dpk_frame_sum
    energy                 -> dpk_frame_sum (fixed-order FP64 sums per frame)

Parameters use the reference's serialized dict layout (the "model" block of
``source/tests/infer/deeppot_sea.yaml``), so models written by the reference's
``serialize()`` load directly and seeded random models can be written out in
the same form.
"""

from __future__ import annotations

import ctypes
import functools
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import (ACTIVATIONS, DPK_PREC_BF16, DPK_PREC_F16, DPK_PREC_F16X2, DPK_PREC_F32,
                   DPK_PREC_F64, call, lib, ptr, stream_ptr)
from .frames import FrameSet, make_frames

PREC = {"f64": DPK_PREC_F64, "f32": DPK_PREC_F32, "bf16": DPK_PREC_BF16, "f16": DPK_PREC_F16,
        "f16x2": DPK_PREC_F16X2}
# torch dtype of the GEMM operands (descriptor, fitting activations) per precision
GEMM_DTYPE = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16,
              "f16": torch.float16, "f16x2": torch.float16}
# 16-bit MFMA operands, FP32 accumulation; f16x2 splits every data operand
# into two FP16 values (hi + lo, ~22 bits) next to FP16 weights
HALF = ("bf16", "f16", "f16x2")
# the net image each 16-bit precision reads (dpk_embed_pack_*)
IMAGE = {"bf16": "bf16", "f16": "f16", "f16x2": "f16"}
_P1, _P2, _P3 = 32, 64, 128


# ------------------------------------------------------------ model dicts --

def _nd(a):
    return {"@class": "np.ndarray", "@is_variable": True, "@version": 1,
            "dtype": "float64", "value": None if a is None else np.asarray(a).tolist()}


def _layer_dict(w, b, idt, act, resnet):
    return {"@class": "Layer", "@version": 1, "activation_function": act, "bias": True,
            "precision": "float64", "resnet": resnet, "use_timestep": idt is not None,
            "@variables": {"w": _nd(w), "b": _nd(b), "idt": None if idt is None else _nd(idt)}}


def _rand_net(rng, dims, resnet_dt, out_layer, act="tanh"):
    layers = []
    for i in range(len(dims) - 1):
        nin, nout = dims[i], dims[i + 1]
        last = out_layer and i == len(dims) - 2
        w = rng.normal(0.0, 1.0 / math.sqrt(nin), size=(nin, nout))
        b = rng.normal(0.0, 1.0 / math.sqrt(nin), size=(nout,))
        use_ts = resnet_dt and not last and (nout == nin or nout == 2 * nin)
        idt = rng.normal(0.1, 0.001, size=(nout,)) if use_ts else None
        layers.append(_layer_dict(w, b, idt, "none" if last else act, not last))
    return {"layers": layers}


def random_model_dict(seed, rcut, rcut_smth, sel, neuron, axis_neuron, fitting_neuron,
                      type_one_side=True, resnet_dt_emb=False, resnet_dt_fit=True,
                      type_map=None, activation="tanh", fitting_activation=None):
    """Seeded model in the reference's serialized layout.

    Weights W ~ N(0, 1/sqrt(fan_in)), biases ~ N(0, 1/sqrt(fan_in)), resnet
    time steps idt ~ N(0.1, 0.001) (deepmd/pt/model/network/mlp.py:162-163),
    davg = 0, dstd = 1, bias_atom_e = out_bias = 0.  numpy PCG64 order:
    embedding nets (NetworkCollection order) then fitting nets by type."""
    rng = np.random.default_rng(seed)
    ntypes = len(sel)
    nnei = int(sum(sel))
    nemb = ntypes if type_one_side else ntypes * ntypes
    fact = activation if fitting_activation is None else fitting_activation
    emb = [_rand_net(rng, [1] + list(neuron), resnet_dt_emb, False, activation)
           for _ in range(nemb)]
    din = neuron[-1] * axis_neuron
    fit = [_rand_net(rng, [din] + list(fitting_neuron) + [1], resnet_dt_fit, True, fact)
           for _ in range(ntypes)]
    tm = type_map or [str(i) for i in range(ntypes)]
    return {"model": {
        "@class": "Model", "type": "standard", "type_map": tm,
        "@variables": {"out_bias": _nd(np.zeros((1, ntypes, 1))),
                       "out_std": _nd(np.ones((1, ntypes, 1)))},
        "descriptor": {
            "@class": "Descriptor", "type": "se_e2_a", "rcut": float(rcut),
            "rcut_smth": float(rcut_smth), "sel": list(sel), "neuron": list(neuron),
            "axis_neuron": int(axis_neuron), "type_one_side": bool(type_one_side),
            "resnet_dt": bool(resnet_dt_emb), "activation_function": activation,
            "env_protection": 0.0, "exclude_types": [], "type_map": tm,
            "@variables": {"davg": _nd(np.zeros((ntypes, nnei, 4))),
                           "dstd": _nd(np.ones((ntypes, nnei, 4)))},
            "embeddings": {"@class": "NetworkCollection", "ndim": 1 if type_one_side else 2,
                           "ntypes": ntypes, "networks": emb},
        },
        "fitting": {
            "@class": "Fitting", "type": "ener", "neuron": list(fitting_neuron),
            "activation_function": fact,
            "resnet_dt": bool(resnet_dt_fit), "dim_descrpt": din, "mixed_types": False,
            "@variables": {"bias_atom_e": _nd(np.zeros((ntypes, 1)))},
            "nets": {"@class": "NetworkCollection", "ndim": 1, "ntypes": ntypes,
                     "networks": fit},
        },
    }}


def _arr(v):
    if v is None:
        return None
    if isinstance(v, dict):
        v = v.get("value")
    return None if v is None else np.asarray(v, dtype=np.float64)


@dataclass
class Layer:
    w: np.ndarray
    b: np.ndarray
    idt: np.ndarray | None
    act: str
    resnet: bool


def _layers(net):
    out = []
    for ly in net["layers"]:
        var = ly["@variables"]
        act = ly.get("activation_function", "tanh")
        act = "none" if act in (None, "none", "linear") else str(act).lower()
        act_code(act)  # raises for an unknown activation
        out.append(Layer(_arr(var["w"]), _arr(var.get("b")), _arr(var.get("idt")), act,
                         bool(ly.get("resnet", False))))
    return out


# ------------------------------------------------------------ activations --

def act_code(name: str) -> tuple[int, float]:
    """(DPK_ACT_* code, silut threshold) of an activation name of
    deepmd/pt/utils/utils.py:174-212 ("silut:2.5" / "custom_silu:2.5" carry
    their threshold, default 3.0)."""
    n = "none" if name is None else str(name).lower()
    base, _, arg = n.partition(":")
    if base not in ACTIVATIONS:
        raise NotImplementedError(f"activation {name!r} is not supported")
    thr = float(arg) if arg and base in ("silut", "custom_silu") else 3.0
    return ACTIVATIONS[base], thr


def act_np(name: str, z: np.ndarray) -> np.ndarray:
    """The activation in numpy (FP64), the same functions as the kernels."""
    code, thr = act_code(name)
    sig = lambda x: 1.0 / (1.0 + np.exp(-x))  # noqa: E731
    if code == ACTIVATIONS["tanh"]:
        return np.tanh(z)
    if code == ACTIVATIONS["relu"]:
        return np.maximum(z, 0.0)
    if code == ACTIVATIONS["relu6"]:
        return np.clip(z, 0.0, 6.0)
    if code == ACTIVATIONS["softplus"]:
        return np.where(z > 20.0, z, np.log1p(np.exp(np.minimum(z, 20.0))))
    if code == ACTIVATIONS["sigmoid"]:
        return sig(z)
    if code == ACTIVATIONS["silu"]:
        return z * sig(z)
    if code == ACTIVATIONS["gelu"]:
        return 0.5 * z * (1.0 + np.tanh(0.7978845608028654 * (z + 0.044715 * z ** 3)))
    if code == ACTIVATIONS["silut"]:
        s = sig(thr)
        slope, cst = s + thr * s * (1.0 - s), thr * s
        return np.where(z >= thr, np.tanh(slope * (z - thr)) + cst, z * sig(z))
    return z


def uniform_act(nets, what: str, hidden_only: bool = False) -> str:
    """The one activation of a set of nets' (hidden) layers."""
    acts = {ly.act for layers in nets for ly in (layers[:-1] if hidden_only else layers)}
    if len(acts) > 1:
        raise NotImplementedError(f"{what} layers with different activations {sorted(acts)}")
    return acts.pop() if acts else "tanh"


# ------------------------------------------------------- embedding packing --

def _res_mode(nin, nout, resnet):
    if not resnet:
        return 0
    if nout == nin:
        return 1
    if nout == 2 * nin:
        return 2
    return 0


def channel_plans(neuron):
    """Internal (16-padded) channel position of every true channel, per layer,
    such that resnet copies (y + x, y + (x, x)) stay lane-local."""
    d1, d2, d3 = neuron
    if d1 > _P1:
        raise NotImplementedError(f"embedding width d1={d1} > {_P1}")
    p1 = np.arange(d1)
    r2 = _res_mode(d1, d2, True)
    if r2 == 2:
        p2 = np.array([h * _P1 + p1[c] for h in range(2) for c in range(d1)])
    elif r2 == 1:
        raise NotImplementedError("embedding layer 2 with d2 == d1 is not supported")
    else:
        if d2 > _P2:
            raise NotImplementedError(f"embedding width d2={d2} > {_P2}")
        p2 = np.arange(d2)
    r3 = _res_mode(d2, d3, True)
    if r3 == 2:
        p3 = np.array([h * _P2 + p2[c] for h in range(2) for c in range(d2)])
    elif r3 == 1:
        raise NotImplementedError("embedding layer 3 with d3 == d2 is not supported")
    else:
        if d3 > _P3:
            raise NotImplementedError(f"embedding width d3={d3} > {_P3}")
        p3 = np.arange(d3)
    return p1, p2, p3, r2, r3


def pack_embedding(nets, neuron):
    """Packed real parameters [n_nets, elems] (float64) and the internal ->
    true output channel map (include/dpk.h)."""
    p1, p2, p3, r2, r3 = channel_plans(neuron)
    d1, d2, d3 = neuron
    elems = 2 * _P1 + _P2 * _P1 + 2 * _P2 + _P3 * _P2 + 2 * _P3
    out = np.zeros((len(nets), elems))
    for k, layers in enumerate(nets):
        if len(layers) != 3:
            raise NotImplementedError("embedding nets must have 3 layers")
        l1, l2, l3 = layers
        if l1.w.shape != (1, d1) or l2.w.shape != (d1, d2) or l3.w.shape != (d2, d3):
            raise ValueError("embedding weights do not match neuron sizes")
        if l1.idt is not None or _res_mode(1, d1, l1.resnet) != 0:
            raise NotImplementedError("embedding layer 1 with resnet / time step")
        v = out[k]
        o = 0
        v[o + p1] = l1.w[0]; o += _P1
        v[o + p1] = l1.b; o += _P1
        w2t = np.zeros((_P2, _P1)); w2t[np.ix_(p2, p1)] = l2.w.T
        v[o:o + _P2 * _P1] = w2t.ravel(); o += _P2 * _P1
        v[o + p2] = l2.b; o += _P2
        v[o:o + _P2] = 1.0
        if l2.idt is not None:
            v[o + p2] = l2.idt
        o += _P2
        w3t = np.zeros((_P3, _P2)); w3t[np.ix_(p3, p2)] = l3.w.T
        v[o:o + _P3 * _P2] = w3t.ravel(); o += _P3 * _P2
        v[o + p3] = l3.b; o += _P3
        v[o:o + _P3] = 1.0
        if l3.idt is not None:
            v[o + p3] = l3.idt
        o += _P3
        assert o == elems
    ch_out = -np.ones(_P3, dtype=np.int32)
    ch_out[p3] = np.arange(d3, dtype=np.int32)
    return out, ch_out, r2, r3


def embed_internal(packed_k, s, r2, r3, act="tanh"):
    """The packed net k (pack_embedding layout, internal channels) evaluated
    in FP64 at the scalars s [n] -> G [n, 128]."""
    o = 0
    w1 = packed_k[o:o + _P1]; o += _P1
    b1 = packed_k[o:o + _P1]; o += _P1
    w2t = packed_k[o:o + _P2 * _P1].reshape(_P2, _P1); o += _P2 * _P1
    b2 = packed_k[o:o + _P2]; o += _P2
    i2 = packed_k[o:o + _P2]; o += _P2
    w3t = packed_k[o:o + _P3 * _P2].reshape(_P3, _P2); o += _P3 * _P2
    b3 = packed_k[o:o + _P3]; o += _P3
    i3 = packed_k[o:o + _P3]
    y1 = act_np(act, np.outer(s, w1) + b1)
    y2 = act_np(act, y1 @ w2t.T + b2) * i2 + (np.tile(y1, 2) if r2 == 2 else 0.0)
    return act_np(act, y2 @ w3t.T + b3) * i3 + (np.tile(y2, 2) if r3 == 2 else 0.0)


def padding_table(packed, net_of, davg, dstd, sec, r2, r3, act="tanh"):
    """pad[ti, jj, c, k]: sum over the slots jj..end of jj's type block of
    R[c] G[k] for the padding rows R = (0 - davg) / dstd of centre type ti
    (G = their net at s = R[0], internal channel k), or None when every
    padding row is zero (include/dpk.h, dpk_embed_fwd)."""
    nt, nnei = davg.shape[0], davg.shape[1]
    R = -davg / dstd  # [nt, nnei, 4]
    if not np.any(R):
        return None
    pad = np.zeros((nt, nnei, 4, _P3))
    for ti in range(nt):
        for t in range(len(sec) - 1):
            a, b = int(sec[t]), int(sec[t + 1])
            if b <= a:
                continue
            G = embed_internal(packed[net_of[ti, t]], R[ti, a:b, 0], r2, r3, act)  # [b-a, 128]
            C = R[ti, a:b, :, None] * G[:, None, :]
            pad[ti, a:b] = np.cumsum(C[::-1], axis=0)[::-1]
    return pad.astype(np.float32)


# ---------------------------------------------------------------- fitting --

class Fitting:
    """Energy fitting nets, one per type, each evaluated only on the atoms of
    its type (the reference runs every net on every atom and masks,
    fitting.py:877-893).  Atoms are processed in "slot" order (grouped by
    type, ``System.order``) so that each type's layer is one contiguous
    hipBLASLt GEMM; bias, tanh, time step, resnet and the derivative factors
    are fused in libdpk kernels (dpk_fit_layer_fwd / _bwd), and the energy dot
    product and the descriptor in/out in dpk_fit_energy /
    dpk_fit_descriptor_fwd / _bwd.  Hidden layers: tanh, optional idt, resnet
    0 / 1 / 2; output layer: linear to one value (the reference's
    EnergyFittingNet, network.py:935-1006).  16-bit (BF16 / FP16) nets of 2
    or 3 hidden layers of (padded) width 256 run layers 2..L, the energy and
    the backward down to dE/dz_1 in one kernel (dpk_fit_tail_bf16 / _f16)
    between the layer-1
    GEMM and the dE/dD GEMM."""

    # Uniform hidden widths above 128 are padded to a multiple of 128 (240 ->
    # 256): zero weight columns / rows, bias and output weight, so the padded
    # activations and their gradients are exactly 0, and the GEMMs get
    # tile-aligned N and K (hipBLASLt runs 240-wide BF16 products at ~60% of
    # the 256-wide rate).
    PAD = 128
    # fused 16-bit tail (dpk_fit_tail_bf16 / _f16) where the shapes allow it
    FUSED_TAIL = True
    # split FP16: descriptor + layer 1 (and their backward) in one kernel each
    # (dpk_fit_l1_fwd / _bwd) instead of writing D and dE/dD around two GEMMs
    FUSED_L1 = True

    def __init__(self, nets, bias_atom_e, out_bias, prec: str, device, nx: int = 0):
        self.prec = prec
        # nx trailing rows of every first-layer W hold the frame / atomic
        # parameter weights (the reference concatenates the normalised
        # parameters to the descriptor, fitting.py:774-830): they enter as
        # z1 = D W1[:Din] + X W1[Din:], so the descriptor GEMM and its backward
        # never see them
        self.nx = int(nx)
        self.prec_code = PREC[prec]
        self.device = device
        self.gemm_dtype = GEMM_DTYPE[prec]
        self.real_dtype = torch.float64 if prec == "f64" else torch.float32
        rd, gd = self.real_dtype, self.gemm_dtype
        self.nets = []
        widths = None
        # hidden-layer activation (utils.py:174-212): tanh in every precision,
        # the others in f32 / f64 (the 16-bit tails fold tanh in)
        self.act_name = uniform_act(nets, "fitting hidden", hidden_only=True)
        self.act, self.act_param = act_code(self.act_name)
        if self.act != ACTIVATIONS["tanh"] and prec not in ("f32", "f64"):
            raise NotImplementedError(f"fitting activation {self.act_name!r} needs an f32 or "
                                      f"f64 fitting (got {prec})")
        hs = {ly.w.shape[1] for layers in nets for ly in layers[:-1]}
        pw = None
        if len(hs) == 1 and min(hs) > self.PAD:
            pw = -(-min(hs) // self.PAD) * self.PAD

        def padded(a, rows, cols):
            out = np.zeros((rows, cols), dtype=np.float64)
            out[:a.shape[0], :a.shape[1]] = a
            return out

        for layers in nets:
            if len(layers) < 2:
                raise NotImplementedError("fitting nets need at least one hidden layer")
            hidden, out = layers[:-1], layers[-1]
            if out.w.shape[1] != 1 or out.act != "none" or out.idt is not None:
                raise NotImplementedError("fitting output layer must be linear to one value")
            tl = []
            for i, ly in enumerate(hidden):
                res = _res_mode(ly.w.shape[0], ly.w.shape[1], ly.resnet)
                if i == 0 and res != 0:
                    raise NotImplementedError("resnet on the first fitting layer")
                hin, h = ly.w.shape
                wv, bv, iv = ly.w, ly.b, ly.idt
                if pw is not None:
                    hin = hin if i == 0 else pw
                    h = pw
                    wv = padded(ly.w, hin, h)
                    bv = padded(ly.b.reshape(1, -1), 1, h)[0]
                    iv = None if ly.idt is None else padded(ly.idt.reshape(1, -1), 1, h)[0]
                wx = None
                if i == 0 and self.nx:
                    if hin <= self.nx:
                        raise ValueError(f"first fitting layer has {hin} inputs for {self.nx} "
                                         "frame / atomic parameters")
                    # split FP16: the parameter term is a small FP32 GEMM into FP32 z1
                    wdt = torch.float32 if prec == "f16x2" else gd
                    wx = torch.tensor(wv[hin - self.nx:], dtype=wdt, device=device).contiguous()
                    wv, hin = wv[:hin - self.nx], hin - self.nx
                w = torch.tensor(wv, dtype=gd, device=device).contiguous()
                tl.append(dict(
                    w=w, wt=w.t().contiguous(), hin=hin, h=h, res=res, wx=wx,
                    b=torch.tensor(bv, dtype=rd, device=device),
                    idt=None if iv is None else torch.tensor(iv, dtype=rd, device=device)))
            ws = [ly["h"] for ly in tl]
            if widths is not None and ws != widths:
                raise NotImplementedError("fitting nets of different shapes per type")
            widths = ws
            wo = out.w[:, 0] if pw is None else padded(out.w[:, 0].reshape(1, -1), 1, pw)[0]
            self.nets.append(dict(hidden=tl, wout=torch.tensor(wo, dtype=rd, device=device),
                                  bout=float(out.b[0])))
        self.widths = widths
        self.din = self.nets[0]["hidden"][0]["hin"]
        self.ebias = (bias_atom_e.reshape(-1) + out_bias.reshape(-1)).astype(np.float64)
        self.tail = None
        L = len(widths)
        self.split = prec == "f16x2"
        if self.split and not (self.FUSED_TAIL and widths == [256] * L and L in (2, 3)):
            raise NotImplementedError("split FP16 fitting needs 2 or 3 hidden layers of width "
                                      "129..256 (use an f32 fitting)")
        if (self.FUSED_TAIL and prec in HALF and widths == [256] * L and L in (2, 3)
                and lib().dpk_fit_tail_bytes(L) > 0):
            self.tail = [torch.from_numpy(pack_fit_tail(net["hidden"], net["wout"],
                                                        self.gemm_dtype)).to(device)
                         for net in self.nets]
            self.tail_fn = "dpk_fit_tail_" + prec
            if self.split:  # layer 1 as one K = 2 Din GEMM of [D_hi | D_lo] and [W1; W1]
                for net in self.nets:
                    ly = net["hidden"][0]
                    ly["w2"] = torch.cat([ly["w"], ly["w"]], 0).contiguous()
                    ly["wt2"] = torch.cat([ly["wt"], ly["wt"]], 0).contiguous()
            self.tail_res = [sum(1 << (t - 1) for t in range(1, L) if net["hidden"][t]["res"] == 1)
                             for net in self.nets]
            assert all(t.numel() == lib().dpk_fit_tail_bytes(L) for t in self.tail)
        self.l1 = None
        if self.split and self.FUSED_L1 and self.din % 16 == 0:
            d3 = self.din // 16
            nb = lib().dpk_fit_l1_image_bytes(d3, self.widths[0])
            if nb:
                self.l1 = []
                for net in self.nets:
                    img = torch.empty(nb, dtype=torch.uint8, device=device)
                    w1 = net["hidden"][0]["w"].float().contiguous()
                    call("dpk_fit_l1_pack", ptr(img), DPK_PREC_F16X2, ptr(w1), d3, 16,
                         self.widths[0], stream_ptr())
                    self.l1.append(img)

    @staticmethod
    def tail_shape(fit_layers) -> bool:
        """Whether nets of these layers take the fused 16-bit tail (hidden
        widths uniform in 129..256 -- padded to 256 -- and 2 or 3 layers)."""
        hs = [[ly.w.shape[1] for ly in layers[:-1]] for layers in fit_layers]
        flat = {h for w in hs for h in w}
        return (len(flat) == 1 and 128 < min(flat) <= 256 and len({len(w) for w in hs}) == 1
                and len(hs[0]) in (2, 3))

    def buffers(self, n: int):
        """Per-system work buffers for n slots (flat; viewed per layer)."""
        dev, rd, gd = self.device, self.real_dtype, self.gemm_dtype
        hmax = max(self.widths)
        nl = len(self.widths)
        if self.split:  # split FP16: [D_hi | D_lo], FP32 z1 and dE/dD, [gz_hi | gz_lo]
            nd = 1 if self.l1 is not None else n  # the fused layer 1 writes no D, dE/dD
            return {"z": torch.empty(n * hmax, dtype=torch.float32, device=dev),
                    "gz": torch.empty(n * 2 * hmax, dtype=gd, device=dev),
                    "D": torch.empty((nd, 2 * self.din), dtype=gd, device=dev),
                    "GD": torch.empty((nd, self.din), dtype=torch.float32, device=dev),
                    "cx": None}
        bf = {
            "z": torch.empty(n * hmax, dtype=gd, device=dev),
            "y": [torch.empty(n * hmax, dtype=rd, device=dev) for _ in range(2)],
            "yg": torch.empty(n * hmax, dtype=gd, device=dev) if gd != rd else None,
            "sd": [torch.empty(n * hmax, dtype=gd, device=dev) for _ in range(nl)],
            "gy": [torch.empty(n * hmax, dtype=rd, device=dev) for _ in range(2)],
            "gz": torch.empty(n * hmax, dtype=gd, device=dev),
            "mm": torch.empty(n * hmax, dtype=gd, device=dev),
            "D": torch.empty((n, self.din), dtype=gd, device=dev),
            "GD": torch.empty((n, self.din), dtype=gd, device=dev),
            "cx": torch.empty(n * self.widths[0], dtype=gd, device=dev) if self.nx else None,
        }
        return bf

    def first_layer(self, z, D, ly, xe, bf, n, a, b):
        """z1 = D W1 (+ X W1x for the frame / atomic parameters X, slot order)
        for the slots a..b of one type."""
        if self.split:
            torch.mm(D, ly["w2"], out_dtype=torch.float32, out=z)
            if xe is not None and ly["wx"] is not None:
                z.addmm_(xe[a:b], ly["wx"])
            return
        if xe is None or ly["wx"] is None:
            torch.mm(D, ly["w"], out=z)
            return
        cx = bf["cx"][:n * ly["h"]].view(n, ly["h"])[a:b]
        torch.mm(xe[a:b], ly["wx"], out=cx)
        torch.addmm(cx, D, ly["w"], out=z)

    def run(self, atom_e, gxyz, xyz, order, segments, n: int, bf, d3: int, m2: int, st: int,
            xe=None):
        """atom_e [natoms] (FP64, per atom) and gxyz = dE/dxyz [natoms, 4, d3]
        for the n real atoms listed in ``order`` (slot order), ``segments`` =
        [(type, start, end)] slot ranges; ``xe`` [n, nx] the normalised frame /
        atomic parameters in slot order (GEMM dtype) when the nets take them."""
        pc = self.prec_code
        if self.nx and xe is None:
            raise ValueError("the fitting nets take frame / atomic parameters: xe is required")
        fused = self.l1 is not None and m2 == 16 and d3 * 16 == self.din
        if not fused:
            call("dpk_fit_descriptor_fwd", ptr(bf["D"]), pc, ptr(xyz), ptr(order), n, d3, m2, st)

        def view(flat, h, a, b):
            return flat[:n * h].view(n, h)[a:b]

        for t, a, b in segments:
            if b <= a:
                continue
            net = self.nets[t]
            hid = net["hidden"]
            nt = b - a
            if self.tail is not None:
                h1 = hid[0]["h"]
                z = view(bf["z"], h1, a, b)
                if fused:
                    call("dpk_fit_l1_fwd", ptr(z), pc, ptr(xyz), ptr(order[a:b]), nt, d3, m2,
                         ptr(self.l1[t]), h1, st)
                    if xe is not None and hid[0]["wx"] is not None:
                        z.addmm_(xe[a:b], hid[0]["wx"])
                else:
                    self.first_layer(z, bf["D"][a:b], hid[0], xe, bf, n, a, b)
                gz = view(bf["gz"], h1 * (2 if self.split else 1), a, b)
                call(self.tail_fn, ptr(atom_e), ptr(gz), ptr(z), ptr(self.tail[t]),
                     len(hid), self.tail_res[t], net["bout"] + float(self.ebias[t]),
                     ptr(order[a:b]), nt, st)
                if fused:
                    call("dpk_fit_l1_bwd", ptr(gxyz), pc, ptr(gz), ptr(xyz), ptr(order[a:b]), nt,
                         d3, m2, ptr(self.l1[t]), h1, st)
                elif self.split:
                    torch.mm(gz, hid[0]["wt2"], out_dtype=torch.float32, out=bf["GD"][a:b])
                else:
                    torch.mm(gz, hid[0]["wt"], out=bf["GD"][a:b])
                continue
            # forward
            x = bf["D"][a:b]
            yi = 0
            for l, ly in enumerate(hid):
                h = ly["h"]
                z = view(bf["z"], h, a, b)
                if l == 0:
                    self.first_layer(z, x, ly, xe, bf, n, a, b)
                else:
                    torch.mm(x, ly["w"], out=z)
                y = view(bf["y"][yi], h, a, b)
                yprev = view(bf["y"][1 - yi], ly["hin"], a, b) if ly["res"] else None
                yg = view(bf["yg"], h, a, b) if bf["yg"] is not None else None
                sd = view(bf["sd"][l], h, a, b)
                call("dpk_fit_layer_fwd", ptr(y), ptr(yg), ptr(sd), pc, ptr(z), ptr(ly["b"]),
                     ptr(ly["idt"]), ptr(yprev), nt, h, ly["hin"], ly["res"], self.act,
                     self.act_param, st)
                x = yg if yg is not None else y
                yi = 1 - yi
            hl = hid[-1]["h"]
            y = view(bf["y"][1 - yi], hl, a, b)
            gz = view(bf["gz"], hl, a, b)
            call("dpk_fit_energy", ptr(atom_e), ptr(gz), pc, ptr(y), ptr(view(bf["sd"][-1], hl, a, b)),
                 ptr(net["wout"]), net["bout"] + float(self.ebias[t]), ptr(order[a:b]), nt, hl, st)
            # backward: g_y_{l-1} = g_z_l W_l^T + res_l^T(g_y_l)
            gi = 0
            for l in range(len(hid) - 1, 0, -1):
                ly = hid[l]
                hp = ly["hin"]
                mm = view(bf["mm"], hp, a, b)
                torch.mm(gz, ly["wt"], out=mm)
                gy = view(bf["gy"][gi], hp, a, b)
                gnext = None if l == len(hid) - 1 else view(bf["gy"][1 - gi], ly["h"], a, b)
                wvec = net["wout"] if l == len(hid) - 1 else None
                gz = view(bf["gz"], hp, a, b)
                call("dpk_fit_layer_bwd", ptr(gy) if l > 1 else None, ptr(gz), pc, ptr(mm),
                     ptr(gnext), ptr(wvec), ptr(view(bf["sd"][l - 1], hp, a, b)), nt, hp,
                     ly["h"], ly["res"], st)
                gi = 1 - gi
            torch.mm(gz, hid[0]["wt"], out=bf["GD"][a:b])
        if not fused:
            call("dpk_fit_descriptor_bwd", ptr(gxyz), pc, ptr(bf["GD"]), ptr(xyz), ptr(order), n,
                 d3, m2, st)


def _tail_fragment_index():
    """(row, column) of a 256 x 256 matrix for every entry of its packed
    BF16 MFMA image: entry ((mb*8 + kc)*64 + 16*g + m)*8 + j holds row
    16*mb + m, column 32*kc + 16*(j//4) + 4*g + j%4."""
    mb = np.arange(16)[:, None, None, None, None]
    kc = np.arange(8)[None, :, None, None, None]
    g = np.arange(4)[None, None, :, None, None]
    m = np.arange(16)[None, None, None, :, None]
    j = np.arange(8)[None, None, None, None, :]
    row = np.broadcast_to(16 * mb + m, (16, 8, 4, 16, 8)).reshape(-1)
    col = np.broadcast_to(32 * kc + 16 * (j // 4) + 4 * g + j % 4, (16, 8, 4, 16, 8)).reshape(-1)
    return row, col


def pack_fit_tail(hidden, wout, dtype=torch.bfloat16) -> np.ndarray:
    """Device image of dpk_fit_tail_bf16 / _f16 for one net (hidden layers of
    width 256 with 16-bit weights of ``dtype``): W_l^T then W_l for l = 2..L as
    fragments, then FP32 b_1, idt_1, (b_l, idt_l), w_out."""
    row, col = _tail_fragment_index()

    def h16_bytes(a):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(
            dtype).view(torch.int16).numpy().tobytes()

    ws = [ly["w"].float().cpu().numpy() for ly in hidden[1:]]  # exact 16-bit values, [in, out]
    fwd = [h16_bytes(w.T[row, col]) for w in ws]  # A = W^T: rows = outputs, K = inputs
    bwd = [h16_bytes(w[row, col]) for w in ws]    # A = W:   rows = inputs,  K = outputs
    vecs = []
    for ly in hidden:
        vecs.append(ly["b"].float().cpu().numpy())
        vecs.append(np.ones(ly["h"], np.float32) if ly["idt"] is None
                    else ly["idt"].float().cpu().numpy())
    vecs.append(wout.float().cpu().numpy())
    vb = np.concatenate(vecs).astype(np.float32).tobytes()
    return np.frombuffer(b"".join(fwd + bwd) + vb, dtype=np.uint8).copy()


# ------------------------------------------------------------------ model --

@dataclass
class System:
    """A packed batch (frames may differ in size) with preallocated buffers."""
    coord: torch.Tensor          # [N, 3] float64
    atype: torch.Tensor          # [N] int32
    fs: FrameSet
    order: torch.Tensor          # [nreal] int32: real atoms grouped by type (slot order)
    segments: list               # [(type, slot start, slot end)]
    nreal: int
    nlist: torch.Tensor = None   # [N, nnei] int32
    wrapped: torch.Tensor = None
    status: torch.Tensor = None
    ws_nlist: torch.Tensor = None
    em: torch.Tensor = None
    xyz: torch.Tensor = None
    desc: torch.Tensor = None
    gxyz: torch.Tensor = None
    nd: torch.Tensor = None
    force: torch.Tensor = None
    vrow: torch.Tensor = None
    virial: torch.Tensor = None
    energy: torch.Tensor = None
    atom_e: torch.Tensor = None
    atom_virial: torch.Tensor = None
    fit: dict = None
    max_nloc: int = 0
    ws_fsum: torch.Tensor = None
    xe: torch.Tensor = None      # [nreal, nx] normalised fitting parameters (slot order)
    extra: dict = field(default_factory=dict)


def _tuned(fn):
    """Run a model method with the model's launch tunables current."""
    @functools.wraps(fn)
    def wrapper(self, *args, **kwargs):
        with self.tuning.active():
            return fn(self, *args, **kwargs)
    return wrapper


class SeAModel:
    """Deep Potential se_e2_a energy model evaluated by libdpk kernels.

    Launch tunables (kernel organisations chosen by deepmd_amd.autotune) are
    the model's own (``self.tuning``), current around its launches."""

    # 16-bit / FP32 embedding: compute each type block only up to its last real
    # neighbour (padding rows from a table), dpk_embed_fwd / _bwd with nlist
    SKIP_PADDING = True

    def __init__(self, model_dict: dict, precision: str = "f64", device="cuda",
                 fitting_precision: str | None = None):
        """``precision`` of the embedding (and, by default, the fitting);
        ``fitting_precision`` may differ among the FP32-activation modes
        (f32 / bf16 / f16), e.g. FP32 fitting GEMMs behind a 16-bit embedding."""
        if precision not in PREC:
            raise ValueError(f"precision must be one of {list(PREC)}")
        self.tuning = _lib.Tuning()
        if fitting_precision is None and precision == "f16x2":
            m0 = model_dict["model"] if "model" in model_dict else model_dict
            fl = [_layers(n) for n in m0["fitting"]["nets"]["networks"]]
            tanh_fit = uniform_act(fl, "fitting hidden", hidden_only=True) == "tanh"
            fitting_precision = "f16x2" if Fitting.tail_shape(fl) and tanh_fit else "f32"
        fitting_precision = fitting_precision or precision
        if fitting_precision not in PREC or ((fitting_precision == "f64") != (precision == "f64")):
            raise ValueError(f"fitting precision {fitting_precision!r} with {precision!r}: both "
                             "f64, or both among f32 / bf16 / f16")
        self.fitting_precision = fitting_precision
        m = model_dict["model"] if "model" in model_dict else model_dict
        de, fi = m["descriptor"], m["fitting"]
        if de.get("type", "se_e2_a") not in ("se_e2_a", "se_a"):
            raise NotImplementedError(f"descriptor {de.get('type')!r}")
        self.type_map = list(m.get("type_map", []))
        self.rcut = float(de["rcut"])
        self.rcut_smth = float(de["rcut_smth"])
        self.sel = [int(x) for x in de["sel"]]
        self.neuron = [int(x) for x in de["neuron"]]
        self.axis_neuron = int(de["axis_neuron"])
        self.type_one_side = bool(de["type_one_side"])
        # u = r + env_protection in the env matrix (env_mat.py:38-45)
        self.env_protection = float(de.get("env_protection", 0.0))
        if self.env_protection < 0.0:
            raise ValueError("env_protection must be >= 0")
        self.ntypes = len(self.sel)
        self.exclude_types = [tuple(int(t) for t in p) for p in de.get("exclude_types", [])]
        for a, b in self.exclude_types:
            if not (0 <= a < self.ntypes and 0 <= b < self.ntypes):
                raise ValueError(f"exclude_types pair {(a, b)} outside the {self.ntypes} types")
        # model-level exclusions (base_atomic_model.py:62-90,380-399): excluded
        # centre types get no energy (also the fitting's own exclude_types,
        # fitting.py:551-556,895-897); excluded type pairs leave the
        # formatted neighbour list (their slots become padding)
        self.atom_exclude_types = sorted(
            {int(t) for t in (m.get("atom_exclude_types") or [])}
            | {int(t) for t in (fi.get("exclude_types") or [])})
        for t in self.atom_exclude_types:
            if not 0 <= t < self.ntypes:
                raise ValueError(f"atom_exclude_types entry {t} outside the {self.ntypes} types")
        self.pair_exclude_types = [tuple(int(t) for t in p)
                                   for p in (m.get("pair_exclude_types") or [])]
        for a, b in self.pair_exclude_types:
            if not (0 <= a < self.ntypes and 0 <= b < self.ntypes):
                raise ValueError(f"pair_exclude_types pair {(a, b)} outside the {self.ntypes} types")
        self.nnei = int(sum(self.sel))
        self.sec = np.concatenate([[0], np.cumsum(self.sel)]).astype(np.int32)
        self.davg = _arr(de["@variables"]["davg"]).reshape(self.ntypes, self.nnei * 4)
        self.dstd = _arr(de["@variables"]["dstd"]).reshape(self.ntypes, self.nnei * 4)
        self.emb_layers = [_layers(n) for n in de["embeddings"]["networks"]]
        # embedding activation (utils.py:174-212; the descriptor's
        # activation_function): tanh in every precision, the others in f32 / f64
        self.emb_act_name = uniform_act(self.emb_layers, "embedding")
        self.emb_act, self.emb_act_param = act_code(self.emb_act_name)
        if self.emb_act != ACTIVATIONS["tanh"] and precision in HALF:
            raise NotImplementedError(f"embedding activation {self.emb_act_name!r} needs the f32 "
                                      f"or f64 precision (the {precision} kernels are built for "
                                      "tanh)")
        self.fit_layers = [_layers(n) for n in fi["nets"]["networks"]]
        self.bias_atom_e = _arr(fi["@variables"]["bias_atom_e"]).reshape(self.ntypes)
        # frame / atomic parameters (fitting.py:749-830): normalised with the
        # stored avg / inv_std and appended to the descriptor
        fv = fi["@variables"]
        if int(fi.get("dim_case_embd", 0) or 0):
            raise NotImplementedError("case embeddings (multi-task fitting) are not supported")
        self.numb_fparam = int(fi.get("numb_fparam", 0) or 0)
        self.numb_aparam = int(fi.get("numb_aparam", 0) or 0)
        self.use_aparam_as_mask = bool(fi.get("use_aparam_as_mask", False))
        self.fparam_stats = None
        self.aparam_stats = None
        if self.numb_fparam:
            self.fparam_stats = (_arr(fv["fparam_avg"]).reshape(self.numb_fparam),
                                 _arr(fv["fparam_inv_std"]).reshape(self.numb_fparam))
        if self.numb_aparam and not self.use_aparam_as_mask:
            self.aparam_stats = (_arr(fv["aparam_avg"]).reshape(self.numb_aparam),
                                 _arr(fv["aparam_inv_std"]).reshape(self.numb_aparam))
        dfp = fi.get("default_fparam")
        self.default_fparam = None if dfp is None else np.asarray(dfp, dtype=np.float64).reshape(-1)
        self.nx = self.numb_fparam + (self.numb_aparam if self.aparam_stats is not None else 0)
        self.out_bias = _arr(m["@variables"]["out_bias"]).reshape(-1, self.ntypes)[0]
        self.precision = precision
        self.prec = PREC[precision]
        self.device = torch.device(device)
        self.d3 = self.neuron[-1]
        self.compressed = False
        self._pack()

    # -- packing -----------------------------------------------------------
    def _pack(self):
        dev = self.device
        real = torch.float64 if self.precision == "f64" else torch.float32
        packed, ch_out, r2, r3 = pack_embedding(self.emb_layers, self.neuron)
        self.emb_params = torch.tensor(packed, dtype=real, device=dev).contiguous()
        self.nnets = len(self.emb_layers)
        self.emb_has_idt = int(any(ly.idt is not None for net in self.emb_layers for ly in net))
        self.emb_image = None
        if self.precision in HALF:
            nb = lib().dpk_embed_image_bytes(_P1, _P2, _P3)
            self.emb_image = torch.empty(self.nnets * nb, dtype=torch.uint8, device=dev)
            call("dpk_embed_pack_" + IMAGE[self.precision], ptr(self.emb_image),
                 ptr(self.emb_params), self.nnets, _P1, _P2, _P3, stream_ptr())
        self.ch_out = torch.tensor(ch_out, dtype=torch.int32, device=dev)
        self.res2, self.res3 = r2, r3
        nt = self.ntypes
        net_of = np.zeros((nt, nt), dtype=np.int32)  # [centre type][neighbour block]
        for ti in range(nt):
            for tb in range(nt):
                net_of[ti, tb] = tb if self.type_one_side else tb * nt + ti
        self.net_of = torch.tensor(net_of.ravel(), dtype=torch.int32, device=dev)
        # 16-bit / FP32: type blocks end at their last real neighbour, the
        # padding tail comes from a per-(centre type, slot) table
        self.emb_pad = None
        self.skip_padding = self.precision in HALF + ("f32",) and self.SKIP_PADDING
        if self.skip_padding:
            pad = padding_table(packed, net_of, self.davg.reshape(nt, self.nnei, 4),
                                self.dstd.reshape(nt, self.nnei, 4), self.sec, r2, r3,
                                self.emb_act_name)
            if pad is not None:
                self.emb_pad = torch.from_numpy(pad).to(dev).contiguous()
        self.pair_mask = None  # (ntypes+1)^2 type-pair table, exclude_mask.py:64-140
        if self.exclude_types:
            tab = np.ones((nt + 1, nt + 1), dtype=np.int32)
            for a, b in self.exclude_types:
                tab[a, b] = tab[b, a] = 0
            self.pair_mask = torch.tensor(tab.ravel(), dtype=torch.int32, device=dev)
        self.model_pair_mask = None  # pair_exclude_types of the atomic model
        if self.pair_exclude_types:
            tab = np.ones((nt + 1, nt + 1), dtype=np.int32)
            for a, b in self.pair_exclude_types:
                tab[a, b] = tab[b, a] = 0
            self.model_pair_mask = torch.tensor(tab.ravel(), dtype=torch.int32, device=dev)
        self.davg_dev = torch.tensor(self.davg, dtype=torch.float64, device=dev)
        self.dstd_dev = torch.tensor(self.dstd, dtype=torch.float64, device=dev)
        self.sec_host = ctypes_int_array(self.sec)   # (array, address): keep the array alive
        self.sel_host = ctypes_int_array(self.sel)
        self.fitting = Fitting(self.fit_layers, self.bias_atom_e, self.out_bias,
                               self.fitting_precision, dev,
                               nx=self.nx)

    def _desc_dtype(self):
        """dtype of ``System.desc`` (the descriptor of :meth:`descriptor`):
        the fitting's GEMM type, FP32 for split FP16."""
        if self.fitting_precision == "f16x2":
            return torch.float32
        return GEMM_DTYPE[self.fitting_precision]

    @property
    def real_dtype(self):
        return torch.float64 if self.precision == "f64" else torch.float32

    # -- frame / atomic parameters -------------------------------------------
    def get_dim_fparam(self) -> int:
        return self.numb_fparam

    def get_dim_aparam(self) -> int:
        return self.numb_aparam

    def has_default_fparam(self) -> bool:
        return self.default_fparam is not None

    def set_params(self, s: System, fparam=None, aparam=None):
        """Normalised fitting inputs of the system's atoms (fitting.py:774-830):
        ``fparam`` [nframes, numb_fparam] (the model's default_fparam when
        None), ``aparam`` [natoms, numb_aparam] in the system's atom order.
        Written in place into ``s.xe`` (slot order), so captured graphs see
        new values."""
        if not self.nx:
            if fparam is not None or aparam is not None:
                raise ValueError("the model takes no frame / atomic parameters")
            return
        dev = self.device
        N, nf = s.fs.natoms, s.fs.nframes
        cols = []
        if self.numb_fparam:
            if fparam is None:
                if self.default_fparam is None:
                    raise ValueError("fparam is required: the model has no default_fparam")
                fparam = np.tile(self.default_fparam, (nf, 1))
            fp = torch.as_tensor(fparam, dtype=torch.float64).to(dev)
            if fp.numel() != nf * self.numb_fparam:
                raise ValueError(f"input fparam: cannot reshape {list(fp.shape)} into "
                                 f"({nf}, {self.numb_fparam}).")
            avg, istd = (torch.tensor(v, dtype=torch.float64, device=dev) for v in self.fparam_stats)
            fp = (fp.view(nf, -1) - avg) * istd
            cols.append(fp[s.fs.frame_of[:N].long()] if N else fp[:0])
        if self.aparam_stats is not None:
            if aparam is None:
                raise ValueError("aparam is required by the model")
            ap = torch.as_tensor(aparam, dtype=torch.float64).to(dev)
            if ap.numel() != N * self.numb_aparam:
                raise ValueError(f"input aparam: cannot reshape {list(ap.shape)} into "
                                 f"({N}, {self.numb_aparam}).")
            avg, istd = (torch.tensor(v, dtype=torch.float64, device=dev) for v in self.aparam_stats)
            cols.append((ap.view(N, -1) - avg) * istd)
        xe = torch.cat(cols, dim=1)
        xdt = torch.float32 if self.fitting.split else self.fitting.gemm_dtype
        if s.xe is None or s.xe.shape[0] != max(s.nreal, 1):
            s.xe = torch.zeros((max(s.nreal, 1), self.nx), dtype=xdt, device=dev)
        if s.nreal:
            s.xe[:s.nreal].copy_(xe[s.order[:s.nreal].long()])
        s.extra["xe_set"] = True

    def _params_ready(self, s: System):
        if self.nx and not s.extra.get("xe_set"):
            if self.aparam_stats is None and self.default_fparam is not None:
                self.set_params(s)
            else:
                raise ValueError("the model takes frame / atomic parameters: call set_params "
                                 "(or pass fparam / aparam) before evaluating")
        return s.xe if self.nx else None

    # -- systems -----------------------------------------------------------
    @_tuned
    def system(self, coords, atypes, boxes=None, nlocs=None) -> System:
        """Pack frames: coords [nf, nloc, 3] (or a list of [nloc_f, 3] arrays),
        atypes likewise, boxes [nf, 9] / list of [3,3] or None (open)."""
        if nlocs is None:
            coords = np.asarray(coords, dtype=np.float64)
            atypes = np.asarray(atypes)
            if coords.ndim == 2:
                coords = coords[None]
            nf = coords.shape[0]
            if atypes.ndim == 1:
                atypes = np.tile(atypes, (nf, 1))
            nlocs = [coords.shape[1]] * nf
            xyz = coords.reshape(-1, 3)
            at = atypes.reshape(-1)
            shared = True
        else:
            xyz = np.concatenate([np.asarray(c, dtype=np.float64).reshape(-1, 3) for c in coords])
            at = np.concatenate([np.asarray(a).reshape(-1) for a in atypes])
            nf = len(nlocs)
            shared = False
        if boxes is None:
            bl = [None] * nf
        else:
            bl = [None if b is None else np.asarray(b, dtype=np.float64).reshape(3, 3) for b in boxes]
        if at.size and int(at.max()) >= self.ntypes:
            raise ValueError(f"atom type {int(at.max())} >= ntypes {self.ntypes}")
        dev = self.device
        fs = make_frames(nlocs, bl, self.rcut, dev, coords_host=xyz, shared_images=shared)
        at_dev = torch.tensor(at, dtype=torch.int32, device=dev)
        order, segments = self._slot_order(at)
        N = fs.natoms
        S = self.nnei
        real = self.real_dtype
        sysd = System(coord=torch.tensor(xyz, dtype=torch.float64, device=dev), atype=at_dev,
                      fs=fs, order=torch.tensor(order, dtype=torch.int32, device=dev),
                      segments=segments, nreal=int(order.size))
        sysd.nlist = torch.empty((N, S), dtype=torch.int32, device=dev)
        sysd.wrapped = torch.empty((N, 3), dtype=torch.float64, device=dev)
        sysd.status = torch.zeros(4, dtype=torch.int32, device=dev)
        lib = _lib.lib()
        sysd.ws_nlist = torch.empty(int(lib.dpk_nlist_workspace_bytes(N, fs.ncells)),
                                    dtype=torch.uint8, device=dev)
        sysd.em = torch.empty((N, S, 4), dtype=real, device=dev)
        sysd.xyz = torch.empty((N, 4, self.d3), dtype=real, device=dev)
        ddt = self._desc_dtype()
        sysd.desc = torch.empty((N, self.d3 * self.axis_neuron), dtype=ddt, device=dev)
        sysd.gxyz = torch.empty((N, 4, self.d3), dtype=real, device=dev)
        sysd.atom_e = torch.empty(N, dtype=torch.float64, device=dev)
        sysd.fit = self.fitting.buffers(sysd.nreal)
        sysd.nd = torch.empty((N, S, 4), dtype=real, device=dev)
        sysd.force = torch.empty((N, 3), dtype=torch.float64, device=dev)
        sysd.vrow = torch.empty(int(lib.dpk_force_virial_workspace_bytes(N)), dtype=torch.uint8,
                                device=dev)
        sysd.virial = torch.empty((nf, 9), dtype=torch.float64, device=dev)
        sysd.energy = torch.empty(nf, dtype=torch.float64, device=dev)
        sysd.max_nloc = int(max(nlocs)) if len(nlocs) else 0
        sysd.ws_fsum = torch.empty(int(lib.dpk_frame_sum_workspace_bytes(nf, sysd.max_nloc, 9)),
                                   dtype=torch.uint8, device=dev)
        return sysd

    # -- forward_lower: caller's extended region and neighbour list ----------
    def lower_system(self, nloc: int, nall: int, nnei_in: int) -> System:
        """Buffers for :meth:`forward_lower` on one frame of nloc centres in an
        extended region of nall atoms with neighbour rows nnei_in wide."""
        dev = self.device
        lf = LowerFrames(natoms=nloc, nall=nall, nframes=1,
                         frame_of=torch.zeros(max(nloc, 1), dtype=torch.int32, device=dev),
                         frames_dev=_frame_bytes(nall, dev), sums_dev=_frame_bytes(nloc, dev),
                         shift_vec=torch.zeros((1, 3), dtype=torch.float64, device=dev))
        S = self.nnei
        real = self.real_dtype
        lib = _lib.lib()
        ddt = self._desc_dtype()
        sysd = System(coord=torch.empty((nall, 3), dtype=torch.float64, device=dev),
                      atype=torch.empty(nall, dtype=torch.int32, device=dev), fs=lf,
                      order=torch.empty(nloc, dtype=torch.int32, device=dev), segments=[],
                      nreal=0)
        sysd.wrapped = sysd.coord
        sysd.nlist = torch.empty((nloc, S), dtype=torch.int32, device=dev)
        sysd.status = torch.zeros(4, dtype=torch.int32, device=dev)
        sysd.em = torch.empty((nloc, S, 4), dtype=real, device=dev)
        sysd.xyz = torch.empty((nloc, 4, self.d3), dtype=real, device=dev)
        sysd.desc = torch.empty((nloc, self.d3 * self.axis_neuron), dtype=ddt, device=dev)
        sysd.gxyz = torch.empty((nloc, 4, self.d3), dtype=real, device=dev)
        sysd.atom_e = torch.empty(nloc, dtype=torch.float64, device=dev)
        sysd.nd = torch.empty((nloc, S, 4), dtype=real, device=dev)
        sysd.force = torch.empty((nall, 3), dtype=torch.float64, device=dev)
        sysd.vrow = torch.empty(int(lib.dpk_force_virial_workspace_bytes(nloc)), dtype=torch.uint8,
                                device=dev)
        # inverted list + per-slot force shares of dpk_force_virial_gather
        sysd.extra["ws_gather"] = torch.empty(
            int(lib.dpk_force_virial_gather_workspace_bytes(nloc, S, nall)), dtype=torch.uint8,
            device=dev)
        sysd.virial = torch.empty((1, 9), dtype=torch.float64, device=dev)
        sysd.energy = torch.empty(1, dtype=torch.float64, device=dev)
        sysd.max_nloc = nloc
        sysd.ws_fsum = torch.empty(int(lib.dpk_frame_sum_workspace_bytes(1, nloc, 9)),
                                   dtype=torch.uint8, device=dev)
        sysd.extra["nnei_in"] = nnei_in
        sysd.extra["nlist_in"] = torch.empty((nloc, max(nnei_in, 1)), dtype=torch.int32,
                                             device=dev)
        sysd.extra["types_key"] = None
        return sysd

    @_tuned
    def forward_lower(self, extended_coord, extended_atype, nlist, mapping=None,
                      do_atomic_virial: bool = False, extra_nlist_sort: bool = False,
                      system: System | None = None, fparam=None, aparam=None, ago: int = 0,
                      copy_outputs: bool = True):
        """``forward_common_lower`` (make_model.py:385-466) of the energy model
        on a caller's extended region: extended_coord [nf, nall, 3] (or
        [nf, nall*3]), extended_atype [nf, nall], nlist [nf, nloc, nnei_in]
        (extended indices, -1 padding; any width, unsorted, possibly built with
        a larger cutoff: it is formatted as make_model.py:592-658 does, with
        ``extra_nlist_sort`` forcing the sort).  ``mapping`` is accepted for
        the signature and not needed (se_e2_a has no message passing).

        ``fparam`` [nf, numb_fparam] and ``aparam`` [nf, nloc, numb_aparam]
        are the fitting's frame / atomic parameters (fitting.py:749-830).

        Returns ``energy`` [nf], ``atom_energy`` [nf, nloc], ``extended_force``
        [nf, nall, 3], ``virial`` [nf, 9] and, with ``do_atomic_virial``,
        ``extended_virial`` [nf, nall, 9] (the per-atom virial with the
        reference's atomic-virial correction, transform_output.py:88-94).

        ``ago`` follows DeepPotPT::compute (DeepPotPT.cc:171-378): with ``ago
        == 0`` the neighbour list, the atom types and their slot order are
        taken in and checked (host synchronisations); with ``ago > 0`` on the
        ``system`` of an earlier call (one frame) the stored list and order are
        reused -- the local atoms and their types are those of that call, as
        between an MD engine's re-neighbourings -- ``nlist`` may be None, and
        the step runs without any host synchronisation (it can be captured in
        a HIP graph).  The list is re-formatted (sorted by distance, cut to
        sel) on every call, as forward_common_lower does.  ``copy_outputs=False``
        returns views of the system's buffers, overwritten by its next call."""
        dev = self.device
        ec = torch.as_tensor(extended_coord, dtype=torch.float64).to(dev)
        s = system
        if ago > 0:  # types and list of the earlier call: extended_atype is not read
            if s is None or s.extra.get("types_key") is None:
                raise ValueError("ago > 0 needs the system of an earlier call with ago == 0")
            if ec.numel() != 3 * s.fs.nall:
                raise ValueError("ago > 0 reuses one frame's system: the extended region changed")
            if fparam is not None or aparam is not None:
                raise ValueError("ago > 0 keeps the frame / atomic parameters of the last call")
            self._last_lower_system = s
            out = self._lower_frame(s, ec.view(-1, 3), None, None, do_atomic_virial,
                                    extra_nlist_sort, reuse=True, copy_outputs=copy_outputs)
            return self._lower_stack([out])
        ea = torch.as_tensor(extended_atype).to(dev)
        nf = ea.shape[0] if ea.dim() == 2 else 1
        ea = ea.view(nf, -1)
        nall = ea.shape[1]
        ec = ec.view(nf, nall, 3)
        if nlist is None:
            raise ValueError("ago == 0 needs a neighbour list")
        nl = torch.as_tensor(nlist).to(dev)
        nl = nl.view(nf, -1, nl.shape[-1])
        nloc, nnei_in = nl.shape[1], nl.shape[2]
        if nloc > nall:
            raise ValueError(f"nlist has {nloc} rows but the extended region {nall} atoms")
        if ea.numel() and int(ea.max()) >= self.ntypes:
            raise ValueError(f"atom type {int(ea.max())} >= ntypes {self.ntypes}")
        if s is None or s.fs.natoms != nloc or s.fs.nall != nall or s.extra["nnei_in"] != nnei_in:
            s = self.lower_system(nloc, nall, nnei_in)
        self._last_lower_system = s
        fp = ap = None
        if fparam is not None:
            fp = torch.as_tensor(fparam, dtype=torch.float64)
            if fp.numel() != nf * self.numb_fparam:
                raise ValueError(f"input fparam: cannot reshape {list(fp.shape)} into "
                                 f"({nf}, {self.numb_fparam}).")
            fp = fp.reshape(nf, -1)
        if aparam is not None:
            ap = torch.as_tensor(aparam, dtype=torch.float64)
            if ap.numel() != nf * nloc * self.numb_aparam:
                raise ValueError(f"input aparam: cannot reshape {list(ap.shape)} into "
                                 f"({nf}, {nloc}, {self.numb_aparam}).")
            ap = ap.reshape(nf, nloc, -1)
        outs = []
        for f in range(nf):
            outs.append(self._lower_frame(s, ec[f], ea[f], nl[f], do_atomic_virial,
                                          extra_nlist_sort, None if fp is None else fp[f:f + 1],
                                          None if ap is None else ap[f],
                                          copy_outputs=copy_outputs or nf > 1))
        return self._lower_stack(outs)

    @staticmethod
    def _lower_stack(outs):
        if len(outs) == 1:  # no copy: the frame's tensors with a leading frame axis
            o = outs[0]
            out = {"energy": o["energy"], "atom_energy": o["atom_energy"][None],
                   "extended_force": o["extended_force"][None], "virial": o["virial"],
                   "nlist": o["nlist"][None]}
            if "extended_virial" in o:
                out["extended_virial"] = o["extended_virial"][None]
            return out
        do_atomic_virial = "extended_virial" in outs[0]
        out = {"energy": torch.cat([o["energy"] for o in outs]),
               "atom_energy": torch.stack([o["atom_energy"] for o in outs]),
               "extended_force": torch.stack([o["extended_force"] for o in outs]),
               "virial": torch.cat([o["virial"] for o in outs])}
        if do_atomic_virial:
            out["extended_virial"] = torch.stack([o["extended_virial"] for o in outs])
        out["nlist"] = torch.stack([o["nlist"] for o in outs])
        return out

    def _lower_frame(self, s: System, ec, ea, nl, atomic_virial, extra_sort, fparam=None,
                     aparam=None, reuse: bool = False, copy_outputs: bool = True):
        fs = s.fs
        nloc, nall = fs.natoms, fs.nall
        st = stream_ptr()
        s.coord.copy_(ec)
        nnei_in = s.extra["nnei_in"]
        if not reuse:
            s.atype.copy_(ea.to(torch.int32))
            if nnei_in:
                s.extra["nlist_in"].copy_(nl.to(torch.int32))
            # slot order of the centre atoms (real ones grouped by type)
            at = ea[:nloc].to(torch.int64)
            key = at.cpu().numpy().tobytes()
            if s.extra["types_key"] != key:
                order, s.segments = self._slot_order(at.cpu().numpy())
                s.order[:order.size].copy_(torch.from_numpy(order))
                s.nreal = int(order.size)
                s.fit = self.fitting.buffers(max(s.nreal, 1))
                s.extra["types_key"] = key
            if self.nx:
                self.set_params(s, fparam, aparam)
            elif fparam is not None or aparam is not None:
                raise ValueError("the model takes no frame / atomic parameters")
        call("dpk_nlist_format", ptr(s.nlist), ptr(s.status), ptr(s.extra["nlist_in"]), nnei_in,
             ptr(s.coord), ptr(s.atype), 1, nloc, nall, self.rcut, self.sel_host[1], self.ntypes,
             int(bool(extra_sort)), st)
        self._pair_exclude(s, st)
        if nloc:
            call("dpk_env_mat_fused", ptr(s.em), self.prec, ptr(s.coord), ptr(s.atype),
                 ptr(fs.frame_of), ptr(fs.frames_dev), ptr(s.nlist), ptr(fs.shift_vec),
                 ptr(self.davg_dev), ptr(self.dstd_dev), nloc, self.nnei, self.rcut,
                 self.rcut_smth, self.env_protection, ptr(self.pair_mask), self.ntypes, st)
            self._embed_forward(s, st)
            if s.nreal < nloc:
                s.atom_e.zero_()
                s.gxyz.zero_()
            if s.nreal:
                self.fitting.run(s.atom_e, s.gxyz, s.xyz, s.order, s.segments, s.nreal, s.fit,
                                 self.d3, self.axis_neuron, st, xe=s.xe if self.nx else None)
            self._embed_backward(s, st)
        av = None
        if atomic_virial:
            av = torch.empty((nall, 9), dtype=torch.float64, device=s.force.device)
        if nloc:
            # neighbour forces gathered per extended atom (no atomics)
            wg = s.extra["ws_gather"]
            call("dpk_force_virial_gather", ptr(s.force), ptr(av), ptr(s.vrow), ptr(s.nd),
                 self.prec, ptr(s.coord), ptr(s.atype), ptr(s.nlist), ptr(self.dstd_dev), nloc,
                 self.nnei, nall, self.rcut, self.rcut_smth, self.env_protection,
                 ptr(self.pair_mask), self.ntypes, ptr(wg), wg.numel(), st)
            call("dpk_frame_sum", ptr(s.virial), ptr(s.vrow), 9, ptr(fs.sums_dev), 1, nloc,
                 ptr(s.ws_fsum), s.ws_fsum.numel(), st)
            call("dpk_frame_sum", ptr(s.energy), ptr(s.atom_e), 1, ptr(fs.sums_dev), 1, nloc,
                 ptr(s.ws_fsum), s.ws_fsum.numel(), st)
        else:
            s.force.zero_()
            if av is not None:
                av.zero_()
            s.virial.zero_()
            s.energy.zero_()
        if not reuse and int(s.status[1].item()) != 0:
            raise ValueError("nlist holds an index outside the extended region")
        out = {"energy": s.energy, "atom_energy": s.atom_e, "extended_force": s.force,
               "virial": s.virial, "nlist": s.nlist}
        if copy_outputs:
            out = {k: v.clone() for k, v in out.items()}
        if atomic_virial:
            out["extended_virial"] = av
        return out

    def _slot_order(self, at):
        """Slot order: the atoms that get an energy (type >= 0 and not in
        atom_exclude_types) grouped by type (stable), and one (type, start,
        end) segment per such type."""
        keep = [t for t in range(self.ntypes) if t not in self.atom_exclude_types]
        order = np.concatenate([np.nonzero(at == t)[0] for t in keep]
                               + [np.zeros(0, dtype=np.int64)]).astype(np.int32)
        segments, a = [], 0
        for t in keep:
            n = int(np.count_nonzero(at == t))
            segments.append((t, a, a + n))
            a += n
        return order, segments

    def _pair_exclude(self, s: System, st):
        if self.model_pair_mask is not None and s.fs.natoms:
            call("dpk_nlist_pair_exclude", ptr(s.nlist), ptr(s.atype), ptr(s.fs.frame_of),
                 ptr(s.fs.frames_dev), s.fs.natoms, self.nnei, ptr(self.model_pair_mask),
                 self.ntypes, st)

    # -- kernels -------------------------------------------------------------
    @_tuned
    def build_nlist(self, s: System, check: bool = True):
        fs = s.fs
        call("dpk_nlist_build", ptr(s.nlist), ptr(s.wrapped), ptr(s.status), ptr(s.coord),
             ptr(s.atype), ptr(fs.frame_of), ptr(fs.frames_dev), fs.nframes, fs.natoms,
             fs.ncells, ptr(fs.img_pos), ptr(fs.shift_vec), self.rcut, self.sel_host[1],
             self.ntypes, fs.max_cand, ptr(s.ws_nlist), s.ws_nlist.numel(), stream_ptr())
        self._pair_exclude(s, stream_ptr())
        if check:
            self.check_capacity(s)

    @_tuned
    def check_capacity(self, s: System, rebuild: bool = True):
        """Re-run the search with a larger candidate buffer if any atom had
        more within-rcut neighbours than it held (synchronises).  With
        ``rebuild=False`` (captured graphs, whose buffer size is fixed) an
        overflow raises instead."""
        n = int(s.status[0].item())
        if n > s.fs.max_cand and not rebuild:
            raise _lib.DpkError(f"{n} neighbour candidates exceed the captured capacity "
                                f"{s.fs.max_cand}; capture again")
        if n > s.fs.max_cand:
            if n > 2048:
                raise _lib.DpkError(f"{n} neighbours within rcut exceed the 2048 capacity")
            for c in (512, 1024, 2048):
                if n <= c:
                    s.fs.max_cand = c
                    break
            self.build_nlist(s, check=False)
        return n

    # -- compression -----------------------------------------------------------
    def enable_compression(self, min_nbor_dist: float, extrapolate: float = 5.0,
                           stride0: float = 0.01, stride1: float = 0.1,
                           check_frequency: int = -1):
        """Replace the embedding nets by quintic tables (the reference's
        DescrptSeA.enable_compression, se_a.py:269-314): tables built in FP64
        on the device, stored in the model's precision (FP64 or FP32)."""
        from . import compress
        if self.precision not in ("f64", "f32"):
            raise NotImplementedError("compressed tables are FP64 or FP32 (precision 'f64'/'f32')")
        if self.compressed:
            raise ValueError("Compression is already enabled.")
        tables, infos = compress.build_tables(
            self.emb_layers, self.davg.reshape(self.ntypes, self.nnei, 4),
            self.dstd.reshape(self.ntypes, self.nnei, 4), self.rcut, self.rcut_smth, self.sel,
            self.type_one_side, min_nbor_dist, extrapolate, stride0, stride1, check_frequency,
            device=self.device)
        if len(tables) > 16:
            raise NotImplementedError(f"{len(tables)} tables (at most 16)")
        real = torch.float64 if self.precision == "f64" else torch.float32
        offs, n = [], 0
        for t in tables:
            offs.append(n)
            n += t.numel()
        self.tab_tables = torch.cat([t.reshape(-1) for t in tables]).to(real).contiguous()
        self.tab_info_np = np.stack(infos)
        self.tab_off = ctypes_array(offs, "longlong")
        self.tab_nspline = ctypes_int_array([t.shape[0] for t in tables])
        self.tab_info = ctypes_array(self.tab_info_np.ravel().tolist(), "double")
        self.tab_raw = [t for t in tables]  # FP64, for inspection and tests
        self.compressed = True

    def _tab_args(self):
        return (ptr(self.tab_tables), self.tab_off[1], self.tab_nspline[1], self.tab_info[1],
                ptr(self.net_of), self.ntypes, self.nnets, self.sec_host[1], self.d3)

    def _embed_forward(self, s: System, st):
        fs = s.fs
        if self.compressed:
            tb, to, tn, ti, no, nt, nn, sec, L = self._tab_args()
            call("dpk_tab_embed_fwd", ptr(s.xyz), self.prec, ptr(s.em), ptr(s.atype), tb, to, tn,
                 ti, no, nt, nn, sec, L, fs.natoms, self.nnei, st)
            return
        (pp, pw, no, co, nt, nn, p1, p2, p3, r2, r3, hi, ac, ap, d3) = self._embed_args()
        nl = ptr(s.nlist) if self._skip_padding() else None
        pad = ptr(self.emb_pad) if nl is not None and self.emb_pad is not None else None
        call("dpk_embed_fwd", ptr(s.xyz), self.prec, ptr(s.em), ptr(s.atype), pp, pw, no, co,
             nt, nn, p1, p2, p3, r2, r3, hi, ac, ap, d3, self.sec_host[1], fs.natoms, self.nnei,
             nl, pad, st)

    def _skip_padding(self):
        return self.skip_padding

    def _embed_backward(self, s: System, st):
        fs = s.fs
        if self.compressed:
            tb, to, tn, ti, no, nt, nn, sec, L = self._tab_args()
            call("dpk_tab_embed_bwd", ptr(s.nd), self.prec, ptr(s.em), ptr(s.gxyz), ptr(s.atype),
                 tb, to, tn, ti, no, nt, nn, sec, L, fs.natoms, self.nnei, st)
            return
        (pp, pw, no, co, nt, nn, p1, p2, p3, r2, r3, hi, ac, ap, d3) = self._embed_args()
        nl = ptr(s.nlist) if self._skip_padding() else None
        call("dpk_embed_bwd", ptr(s.nd), self.prec, ptr(s.em), ptr(s.gxyz), ptr(s.atype), pp, pw,
             no, co, nt, nn, p1, p2, p3, r2, r3, hi, ac, ap, d3, self.sec_host[1], fs.natoms,
             self.nnei, nl, st)

    def capture(self, s: System, atomic_virial: bool = False) -> GraphStep:
        """Capture one E+F+V step of ``s`` in a HIP graph (see GraphStep)."""
        return GraphStep(self, s, atomic_virial)

    def _embed_args(self):
        return (ptr(self.emb_params), ptr(self.emb_image), ptr(self.net_of), ptr(self.ch_out),
                self.ntypes, self.nnets, _P1, _P2, _P3, self.res2, self.res3, self.emb_has_idt,
                self.emb_act, self.emb_act_param, self.d3)

    @_tuned
    def descriptor(self, s: System):
        """Env matrix -> embedding -> descriptor (kernels only)."""
        fs = s.fs
        st = stream_ptr()
        call("dpk_env_mat_fused", ptr(s.em), self.prec, ptr(s.wrapped), ptr(s.atype),
             ptr(fs.frame_of), ptr(fs.frames_dev), ptr(s.nlist), ptr(fs.shift_vec),
             ptr(self.davg_dev), ptr(self.dstd_dev), fs.natoms, self.nnei, self.rcut,
             self.rcut_smth, self.env_protection, ptr(self.pair_mask), self.ntypes, st)
        self._embed_forward(s, st)
        out_prec = DPK_PREC_F32 if self.fitting_precision == "f16x2" else PREC[self.fitting_precision]
        call("dpk_descriptor_fwd", ptr(s.desc), out_prec, ptr(s.xyz), self.prec, fs.natoms,
             self.d3, self.axis_neuron, st)
        return s.desc

    @_tuned
    def evaluate(self, s: System, build_nlist: bool = True, atomic_virial: bool = False,
                 check_nlist: bool = True, timers: dict | None = None):
        """Energy [nf], atom energy [N], force [N,3], virial [nf,9] (FP64).

        ``check_nlist=False`` skips the candidate-capacity check (a device
        sync); the caller then checks ``s.status`` itself.  ``timers`` (a dict)
        receives (start, end) HIP event pairs per stage, recorded on the
        launch stream."""
        fs = s.fs

        def mark(name):
            if timers is None:
                return None
            ev = torch.cuda.Event(enable_timing=True)
            ev.record()
            timers.setdefault(name, []).append([ev, None])
            return name

        def done(name):
            if timers is not None and name is not None:
                ev = torch.cuda.Event(enable_timing=True)
                ev.record()
                timers[name][-1][1] = ev

        if build_nlist:
            t = mark("nlist")
            self.build_nlist(s, check=check_nlist)
            done(t)
        fs_ = fs
        st = stream_ptr()
        t = mark("env_mat")
        call("dpk_env_mat_fused", ptr(s.em), self.prec, ptr(s.wrapped), ptr(s.atype),
             ptr(fs_.frame_of), ptr(fs_.frames_dev), ptr(s.nlist), ptr(fs_.shift_vec),
             ptr(self.davg_dev), ptr(self.dstd_dev), fs_.natoms, self.nnei, self.rcut,
             self.rcut_smth, self.env_protection, ptr(self.pair_mask), self.ntypes, st)
        done(t)
        t = mark("embed_fwd")
        self._embed_forward(s, st)
        done(t)
        t = mark("fitting")
        if s.nreal < fs.natoms:  # virtual atoms: no energy, no gradient
            s.atom_e.zero_()
            s.gxyz.zero_()
        self.fitting.run(s.atom_e, s.gxyz, s.xyz, s.order, s.segments, s.nreal, s.fit, self.d3,
                         self.axis_neuron, st, xe=self._params_ready(s))
        done(t)
        t = mark("embed_bwd")
        self._embed_backward(s, st)
        done(t)
        t = mark("force_virial")
        s.force.zero_()
        av = None
        if atomic_virial:
            if s.atom_virial is None:
                s.atom_virial = torch.empty((fs.natoms, 9), dtype=torch.float64,
                                            device=s.force.device)
            av = s.atom_virial
            av.zero_()
        call("dpk_force_virial_cells", ptr(s.force), ptr(av), ptr(s.nd), self.prec,
             ptr(s.wrapped), ptr(s.atype), ptr(fs.frame_of), ptr(fs.frames_dev), ptr(s.nlist),
             ptr(fs.shift_vec), ptr(fs.img_shift), ptr(self.dstd_dev), fs.natoms, fs.ncells,
             self.nnei, self.rcut, self.rcut_smth, self.env_protection, ptr(s.ws_nlist),
             s.ws_nlist.numel(),
             ptr(s.vrow), s.vrow.numel(), ptr(self.pair_mask), self.ntypes, st)
        call("dpk_frame_sum", ptr(s.virial), ptr(s.vrow), 9, ptr(fs.frames_dev), fs.nframes,
             s.max_nloc, ptr(s.ws_fsum), s.ws_fsum.numel(), st)
        done(t)
        call("dpk_frame_sum", ptr(s.energy), ptr(s.atom_e), 1, ptr(fs.frames_dev), fs.nframes,
             s.max_nloc, ptr(s.ws_fsum), s.ws_fsum.numel(), st)
        out = {"energy": s.energy, "atom_energy": s.atom_e, "force": s.force, "virial": s.virial}
        if atomic_virial:
            out["atom_virial"] = av
        return out


@dataclass
class LowerFrames:
    """Frame tables of one extended region handed in by the caller (the
    forward_lower path): the extended index of a neighbour is its position in
    the caller's arrays, so the index frame has nloc = nall, a single image and
    a zero shift; the sum frame covers the nloc centre atoms."""
    natoms: int                  # centre atoms (nloc)
    nall: int
    nframes: int
    frame_of: torch.Tensor       # [nloc] int32 zeros
    frames_dev: torch.Tensor     # index frame
    sums_dev: torch.Tensor       # sum frame
    shift_vec: torch.Tensor      # [1, 3] zeros


def _frame_bytes(nloc: int, device) -> torch.Tensor:
    fr = _lib.Frame()
    fr.nloc = nloc
    fr.ns = 1
    arr = (_lib.Frame * 1)(fr)
    raw = bytes(ctypes.string_at(ctypes.addressof(arr), ctypes.sizeof(arr)))
    return torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(device)


class GraphStep:
    """One E+F+V step of a packed system captured in a HIP graph.

    ``replay()`` re-runs every kernel of ``SeAModel.evaluate`` (neighbour list
    included) on the system's current device buffers, so an MD-style caller
    updates ``system.coord`` in place and replays; ``out`` holds the output
    tensors the replays write.  Capacity of the neighbour candidate buffer is
    checked once at capture; ``check()`` re-checks it after replays (the
    kernels record the largest count seen in ``system.status``).
    """

    def __init__(self, model: "SeAModel", system: System, atomic_virial: bool = False):
        self.model, self.system = model, system
        side = torch.cuda.Stream()  # default priority
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            # first runs: candidate capacity check, occupancy / GEMM heuristics
            for _ in range(2):
                model.evaluate(system, check_nlist=True, atomic_virial=atomic_virial)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.out = model.evaluate(system, check_nlist=False, atomic_virial=atomic_virial)

    def replay(self):
        self.graph.replay()
        return self.out

    def check(self):
        """Synchronises; raises if any atom had more candidates than the
        buffer held in the last replay."""
        return self.model.check_capacity(self.system, rebuild=False)


def ctypes_array(a, kind: str):
    """(ctypes array, address) pair for a small host array of c_<kind>."""
    ct = {"longlong": ctypes.c_longlong, "double": ctypes.c_double, "int": ctypes.c_int}[kind]
    conv = float if kind == "double" else int
    arr = (ct * len(a))(*[conv(x) for x in a])
    return arr, ctypes.addressof(arr)


def ctypes_int_array(a):
    """(ctypes array, address) pair for a small host int32 array."""
    arr = (ctypes.c_int * len(a))(*[int(x) for x in a])
    return arr, ctypes.addressof(arr)

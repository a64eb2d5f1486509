"""Training loop for se_e2_a energy models: data systems, statistics,
learning-rate schedule, prefactor annealing, Adam steps, learning curve and
checkpoints.

It follows the reference's PyTorch trainer for a single-task energy model
(deepmd/pt/train/training.py:1329-1430, 2256-2323): at step t the learning
rate is ``lr(t)`` of the schedule, the loss prefactors move from their start
to their limit values with lr / start_lr (deepmd/pt/loss/ener.py:245-251),
a batch of frames is drawn from one training system, and Adam takes one step
on the E + F (+ V) loss whose force term is differentiated through the
hand-written prod_force / prod_virial kernels (``train.SeATrainer``).

* Data: the ``set.*/{coord,box,energy,force,virial}.npy`` systems with
  ``type.raw`` / ``type_map.raw`` / ``nopbc`` (deepmd/utils/data.py:65-150,
  405-430), batch sizes ``"auto"`` / ``"auto:N"`` (data_system.py:659-667),
  systems drawn with probability proportional to their batch count
  (``prob_sys_size``, data_system.py:799-815).
* A model built from the ``model`` section starts from the reference's
  initialisation (mlp.py:149-163: W ~ N(0, 1/sqrt(n_in + n_out)),
  b ~ N(0, 1), idt ~ N(0.1, 0.001)) and the input statistics of the
  reference: per centre type, the mean of the radial column and the standard
  deviations of the radial and angular columns of the raw environment matrix
  (dpmodel/utils/env_mat_stat.py:133-420, deepmd/utils/env_mat_stat.py:
  StatItem), and the per-type energy bias by least squares on the type counts
  (deepmd/utils/out_stat.py:compute_stats_from_redu).

Everything numeric runs on the device through ``libdpk`` (neighbour lists,
environment matrices, forces, virials) and torch (nets, Adam).  Under an
initialised ``torch.distributed`` process group each rank draws its own
batches (seed + rank) and the gradients are averaged over the ranks before
every optimiser step (``allreduce_gradients``; the reference wraps its model
in DistributedDataParallel, training.py:1057), and only rank 0 writes the
learning curve and checkpoints.
"""

from __future__ import annotations

import glob
import math
import os
import types

import numpy as np
import torch

from .checkpoint import save_checkpoint
from .model import _layer_dict, _nd
from .train import SeATrainer, TrainBatch


# ------------------------------------------------------------ learning rate --

class LearningRateExp:
    """Exponential decay with optional linear warmup
    (dpmodel/utils/learning_rate.py:36-135,263-400).

    lr(t) = start_lr r^floor(t' / s) (``smooth``: r^(t' / s)), t' = t - warmup,
    clipped below at stop_lr; r = (stop_lr / start_lr)^(s / T) with T the
    decay steps, unless ``decay_rate`` is given; ``decay_steps`` >= T without
    a decay rate becomes 100 (T // 10 > 100) or T // 100 + 1."""

    def __init__(self, start_lr: float, num_steps: int, stop_lr: float | None = None,
                 stop_lr_ratio: float | None = None, decay_steps: int = 5000,
                 decay_rate: float | None = None, warmup_steps: int = 0,
                 warmup_ratio: float | None = None, warmup_start_factor: float = 0.0,
                 smooth: bool = False, **_):
        _base_init(self, start_lr, num_steps, stop_lr, stop_lr_ratio, warmup_steps,
                   warmup_ratio, warmup_start_factor)
        total = self.decay_num_steps
        if decay_steps <= 0:
            raise ValueError(f"decay_steps ({decay_steps}) must be positive.")
        self.decay_steps = decay_steps
        if decay_rate is None and self.decay_steps >= total:
            self.decay_steps = 100 if total // 10 > 100 else total // 100 + 1
        self.min_lr = self.stop_lr
        if decay_rate is not None:
            self.decay_rate = float(decay_rate)
        elif total == 0:
            self.decay_rate = 1.0
        else:
            self.decay_rate = float(np.exp(np.log(max(self.stop_lr, 1e-10) / self.start_lr)
                                           / (total / self.decay_steps)))
        self.smooth = bool(smooth)

    def _decay(self, t: float) -> float:
        e = t / self.decay_steps if self.smooth else float(int(t) // self.decay_steps)
        return max(self.start_lr * self.decay_rate ** e, self.min_lr)

    def value(self, step: int) -> float:
        return _warmup_value(self, step)


class LearningRateCosine:
    """Cosine annealing from start_lr to stop_lr over the decay steps, then
    stop_lr, with optional linear warmup (learning_rate.py:575-697)."""

    def __init__(self, start_lr: float, num_steps: int, stop_lr: float | None = None,
                 stop_lr_ratio: float | None = None, warmup_steps: int = 0,
                 warmup_ratio: float | None = None, warmup_start_factor: float = 0.0, **_):
        _base_init(self, start_lr, num_steps, stop_lr, stop_lr_ratio, warmup_steps,
                   warmup_ratio, warmup_start_factor)
        self.lr_min_factor = self.stop_lr / self.start_lr

    def _decay(self, t: float) -> float:
        if self.decay_num_steps == 0:
            return self.start_lr
        if t >= self.decay_num_steps:
            return self.start_lr * self.lr_min_factor
        return self.start_lr * (self.lr_min_factor + 0.5 * (1 - self.lr_min_factor)
                                * (1 + math.cos(math.pi * t / self.decay_num_steps)))

    def value(self, step: int) -> float:
        return _warmup_value(self, step)


class LearningRateWSD:
    """Warmup-stable-decay (learning_rate.py:403-572): start_lr until the
    last ``decay_phase_ratio`` of the steps, then inverse-linear (default),
    cosine or linear decay to stop_lr with tau = (t - stable) / decay steps;
    stop_lr from the last step on."""

    def __init__(self, start_lr: float, num_steps: int, stop_lr: float | None = None,
                 stop_lr_ratio: float | None = None, warmup_steps: int = 0,
                 warmup_ratio: float | None = None, warmup_start_factor: float = 0.0,
                 decay_phase_ratio: float = 0.1, decay_type: str = "inverse_linear", **_):
        _base_init(self, start_lr, num_steps, stop_lr, stop_lr_ratio, warmup_steps,
                   warmup_ratio, warmup_start_factor)
        if self.stop_lr <= 0:
            raise ValueError(f"stop_lr ({self.stop_lr}) must be positive.")
        if decay_phase_ratio <= 0 or decay_phase_ratio > 1:
            raise ValueError(f"decay_phase_ratio ({decay_phase_ratio}) must be in (0, 1].")
        if decay_type not in ("inverse_linear", "cosine", "linear"):
            raise ValueError(f"decay_type {decay_type!r}")
        self.decay_type = decay_type
        self.decay_phase_steps = max(1, min(int(decay_phase_ratio * self.num_steps),
                                            self.decay_num_steps))
        self.stable_steps = self.decay_num_steps - self.decay_phase_steps

    def _decay(self, t: float) -> float:
        if t >= self.decay_num_steps:
            return self.stop_lr
        if t < self.stable_steps:
            return self.start_lr
        tau = min(max((t - self.stable_steps) / self.decay_phase_steps, 0.0), 1.0)
        a, b = self.start_lr, self.stop_lr
        if self.decay_type == "inverse_linear":
            return 1.0 / (tau / b + (1.0 - tau) / a)
        if self.decay_type == "cosine":
            return b + (a - b) * 0.5 * (1.0 + math.cos(math.pi * tau))
        return a + (b - a) * tau

    def value(self, step: int) -> float:
        return _warmup_value(self, step)


def _base_init(self, start_lr, num_steps, stop_lr, stop_lr_ratio, warmup_steps, warmup_ratio,
               warmup_start_factor):
    """Argument checks and warmup set-up shared by the schedules
    (learning_rate.py:36-133)."""
    if start_lr <= 0 or not np.isfinite(start_lr):
        raise ValueError(f"start_lr ({start_lr}) must be positive and finite.")
    if stop_lr is not None and stop_lr_ratio is not None:
        raise ValueError("stop_lr and stop_lr_ratio are mutually exclusive.")
    if stop_lr is None and stop_lr_ratio is None:
        raise ValueError("Either stop_lr or stop_lr_ratio must be provided.")
    self.stop_lr = start_lr * stop_lr_ratio if stop_lr_ratio is not None else float(stop_lr)
    if warmup_steps != 0 and warmup_ratio is not None:
        raise ValueError("warmup_steps and warmup_ratio are mutually exclusive.")
    self.warmup_steps = int(warmup_ratio * num_steps) if warmup_ratio is not None \
        else int(warmup_steps)
    if num_steps < 0:
        raise ValueError("num_steps must be non-negative")
    if self.warmup_steps < 0:
        raise ValueError("warmup_steps must be non-negative")
    if num_steps > 0 and self.warmup_steps >= num_steps:
        raise ValueError("warmup_steps must be smaller than num_steps")
    if num_steps == 0 and self.warmup_steps != 0:
        raise ValueError("warmup_steps must be 0 when num_steps is 0")
    self.warmup_start_lr = warmup_start_factor * start_lr
    self.start_lr = float(start_lr)
    self.num_steps = int(num_steps)
    self.decay_num_steps = self.num_steps - self.warmup_steps


def _warmup_value(self, step: int) -> float:
    """learning_rate.py:169-219: linear warmup, then the decay of step - warmup."""
    if self.warmup_steps == 0:
        return float(self._decay(float(step)))
    if step < self.warmup_steps:
        return float(self.warmup_start_lr
                     + (self.start_lr - self.warmup_start_lr) * step / self.warmup_steps)
    return float(self._decay(float(max(step - self.warmup_steps, 0))))


def make_learning_rate(params: dict, num_steps: int):
    """Schedule of a ``learning_rate`` section (type ``exp``, ``cosine`` or ``wsd``;
    stop_lr defaults to 5e-8 as in argcheck.py:3765-3776)."""
    p = dict(params)
    kind = p.pop("type", "exp")
    p = {k: v for k, v in p.items() if not k.startswith("_")}
    if p.get("stop_lr") is None and p.get("stop_lr_ratio") is None:
        p["stop_lr"] = 5e-8
    if "start_lr" not in p:
        raise ValueError("learning_rate.start_lr is required")
    if kind == "exp":
        return LearningRateExp(num_steps=num_steps, **p)
    if kind == "cosine":
        return LearningRateCosine(num_steps=num_steps, **p)
    if kind == "wsd":
        return LearningRateWSD(num_steps=num_steps, **p)
    raise NotImplementedError(f"learning_rate type {kind!r} (exp, cosine and wsd only)")


# --------------------------------------------------------------------- data --

class DeepmdData:
    """One training system: ``type.raw`` (per-atom types), optional
    ``type_map.raw``, ``nopbc`` (open boundaries), and sorted ``set.*``
    directories of ``coord`` [nf, natoms*3], ``box`` [nf, 9], ``energy``
    [nf], ``force`` [nf, natoms*3], ``virial`` [nf, 9] arrays
    (deepmd/utils/data.py:65-150).  Types are remapped to ``type_map`` by
    element name when both maps are known (data.py:101-124).  A property
    missing from a set has ``find_<key>`` = 0 for that set's frames."""

    KEYS = (("energy", 1), ("force", 3), ("virial", 9))

    def __init__(self, path: str, type_map: list[str] | None = None):
        if not os.path.isdir(path):
            raise FileNotFoundError(f"System {path} is not found!")
        self.path = path
        dirs = sorted(glob.glob(os.path.join(path, "set.*")))
        if not dirs:
            raise FileNotFoundError(f"No set.* is found in {path}")
        self.atom_type = np.loadtxt(os.path.join(path, "type.raw"), dtype=np.int32, ndmin=1)
        self.natoms = len(self.atom_type)
        tm_file = os.path.join(path, "type_map.raw")
        own_map = open(tm_file).read().split() if os.path.isfile(tm_file) else None
        if own_map is not None and len(own_map) < self.atom_type.max() + 1:
            raise ValueError(f"{tm_file} names fewer types than type.raw uses")
        if type_map is not None and own_map is not None and len(type_map):
            missing = [e for e in own_map if e not in type_map]
            if missing:
                raise ValueError(f"Elements {missing} are not present in the provided `type_map`.")
            remap = np.array([type_map.index(e) for e in own_map], dtype=np.int32)
            self.atom_type = remap[self.atom_type]
        self.pbc = not os.path.isfile(os.path.join(path, "nopbc"))
        coords, boxes, vals, finds, sizes = [], [], {k: [] for k, _ in self.KEYS}, \
            {k: [] for k, _ in self.KEYS}, []
        for d in dirs:
            c = np.load(os.path.join(d, "coord.npy")).astype(np.float64)
            nf = c.shape[0]
            coords.append(c.reshape(nf, self.natoms, 3))
            sizes.append(nf)
            if self.pbc:
                boxes.append(np.load(os.path.join(d, "box.npy")).astype(np.float64).reshape(nf, 9))
            for key, ndof in self.KEYS:
                f = os.path.join(d, key + ".npy")
                shape = (nf, self.natoms, 3) if key == "force" else (nf, ndof)
                if os.path.isfile(f):
                    vals[key].append(np.load(f).astype(np.float64).reshape(shape))
                    finds[key].append(np.ones(nf))
                else:
                    vals[key].append(np.zeros(shape))
                    finds[key].append(np.zeros(nf))
        self.set_sizes = sizes
        self.coord = np.concatenate(coords)
        self.box = np.concatenate(boxes) if self.pbc else None
        self.labels = {k: np.concatenate(v) for k, v in vals.items()}
        self.finds = {k: np.concatenate(v) for k, v in finds.items()}
        self.nframes = self.coord.shape[0]

    def numb_batch(self, batch_size: int) -> int:
        """Batches per epoch: sum over sets of max(1, nframes // batch_size)
        (data.py:411-426)."""
        return sum(max(1, n // batch_size) for n in self.set_sizes)

    def frames(self, idx) -> dict:
        idx = np.asarray(idx)
        out = {"coord": self.coord[idx], "atype": np.tile(self.atom_type, (len(idx), 1)),
               "box": None if self.box is None else self.box[idx]}
        for key, _ in self.KEYS:
            out[key] = self.labels[key][idx]
            # a batch either has the property in every frame or is treated as lacking it
            out["find_" + key] = float(np.all(self.finds[key][idx] > 0))
        return out


class DataSystem:
    """Training systems with their batch sizes and sampling probabilities.

    ``batch_size`` is an int, a list (one per system), ``"auto"`` or
    ``"auto:N"`` (the smallest size with size * natoms >= N, N = 32 by
    default; data_system.py:659-667).  ``get_batch`` draws a system with
    probability proportional to its batch count (prob_sys_size) and returns
    its next batch from a per-system shuffled order, reshuffled each epoch."""

    def __init__(self, systems, batch_size="auto", type_map=None, seed=None):
        if isinstance(systems, str):
            systems = [systems]
        self.systems = [DeepmdData(p, type_map) for p in systems]
        n = len(self.systems)
        if isinstance(batch_size, str):
            words = batch_size.split(":")
            if words[0] != "auto":
                raise ValueError(f"batch_size {batch_size!r}")
            rule = int(words[1]) if len(words) > 1 else 32
            self.batch_size = [-(-rule // s.natoms) for s in self.systems]
        elif isinstance(batch_size, (list, tuple)):
            if len(batch_size) != n:
                raise ValueError("one batch size per system")
            self.batch_size = [int(b) for b in batch_size]
        else:
            self.batch_size = [int(batch_size)] * n
        nb = np.array([s.numb_batch(b) for s, b in zip(self.systems, self.batch_size)], float)
        self.probs = nb / nb.sum()
        self.rng = np.random.default_rng(seed)
        self._order = [None] * n
        self._pos = [0] * n

    def get_batch(self, sys_idx: int | None = None) -> tuple[int, dict]:
        i = int(self.rng.choice(len(self.systems), p=self.probs)) if sys_idx is None else sys_idx
        s, bs = self.systems[i], self.batch_size[i]
        if self._order[i] is None or self._pos[i] >= s.nframes:
            self._order[i] = self.rng.permutation(s.nframes)
            self._pos[i] = 0
        idx = self._order[i][self._pos[i]:self._pos[i] + bs]
        self._pos[i] += bs
        return i, s.frames(idx)

    def stat_frames(self, nbatch: int) -> list[dict]:
        """The first ``nbatch`` batches of every system (its first
        nbatch * batch_size frames), for the input statistics."""
        return [s.frames(np.arange(min(s.nframes, nbatch * b)))
                for s, b in zip(self.systems, self.batch_size)]


# --------------------------------------------------------------- model init --

def _init_net(rng, dims, resnet_dt, out_layer):
    """mlp.py:149-163 (stddev 1, bavg 0); a time step only where the layer
    has a residual connection (n_out == n_in or 2 n_in, mlp.py:89-93)."""
    layers = []
    for i in range(len(dims) - 1):
        nin, nout = dims[i], dims[i + 1]
        last = out_layer and i == len(dims) - 2
        w = rng.normal(0.0, 1.0 / math.sqrt(nin + nout), size=(nin, nout))
        b = rng.normal(0.0, 1.0, size=(nout,))
        use_ts = resnet_dt and not last and (nout == nin or nout == 2 * nin)
        idt = rng.normal(0.1, 0.001, size=(nout,)) if use_ts else None
        layers.append(_layer_dict(w, b, idt, "none" if last else "tanh", not last))
    return {"layers": layers}


def init_model_dict(model_params: dict, seed=None) -> dict:
    """A freshly initialised se_e2_a energy model (serialized layout) from
    an input.json ``model`` section; davg = 0, dstd = 1 and zero energy bias
    until ``compute_input_stats`` fills them.  Defaults as argcheck.py
    (descriptor: rcut 6, rcut_smth 0.5, neuron [10, 20, 40], axis_neuron 4,
    type_one_side False, resnet_dt False; fitting: neuron [120, 120, 120],
    resnet_dt True)."""
    de = dict(model_params.get("descriptor", {}))
    fi = dict(model_params.get("fitting_net", {}))
    if de.get("type", "se_e2_a") not in ("se_e2_a", "se_a"):
        raise NotImplementedError(f"descriptor {de.get('type')!r} (se_e2_a only)")
    if fi.get("type", "ener") != "ener":
        raise NotImplementedError(f"fitting {fi.get('type')!r} (energy only)")
    for key, d in (("activation_function", de), ("activation_function", fi)):
        if d.get(key, "tanh") != "tanh":
            raise NotImplementedError(f"{key} {d[key]!r} (tanh only)")
    type_map = list(model_params["type_map"])
    sel = de.get("sel")
    if not isinstance(sel, (list, tuple)) or len(sel) != len(type_map):
        raise ValueError("descriptor.sel must list one count per type of type_map")
    sel = [int(x) for x in sel]
    ntypes, nnei = len(sel), int(sum(sel))
    neuron = [int(x) for x in de.get("neuron", [10, 20, 40])]
    axis = int(de.get("axis_neuron", 4))
    tos = bool(de.get("type_one_side", False))
    rdt_e = bool(de.get("resnet_dt", False))
    fneuron = [int(x) for x in fi.get("neuron", [120, 120, 120])]
    rdt_f = bool(fi.get("resnet_dt", True))
    rng = np.random.default_rng(seed)
    nemb = ntypes if tos else ntypes * ntypes
    emb = [_init_net(rng, [1] + neuron, rdt_e, False) for _ in range(nemb)]
    din = neuron[-1] * axis
    fit = [_init_net(rng, [din] + fneuron + [1], rdt_f, True) for _ in range(ntypes)]
    return {"model": {
        "@class": "Model", "type": "standard", "type_map": type_map,
        "@variables": {"out_bias": _nd(np.zeros((1, ntypes, 1))),
                       "out_std": _nd(np.ones((1, ntypes, 1)))},
        "descriptor": {
            "@class": "Descriptor", "type": "se_e2_a", "rcut": float(de.get("rcut", 6.0)),
            "rcut_smth": float(de.get("rcut_smth", 0.5)), "sel": sel, "neuron": neuron,
            "axis_neuron": axis, "type_one_side": tos, "resnet_dt": rdt_e,
            "activation_function": "tanh", "env_protection": float(de.get("env_protection", 0.0)),
            "exclude_types": [list(p) for p in de.get("exclude_types", [])],
            "type_map": type_map,
            "@variables": {"davg": _nd(np.zeros((ntypes, nnei, 4))),
                           "dstd": _nd(np.ones((ntypes, nnei, 4)))},
            "embeddings": {"@class": "NetworkCollection", "ndim": 1 if tos else 2,
                           "ntypes": ntypes, "networks": emb},
        },
        "fitting": {
            "@class": "Fitting", "type": "ener", "neuron": fneuron, "resnet_dt": rdt_f,
            "dim_descrpt": din, "mixed_types": False,
            "@variables": {"bias_atom_e": _nd(np.zeros((ntypes, 1)))},
            "nets": {"@class": "NetworkCollection", "ndim": 1, "ntypes": ntypes,
                     "networks": fit},
        },
    }}


def _std(n, s, s2, default=1e-1, protection=1e-2):
    """StatItem.compute_std (deepmd/utils/env_mat_stat.py)."""
    if n == 0:
        return default
    val = math.sqrt(max(s2 / n - (s / n) * (s / n), 0.0))
    return protection if abs(val) < protection else val


def env_mat_sums(model_dict: dict, frames: list[dict], device="cuda") -> np.ndarray:
    """Per centre type t: [n_r, sum r, sum r^2, n_a, sum a, sum a^2] over the
    raw environment matrix (davg 0, dstd 1; padding and excluded slots count
    as zeros) of the given frames (env_mat_stat.py:273-372)."""
    from .model import SeAModel
    m = model_dict["model"] if "model" in model_dict else model_dict
    de = m["descriptor"]
    nt, nnei = len(de["sel"]), int(sum(de["sel"]))
    raw = {"model": dict(m, descriptor=dict(de, **{"@variables": {
        "davg": _nd(np.zeros((nt, nnei, 4))), "dstd": _nd(np.ones((nt, nnei, 4)))}}))}
    geo = types.SimpleNamespace(geo=SeAModel(raw, "f64", device))
    out = np.zeros((nt, 6))
    for fr in frames:
        for k in range(fr["coord"].shape[0]):  # one frame at a time: boxes may differ
            box = None if fr["box"] is None else fr["box"][k:k + 1]
            b = TrainBatch(geo, fr["coord"][k:k + 1], fr["atype"][k:k + 1], box)
            em = b.em.reshape(-1, nnei, 4)
            at = b.atype.reshape(-1)
            for t in range(nt):
                rows = em[at == t]
                r, a = rows[..., :1], rows[..., 1:]
                out[t] += [r.numel(), r.sum().item(), (r * r).sum().item(),
                           a.numel(), a.sum().item(), (a * a).sum().item()]
    return out


def energy_bias(frames: list[dict], ntypes: int, rcond=None) -> np.ndarray:
    """Per-type energy bias: least squares of frame energies on type counts
    (out_stat.py compute_stats_from_redu); frames without energies are left
    out."""
    cnt, e = [], []
    for fr in frames:
        if not fr["find_energy"]:
            continue
        c = np.stack([np.bincount(a, minlength=ntypes)[:ntypes] for a in fr["atype"]])
        cnt.append(c.astype(np.float64))
        e.append(fr["energy"].reshape(-1))
    if not cnt:
        return np.zeros(ntypes)
    bias, *_ = np.linalg.lstsq(np.concatenate(cnt), np.concatenate(e), rcond=rcond)
    return bias


def compute_input_stats(model_dict: dict, frames: list[dict], device="cuda") -> dict:
    """model_dict with davg / dstd (env_mat_stat.py:386-420: davg = [avg_r, 0,
    0, 0], dstd = [std_r, std_a, std_a, std_a] for every neighbour slot of a
    centre type) and out_bias from ``frames``."""
    m = model_dict["model"] if "model" in model_dict else model_dict
    de = m["descriptor"]
    nt, nnei = len(de["sel"]), int(sum(de["sel"]))
    sums = env_mat_sums(model_dict, frames, device)
    davg = np.zeros((nt, nnei, 4))
    dstd = np.ones((nt, nnei, 4))
    for t in range(nt):
        nr, sr, sr2, na, sa, sa2 = sums[t]
        davg[t, :, 0] = sr / nr if nr else 0.0
        dstd[t, :, 0] = _std(nr, sr, sr2)
        dstd[t, :, 1:] = _std(na, sa, sa2)
    de["@variables"] = {"davg": _nd(davg), "dstd": _nd(dstd)}
    m["@variables"]["out_bias"] = _nd(energy_bias(frames, nt).reshape(1, nt, 1))
    return model_dict


def scale_output_layer(model_dict: dict, factor: float) -> dict:
    """The model with every fitting net's output layer (weights and bias)
    multiplied by ``factor``: atom energies minus their bias, forces and
    virials scale by ``factor`` exactly, relative errors are unchanged."""
    import copy
    d = copy.deepcopy(model_dict)
    m = d["model"] if "model" in d else d
    for net in m["fitting"]["nets"]["networks"]:
        var = net["layers"][-1]["@variables"]
        for key in ("w", "b"):
            v = var.get(key)
            if v is not None and v.get("value") is not None:
                v["value"] = (np.asarray(v["value"], dtype=np.float64) * factor).tolist()
    return d


def physical_scale_model(model_dict: dict, frames: list[dict], force_rms: float = 1.0,
                         device="cuda") -> tuple[dict, float]:
    """A randomly initialised model brought to a physical force scale: input
    statistics (davg / dstd) from ``frames`` (compute_input_stats), then the
    output layers scaled so that the FP64 force RMS over the frames is
    ``force_rms`` (eV/A).  Returns the model dict and the scale factor."""
    from .model import SeAModel
    d = compute_input_stats(model_dict, frames, device)
    m = SeAModel(d, "f64", device)
    ss = []
    for fr in frames:
        for k in range(fr["coord"].shape[0]):
            box = None if fr["box"] is None else fr["box"][k:k + 1]
            s = m.system(fr["coord"][k:k + 1], fr["atype"][k:k + 1], box)
            ss.append(m.evaluate(s)["force"].double().pow(2).sum().item() / max(s.fs.natoms, 1) / 3)
    rms = math.sqrt(sum(ss) / max(len(ss), 1))
    factor = force_rms / rms if rms > 0 else 1.0
    return scale_output_layer(d, factor), factor


# ------------------------------------------------------------------ trainer --

def _rank_world():
    import torch.distributed as dist
    if dist.is_available() and dist.is_initialized():
        return dist.get_rank(), dist.get_world_size()
    return 0, 1


def allreduce_gradients(params, group=None) -> None:
    """Average ``p.grad`` over the ranks of an initialised process group (no-op
    without one): all gradients in one flat FP64 buffer, one all-reduce, then
    divided by the world size.  A gloo group reduces a host copy (gloo is
    given host tensors only); a missing gradient counts as zeros."""
    import torch.distributed as dist
    if not (dist.is_available() and dist.is_initialized()):
        return
    ws = dist.get_world_size(group)
    if ws == 1:
        return
    params = list(params)
    grads = [p.grad if p.grad is not None else torch.zeros_like(p) for p in params]
    flat = torch.cat([g.detach().reshape(-1).to(torch.float64) for g in grads])
    host = dist.get_backend(group) == "gloo" and flat.device.type != "cpu"
    buf = flat.cpu() if host else flat
    dist.all_reduce(buf, group=group)
    buf /= ws
    if host:
        flat.copy_(buf)
    off = 0
    for p, g in zip(params, grads):
        n = g.numel()
        val = flat[off:off + n].view_as(g).to(g.dtype)
        if p.grad is None:
            p.grad = val.clone()
        else:
            p.grad.copy_(val)
        off += n


class Trainer:
    """Single-task energy training from an input.json-style ``config``
    (sections ``model``, ``learning_rate``, ``loss``, ``optimizer``,
    ``training``), starting from ``init_model`` (a serialized model dict or a
    path) or from a fresh initialisation plus input statistics.

    Supported: loss ``ener`` (mse), optimizer ``Adam`` (beta1, beta2,
    weight_decay), ``training.gradient_max_norm`` clipping, ``disp_file`` /
    ``disp_freq`` learning curve (``*_val`` columns from
    ``training.validation_data`` when given), ``save_ckpt`` / ``save_freq``
    checkpoints in the reference's ``.pt`` layout."""

    def __init__(self, config: dict, init_model=None, precision: str = "f64",
                 device="cuda", output_dir: str = ".", restart: bool = False):
        mp = config["model"]
        tr = config["training"]
        lp = dict(config.get("loss", {}))
        op = dict(config.get("optimizer", {}))
        if lp.get("type", "ener") != "ener" or lp.get("loss_func", "mse") != "mse":
            raise NotImplementedError("loss: type ener with mse only")
        if op.get("type", "Adam") != "Adam":
            raise NotImplementedError("optimizer: Adam only")
        self.output_dir = output_dir
        self.rank, self.world = _rank_world()
        td = tr["training_data"]
        self.seed = tr.get("seed")
        # every rank starts from the same weights and statistics (seed) but
        # draws its own batches (seed + rank)
        self.data = DataSystem(td["systems"], td.get("batch_size", "auto"),
                               list(mp["type_map"]),
                               seed=None if self.seed is None else self.seed + self.rank)
        vd = tr.get("validation_data")
        self.valid = None
        if vd:
            # validation batches: numb_btch per display, averaged with natoms
            # weights (training.py:1576-1605)
            self.valid = DataSystem(vd["systems"], vd.get("batch_size", "auto"),
                                    list(mp["type_map"]),
                                    # its own random stream, apart from the training batches
                                    seed=None if self.seed is None else self.seed + 7919)
            self.valid_numb_batch = int(vd.get("numb_btch", 1))
        self.num_steps = int(tr["numb_steps"])
        self.lr = make_learning_rate(config.get("learning_rate", {"start_lr": 1e-3}),
                                     self.num_steps)
        if init_model is None:
            md = init_model_dict(mp, seed=self.seed)
            md = compute_input_stats(md, self.data.stat_frames(int(mp.get("data_stat_nbatch", 10))),
                                     device)
        elif isinstance(init_model, str):
            from .deep_pot import load_model_dict
            md = load_model_dict(init_model)
        else:
            md = init_model
        pref = {k: float(lp.get(k, d)) for k, d in (
            ("start_pref_e", 0.02), ("limit_pref_e", 1.0), ("start_pref_f", 1000.0),
            ("limit_pref_f", 1.0), ("start_pref_v", 0.0), ("limit_pref_v", 0.0))}
        self.model = SeATrainer(md, precision, device, pref_e=pref["start_pref_e"],
                                pref_f=pref["start_pref_f"], pref_v=pref["start_pref_v"],
                                limit_pref_e=pref["limit_pref_e"],
                                limit_pref_f=pref["limit_pref_f"],
                                limit_pref_v=pref["limit_pref_v"], start_lr=self.lr.start_lr,
                                lr=self.lr.start_lr)
        self.model.opt = torch.optim.Adam(
            self.model.parameters(), lr=self.lr.start_lr,
            betas=(float(op.get("adam_beta1", 0.9)), float(op.get("adam_beta2", 0.999))),
            weight_decay=float(op.get("weight_decay", 0.0)))
        self.gradient_max_norm = float(tr.get("gradient_max_norm", 0.0))
        self.disp_file = tr.get("disp_file", "lcurve.out")
        self.disp_freq = int(tr.get("disp_freq", 1000))
        self.save_freq = int(tr.get("save_freq", 1000))
        self.save_ckpt = tr.get("save_ckpt", "model.ckpt")
        # restart: continue at the checkpoint's step index with its optimiser
        # state (training.py:781-785 reads train_infos.step as start_step)
        self.step_id = 0
        self.start_step = 0
        if restart:
            if not isinstance(init_model, str) or not init_model.endswith(".pt"):
                raise ValueError("restart needs a .pt training checkpoint as init_model")
            raw = torch.load(init_model, map_location="cpu", weights_only=True)
            self.step_id = self.start_step = int(
                raw["model"]["_extra_state"]["train_infos"]["step"])
            if "optimizer" in raw:
                self.model.opt.load_state_dict(raw["optimizer"])

    def _label(self, fr: dict) -> dict:
        dev = self.model.device
        lab = {k: torch.tensor(fr[k], dtype=torch.float64, device=dev)
               for k in ("energy", "force", "virial")}
        for k in ("energy", "force", "virial"):
            lab["find_" + k] = fr["find_" + k]
        return lab

    def step(self) -> dict:
        """One optimiser step at the current step index; returns the display
        terms (floats) and the learning rate used."""
        cur_lr = self.lr.value(self.step_id)
        m = self.model
        for g in m.opt.param_groups:
            g["lr"] = cur_lr
        _, fr = self.data.get_batch()
        b = m.batch(fr["coord"], fr["atype"], fr["box"])
        m.opt.zero_grad(set_to_none=True)
        pred = m.predict(b, create_graph=True)
        loss, more = m.loss_terms(pred, self._label(fr), b.nloc, cur_lr)
        loss.backward()
        allreduce_gradients(m.parameters())
        if self.gradient_max_norm > 0.0:
            norm = torch.nn.utils.clip_grad_norm_(m.parameters(), self.gradient_max_norm)
            if not torch.isfinite(norm):
                raise FloatingPointError(f"non-finite gradient norm {norm.item()}")
        m.opt.step()
        self.step_id += 1
        out = {k: float(v) for k, v in more.items()}
        out["lr"] = cur_lr
        return out

    def validate(self, lr: float) -> dict:
        """Display terms averaged over ``numb_btch`` validation batches, each
        weighted by its atom count (training.py:1576-1605)."""
        m = self.model
        acc, tot = {}, 0
        for _ in range(self.valid_numb_batch):
            _, fr = self.valid.get_batch()
            b = m.batch(fr["coord"], fr["atype"], fr["box"])
            pred = m.predict(b, create_graph=False)
            _, more = m.loss_terms(pred, self._label(fr), b.nloc, lr)
            tot += b.nloc
            for k, v in more.items():
                acc[k] = acc.get(k, 0.0) + float(v) * b.nloc
        return {k: v / tot for k, v in acc.items()}

    def save(self, path: str | None = None, lr: float | None = None) -> str | None:
        """Checkpoint of the current weights (rank 0 only; None elsewhere)."""
        if self.rank != 0:
            return None
        path = path or os.path.join(self.output_dir, f"{self.save_ckpt}-{self.step_id}.pt")
        # train_infos.step is the 0-based index of the last step taken (the
        # reference saves step=_step_id = display step - 1, training.py:1760)
        last = max(self.step_id - 1, 0)
        lr = self.lr.value(last) if lr is None else lr
        save_checkpoint(path, self.model.model_dict(), lr=lr, step=last,
                        optimizer_state=self.model.opt.state_dict())
        with open(os.path.join(os.path.dirname(path) or ".", "checkpoint"), "w") as f:
            f.write(os.path.basename(path))
        return path

    def run(self, num_steps: int | None = None) -> list[tuple[int, dict]]:
        """Train ``num_steps`` steps (default: the configured total); writes
        the learning curve (training.py:2256-2323 format: ``# step`` then the
        sorted display keys with suffix ``_trn``, then ``lr``) and
        checkpoints; returns the displayed (step, terms) records."""
        n = self.num_steps - self.step_id if num_steps is None else num_steps
        records = []
        fout = open(os.path.join(self.output_dir, self.disp_file), "a") if self.rank == 0 \
            else open(os.devnull, "w")
        try:
            header = self.step_id == 0
            for _ in range(n):
                res = self.step()
                disp = self.step_id
                if disp % self.disp_freq == 0 or disp == 1:
                    keys = sorted(k for k in res if k != "lr")
                    val = self.validate(res["lr"]) if self.valid is not None else None
                    if header:
                        cols = ("".join("   %11s %11s" % (k + "_val", k + "_trn") for k in keys)
                                if val else "".join("   %11s" % (k + "_trn") for k in keys))
                        fout.write("# {:5s}".format("step") + cols
                                   + "   {:8s}\n".format("lr"))
                        fout.write("# If there is no available reference data, "
                                   "rmse_*_{val,trn} will print nan\n")
                        header = False
                    if val:
                        body = "".join("   %11.2e %11.2e" % (val.get(k, float("nan")), res[k])
                                       for k in keys)
                        res = dict(res, **{k + "_val": v for k, v in val.items()})
                    else:
                        body = "".join("   %11.2e" % res[k] for k in keys)
                    fout.write(f"{disp:7d}" + body + f"   {res['lr']:8.1e}\n")
                    fout.flush()
                    records.append((disp, res))
                # the last step always writes a checkpoint (training.py:1744-1748);
                # none at the step a restart began from
                if disp == self.num_steps or (self.save_freq > 0 and disp % self.save_freq == 0
                                              and disp - 1 != self.start_step):
                    self.save()
        finally:
            fout.close()
        return records

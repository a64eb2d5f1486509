"""Autotuner for libdpk's launch tunables.

The kernels have equivalent organisations whose speed depends on the system
(atoms per cell, neighbour count, number of atoms vs the device's resident
wavefronts): the neighbour search with or without the cell neighbourhood
staged in LDS and its staging capacity, the cell-blocked or per-atom force
kernel, and the size of the persistent BF16 embedding grid.  ``autotune``
times whole E+F+V steps of a real system for each candidate value (HIP
events on the launch stream), one tunable at a time, keeps the fastest and
leaves it in the model's own tuning set (``model.tuning``, made current
around the model's launches), so models tuned for different systems keep
their own choices.  Every candidate gives the same results; only the time
differs.
"""

from __future__ import annotations

import torch

from . import _lib

# name -> (DPK_TUNE_* key, candidate values); the first value is the default
TUNABLES = {
    "nlist_mode": (0, [0, 1]),
    "nlist_stage_pct": (1, [125, 100, 150, 200]),
    "force_mode": (2, [0, 1]),
    "embed_grid_pct": (3, [100, 200, 50, 0]),
    "nlist_waves": (4, [0, 4, 8, 16]),
    "force_waves": (5, [0, 4, 8, 16]),
    "nlist_sort": (6, [0, 1]),
}

_CACHE: dict = {}


def get(model=None) -> dict:
    """The model's tunables, or the process-wide defaults without a model."""
    L = _lib.lib()
    if model is None:
        return {name: int(L.dpk_tune_get(key)) for name, (key, _) in TUNABLES.items()}
    return {name: model.tuning.get(key) for name, (key, _) in TUNABLES.items()}


def set_values(values: dict, model=None) -> None:
    """Set tunables in the model's set, or the process-wide defaults."""
    for name, v in values.items():
        key, _ = TUNABLES[name]
        if model is None:
            _lib.call("dpk_tune_set", key, int(v))
        else:
            model.tuning.set(key, int(v))


def cache_key(model, system) -> tuple:
    """What the choice depends on: the device, the system's size and
    density of neighbours (natoms, frames, sel per type), and the model's
    kernels (precision, compression, type_one_side, ntypes)."""
    return (torch.cuda.get_device_name(), system.fs.natoms, system.fs.nframes, model.ntypes,
            tuple(int(x) for x in model.sel), model.nnei, bool(model.type_one_side),
            model.precision, model.compressed)


def _time_steps(model, system, reps: int, rounds: int = 3) -> float:
    """Fastest of ``rounds`` timings of ``reps`` steps, per step (the minimum
    drops one-off stalls, so ~1 % differences between candidates show)."""
    model.evaluate(system, check_nlist=False)  # warm-up with the new setting
    st = torch.cuda.current_stream()
    best = float("inf")
    for _ in range(rounds):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        for _ in range(reps):
            model.evaluate(system, check_nlist=False)
        b.record(st)
        b.synchronize()
        best = min(best, a.elapsed_time(b) / reps)
    return best


def autotune(model, system, reps: int = 3, names=None) -> dict:
    """Choose every tunable for (model, system) in ``model.tuning``; returns
    {name: value, "ms": t}.  Results are cached per ``cache_key``."""
    key = cache_key(model, system)
    if key in _CACHE:
        set_values({k: v for k, v in _CACHE[key].items() if k in TUNABLES}, model)
        return dict(_CACHE[key])
    model.evaluate(system, check_nlist=True)  # capacity settled before timing
    torch.cuda.synchronize()
    best = get(model)
    best_ms = _time_steps(model, system, reps)
    for name in (names or TUNABLES):
        _, values = TUNABLES[name]
        for v in values:
            if v == best[name]:
                continue
            trial = dict(best, **{name: v})
            set_values(trial, model)
            ms = _time_steps(model, system, reps)
            if ms < best_ms * 0.99:  # keep a change only when clearly faster
                best, best_ms = trial, ms
        set_values(best, model)
    out = dict(best, ms=best_ms)
    _CACHE[key] = out
    return dict(out)

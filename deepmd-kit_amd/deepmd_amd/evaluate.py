"""Model test on labelled data systems (the reference's ``dp test`` for
energy models): energy, energy per atom, force, virial and virial per atom
MAE / RMSE of a ``DeepPot`` over ``set.*/*.npy`` systems.

Metric rules follow deepmd/utils/eval_metrics.py:41-195 (MAE and RMSE of the
prediction - label differences; the per-atom forms scale by 1 / natoms; a
metric exists only when the system has that label, virials only with
periodic boxes) and the combination over systems follows
deepmd/utils/weight_avg.py (MAE averaged with weights = number of compared
values, RMSE as the root of the weighted mean square).
"""

from __future__ import annotations

import numpy as np

from .training import DeepmdData

KEYS = (("energy", "mae_e", "rmse_e"), ("energy_per_atom", "mae_ea", "rmse_ea"),
        ("force", "mae_f", "rmse_f"), ("virial", "mae_v", "rmse_v"),
        ("virial_per_atom", "mae_va", "rmse_va"))


def error_stat(pred: np.ndarray, label: np.ndarray, scale: float = 1.0) -> tuple:
    """(mae, rmse, weight) of pred - label (eval_metrics.py:114-127)."""
    d = np.asarray(pred, dtype=np.float64) - np.asarray(label, dtype=np.float64)
    return (float(np.mean(np.abs(d))) * scale, float(np.sqrt(np.mean(d * d))) * scale,
            float(d.size))


def system_metrics(pred: dict, data: dict, natoms: int, pbc: bool) -> dict:
    """{metric: (mae, rmse, weight)} of one system (eval_metrics.py:146-193)."""
    out = {}
    if data.get("find_energy", 0.0):
        out["energy"] = error_stat(pred["energy"].reshape(-1, 1), data["energy"].reshape(-1, 1))
        out["energy_per_atom"] = error_stat(pred["energy"].reshape(-1, 1),
                                            data["energy"].reshape(-1, 1), 1.0 / natoms)
    if data.get("find_force", 0.0):
        out["force"] = error_stat(pred["force"].reshape(-1), data["force"].reshape(-1))
    if pbc and data.get("find_virial", 0.0):
        out["virial"] = error_stat(pred["virial"].reshape(-1, 9), data["virial"].reshape(-1, 9))
        out["virial_per_atom"] = error_stat(pred["virial"].reshape(-1, 9),
                                            data["virial"].reshape(-1, 9), 1.0 / natoms)
    return out


def merge(per_system: list[dict]) -> dict:
    """Weighted combination over systems (weight_avg.py)."""
    out = {}
    for name, kmae, krmse in KEYS:
        parts = [m[name] for m in per_system if name in m]
        if not parts:
            continue
        w = sum(p[2] for p in parts)
        out[kmae] = sum(p[0] * p[2] for p in parts) / w
        out[krmse] = float(np.sqrt(sum(p[1] * p[1] * p[2] for p in parts) / w))
    return out


def test_model(dp, systems, numb_test: int = 0) -> dict:
    """Evaluate ``dp`` (a ``deep_pot.DeepPot``) on labelled systems; returns
    {"systems": [per-system {key: value}], "average": {key: value}} with the
    keys mae_e, rmse_e, mae_ea, rmse_ea, mae_f, rmse_f (and the virial ones
    where labels exist).  ``numb_test`` > 0 takes the first frames only."""
    if isinstance(systems, str):
        systems = [systems]
    tm = dp.get_type_map()
    per, rows = [], []
    for path in systems:
        d = DeepmdData(path, type_map=tm or None)
        n = d.nframes if numb_test <= 0 else min(numb_test, d.nframes)
        fr = d.frames(np.arange(n))
        e, f, v = dp.eval(fr["coord"].reshape(n, -1),
                          None if fr["box"] is None else fr["box"], d.atom_type)
        pred = {"energy": np.asarray(e), "force": np.asarray(f), "virial": np.asarray(v)}
        m = system_metrics(pred, fr, d.natoms, d.pbc)
        per.append(m)
        rows.append(merge([m]))
    return {"systems": rows, "average": merge(per)}

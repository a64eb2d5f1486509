"""Benchmark: atom-steps/s of one E+F+V evaluation of the se_e2_a model on MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c4] [--precision bf16]

Workload (default, the headline configuration): FCC Cu, 32x32x32 conventional
cells (131,072 atoms), a = 3.615 A, Gaussian jitter sigma = 0.05 A, one type;
model rcut 8, rcut_smth 2, sel [200], embedding [32, 64, 128], axis 16,
fitting [240, 240, 240], seeded random weights.  One step = neighbour list
(cell list) + env matrix + embedding + descriptor + fitting + backward +
forces + virial for the whole frame, with inputs resident in HBM.  Mixed
precision "f16x2": FP16 MFMA weights with every data operand carried as a
hi + lo pair of FP16 values, FP32 accumulation; FP64 geometry, energy sum and
virial.

Prints one JSON line (rank 0).  Besides the throughput it reports, for the
dominant kernel, the roofline fraction (algorithmic FLOPs per launch / launch
time measured with HIP events on the launch stream, against the BF16 dense
MFMA peak), the error of forces / energy / virial against this framework's
FP64 path on the same input -- for the random-init model and for the same
architecture brought to a physical force scale (input statistics, forces
scaled to 1 eV/A RMS), the latter being the force-RMSE <= 1e-3 eV/A gate --,
the FP64 path's throughput, and a CPU baseline (the FP64 CPU oracle on a
bounded sample, timed on this host's cores).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "deepmd-kit_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

PEAK = {"bf16": 2.5e15, "f16": 2.5e15, "f16x2": 2.5e15, "f32": 157.3e12,
        "f64": 78.6e12}   # dense MFMA of the operand type, FLOP/s
FORCE_RMSE_GATE = 1e-3   # eV/A, mixed vs FP64 at a physical force scale
METRIC = "atom-steps/s E+F+V on 1 MI355X, mixed vs FP64 at force RMSE<=1e-3 eV/Å"


def configs():
    return {
        "c4": dict(workload="FCC Cu 32x32x32 (131072 atoms), rcut 8, sel [200], emb [32,64,128], "
                            "axis 16, fit [240,240,240], E+F+V one frame",
                   kind="fcc", ncell=32, a=3.615, sigma=0.05, geom_seed=40, weight_seed=1,
                   rcut=8.0, rcut_smth=2.0, sel=[200], neuron=[32, 64, 128], axis=16,
                   fit=[240, 240, 240], precision="f16x2", phys_ncell=8),
        "c0": dict(workload="64 H2O (192 atoms), 12.44 A box, rcut 6, sel [46,92], emb [25,50,100], "
                            "axis 16, fit [240,240,240], E+F+V one frame",
                   kind="water", nside=4, box=12.44, geom_seed=0, weight_seed=1, rcut=6.0,
                   rcut_smth=0.5, sel=[46, 92], neuron=[25, 50, 100], axis=16,
                   fit=[240, 240, 240], precision="f64"),
        "c3": dict(workload="256 ragged frames (24..768 atoms, FCC/BCC a=3.6 A, sigma 0.08 A, "
                            "4 species), rcut 6, sel [60,60,60,60], emb [32,64,128], axis 16, "
                            "fit [240,240,240], E+F+V, one hipGraph-captured step",
                   kind="ragged", nframes=256, nmin=24, nmax=768, a=3.6, sigma=0.08, ntypes=4,
                   geom_seed=30, weight_seed=1, rcut=6.0, rcut_smth=0.5, sel=[60, 60, 60, 60],
                   neuron=[32, 64, 128], axis=16, fit=[240, 240, 240], type_one_side=False,
                   precision="f32"),
        "c2": dict(workload="training step: 8 frames x 64 H2O (1536 atoms), 12.44 A box, "
                            "energy+force+virial MSE loss (prefactors 1) vs a second seeded "
                            "model, BF16 embedding GEMMs, Adam; rcut 6, sel [46,92], "
                            "emb [25,50,100], axis 16, fit [240,240,240]",
                   kind="train", nside=4, box=12.44, nframes=8, geom_seeds=list(range(20, 28)),
                   weight_seed=1, label_seed=2, rcut=6.0, rcut_smth=0.5, sel=[46, 92],
                   neuron=[25, 50, 100], axis=16, fit=[240, 240, 240], precision="bf16"),
        "c1": dict(workload="4096 H2O (12288 atoms), 49.76 A box, rcut 6, sel [46,92], "
                            "emb [25,50,100] compressed (quintic tables, stride 0.01/0.1, "
                            "extrapolate 5), axis 16, fit [240,240,240], E+F+V one frame",
                   kind="water", nside=16, box=49.76, geom_seed=10, weight_seed=1, rcut=6.0,
                   rcut_smth=0.5, sel=[46, 92], neuron=[25, 50, 100], axis=16,
                   fit=[240, 240, 240], precision="f32",
                   compress=dict(min_nbor_dist=0.5, extrapolate=5.0, stride0=0.01, stride1=0.1)),
    }


def make_system(cfg):
    """List of frames (coords [n,3], atype [n], box [3,3])."""
    from deepmd_amd import systems
    if cfg["kind"] == "fcc":
        return [systems.fcc(cfg["ncell"], cfg["a"], cfg["sigma"], cfg["geom_seed"])]
    if cfg["kind"] == "ragged":
        return systems.ragged_metal(cfg["nframes"], cfg["nmin"], cfg["nmax"], cfg["a"],
                                    cfg["sigma"], cfg["ntypes"], cfg["geom_seed"])
    return [systems.water(cfg["nside"], cfg["box"], cfg["geom_seed"])]


def run_train(cfg, precision, steps, warmup):
    """Time `steps` training steps (forward, force/virial, loss, backward
    through the second-order kernels, Adam) on one batch, replayed from a
    HIP graph; the eager step time is reported beside it."""
    import numpy as np
    import torch
    from deepmd_amd import systems
    from deepmd_amd.model import SeAModel, random_model_dict
    from deepmd_amd.train import SeATrainer
    fr = [systems.water(cfg["nside"], cfg["box"], sd) for sd in cfg["geom_seeds"]]
    coords = np.stack([f[0] for f in fr])
    atypes = np.stack([f[1] for f in fr])
    boxes = np.stack([f[2].reshape(9) for f in fr])
    args = (cfg["rcut"], cfg["rcut_smth"], cfg["sel"], cfg["neuron"], cfg["axis"], cfg["fit"])
    d, d2 = random_model_dict(cfg["weight_seed"], *args), random_model_dict(cfg["label_seed"], *args)
    lm = SeAModel(d2, "f64", "cuda")
    out = lm.evaluate(lm.system(coords, atypes, boxes))
    nf, n = coords.shape[:2]
    label = {"energy": out["energy"].clone(), "force": out["force"].view(nf, n, 3).clone(),
             "virial": out["virial"].clone()}
    tr = SeATrainer(d, precision, "cuda", capturable=True)
    b = tr.batch(coords, atypes, boxes)
    ref = SeATrainer(d, "f64", "cuda")
    l64, g64, p64 = ref.loss_and_grads(ref.batch(coords, atypes, boxes), label)
    lmx, gmx, pmx = tr.loss_and_grads(b, label)
    a = torch.cat([x.detach().double().reshape(-1) for x in gmx])
    r = torch.cat([x.detach().reshape(-1) for x in g64])
    err = {"loss_rel_err": float(abs(lmx.item() - l64.item()) / abs(l64.item())),
           "grad_rel_err": float((a - r).norm() / r.norm()),
           "grad_cosine": float(torch.dot(a, r) / (a.norm() * r.norm()))}
    # no autograd graph of the trainer may outlive this point: graph capture
    # needs the weights' gradient accumulators to be created on its stream
    del lmx, gmx, pmx, l64, g64, p64, a, r, ref
    # eager steps (for reference), then the step captured in a HIP graph
    for _ in range(max(warmup, 1)):
        tr.step(b, label)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        tr.step(b, label)
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) / max(steps, 1) * 1e3
    g = tr.capture_step(b, label, warmup=1)
    g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = g.replay()
    torch.cuda.synchronize()
    err["eager_ms_per_step"] = eager_ms
    return time.perf_counter() - t0, nf * n, err, float(loss.item())


def model_dict(cfg):
    from deepmd_amd import model
    return model.random_model_dict(cfg["weight_seed"], cfg["rcut"], cfg["rcut_smth"], cfg["sel"],
                                   cfg["neuron"], cfg["axis"], cfg["fit"],
                                   type_one_side=cfg.get("type_one_side", True))


def embed_flops(cfg, natoms, slots=None):
    """Algorithmic FLOPs of one embedding launch (forward or backward count the
    same): 2 (d1 + d1 d2 + d2 d3) + 8 d3 per computed slot; `slots` = the
    slots the kernels evaluate (all natoms * S, or, where type blocks stop at
    their last real neighbour, the tiles of 16 up to it)."""
    S = sum(cfg["sel"])
    d1, d2, d3 = cfg["neuron"]
    n = natoms * S if slots is None else slots
    return n * (2 * (d1 + d1 * d2 + d2 * d3) + 8 * d3)


def computed_slots(m, s):
    """Slots the embedding kernels evaluate: per atom and type block, tiles of
    16 up to the block's last real neighbour (None where every slot is)."""
    import numpy as np
    if not getattr(m, "skip_padding", False) or m.compressed:
        return None
    S = sum(m.sel)
    nl = s.nlist.view(-1, S).cpu().numpy()
    total = 0
    for t in range(len(m.sel)):
        blk = nl[:, m.sec[t]:m.sec[t + 1]] >= 0
        w = blk.shape[1]
        end = np.where(blk.any(1), w - np.argmax(blk[:, ::-1], axis=1), 0)
        total += int(np.sum((end + 15) // 16) * 16)
    return total


def step_flops(cfg, natoms):
    """Algorithmic FLOPs per E+F+V step: 2 * (S E_emb + C + D + F_fit) per atom."""
    S = sum(cfg["sel"])
    d1, d2, d3 = cfg["neuron"]
    m2 = cfg["axis"]
    din = d3 * m2
    f = cfg["fit"]
    fit = 2 * (din * f[0] + sum(f[i] * f[i + 1] for i in range(len(f) - 1)) + f[-1])
    fwd = S * 2 * (d1 + d1 * d2 + d2 * d3) + 8 * d3 * S + 8 * d3 * m2 + fit
    return natoms * 2 * fwd


def run_gpu(cfg, precision, steps, warmup, frames, want_timers=True, graph=False,
            autotune=False, d=None, lower=False):
    """Time `steps` E+F+V evaluations of the packed frames.  With `graph` the
    step is one HIP-graph replay; otherwise eager launches with per-stage HIP
    event timers on the launch stream.  `d`: the model dict (default: the
    configuration's seeded model)."""
    import torch
    from deepmd_amd.model import SeAModel
    d = model_dict(cfg) if d is None else d
    m = SeAModel(d, precision, "cuda")
    if cfg.get("compress") and precision != "f64":  # the FP64 reference stays uncompressed
        m.enable_compression(**cfg["compress"])
    if len(frames) == 1:
        x, at, box = frames[0]
        s = m.system(x[None], at[None], box.reshape(1, 9))
    else:
        s = m.system([f[0] for f in frames], [f[1] for f in frames],
                     [f[2].reshape(9) for f in frames], nlocs=[len(f[1]) for f in frames])
    stage_ms = {}
    tuning = None
    if autotune:
        from deepmd_amd import autotune as at
        tuning = at.autotune(m, s)
    if graph:
        g = m.capture(s)
        for _ in range(max(warmup, 1)):
            out = g.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            out = g.replay()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        g.check()
        if want_timers:  # stage split from eager runs (not the timed region)
            timers = {}
            for _ in range(3):
                m.evaluate(s, check_nlist=False, timers=timers)
            torch.cuda.synchronize()
            stage_ms = {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in timers.items()}
            out = g.replay()
            torch.cuda.synchronize()
    else:
        for _ in range(max(warmup, 1)):
            out = m.evaluate(s, check_nlist=True)
        torch.cuda.synchronize()
        timers = {} if want_timers else None
        t0 = time.perf_counter()
        for _ in range(steps):
            out = m.evaluate(s, check_nlist=False, timers=timers)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        if timers:
            stage_ms = {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in timers.items()}
    nmax = int(s.status[0].item())
    if nmax > s.fs.max_cand:
        raise RuntimeError(f"neighbour capacity exceeded during timing ({nmax} > {s.fs.max_cand})")
    res = {k: out[k].detach().double().cpu().numpy() for k in ("energy", "force", "virial")}
    res["computed_slots"] = computed_slots(m, s)
    if lower and len(frames) == 1:
        res["md_lower"] = time_lower(m, s, steps, warmup, float(res["energy"].ravel()[0]))
    return elapsed, stage_ms, res, nmax, s.fs.natoms, tuning


def lower_inputs(s):
    """An MD engine's view of one packed frame: the extended region (local
    atoms first, then the periodic images the neighbour list references) and
    the list in extended indices.  Built on the device from the frame's own
    list, whose extended index is image * natoms + j with image shift
    shift_vec[image]."""
    import torch
    n = s.fs.natoms
    dev = s.nlist.device
    nl = s.nlist.view(n, -1).long()
    sv = s.fs.shift_vec
    home = int(torch.nonzero((sv == 0).all(1))[0])
    used = torch.unique(nl[nl >= 0])
    ghosts = used[used // n != home]
    ext = torch.cat([home * n + torch.arange(n, device=dev), ghosts])
    newidx = torch.full((sv.shape[0] * n,), -1, dtype=torch.long, device=dev)
    newidx[ext] = torch.arange(ext.numel(), device=dev)
    nl2 = torch.where(nl >= 0, newidx[nl.clamp(min=0)], -1).to(torch.int32)
    ec = s.wrapped.view(-1, 3)[ext % n] + sv[ext // n]
    ea = s.atype[ext % n]
    return ec.contiguous(), ea.contiguous(), nl2.contiguous()


def time_lower(m, s, steps, warmup, energy_ref):
    """forward_lower as an MD engine drives it between re-neighbourings
    (ago > 0: stored list and slot order, no host synchronisation), one
    HIP-graph replay per step, on the extended region of the packed frame."""
    import torch
    ec, ea, nl = lower_inputs(s)
    first = m.forward_lower(ec[None], ea[None], nl[None])  # ago == 0: list taken in, checked
    ls = m._last_lower_system
    xin = ec.clone()
    for _ in range(2):
        m.forward_lower(xin[None], ea[None], None, system=ls, ago=1, copy_outputs=False)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        out = m.forward_lower(xin[None], ea[None], None, system=ls, ago=1, copy_outputs=False)
    for _ in range(max(warmup, 1)):
        graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        graph.replay()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    e = float(out["energy"][0])
    n = s.fs.natoms
    return {"api": "SeAModel.forward_lower(ago > 0), the DeepPotPT::compute(nghost, InputNlist, "
                   "ago) path", "natoms_local": n, "nall": int(ec.shape[0]),
            "nnei_in": int(nl.shape[1]), "hip_graph": True, "steps": steps,
            "ms_per_step": el / steps * 1e3, "value": n * steps / el, "unit": "atom-steps/s",
            "energy_rel_vs_packed_path": abs(e - energy_ref) / abs(energy_ref),
            "energy_rel_first_call": abs(float(first["energy"][0]) - e) / abs(e)}


def physical_model(cfg):
    """The configuration's architecture at a physical force scale: input
    statistics from a smaller box of the same material (another seed), output
    layers scaled to a 1 eV/A force RMS there (FP64)."""
    from deepmd_amd import systems, training
    n = cfg.get("phys_ncell", 8)
    x, at, box = systems.fcc(n, cfg["a"], cfg["sigma"], cfg["geom_seed"] + 1)
    fr = [{"coord": x[None], "atype": at[None], "box": box.reshape(1, 9), "find_energy": 0.0}]
    d, factor = training.physical_scale_model(model_dict(cfg), fr, 1.0)
    return d, {"stats_sample": f"FCC {n}x{n}x{n} ({len(at)} atoms), seed {cfg['geom_seed'] + 1}",
               "output_scale": factor}


def force_errors(res, ref, frames):
    import numpy as np
    df = res["force"] - ref["force"]
    nat = np.array([len(f[1]) for f in frames])
    return {
        "force_rmse_eV_per_A": float(np.sqrt(np.mean(df ** 2))),
        "force_rms_eV_per_A": float(np.sqrt(np.mean(ref["force"] ** 2))),
        "force_rel_rmse": float(np.sqrt(np.mean(df ** 2) / max(np.mean(ref["force"] ** 2), 1e-300))),
        "force_max_abs_err_eV_per_A": float(np.max(np.abs(df))),
        "energy_per_atom_err_eV": float(np.max(np.abs(res["energy"] - ref["energy"]) / nat)),
        "virial_max_abs_err_eV": float(np.max(np.abs(res["virial"] - ref["virial"]))),
        "virial_max_abs_eV": float(np.max(np.abs(ref["virial"]))),
        "virial_per_atom_err_eV": float(np.max(np.abs(res["virial"] - ref["virial"]).reshape(
            len(frames), -1) / nat[:, None])),
    }


def host_facts():
    """CPU model (lscpu), logical CPUs available and torch's BLAS / LAPACK."""
    import re
    import subprocess
    import torch
    model = "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        m = re.search(r"Model name:\s*(.+)", out)
        model = m.group(1).strip() if m else model
    except (OSError, subprocess.SubprocessError):
        pass
    cfg = torch.__config__.show()
    blas = re.search(r"BLAS_INFO=(\S+)", cfg)
    lapack = re.search(r"LAPACK_INFO=(\S+)", cfg)
    return {"cpu_model": model, "cpus_available": len(os.sched_getaffinity(0)),
            "blas": blas.group(1).rstrip(",") if blas else "unknown",
            "lapack": lapack.group(1).rstrip(",") if lapack else "unknown"}


def cpu_baseline(cfg, target_s=12.0):
    """FP64 CPU oracle (torch CPU, autograd forces) on a bounded sample of the
    same workload: a smaller box of the same material and model."""
    import numpy as np
    import torch
    from deepmd_amd import systems
    from oracle import sea
    threads = len(os.sched_getaffinity(0))
    torch.set_num_threads(threads)
    d = model_dict(cfg)
    om = sea.model_from_serialized(d)
    if cfg["kind"] == "fcc":
        nc = 6
        frames = [systems.fcc(nc, cfg["a"], cfg["sigma"], cfg["geom_seed"])]
        sample = f"FCC Cu {nc}x{nc}x{nc} ({len(frames[0][1])} atoms), same model, FP64 torch-CPU oracle"
    elif cfg["kind"] == "ragged":
        frames = make_system(cfg)[:8]
        sample = (f"first 8 of the ragged frames ({sum(len(f[1]) for f in frames)} atoms), "
                  "same model, FP64 torch-CPU oracle")
    else:
        frames = [systems.water(cfg["nside"], cfg["box"], cfg["geom_seed"])]
        sample = f"{len(frames[0][1])}-atom water, same model, FP64 torch-CPU oracle"
    natoms = sum(len(f[1]) for f in frames)

    def one_pass():
        for x, at, box in frames:
            sea.forward_chunked(om, torch.tensor(x).view(1, -1), torch.tensor(at).view(1, -1).long(),
                                torch.tensor(box.ravel()).view(1, 9), chunk=1024)

    n, t0 = 0, time.perf_counter()
    while True:
        one_pass()
        n += 1
        el = time.perf_counter() - t0
        if el > target_s or n >= 50:
            break
    hf = host_facts()
    return {"value": natoms * n / el, "unit": "atom-steps/s", "cores": threads, "kind": "port",
            "sample": f"{sample}; {n} E+F+V evaluations in {el:.1f} s, torch {torch.__version__}, "
                      f"{threads} threads on {hf['cpu_model']}, BLAS {hf['blas']}",
            "host": hf}


def copy_calibration(nbytes=1 << 30, reps=3):
    """A plain device copy of known size: under rocprofv3 --pmc its FETCH_SIZE /
    WRITE_SIZE per dispatch give the counters' scale (bytes moved = nbytes each
    way)."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}))


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` from its committed counter summary
    (profiles/<kernel>_c4_traffic.md, line "traffic_bytes_per_launch = X"),
    or None if there is none."""
    import re
    path = os.path.join(ROOT, "profiles", f"{kernel}_c4_traffic.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def limiter_from_profile(kernel):
    """What bounds `kernel` besides its roofline: VALU / MFMA busy percentages
    from the committed utilisation summary (profiles/c4_valu_mfma_utilisation.md,
    line "<kernel>: valu_busy_pct = V, mfma_busy_pct = M"), or None."""
    import re
    path = os.path.join(ROOT, "profiles", "c4_valu_mfma_utilisation.md")
    if not os.path.exists(path):
        return None
    m = re.search(rf"{kernel}: valu_busy_pct = ([0-9.]+), mfma_busy_pct = ([0-9.]+)",
                  open(path).read())
    if not m:
        return None
    return {"valu_busy_pct": float(m.group(1)), "mfma_busy_pct": float(m.group(2)),
            "source": "profiles/c4_valu_mfma_utilisation.md"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c4", choices=sorted(configs()))
    ap.add_argument("--precision", default=None, choices=["f16x2", "f16", "bf16", "f32", "f64"],
                    help="default: the configuration's (f16x2 for c4, f32 for c3, f64 for c0)")
    ap.add_argument("--graph", default="on", choices=["on", "off"],
                    help="time HIP-graph replays of the captured step (on) or eager launches")
    ap.add_argument("--autotune", default="on", choices=["on", "off"],
                    help="choose libdpk launch tunables by timing them on this system first")
    ap.add_argument("--f64-steps", type=int, default=3)
    ap.add_argument("--md-lower", default="on", choices=["on", "off"],
                    help="also time forward_lower (the MD-engine path) on the same frame")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build evaluates one frame on one GPU; --gpus {args.gpus} is not "
              "supported (multi-GPU spatial decomposition is not built)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    if args.copy_calibration:
        copy_calibration()
        return
    cfg = configs()[args.config]
    prec = args.precision or cfg.get("precision", "bf16")
    if cfg["kind"] == "train":
        el, natoms, err, loss = run_train(cfg, prec, args.steps, args.warmup)
        print(json.dumps({
            "metric": "atom-steps/s of the E+F+V training step on 1 MI355X (c2)",
            "value": natoms * args.steps / el, "unit": "atom-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": prec,
            "data": "synthetic", "config": {"workload": cfg["workload"], "natoms": natoms,
                                            "hip_graph": True},
            "error_vs_fp64": err, "final_loss": loss}))
        return
    frames = make_system(cfg)
    graph = args.graph == "on"
    elapsed, stage_ms, res, nmax, natoms, tuning = run_gpu(cfg, prec, args.steps, args.warmup,
                                                           frames, graph=graph,
                                                           autotune=args.autotune == "on",
                                                           lower=args.md_lower == "on")
    value = natoms * args.steps / elapsed
    # error and throughput of the FP64 path on the same input
    f64_el, f64_stage, ref, _, _, _ = run_gpu(cfg, "f64", max(args.f64_steps, 1), 1, frames,
                                              want_timers=False, graph=graph)
    f64_value = natoms * max(args.f64_steps, 1) / f64_el
    err = force_errors(res, ref, frames)
    # the accuracy gate: the same architecture at a physical force scale
    gate = None
    if prec not in ("f64",) and cfg["kind"] == "fcc":
        dph, info = physical_model(cfg)
        _, _, rph, _, _, _ = run_gpu(cfg, prec, 0, 1, frames, want_timers=False, d=dph)
        _, _, rph64, _, _, _ = run_gpu(cfg, "f64", 0, 1, frames, want_timers=False, d=dph)
        eph = force_errors(rph, rph64, frames)
        gate = dict(info, **eph, limit_eV_per_A=FORCE_RMSE_GATE,
                    passed=bool(eph["force_rmse_eV_per_A"] <= FORCE_RMSE_GATE))
    # dominant kernel and its roofline
    kern = max((k for k in stage_ms if k.startswith("embed")), key=lambda k: stage_ms[k])
    # algorithmic work: every one of the S N slots (the model's fixed shape)
    flops = embed_flops(cfg, natoms)
    flops_evaluated = embed_flops(cfg, natoms, res.get("computed_slots"))
    if cfg.get("compress"):  # tabulated: 18 d3 S forward, 36 d3 S backward (VALU)
        flops = natoms * sum(cfg["sel"]) * cfg["neuron"][-1] * (36 if kern == "embed_bwd" else 18)
    achieved = flops / (stage_ms[kern] * 1e-3)
    peak = PEAK[prec]
    out = {
        "metric": METRIC if args.config == "c4" else f"atom-steps/s E+F+V on 1 MI355X ({args.config})",
        "value": value, "unit": "atom-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": prec, "data": "synthetic",
        "config": {"workload": cfg["workload"], "natoms": natoms, "nframes": len(frames),
                   "precision": prec, "nlist_included": True, "hip_graph": graph,
                   "max_within_rcut": nmax, "tuning": tuning},
        "roofline": {"kernel": kern, "bound": "valu" if cfg.get("compress") else "mfma",
                     "achieved": achieved / 1e12,
                     "peak": peak / 1e12, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": traffic_from_profile(kern) if args.config == "c4" else None,
                     "flops_per_launch": flops, "launch_ms": stage_ms[kern],
                     "flops_evaluated_slots": flops_evaluated,
                     "limiter": limiter_from_profile(kern) if args.config == "c4" else None},
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "md_lower": res.get("md_lower"),
        "step_tflops_algorithmic": step_flops(cfg, natoms) / (elapsed / args.steps) / 1e12,
        "error_vs_fp64": err,
        "accuracy_gate_physical_scale": gate,
        "fp64_gpu": {"value": f64_value, "ms_per_step": f64_el / max(args.f64_steps, 1) * 1e3,
                     "speedup_mixed_over_fp64": value / f64_value},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

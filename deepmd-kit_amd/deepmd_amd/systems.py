"""Synthetic systems for the benchmark and parity workloads (numpy PCG64).

* :func:`water` -- O on a g x g x g lattice in a cubic box with uniform
  jitter, each H at 0.9572 A from its O with an H-O-H angle of 104.52 deg in a
  random orientation (QR of a Gaussian 3x3, sign-fixed); atoms ordered
  [O, H, H, ...], types O = 0, H = 1.
* :func:`fcc` / :func:`bcc` -- conventional-cell lattices with Gaussian jitter.
* :func:`ragged_metal` -- frames with random atom counts on FCC/BCC sites,
  random species.
"""

from __future__ import annotations

import numpy as np

OH = 0.9572
HOH = np.deg2rad(104.52)


def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def water(nmol_side: int, box: float, seed: int, jitter: float = 0.3):
    """(coord [3 n^3, 3], atype, box [3,3])."""
    rng = np.random.default_rng(seed)
    g = nmol_side
    idx = np.stack(np.meshgrid(np.arange(g), np.arange(g), np.arange(g), indexing="ij"),
                   -1).reshape(-1, 3)
    o = (idx + 0.5) * box / g + rng.uniform(-jitter, jitter, size=(idx.shape[0], 3))
    h = 0.5 * HOH
    u1 = np.array([np.sin(h), np.cos(h), 0.0])
    u2 = np.array([-np.sin(h), np.cos(h), 0.0])
    out = np.empty((idx.shape[0], 3, 3))
    for k in range(idx.shape[0]):
        q = _rotation(rng)
        out[k, 0] = o[k]
        out[k, 1] = o[k] + OH * (q @ u1)
        out[k, 2] = o[k] + OH * (q @ u2)
    at = np.tile(np.array([0, 1, 1], dtype=np.int32), idx.shape[0])
    return out.reshape(-1, 3), at, np.diag([box] * 3)


def _lattice(kind, ncell, a):
    base = {"fcc": np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]),
            "bcc": np.array([[0, 0, 0], [0.5, 0.5, 0.5]])}[kind]
    n = np.asarray(ncell)
    idx = np.stack(np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]),
                               indexing="ij"), -1).reshape(-1, 1, 3)
    return ((idx + base[None]) * a).reshape(-1, 3)


def fcc(ncell: int, a: float, sigma: float, seed: int, ntypes: int = 1):
    rng = np.random.default_rng(seed)
    x = _lattice("fcc", (ncell,) * 3, a)
    x = x + rng.normal(0.0, sigma, size=x.shape)
    at = (rng.integers(0, ntypes, size=x.shape[0]) if ntypes > 1
          else np.zeros(x.shape[0])).astype(np.int32)
    return x, at, np.diag([ncell * a] * 3)


def ragged_metal(nframes: int, nmin: int, nmax: int, a: float, sigma: float, ntypes: int,
                 seed: int):
    """Frames with N_f ~ U{nmin..nmax} atoms on random FCC or BCC sites of the
    smallest near-cubic supercell holding N_f sites."""
    rng = np.random.default_rng(seed)
    frames = []
    for _ in range(nframes):
        n = int(rng.integers(nmin, nmax + 1))
        kind = "fcc" if rng.random() < 0.5 else "bcc"
        per = 4 if kind == "fcc" else 2
        m = 1
        while per * m ** 3 < n:
            m += 1
        sites = _lattice(kind, (m, m, m), a)
        pick = rng.choice(sites.shape[0], size=n, replace=False)
        x = sites[pick] + rng.normal(0.0, sigma, size=(n, 3))
        at = rng.integers(0, ntypes, size=n).astype(np.int32)
        frames.append((x, at, np.diag([m * a] * 3)))
    return frames

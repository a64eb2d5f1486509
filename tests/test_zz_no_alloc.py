"""The allocation-free neighbour-list entry point (dpk_nlist_build_noalloc,
static device scratch) against dpk_nlist_build with a caller's workspace:
bit-identical lists, wrapped coordinates and status on two calls in a row,
the second on a smaller system, and a clear error above the scratch limit.
Calls go through the C ABI (ctypes)."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda"


@pytest.fixture(scope="module", autouse=True)
def _hip():
    if not torch.cuda.is_available():
        pytest.fail("a HIP device is required for the gpu tests")


def _system(kind):
    from deepmd_amd import systems
    from deepmd_amd.model import SeAModel, random_model_dict
    if kind == "cu":
        d = random_model_dict(1, 8.0, 2.0, [200], [4, 8, 16], 4, [8, 8])
        x, at, box = systems.fcc(12, 3.615, 0.05, 3)   # 6912 atoms
    else:
        d = random_model_dict(1, 6.0, 0.5, [46, 92], [4, 8, 16], 4, [8, 8])
        x, at, box = systems.water(3, 9.33, 5)          # 81 atoms
    m = SeAModel(d, "f64", DEV)
    return m, m.system(x[None], at[None], box.reshape(1, 9))


def _noalloc(m, s):
    from deepmd_amd._lib import call, ptr, stream_ptr
    fs = s.fs
    nl = torch.full_like(s.nlist, -7)
    wr = torch.zeros_like(s.wrapped)
    st = torch.full_like(s.status, -1)
    print(f"dpk_nlist_build_noalloc natoms={fs.natoms} ncells={fs.ncells}", flush=True)
    call("dpk_nlist_build_noalloc", ptr(nl), ptr(wr), ptr(st), ptr(s.coord), ptr(s.atype),
         ptr(fs.frame_of), ptr(fs.frames_dev), fs.nframes, fs.natoms, fs.ncells, ptr(fs.img_pos),
         ptr(fs.shift_vec), m.rcut, m.sel_host[1], m.ntypes, fs.max_cand, stream_ptr())
    torch.cuda.synchronize()
    return nl, wr, st


def test_noalloc_nlist_matches_workspace_entry_twice():
    for kind in ("cu", "water"):   # the second call on a smaller system
        m, s = _system(kind)
        m.build_nlist(s)           # dpk_nlist_build with the system's workspace
        torch.cuda.synchronize()
        nl, wr, st = _noalloc(m, s)
        assert torch.equal(nl, s.nlist), kind
        assert torch.equal(wr, s.wrapped), kind
        assert int(st[0]) == int(s.status[0]), kind
        assert int((nl >= 0).sum()) > 0


def test_noalloc_nlist_rejects_sizes_above_the_scratch():
    from deepmd_amd._lib import DpkError, call, ptr, stream_ptr
    m, s = _system("water")
    fs = s.fs
    # 3 million atoms need 48 MB of cell-list workspace: refused before any launch
    with pytest.raises(DpkError, match="static scratch"):
        call("dpk_nlist_build_noalloc", ptr(s.nlist), ptr(s.wrapped), ptr(s.status),
             ptr(s.coord), ptr(s.atype), ptr(fs.frame_of), ptr(fs.frames_dev), fs.nframes,
             3_000_000, fs.ncells, ptr(fs.img_pos), ptr(fs.shift_vec), m.rcut, m.sel_host[1],
             m.ntypes, fs.max_cand, stream_ptr())

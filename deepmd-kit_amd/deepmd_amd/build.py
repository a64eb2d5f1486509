"""Compile libdpk.so (the gfx950 HIP kernels + C ABI) in-tree with hipcc, and
libdeepmd_op_pt.so (the torch.ops.deepmd operator library over it) and
libdpk_deeppot.so (the C++ MD-engine runner over it).

    python -m deepmd_amd.build          # from deepmd-kit_amd/

Sources: ``deepmd-kit_amd/csrc/*.hip`` and ``*.cpp``; header ``include/dpk.h``;
the operator library from ``csrc/torch_op/deepmd_op_pt.cpp`` (host C++ against
libtorch, linked to libdpk.so).  Outputs in ``deepmd-kit_amd/deepmd_amd/lib/``.
Objects are compiled in parallel; every call rebuilds (the build check must
compile each time).
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)            # deepmd-kit_amd/
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libdpk.so")
OP_LIB = os.path.join(LIBDIR, "libdeepmd_op_pt.so")
OP_SRC = os.path.join(CSRC, "torch_op", "deepmd_op_pt.cpp")
DEEPPOT_LIB = os.path.join(LIBDIR, "libdpk_deeppot.so")
DEEPPOT_SRC = os.path.join(ROOT, "cpp", "deep_pot_lower.cpp")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

ARCH = "gfx950"
HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # no packed FP32 instructions (v_pk_{add,mul,fma}_f32) anywhere
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    # MFMA results in VGPRs (no AGPR round trips before the VALU epilogues)
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm is required to build libdpk)")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


def _compile(src, obj):
    cmd = [hipcc(), *HIPFLAGS, "-I", INCLUDE, "-I", CSRC, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in (r.stderr or "").splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {os.path.basename(src)}:\n{err}")
    return obj


def build_op_library(verbose: bool = True) -> str:
    """libdeepmd_op_pt.so: host C++ against libtorch's headers and libraries
    (the HIP build of torch: c10::hip streams), linked to libdpk.so with an
    $ORIGIN run path, so ``torch.ops.load_library`` needs nothing else."""
    import torch  # noqa: F401  (paths and ABI flag of the installed torch)
    from torch.utils import cpp_extension as ce
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    tlib = ce.library_paths()[0]
    cxx = shutil.which("g++") or "g++"
    with tempfile.TemporaryDirectory(prefix="dpk_op_") as tmp:
        out = os.path.join(tmp, "libdeepmd_op_pt.so")
        cmd = [cxx, "-std=c++17", "-O2", "-fPIC", "-shared", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
               "-DUSE_ROCM=1", "-D__HIP_PLATFORM_AMD__=1",
               *[f"-I{p}" for p in ce.include_paths()], "-I/opt/rocm/include", f"-I{INCLUDE}",
               OP_SRC, "-o", out, f"-L{tlib}", "-ltorch", "-ltorch_cpu", "-lc10", "-lc10_hip",
               f"-L{LIBDIR}", "-ldpk", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{tlib}"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("op library build failed:\n" + r.stderr[-4000:])
        shutil.move(out, OP_LIB)
    if verbose:
        print(f"built {OP_LIB}")
    return OP_LIB


def build_deeppot_library(verbose: bool = True) -> str:
    """libdpk_deeppot.so: the C++ MD-engine runner (host C++ over libdpk's C
    ABI and the HIP runtime), with an $ORIGIN run path to libdpk.so."""
    cxx = shutil.which("g++") or "g++"
    with tempfile.TemporaryDirectory(prefix="dpk_dp_") as tmp:
        out = os.path.join(tmp, "libdpk_deeppot.so")
        cmd = [cxx, "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-D__HIP_PLATFORM_AMD__=1",
               f"-I{ROCM}/include", f"-I{INCLUDE}", DEEPPOT_SRC, "-o", out,
               f"-L{LIBDIR}", "-ldpk", f"-L{ROCM}/lib", "-lamdhip64",
               "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM}/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("runner library build failed:\n" + r.stderr[-4000:])
        shutil.move(out, DEEPPOT_LIB)
    if verbose:
        print(f"built {DEEPPOT_LIB}")
    return DEEPPOT_LIB


def build(verbose: bool = True, jobs: int | None = None) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs) + 1, max(1, min(16, os.cpu_count() or 1)))
    with tempfile.TemporaryDirectory(prefix="dpk_build_") as tmp:
        objs = [os.path.join(tmp, os.path.basename(s) + ".o") for s in srcs]
        with cf.ThreadPoolExecutor(jobs) as ex:
            list(ex.map(_compile, srcs, objs))
        tmp_lib = os.path.join(tmp, "libdpk.so")
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp_lib]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stderr)
        shutil.move(tmp_lib, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    build_op_library(verbose)
    build_deeppot_library(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)

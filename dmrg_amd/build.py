"""Build the in-tree HIP library ``dmrg_amd/libdmrgx.so`` for gfx950.

Every ``*.hip`` / ``*.cpp`` under ``dmrg_amd/csrc`` is compiled with hipcc and
linked into one shared library.  Run ``python -m dmrg_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libdmrgx.so"
OBJ = PKG / "_obj"
ARCH = os.environ.get("DMRGX_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off (hand-written and compiler-formed alike)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
    f"-I{PKG.parent / 'include'}",
]


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr[-8000:]}")
    return obj


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *[str(o) for o in objs], "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-8000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} from {len(srcs)} sources")
    return OUT


CLI_SRC = PKG / "cli" / "dmrgx_main.cpp"
CLI_OUT = PKG / "dmrgx"


def build_cli(verbose: bool = True) -> Path:
    """The command-line driver ``dmrg_amd/dmrgx`` (the reference mains'
    interface), linked against libdmrgx.so in the same directory."""
    tmp = CLI_OUT.with_suffix(".tmp")
    cmd = ["g++", "-O2", "-std=c++17", f"-I{PKG.parent / 'include'}", str(CLI_SRC), f"-L{PKG}", "-ldmrgx",
           "-Wl,-rpath,$ORIGIN", "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"cli build failed:\n{r.stderr[-8000:]}")
    os.replace(tmp, CLI_OUT)
    if verbose:
        print(f"built {CLI_OUT}")
    return CLI_OUT


TORCH_SRC = PKG / "torch_ext" / "torch_heff.cpp"
TORCH_OUT = PKG / "dmrgx_torch.so"


def build_torch(verbose: bool = True) -> Path:
    """The torch::autograd::Function module (``dmrg_amd/dmrgx_torch.so``),
    linked against libdmrgx.so in the same directory."""
    from torch.utils import cpp_extension

    bdir = OBJ / "torch"
    bdir.mkdir(parents=True, exist_ok=True)
    cpp_extension.load(
        name="dmrgx_torch", sources=[str(TORCH_SRC)], build_directory=str(bdir),
        extra_include_paths=[str(PKG.parent / "include")],
        extra_cflags=["-O2"],
        # $ORIGIN escaped for ninja ($$) and for the shell it runs (\\$)
        extra_ldflags=[f"-L{PKG}", "-ldmrgx", "-Wl,-rpath,\\$$ORIGIN", "-Wl,-rpath,\\$$ORIGIN/../.."],
        verbose=False, is_python_module=True)
    built = bdir / "dmrgx_torch.so"
    tmp = TORCH_OUT.with_suffix(".so.tmp")
    tmp.write_bytes(built.read_bytes())
    os.replace(tmp, TORCH_OUT)
    if verbose:
        print(f"built {TORCH_OUT}")
    return TORCH_OUT


if __name__ == "__main__":
    build()
    build_cli()
    if "--torch" in sys.argv:
        build_torch()
    sys.exit(0)

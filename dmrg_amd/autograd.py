"""torch.autograd access to the two-site H_eff matvec: ``heff_apply`` is the
C++ torch::autograd::Function of ``dmrg_amd/dmrgx_torch.so`` (forward =
dmrgx_heff_apply, backward = dmrgx_heff_adjoint).  Inputs are contiguous
device tensors (float64 or complex128) in the reference's flat block layout;
the get_map structure of the step is a ``Structure``.

One HIP runtime per process: torch ships its own libamdhip64 under the same
soname as the system one libdmrgx.so links, and the first one loaded serves
every library.  A process that uses both must import torch before loading
dmrg_amd's library, so that libdmrgx.so binds to torch's runtime."""
from __future__ import annotations

import importlib.util
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_MOD = None


def module():
    """The compiled extension (built by ``python -m dmrg_amd.build --torch``)."""
    global _MOD
    if _MOD is None:
        import torch  # noqa: F401  (libtorch must be loaded first)
        from . import _lib, load_library
        load_library()
        from . import TORCH_LOADED_FIRST
        if not TORCH_LOADED_FIRST:
            raise RuntimeError("libdmrgx.so was loaded before torch: import torch first in a process that uses "
                               "dmrg_amd.autograd (one HIP runtime per process)")
        path = Path(__file__).resolve().parent / "dmrgx_torch.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: run `python -m dmrg_amd.build --torch`")
        spec = importlib.util.spec_from_file_location("dmrgx_torch", path)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _MOD = mod
    return _MOD


@dataclass
class Structure:
    """rmap / lmap: get_map tuples (n x 7); roff, rsize, loff, lsize: block
    offsets and sizes of the environments; n_store: length of the R-pass
    intermediate."""
    rmap: np.ndarray
    roff: np.ndarray
    rsize: np.ndarray
    lmap: np.ndarray
    loff: np.ndarray
    lsize: np.ndarray
    n_store: int

    @classmethod
    def from_case(cls, case):
        return cls(np.asarray(case.rmap, np.int32), np.asarray(case.roff, np.int64), np.asarray(case.rsize, np.int64),
                   np.asarray(case.lmap, np.int32), np.asarray(case.loff, np.int64), np.asarray(case.lsize, np.int64),
                   int(case.n_store))


def heff_apply(psi, renv, lenv, st: Structure, prec: str = "fp64"):
    """H_eff(renv, lenv) psi, differentiable in psi, renv and lenv."""
    import torch
    from . import PREC_NAMES
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    return module().heff_apply(psi, renv, lenv, t(st.rmap), t(st.roff), t(st.rsize), t(st.lmap), t(st.loff),
                               t(st.lsize), int(st.n_store), int(PREC_NAMES[prec]))

"""The C-ABI library loads on a CPU-only host and exports every symbol that
include/*.h declares (no compute calls)."""
import re
from pathlib import Path

import dmrg_amd

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b(dmrgx_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = dmrg_amd.load_library()
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    from dmrg_amd import driver
    assert set(dmrg_amd.EXPORTS + driver.DRIVER_EXPORTS) == names


def test_torch_autograd_module_loads():
    """dmrgx_torch.so (the torch::autograd::Function) imports and links
    libdmrgx.so from its own directory, in a fresh process that imports torch
    first (one HIP runtime per process); no device call."""
    import subprocess
    import sys
    from pathlib import Path
    root = Path(__file__).resolve().parents[1]
    code = ("import torch, sys; sys.path.insert(0, %r); from dmrg_amd import autograd; m = autograd.module(); "
            "assert hasattr(m, 'heff_apply') and hasattr(m, 'init'); print('ok')" % str(root))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "ok" in r.stdout, r.stderr[-2000:]

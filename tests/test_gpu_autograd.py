"""The torch::autograd::Function of the H_eff matvec: torch.autograd.gradcheck
(central finite differences of the forward) on psi and both environments, in
FP64.  Runs in a child process that imports torch before the library, so the
process has a single HIP runtime (torch's)."""
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]

CHECK = r'''
import sys
import numpy as np
import torch
sys.path.insert(0, {root!r})
from dmrg_amd.autograd import Structure, heff_apply, module
from tests.cases import synthetic_case
import dmrg_amd
dmrg_amd.init(0)
module()
dev = torch.device("cuda:0")
for dt in (np.complex128, np.float64):
    print("case gradcheck", np.dtype(dt).name, flush=True)
    case = synthetic_case(2, 3, 8, dtype=dt, uneven=[5, 3])
    st = Structure.from_case(case)
    rng = np.random.default_rng(4)
    psi0 = rng.standard_normal(case.n_psi) + (1j * rng.standard_normal(case.n_psi) if dt == np.complex128 else 0)
    tdt = torch.complex128 if dt == np.complex128 else torch.float64
    psi = torch.tensor(psi0, dtype=tdt, device=dev, requires_grad=True)
    R = torch.tensor(case.rflat, device=dev, requires_grad=True)
    L = torch.tensor(case.lflat, device=dev, requires_grad=True)
    y = heff_apply(psi, R, L, st).detach().cpu().numpy()
    ref = case.oracle_apply(psi0)
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    assert err < 1e-12, err
    ok = torch.autograd.gradcheck(lambda p, r, l: heff_apply(p, r, l, st), (psi, R, L), eps=1e-6, atol=1e-7,
                                  rtol=1e-6, fast_mode=True)
    assert ok
print("gradcheck ok", flush=True)
'''


def test_heff_gradcheck(gpu):
    r = subprocess.run([sys.executable, "-c", CHECK.format(root=str(ROOT))], capture_output=True, text=True,
                       timeout=200, cwd=str(ROOT))
    print(r.stdout[-2000:], r.stderr[-4000:], flush=True)
    assert r.returncode == 0 and "gradcheck ok" in r.stdout

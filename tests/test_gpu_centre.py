"""dmrgx_state_heff_create: the two-site operator at the centre bond of a
finished run, rebuilt from the MPS (environments recomputed site by site).
Its Rayleigh quotient on the run's own two-site state is the run's energy,
the reduced-precision plans agree with the FP64 one, and the structurally
zero slices of the real environments are skipped."""
import ctypes as C

import numpy as np
import pytest

from dmrg_amd import driver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("model,params,L,m", [
    ("xyz", (1.0, 0.8, 0.5), 12, 32),
    ("potts3", (1.0, 1.0, 0.0, 3.14159265359 / 8), 8, 54),
    ("kitaev", (-1.0, -1.0, 0.25), 12, 32),
])
def test_centre_operator_energy(gpu, model, params, L, m):
    print(f"case centre operator {model} L={L} m={m}", flush=True)
    res = driver.run_dmrg(model, L, m, 2, params, schedule="dmrg2", keep_state=True)
    st = res.state
    npd = np.complex128 if model in ("xyz", "potts3") else np.float64
    h64 = st.centre_heff("fp64")
    assert h64.flops_exec < 0.5 * h64.flops
    # the plan's own copy of the two-site state
    psi = np.zeros(h64.n_psi, npd)
    gpu.check(gpu.lib().dmrgx_memcpy_d2h(psi.ctypes.data_as(C.c_void_p), C.c_void_p(h64.psi_ptr), psi.nbytes))
    xd = gpu.DeviceArray.from_numpy(psi)
    y = gpu.DeviceArray(h64.n_psi, npd)
    h64.apply(xd, y)
    y64 = y.numpy()
    e = np.vdot(psi, y64).real / np.vdot(psi, psi).real
    print(f"  Rayleigh {e:.12f} run {res.energy:.12f}", flush=True)
    assert abs(e - res.energy) < 1e-9 * abs(res.energy), (e, res.energy)
    for prec, tol in (("fp32", 2e-5), ("bf16x3", 3e-5)):
        h = st.centre_heff(prec)
        h.apply(xd, y)
        err = np.max(np.abs(y.numpy() - y64)) / np.max(np.abs(y64))
        assert err < tol, (prec, err)
        h.close()
    h64.close()
    st.close()

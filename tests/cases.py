"""Input builders shared by the tests: environments with the block structure
of the reference's two-site step, their get_map task lists (from the oracle)
and flat device-ready arrays."""
from __future__ import annotations

from oracle import qtensor as qt
from oracle.dmrg import Runner
from oracle.models import make_model
from oracle.synth import HeffCase, flat_blocks, synthetic_case, synthetic_env  # noqa: F401


def dmrg_case(model, L, m, sweeps=1, sign_rule="physical", **p):
    """Environments of a real two-site step after a short oracle DMRG."""
    mpo, n, ferm = make_model(model, L, **p)
    R = Runner(mpo, n, ferm, sign_rule=sign_rule)
    R.dmrg2(m, sweeps, ladder=False)
    l, r = L // 2 - 1, L // 2
    lenv = R._lenv(l)
    renv = R._renv(r)
    qt.set_symmetry(n)
    return HeffCase(lenv, renv, fermion=R.alpha_on), R

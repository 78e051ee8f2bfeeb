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


def dmrg_step_case(model, L, m, step, sign_rule="physical", **p):
    """Environments of the `step`-th update_two call of the oracle's two-site
    schedule (counting from 1), taken before that step runs."""
    mpo, n, ferm = make_model(model, L, **p)
    R = Runner(mpo, n, ferm, sign_rule=sign_rule)
    out = []

    class _Stop(Exception):
        pass

    inner = R.update_two
    count = [0]

    def hook(l, r, cutoff):
        count[0] += 1
        if count[0] == step:
            out.append(HeffCase(R._lenv(l), R._renv(r), fermion=R.alpha_on))
            raise _Stop
        return inner(l, r, cutoff)

    R.update_two = hook
    try:
        R.dmrg2(m, 1)
    except _Stop:
        pass
    qt.set_symmetry(n)
    return out[0]

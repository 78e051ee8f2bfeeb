"""Writes tests/golden/ref_vectors.json: input/output vectors of the
reference's own code, compiled from its sources (oracle/Makefile ->
oracle/_ref/ref_harness).

For each case the oracle runs a short two-site DMRG to get realistic
environment stores store(l-1), store(r+1) at the centre bond; the harness
then computes, with the reference's functions, the bulk MPO tensor and its
edges (src/operators.cpp, include/mpo.h), update_two's environments
(src/core.cpp:188-192), generate_dim_sym (src/useful.cpp:151-194) and both
get_map calls (src/qtensor.cpp:997-1119).  The fixture keeps the inputs and
those outputs.  Run from the repository root after `make -C oracle`:

    python tests/golden/make_ref_vectors.py
"""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import qtensor as qt  # noqa: E402
from oracle.dmrg import Runner  # noqa: E402
from oracle.models import make_model  # noqa: E402

HARNESS = ROOT / "oracle" / "_ref" / "ref_harness"
OUT = Path(__file__).resolve().parent / "ref_vectors.json"

# (name, model, parameters for the oracle, harness parameters, L, m, sign rule, sector)
CASES = [
    ("kitaev_L8_m6", "kitaev", dict(t=-1.0, p=-1.0, u=0.25), (-1.0, -1.0, 0.25, 0.0), 8, 6, "physical", 0),
    ("kitaev_L8_m6_s1", "kitaev", dict(t=-0.7, p=-1.3, u=0.4), (-0.7, -1.3, 0.4, 0.0), 8, 6, "physical", 1),
    ("xyz_L8_m6_reference", "xyz", dict(jx=1.0, jy=0.8, jz=0.5), (1.0, 0.8, 0.5, 0.0), 8, 6, "reference", 0),
    ("potts2_L8_m6", "potts2", dict(h=1.0, j=0.7), (1.0, 0.7, 0.0, 0.0), 8, 6, "reference", 0),
]


def tensor_text(t: qt.QTensor) -> str:
    out = [str(len(t.dim))]
    for d in t.dim:
        out += [str(len(d))] + [str(x) for x in d]
    out.append(str(len(t.sym)))
    for s in t.sym:
        out += [str(x) for x in s]
    out += [str(x) for x in t.dir]
    flat = t.flat()
    if np.iscomplexobj(flat):
        out += [f"{float(x.real)!r} {float(x.imag)!r}" for x in flat]
    else:
        out += [repr(float(x)) for x in flat]
    return " ".join(out)


def tensor_json(t: qt.QTensor) -> dict:
    flat = t.flat()
    d = {"dim": t.dim, "sym": t.sym, "dir": list(t.dir)}
    if np.iscomplexobj(flat):
        d["re"] = [float(x) for x in flat.real]
        d["im"] = [float(x) for x in flat.imag]
    else:
        d["re"] = [float(x) for x in flat]
    return d


def parse_output(text: str, cplx: bool) -> dict:
    lines = text.strip().splitlines()
    res, i = {}, 0

    def ints(line):
        return [int(x) for x in line.split()[1:]]

    def dimsym(v):
        n, p, out = v[0], 1, []
        for _ in range(n):
            k = v[p]
            out.append(v[p + 1:p + 1 + k])
            p += 1 + k
        return out

    while i < len(lines):
        w = lines[i].split()
        key = w[0]
        if key in ("W", "Wleft", "Wright", "lenv", "renv"):
            dim = dimsym([int(x) for x in w[2:]])
            sym = dimsym(ints(lines[i + 1]))
            dirs = ints(lines[i + 2])[1:]
            vals = [float(x) for x in lines[i + 3].split()[2:]]
            t = {"dim": dim, "sym": sym, "dir": dirs}
            if cplx:
                t["re"], t["im"] = vals[0::2], vals[1::2]
            else:
                t["re"] = vals
            res[key] = t
            i += 4
        elif key in ("rmap", "lmap"):
            v = ints(lines[i])
            res[key] = [v[1 + 7 * k:8 + 7 * k] for k in range(v[0])]
            i += 1
        else:
            res[key] = dimsym(ints(lines[i]))
            i += 1
    return res


def main():
    if not HARNESS.exists():
        sys.exit(f"{HARNESS} is missing: run `make -C oracle` first")
    cases = []
    for name, model, kw, hp, L, m, rule, sector in CASES:
        mpo, n, ferm = make_model(model, L, **kw)
        R = Runner(mpo, n, ferm, sign_rule=rule)
        R.dmrg2(m, 1)
        l, r = L // 2 - 1, L // 2
        ls, rs = R.mps.store[l - 1], R.mps.store[r + 1]
        cplx = np.iscomplexobj(ls.flat())
        text = (f"model {model} {' '.join(repr(float(x)) for x in hp)}\n"
                f"lstore {tensor_text(ls)}\nrstore {tensor_text(rs)}\nsector {sector}\n")
        run = subprocess.run([str(HARNESS)], input=text, capture_output=True, text=True, timeout=120)
        if run.returncode != 0:
            sys.exit(f"{name}: harness failed: {run.stderr}")
        out = parse_output(run.stdout, cplx)
        cases.append({"name": name, "model": model, "params": kw, "harness_params": list(hp), "L": L, "m": m,
                      "sign_rule": rule, "sector": sector, "l": l, "r": r,
                      "lstore": tensor_json(ls), "rstore": tensor_json(rs), "reference": out})
        print(f"{name}: lenv {len(out['lenv']['sym'])} blocks, renv {len(out['renv']['sym'])}, "
              f"rmap {len(out['rmap'])}, lmap {len(out['lmap'])} tasks")
    OUT.write_text(json.dumps({"source": "oracle/_ref/ref_harness (reference sources, oracle/Makefile)",
                               "cases": cases}, separators=(",", ":")))
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()

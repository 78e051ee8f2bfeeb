"""Writes tests/golden/ref_vectors_s4e.json: vectors of the reference's own
code for the Z2 x Z2 x Z2 group and the 4e model, computed by
oracle/_ref/ref_harness_s4e (oracle/Makefile: the reference's sources and
4e.cpp under the SYMMETRY 8 / S4E / PBC switches its 4e main asks for).

Per case the harness gets a 4e MPO tensor (H_4e1 or H_4e2 with the given
couplings) and, as the stored environments, the chain's two MPO edge tensors,
and returns the tensor, its edges, update_two's environments, the periodic
edge environments through the identity operator, generate_dim_sym, both
get_map task lists and the 4e main's bond-1 operators.  The fixture keeps
the small tensors as their nonzero entries, the large ones as per-block norms
and eight fixed random projections of their values, and the task lists and
sector lists as SHA-256 digests of their canonical text (with their lengths
and first tuples), which pins them bit for bit.

It also records the exact low spectrum of the four-tensor (eight-site) 4e
chain, open and periodic, from the oracle's exact two-site operator
(oracle.dmrg.exact_chain_spectrum), whose pieces the harness vectors pin.

Run from the repository root after `make -C oracle`:

    python tests/golden/make_ref_vectors_s4e.py
"""
from __future__ import annotations

import hashlib
import json
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))

from make_ref_vectors import tensor_text  # noqa: E402
from oracle import models  # noqa: E402
from oracle import qtensor as qt  # noqa: E402
from oracle.dmrg import exact_chain_spectrum  # noqa: E402

HARNESS = ROOT / "oracle" / "_ref" / "ref_harness_s4e"
OUT = Path(__file__).resolve().parent / "ref_vectors_s4e.json"

# (name, second tensor kind, t1, t3, f0, f1, p1, frac of the tensor, sector)
CASES = [
    ("h4e1_s0", False, 1.0, 0.5, 0.3, 0.2, 0.2, 0.25, 0),
    ("h4e2_s5", True, 0.8, -0.4, 0.1, 0.6, 0.6, 0.5, 5),
]
# the chain whose spectrum is recorded: 8 sites, (t1, t3, f0, f1, frac)
CHAIN = dict(t1=1.0, t3=0.5, f0=0.3, f1=0.2, frac=0.25)
SMALL = ("W", "Wleft", "Wright")
LARGE = ("lenv", "renv", "ledge", "redge")
N_PROJ = 8


def digest(obj) -> str:
    return hashlib.sha256(json.dumps(obj, separators=(",", ":")).encode()).hexdigest()


def projections(flat: np.ndarray, seed: int = 20261204) -> list:
    """Eight fixed random projections of the values: sum_i r_k[i] v[i] with
    r_k uniform(-1, 1) from one seeded stream of length len(v) per k."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(N_PROJ):
        r = rng.uniform(-1.0, 1.0, flat.size)
        p = complex(np.dot(r, flat))
        out.append([p.real, p.imag])
    return out


def block_norms(dim, sym, flat) -> list:
    out, pos = [], 0
    for s in sym:
        n = int(np.prod([dim[j][s[j]] for j in range(len(dim))]))
        out.append(float(np.linalg.norm(flat[pos:pos + n])))
        pos += n
    return out


def parse(text: str) -> dict:
    """The harness's printout (tensors: dim / sym / dir / val lines; maps and
    sector lists: one line each)."""
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
        if len(w) > 1 and w[1] == "dim":
            dim = dimsym([int(x) for x in w[2:]])
            sym = dimsym(ints(lines[i + 1]))
            dirs = ints(lines[i + 2])[1:]
            vals = np.array([float(x) for x in lines[i + 3].split()[2:]])
            res[key] = {"dim": dim, "sym": sym, "dir": dirs, "val": vals[0::2] + 1j * vals[1::2]}
            i += 4
        elif key in ("rmap", "lmap"):
            v = ints(lines[i])
            res[key] = [v[1 + 7 * k:8 + 7 * k] for k in range(v[0])]
            i += 1
        else:
            res[key] = dimsym(ints(lines[i]))
            i += 1
    return res


def compress(out: dict) -> dict:
    c = {}
    for k, t in out.items():
        if isinstance(t, dict):
            base = {"dim": t["dim"], "dir": t["dir"], "n_blocks": len(t["sym"]), "sym_sha": digest(t["sym"])}
            v = t["val"]
            if k in LARGE:
                base["block_norms"] = block_norms(t["dim"], t["sym"], v)
                base["proj"] = projections(v)
            else:
                base["sym"] = t["sym"]
                nz = np.nonzero(v)[0]
                base["nz"] = [[int(j), float(v[j].real), float(v[j].imag)] for j in nz]
                base["size"] = int(v.size)
            c[k] = base
        elif k in ("rmap", "lmap"):
            c[k] = {"n": len(t), "sha": digest(t), "head": t[:16]}
        elif k in ("dim", "dim_ret", "ldim"):
            c[k] = t
        else:
            c[k] = {"n": len(t), "sha": digest(t), "head": t[:16]}
    return c


def main():
    if not HARNESS.exists():
        sys.exit(f"{HARNESS} is missing: run `make -C oracle` first")
    qt.set_symmetry(qt.S4E)
    cases = []
    for name, second, t1, t3, f0, f1, p1, frac, sector in CASES:
        ls = models.H_4e(False, t1, t3, f0, f1, p1, 0.0).left()
        rs = models.H_4e(True, t1, t3, f0, f1, p1, frac).right()
        text = (f"model {'4e2' if second else '4e1'} {t1!r} {t3!r} {f0!r} {f1!r} {p1!r} {frac!r}\n"
                f"lstore {tensor_text(ls)}\nrstore {tensor_text(rs)}\nsector {sector}\n")
        run = subprocess.run([str(HARNESS)], input=text, capture_output=True, text=True, timeout=300)
        if run.returncode != 0:
            sys.exit(f"{name}: harness failed: {run.stderr}")
        out = compress(parse(run.stdout))
        cases.append({"name": name, "second": second, "params": [t1, t3, f0, f1, p1, frac], "sector": sector,
                      "reference": out})
        print(f"{name}: rmap {out['rmap']['n']} tasks, lmap {out['lmap']['n']}, lenv {out['lenv']['n_blocks']} "
              f"blocks", flush=True)
    spectra = {}
    for label, sector, pbc in (("open_s0", 0, False), ("open_s3", 3, False), ("pbc_s0", 0, True)):
        mpo, _, _ = models.make_model("4e", 8, **CHAIN)
        edges = None
        if pbc:
            p = CHAIN
            fr = p["frac"] * 3 / 4
            edges = (models.H_4e(False, 0.0, -p["t3"], 0.0, p["f1"], p["f1"], fr).right(),
                     models.H_4e(True, 0.0, -p["t3"], 0.0, p["f1"], p["f1"], fr).left())
        w, n = exact_chain_spectrum(mpo, sector, k=3, edges=edges)
        spectra[label] = {"sector": sector, "pbc": pbc, "n": n, "lowest": [float(x) for x in w]}
        print(f"spectrum {label}: {w} (n={n})", flush=True)
    OUT.write_text(json.dumps({"source": "oracle/_ref/ref_harness_s4e (reference sources and 4e.cpp, oracle/Makefile)",
                               "projection_seed": 20261204, "cases": cases,
                               "chain": {"L": 8, "params": CHAIN, "spectra": spectra}}, separators=(",", ":")))
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()

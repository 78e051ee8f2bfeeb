"""Writes batch_couplings_64.json: the couplings (Jx, Jy, Jz) of the 64
independent XYZ chains of the batched workload, drawn i.i.d. U[0.5, 1.5) as
0.5 + (x >> 11) 2^-53 from one splitmix64 stream started at SEED, chain-major.

Run from the repository root: python -m tests.golden.make_batch_couplings"""
import json
from pathlib import Path

from dmrg_amd.driver import batch_couplings

SEED = 20261203
N_CHAINS = 64


def main():
    J = batch_couplings(N_CHAINS, SEED)
    out = {"seed": SEED, "n_chains": N_CHAINS, "rule": "0.5 + (splitmix64 >> 11) * 2^-53, chain-major (Jx, Jy, Jz)",
           "couplings": [[float(v) for v in row] for row in J]}
    p = Path(__file__).with_name("batch_couplings_64.json")
    p.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {p}")


if __name__ == "__main__":
    main()

"""Writes tests/golden/c1_em_1000_records.json: the oracle's records of the C1 run.

C1: 32^3, L = 16, N = 20, charge -0.5 per bead, counterions, salt 0.05, rho 0.2, B 0.5,
lB 1; fields 0 plus complex Gaussian noise of sigma 1e-3 (initial_fields, cfg = 1);
1000 explicit Euler-Maruyama steps at dt = 0.01 with the Philox noise of seed 0xC1,
H, mu and the full stress tensor recorded every 10 steps.  The run takes about 80 s on
eight host cores.  Run from the repository root after __graft_entry__.build():

    python tests/golden/make_c1_em_records.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from ftspolysol_amd.api import SCHEME_EM  # noqa: E402
from ftspolysol_amd.system import config, initial_fields  # noqa: E402
from oracle_lib import oracle_engine  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "c1_em_1000_records.json")


def c1_system():
    return config(1, scheme=SCHEME_EM, stress_mask=0x7F, seed=0xC1)


def pair(z):
    return [float(np.real(z)), float(np.imag(z))]


def record_json(r):
    return {"step": int(r["step"]), "H": pair(r["H"]), "p_iso": pair(r["p_iso"]),
            "mu_ex_poly": [pair(z) for z in r["mu_ex_poly"]], "mu_ex_ion": [pair(z) for z in r["mu_ex_ion"]],
            "stress": [pair(z) for z in r["stress"]], "Q_poly": [pair(z) for z in r["Q_poly"]]}


def main():
    s = c1_system()
    o = oracle_engine(s)
    o.set_fields(*initial_fields(s, sigma=1e-3, cfg=1))
    recs = o.step(1000, record_every=10)
    data = {"workload": "C1: 32^3 L=16 N=20 z=-0.5 salt=0.05 rho=0.2 B=0.5 lB=1, EM dt=0.01, seed 0xC1, "
                        "initial_fields(sigma=1e-3, cfg=1), 1000 steps, records every 10",
            "source": "oracle (oracle/libfts_oracle.so), tests/golden/make_c1_em_records.py",
            "records": [record_json(r) for r in recs]}
    with open(OUT, "w") as f:
        f.write("{\n")
        f.write(f'"workload": {json.dumps(data["workload"])},\n"source": {json.dumps(data["source"])},\n"records": [\n')
        f.write(",\n".join(json.dumps(r, separators=(",", ":")) for r in data["records"]))
        f.write("\n]}\n")
    print(f"wrote {len(recs)} records to {OUT}")


if __name__ == "__main__":
    main()

"""The C1 golden records (tests/golden/c1_em_1000_records.json, written by
tests/golden/make_c1_em_records.py from the oracle) against a fresh oracle run of the
first 100 steps: the fixture is what the oracle computes for C1.  The GPU engine is held
to all 100 records of the 1000-step run in test_gpu_c1.py."""
import json
import os

import numpy as np

from ftspolysol_amd.api import SCHEME_EM
from ftspolysol_amd.system import config, initial_fields
from oracle_lib import oracle_engine

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_em_1000_records.json")


def load_golden():
    return json.load(open(GOLDEN))["records"]


def cx(p):
    return np.asarray(p, dtype=float)[..., 0] + 1j * np.asarray(p, dtype=float)[..., 1]


def compare(rec, ref, tol):
    """Relative differences of one record against a golden one: H and p_iso by their own
    size, the stress tensor by its largest component, mu by max(1, |mu|)."""
    assert rec["step"] == ref["step"]
    out = {"H": abs(rec["H"] - cx(ref["H"])) / abs(cx(ref["H"])),
           "p_iso": abs(rec["p_iso"] - cx(ref["p_iso"])) / abs(cx(ref["p_iso"])),
           "stress": np.abs(rec["stress"] - cx(ref["stress"])).max() / np.abs(cx(ref["stress"])).max()}
    for key in ("mu_ex_poly", "mu_ex_ion"):
        r = cx(ref[key])
        out[key] = np.abs(rec[key] - r).max() / max(1.0, np.abs(r).max())
    assert max(out.values()) < tol, (rec["step"], out)
    return out


def test_golden_records_are_the_oracle_run():
    golden = load_golden()
    assert len(golden) == 100 and golden[-1]["step"] == 1000
    s = config(1, scheme=SCHEME_EM, stress_mask=0x7F, seed=0xC1)
    o = oracle_engine(s)
    o.set_fields(*initial_fields(s, sigma=1e-3, cfg=1))
    recs = o.step(100, record_every=10)
    for rec, ref in zip(recs, golden[:10]):
        compare(rec, ref, 1e-12)

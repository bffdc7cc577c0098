"""GPU: C1 as the configuration states it: 32^3, N = 20, FP64, 1000 explicit
Euler-Maruyama steps at dt = 0.01 from seeded initial fields, H, mu and the stress tensor
recorded every 10 steps, against the oracle's records of the same run
(tests/golden/c1_em_1000_records.json).  The dynamics contracts (a 1e-13 relative change
of the initial fields moves no record by more than 3e-15 over the 1000 steps in the
oracle), so the engine is held to every record at 1e-9."""
import pytest

from ftspolysol_amd.api import SCHEME_EM, hip_engine
from ftspolysol_amd.system import config, initial_fields
from test_c1_records import compare, load_golden

pytestmark = pytest.mark.gpu


def test_c1_thousand_em_steps_match_the_oracle_records():
    print("C1 1000 EM steps FP64", flush=True)
    golden = load_golden()
    s = config(1, scheme=SCHEME_EM, stress_mask=0x7F, seed=0xC1)
    e = hip_engine(s)
    e.set_fields(*initial_fields(s, sigma=1e-3, cfg=1))
    recs = e.step(1000, record_every=10)
    assert len(recs) == len(golden) == 100
    worst = {}
    for rec, ref in zip(recs, golden):
        for k, v in compare(rec, ref, 1e-9).items():
            worst[k] = max(worst.get(k, 0.0), v)
    print("largest relative differences over the 100 records:", worst, flush=True)
    e.close()

"""GPU: replica batches (fts_batch_*) against the same systems stepped one by one.

A batch runs the bead steps of all replicas of a group in shared launches and
everything else per replica, so each replica must reproduce the single-engine
trajectory of its own system (same Philox stream): identical arithmetic, checked
at 1e-12 relative.  The single engines themselves are pinned to the oracle by
test_gpu_parity.py.
"""
import numpy as np
import pytest

from ftspolysol_amd.api import (FIELD_OMEGA, FIELD_PSI, PREC_FP32, PREC_FP64, SCHED_CONV, SCHED_LINE, SCHED_PLANE, SCHED_SEPARABLE, SCHEME_EM,
                                SCHEME_PC_REFERENCE, SCHEME_SEMI_IMPLICIT, SCHEME_SIPC, SYMMETRY_OFF, hip_batch,
                                hip_engine)
from ftspolysol_amd.system import initial_fields, make_system

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def replica_systems(scheme, precision, sched):
    """Four replicas in two groups: three 32^3 systems (N = 20, 12 and 16, different
    salt and lB, one with both sweeps forced) and one 16^3 system."""
    kw = dict(scheme=scheme, precision=precision, fft_schedule=sched, stress_mask=0x7F, dt=0.01)
    return [
        make_system(32, 20, salt=0.05, lB=1.0, seed=101, **kw),
        make_system(32, 12, salt=0.2, lB=0.7, seed=102, **kw),
        make_system(32, 16, salt=0.01, lB=1.5, seed=103, chain_symmetry=SYMMETRY_OFF, **kw),
        make_system(16, 10, salt=0.1, lB=1.0, seed=104, **kw),
    ]


def start_fields(engines, systems):
    for r, (e, s) in enumerate(zip(engines, systems)):
        om, ps = initial_fields(s, sigma=0.05, seed=500 + r)
        e.set_fields(om, ps)


@pytest.mark.parametrize("scheme,precision,sched", [(SCHEME_PC_REFERENCE, PREC_FP64, SCHED_PLANE),
                                                    (SCHEME_EM, PREC_FP64, SCHED_LINE),
                                                    (SCHEME_EM, PREC_FP32, SCHED_PLANE),
                                                    (SCHEME_PC_REFERENCE, PREC_FP32, SCHED_LINE),
                                                    (SCHEME_EM, PREC_FP64, SCHED_CONV),
                                                    (SCHEME_PC_REFERENCE, PREC_FP32, SCHED_CONV),
                                                    (SCHEME_SEMI_IMPLICIT, PREC_FP64, SCHED_SEPARABLE),
                                                    (SCHEME_PC_REFERENCE, PREC_FP32, SCHED_SEPARABLE),
                                                    (SCHEME_SIPC, PREC_FP32, SCHED_CONV)])
def test_batch_matches_single_engines(scheme, precision, sched):
    print(f"batch scheme={scheme} precision={precision} sched={sched}", flush=True)
    systems = replica_systems(scheme, precision, sched)
    b = hip_batch(systems)
    assert b.info.n_replicas == 4 and b.info.n_groups == 2
    assert b.group_of == [0, 0, 0, 1]
    assert list(b.info.group_size[:2]) == [3, 1]
    assert [b.replicas[r].info.mirrored_species for r in range(4)] == [1, 1, 0, 1]
    singles = [hip_engine(s) for s in replica_systems(scheme, precision, sched)]
    start_fields(b.replicas, systems)
    start_fields(singles, systems)
    rb = b.step(4, record_every=2)
    for r, e in enumerate(singles):
        rs = e.step(4, record_every=2)
        assert len(rb[r]) == len(rs) == 2
        for x, y in zip(rb[r], rs):
            assert x["step"] == y["step"]
            assert abs(x["H"] - y["H"]) <= 1e-12 * abs(y["H"]), (r, x["H"], y["H"])
            assert rel(x["stress"], y["stress"]) <= 1e-12
            assert rel(x["mu_ex_poly"], y["mu_ex_poly"]) <= 1e-12
        assert rel(b.replicas[r].get_field(FIELD_OMEGA), e.get_field(FIELD_OMEGA)) <= 1e-12
        assert rel(b.replicas[r].get_field(FIELD_PSI), e.get_field(FIELD_PSI)) <= 1e-12
    # a second call continues the noise stream of every replica
    rb2 = b.step(2, record_every=1)
    for r, e in enumerate(singles):
        rs = e.step(2, record_every=1)
        assert [x["step"] for x in rb2[r]] == [5, 6]
        for x, y in zip(rb2[r], rs):
            assert abs(x["H"] - y["H"]) <= 1e-12 * abs(y["H"])


def test_batch_headline_autotuned():
    """Four FP32 replicas at 96^3, N=100: the autotuner times both schedules with
    every replica in each launch, and the batch reproduces a single engine."""
    print("batch 96^3 x4 autotuned", flush=True)
    systems = [make_system(96, 100, scheme=SCHEME_EM, precision=PREC_FP32, seed=7 + r) for r in range(4)]
    b = hip_batch(systems)
    sched = b.info.group_schedule[0]
    print("batch tune ms (conv, separable):", list(b.info.group_tune_ms[0]), "chosen", sched, flush=True)
    assert sched in (SCHED_CONV, SCHED_SEPARABLE)
    tune = list(b.info.group_tune_ms[0])
    assert min(tune) > 0 and tune[0 if sched == SCHED_CONV else 1] == min(tune)
    start_fields(b.replicas, systems)
    b.set_timing(True)
    rb = b.step(3, record_every=1)
    t = b.get_timing(0)
    assert t["step_launches"] == 3 * 99 and t["bead_steps"] == 4 * 3 * 99 and t["sweep_ms"] > 0
    single = hip_engine(make_system(96, 100, scheme=SCHEME_EM, precision=PREC_FP32, seed=9, fft_schedule=sched))
    om, ps = initial_fields(systems[2], sigma=0.05, seed=502)
    single.set_fields(om, ps)
    rs = single.step(3, record_every=1)
    for x, y in zip(rb[2], rs):
        assert abs(x["H"] - y["H"]) <= 1e-12 * abs(y["H"])


@pytest.mark.parametrize("scheme,sched", [(SCHEME_PC_REFERENCE, SCHED_PLANE), (SCHEME_PC_REFERENCE, SCHED_LINE),
                                          (SCHEME_SEMI_IMPLICIT, SCHED_CONV), (SCHEME_EM, SCHED_CONV),
                                          (SCHEME_EM, SCHED_SEPARABLE), (SCHEME_SIPC, SCHED_CONV)])
def test_graph_replays_match_eager_launches(scheme, sched):
    """Whole-step graphs (level 2), sweep graphs (1) and launches issued one by one (0):
    several step calls with records on every step, on none and on every second step (so
    the captured steps change between calls and replays follow captures) must give
    bit-identical fields and records, and the same timing counters."""
    print(f"graph levels scheme={scheme} sched={sched}", flush=True)
    out = []
    for level in (2, 1, 0):
        systems = replica_systems(scheme, PREC_FP32, sched)
        b = hip_batch(systems)
        b.set_graph(level)
        b.set_timing(True)
        start_fields(b.replicas, systems)
        recs = [b.step(3, record_every=1), b.step(4), b.step(4, record_every=2), b.step(3, record_every=1)]
        t = b.get_timing(0)
        assert t["sweep_ms"] > 0.0
        out.append(([e.get_field(FIELD_OMEGA) for e in b.replicas], [e.get_field(FIELD_PSI) for e in b.replicas],
                    [[(x["step"], x["H"], tuple(x["stress"])) for call in recs for x in call[r]] for r in range(len(systems))],
                    (t["sweeps"], t["bead_steps"], t["step_launches"], t["algorithmic_bytes"])))
        b.close()
    for level, (w, p, h, t) in zip((1, 0), out[1:]):
        for r in range(len(w)):
            assert np.array_equal(w[r], out[0][0][r]) and np.array_equal(p[r], out[0][1][r]), (level, r)
            assert h[r] == out[0][2][r], (level, r)
        assert t == out[0][3], (level, t, out[0][3])


@pytest.mark.parametrize("scheme,precision", [(SCHEME_PC_REFERENCE, PREC_FP32), (SCHEME_EM, PREC_FP64)])
def test_side_streams_do_not_change_results(scheme, precision):
    """The replicas' own work on side streams (forked from and joined into the group
    stream, records finished on the host after the step's one sync) against the same
    work on the group stream alone: fields and every recorded observable, stress
    included, must be bit-identical for 0, 2 and the default number of side streams."""
    print(f"side streams scheme={scheme} precision={precision}", flush=True)
    out = []
    for n in (0, 2, None):
        systems = replica_systems(scheme, precision, SCHED_PLANE)
        b = hip_batch(systems)
        if n is not None:
            b.set_streams(n)
        start_fields(b.replicas, systems)
        recs = [b.step(2, record_every=1) for _ in range(2)]
        out.append(([e.get_field(FIELD_OMEGA) for e in b.replicas], [e.get_field(FIELD_PSI) for e in b.replicas],
                    [[(x["H"], tuple(x["mu_ex_poly"]), tuple(x["stress"]), x["p_iso"]) for call in recs for x in call[r]]
                     for r in range(len(systems))]))
        b.close()
    for w, p, h in out[1:]:
        for r in range(len(w)):
            assert np.array_equal(w[r], out[0][0][r]) and np.array_equal(p[r], out[0][1][r]), r
            assert h[r] == out[0][2][r], r

"""Checkpoint and non-finite restart (complexLangevin.cu:518-667).

CPU: the oracle's restart semantics, pinned by construction.  A run with a NaN
injected at step 5 and a checkpoint every 3 steps must equal
  steps 1-3  the run without injection (same key), and
  steps 4-8  a fresh engine with the next noise key (seed + 0x9E3779B97F4A7C15)
             started from the step-3 fields at step counter 3,
with the records of the failed steps 4-5 dropped.
GPU: the engine (single and in a batch) against the oracle on the same injection.
"""
import numpy as np
import pytest

from ftspolysol_amd.api import FIELD_OMEGA, FIELD_PSI, SCHEME_EM, SCHEME_PC_REFERENCE, SCHEME_SIPC, FtsError, hip_batch, hip_engine
from ftspolysol_amd.system import initial_fields, make_system
from oracle_lib import oracle_engine

GOLDEN = 0x9E3779B97F4A7C15


def small(scheme, seed=77, **kw):
    return make_system(16, 8, rho=0.2, salt=0.05, B=0.5, lB=1.0, dt=0.01, scheme=scheme, seed=seed, **kw)


def start(e, s):
    om, ps = initial_fields(s, sigma=0.05, seed=4242)
    e.set_fields(om, ps)


def H_of(recs):
    return np.array([r["H"] for r in recs])


@pytest.mark.parametrize("scheme", [SCHEME_EM, SCHEME_PC_REFERENCE, SCHEME_SIPC])
def test_oracle_restart_semantics(scheme):
    s = small(scheme)
    a = oracle_engine(s)
    start(a, s)
    a.set_restart(3, 2)
    a.inject_nonfinite(5)
    ra = a.step(8, record_every=1)
    assert a.get_info().restarts == 1 and a.get_info().noise_epoch == 1
    assert a.get_info().step == 8
    assert [r["step"] for r in ra] == list(range(1, 9))
    b = oracle_engine(s)
    start(b, s)
    rb = b.step(3, record_every=1)
    om, ps = b.get_field(FIELD_OMEGA), b.get_field(FIELD_PSI)[0]
    c = oracle_engine(small(scheme, seed=(77 + GOLDEN) % 2 ** 64))
    c.set_fields(om, ps)
    c.set_step(3)
    rc = c.step(5, record_every=1)
    np.testing.assert_array_equal(H_of(ra[:3]), H_of(rb))
    # the fresh engine re-removes the (1e-18) psi mean on set_fields: equal to rounding
    assert np.abs(H_of(ra[3:]) - H_of(rc)).max() <= 1e-14 * np.abs(H_of(rc)).max()
    w_a, w_c = a.get_field(FIELD_OMEGA), c.get_field(FIELD_OMEGA)
    assert np.abs(w_a - w_c).max() <= 1e-14 * np.abs(w_c).max()
    # and the restarted segment differs from the uninjected trajectory (new noise)
    b.step(5)
    assert np.abs(b.get_field(FIELD_OMEGA) - w_a).max() > 1e-3 * np.abs(w_a).max()


def test_oracle_restart_cap_and_off():
    s = small(SCHEME_EM)
    a = oracle_engine(s)
    start(a, s)
    a.inject_nonfinite(2)
    with pytest.raises(FtsError) as ei:  # restarts off: the failure is returned
        a.step(4)
    assert ei.value.code == 1
    b = oracle_engine(s)
    start(b, s)
    b.set_restart(3, 0)
    b.inject_nonfinite(2)
    with pytest.raises(FtsError) as ei:
        b.step(4)
    assert "maximum" in str(ei.value)


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", [SCHEME_EM, SCHEME_PC_REFERENCE, SCHEME_SIPC])
def test_engine_restart_matches_oracle(scheme):
    print(f"restart engine scheme={scheme}", flush=True)
    s = small(scheme)
    g, o = hip_engine(s), oracle_engine(s)
    for e in (g, o):
        start(e, s)
        e.set_restart(3, 2)
        e.inject_nonfinite(5)
    rg, ro = g.step(8, record_every=1), o.step(8, record_every=1)
    assert g.get_info().restarts == 1 and g.get_info().noise_epoch == 1
    assert [r["step"] for r in rg] == [r["step"] for r in ro] == list(range(1, 9))
    hg, ho = H_of(rg), H_of(ro)
    assert np.abs(hg - ho).max() <= 1e-10 * np.abs(ho).max()
    w_g, w_o = g.get_field(FIELD_OMEGA), o.get_field(FIELD_OMEGA)
    assert np.abs(w_g - w_o).max() <= 1e-10 * np.abs(w_o).max()


@pytest.mark.gpu
def test_batch_restart_only_rewinds_the_failed_replica():
    print("restart batch", flush=True)
    systems = [small(SCHEME_EM, seed=90 + r) for r in range(3)]
    b = hip_batch(systems)
    singles = [hip_engine(small(SCHEME_EM, seed=90 + r)) for r in range(3)]
    for e, s in zip(b.replicas + singles, systems + systems):
        start(e, s)
        e.set_restart(3, 1)
    b.replicas[1].inject_nonfinite(5)
    singles[1].inject_nonfinite(5)
    rb = b.step(8, record_every=1)
    for r, e in enumerate(singles):
        rs = e.step(8, record_every=1)
        assert [x["step"] for x in rb[r]] == list(range(1, 9))
        assert np.abs(H_of(rb[r]) - H_of(rs)).max() <= 1e-12 * np.abs(H_of(rs)).max()
        assert b.replicas[r].get_info().restarts == (1 if r == 1 else 0)
        assert b.replicas[r].get_info().step == 8

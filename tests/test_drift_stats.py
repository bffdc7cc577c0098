"""Drift statistics (the rows of drift_max.dat: complexLangevin.cu:318-324, 466-497).

CPU: the oracle's values are pinned by the step itself.  With the noise off, a
predictor-corrector step moves w by -dt lambda/2 (u + u_bar) and psi by
-dt lambda_psi/2 (u_psi + u_psi_bar) (the psi drifts are mean-free, so the mean
removal after the update changes nothing), so the row of a step must equal
max |w_(t+1) - w_t| / dt and max |psi_(t+1) - psi_t| / dt, read from the fields
before and after it; an EM step moves them by -dt lambda u, the statistic
reported for EM.  A restart from the checkpoint drops the rows after it.
GPU: the engine's rows against the oracle's on the same fields and noise, and the
rows of a batch's replicas against single engines of the same systems.
"""
import numpy as np
import pytest

from ftspolysol_amd.api import FIELD_OMEGA, FIELD_PSI, SCHEME_EM, SCHEME_PC_REFERENCE, SCHEME_SIPC, hip_batch, hip_engine
from ftspolysol_amd.system import initial_fields, make_system
from oracle_lib import oracle_engine


def small(scheme, seed=91, **kw):
    return make_system(12, 6, rho=0.2, salt=0.05, B=0.5, lB=1.0, dt=0.01, scheme=scheme, seed=seed, **kw)


def fields(e):
    return e.get_field(FIELD_OMEGA), e.get_field(FIELD_PSI)[0]


@pytest.mark.parametrize("scheme", [SCHEME_PC_REFERENCE, SCHEME_EM])
def test_oracle_drift_rows_equal_field_increments(scheme):
    s = small(scheme, noise=False)
    o = oracle_engine(s)
    o.set_fields(*initial_fields(s, sigma=0.05, seed=12))
    o.set_drift_stats(True)
    o.step(2)  # leave the initial state
    o.get_drift_rows()
    w0, p0 = fields(o)
    o.step(1)
    w1, p1 = fields(o)
    rows = o.get_drift_rows()
    assert rows.shape == (1, o.info.n_omega + 2) and rows[0, 0] == 3
    dt = 0.01
    np.testing.assert_allclose(rows[0, 1], np.abs(w1[0] - w0[0]).max() / dt, rtol=1e-11)
    np.testing.assert_allclose(rows[0, 2], np.abs(p1 - p0).max() / dt, rtol=1e-9)
    assert rows[0, 1] > 0 and rows[0, 2] > 0


def test_oracle_drift_rows_follow_steps_and_restarts():
    s = small(SCHEME_PC_REFERENCE)
    o = oracle_engine(s)
    o.set_fields(*initial_fields(s, sigma=0.05, seed=12))
    o.set_drift_stats(True)
    o.step(4)
    rows = o.get_drift_rows(capacity=3)
    assert list(rows[:, 0]) == [1, 2, 3]
    assert list(o.get_drift_rows()[:, 0]) == [4]
    assert o.get_drift_rows().shape[0] == 0
    # a NaN at step 7 with a checkpoint every 3 steps: rows of 7.. re-run after the restart,
    # none of the failed attempt is kept
    o.set_restart(3, 2)
    o.inject_nonfinite(7)
    o.step(5)
    rows = o.get_drift_rows()
    assert list(rows[:, 0]) == [5, 6, 7, 8, 9]
    assert np.isfinite(rows).all()
    o.set_drift_stats(False)
    o.step(1)
    assert o.get_drift_rows().shape[0] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", [SCHEME_PC_REFERENCE, SCHEME_EM, SCHEME_SIPC])
def test_engine_drift_rows_match_oracle(scheme):
    print(f"drift rows scheme={scheme}", flush=True)
    s = small(scheme)
    g, o = hip_engine(s), oracle_engine(s)
    for e in (g, o):
        e.set_fields(*initial_fields(s, sigma=0.05, seed=12))
        e.set_drift_stats(True)
        e.step(4)
    a, b = g.get_drift_rows(), o.get_drift_rows()
    assert a.shape == b.shape == (4, 3)
    assert np.array_equal(a[:, 0], b[:, 0])
    np.testing.assert_allclose(a[:, 1:], b[:, 1:], rtol=1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", [SCHEME_PC_REFERENCE, SCHEME_EM, SCHEME_SIPC])
def test_batch_drift_rows_match_single_engines(scheme):
    """Replicas of a batch collect the same rows as single engines of their systems."""
    print(f"batch drift rows scheme={scheme}", flush=True)
    systems = [small(scheme, seed=91 + r) for r in range(3)]
    b = hip_batch(systems)
    singles = [hip_engine(s) for s in systems]
    for r, s in enumerate(systems):
        for e in (b.replicas[r], singles[r]):
            e.set_fields(*initial_fields(s, sigma=0.05, seed=40 + r))
            e.set_drift_stats(True)
    b.step(3)
    b.step(2)
    for r, e in enumerate(singles):
        e.step(5)
        a, c = b.replicas[r].get_drift_rows(), e.get_drift_rows()
        assert a.shape == c.shape == (5, 3)
        assert list(a[:, 0]) == [1, 2, 3, 4, 5]
        np.testing.assert_allclose(a[:, 1:], c[:, 1:], rtol=1e-12)
    b.close()

"""GPU parity at strong fields: the HIP engine against the oracle on the system and
fields of the oracle's electrostatics and stress pins (test_oracle_pins.py), where
the ion and chain partition functions are far from real and every exponential works
far from its linear regime (a charged, non-palindromic AB copolymer with non-uniform
bonds; counterions, salt cations and anions with two smearing lengths).

Tolerance: the FP64 parity bar, 1e-10 relative, for densities, ion densities, Q, H,
mu and all six stress components with p_iso; fields after PC, EM and semi-implicit
steps likewise.
"""
import numpy as np
import pytest

from ftspolysol_amd.api import (FIELD_ION_RHO, FIELD_OMEGA, FIELD_PSI, FIELD_RHO_C, FIELD_RHO_C_ION, FIELD_RHO_C_POLY,
                                FIELD_RHO_P, FIELD_RHO_TYPE, SCHED_CONV, SCHED_LINE, SCHED_PLANE, SCHED_SEPARABLE, SCHEME_EM,
                                SCHEME_PC_REFERENCE, SCHEME_SEMI_IMPLICIT, hip_engine)
from oracle_lib import oracle_engine
from test_oracle_pins import copolymer, random_fields

pytestmark = pytest.mark.gpu

FIELDS = [FIELD_OMEGA, FIELD_PSI, FIELD_RHO_TYPE, FIELD_RHO_P, FIELD_RHO_C, FIELD_RHO_C_POLY, FIELD_RHO_C_ION, FIELD_ION_RHO]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def obs_close(a, b, tol=1e-10):
    for key in ("H", "p_iso"):
        assert abs(a[key] - b[key]) < tol * abs(b[key]), key
    for key in ("stress", "Q_poly", "Q_ion"):
        assert rel(a[key], b[key]) < tol, key
    for key in ("mu_ex_poly", "mu_ex_ion"):
        assert np.abs(a[key] - b[key]).max() < tol * max(np.abs(b[key]).max(), 1.0), key


@pytest.mark.parametrize("sched", [SCHED_PLANE, SCHED_CONV, SCHED_LINE, SCHED_SEPARABLE])
def test_strong_field_densities_and_measure(sched):
    print(f"strong fields densities+measure sched={sched}", flush=True)
    s = copolymer(stress_mask=0x7F, fft_schedule=sched)
    g, o = hip_engine(s), oracle_engine(s)
    assert g.info.fft_schedule == sched and g.info.n_ions == 3
    om, ps = random_fields(s.m, g.info.n_omega, 0.3, 21)
    for e in (g, o):
        e.set_fields(om, ps)
    Qg, Qo = g.compute_densities(), o.compute_densities()
    assert abs(Qo[0].imag) > 0.1 * abs(Qo[0])   # far from real
    assert rel(Qg, Qo) < 1e-11
    for fid in FIELDS:
        assert rel(g.get_field(fid), o.get_field(fid)) < 1e-10, fid
    obs_close(g.measure(), o.measure())


@pytest.mark.parametrize("scheme", [SCHEME_PC_REFERENCE, SCHEME_EM, SCHEME_SEMI_IMPLICIT])
def test_strong_field_steps(scheme):
    print(f"strong fields steps scheme={scheme}", flush=True)
    s = copolymer(stress_mask=0x7F, scheme=scheme, dt=0.002, lambda_psi=0.7, seed=5)
    g, o = hip_engine(s), oracle_engine(s)
    om, ps = random_fields(s.m, g.info.n_omega, 0.25, 22)
    for e in (g, o):
        e.set_fields(om, ps)
    rg, ro = g.step(3, record_every=1), o.step(3, record_every=1)
    assert len(rg) == len(ro) == 3
    for a, b in zip(rg, ro):
        obs_close(a, b)
    assert rel(g.get_field(FIELD_OMEGA), o.get_field(FIELD_OMEGA)) < 1e-10
    assert rel(g.get_field(FIELD_PSI), o.get_field(FIELD_PSI)) < 1e-10

"""GPU parity: the HIP engine (libfts_hip.so, through the C ABI) against the FP64
CPU oracle on the same inputs and the same Philox noise.

Tolerances: FP64 engine vs oracle 1e-10 relative for Q, H, mu, densities and the
stress tensor (the FP64 parity bar); FFTs and a single bead step 1e-13; the FP32
engine 1e-4 relative to FP64 for H, mu and stress, 2e-5 for a bead step.
"""
import numpy as np
import pytest

from ftspolysol_amd.api import (SCHED_CONV, SCHED_CONV_MFMA, FtsError, SCHED_LINE, SCHED_PLANE, SCHED_SEPARABLE, SYMMETRY_OFF, FIELD_ION_RHO, FIELD_OMEGA, FIELD_PSI, FIELD_QB, FIELD_QF, FIELD_RHO_C,
                                FIELD_RHO_C_ION, FIELD_RHO_C_POLY, FIELD_RHO_P, FIELD_RHO_TYPE, PREC_FP32, PREC_FP64,
                                SCHEME_EM, SCHEME_PC_REFERENCE, SCHEME_SEMI_IMPLICIT, SCHEME_SIPC, hip_engine)
from ftspolysol_amd.system import System, initial_fields, make_system
from oracle_lib import oracle_engine

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def bare_system(m, **kw):
    return System(m=m, dL=0.5, species=[dict(num_chains=1, types=[0, 0], charges=[0.0, 0.0])], xi=0.1, **kw)


@pytest.mark.parametrize("m", [(8, 8, 8), (16, 12, 12), (24, 16, 16), (32, 32, 32), (48, 24, 24), (96, 96, 96), (128, 64, 64)])
def test_fft_matches_numpy(m):
    print(f"fft {m}", flush=True)
    e = hip_engine(bare_system(m))
    rng = np.random.default_rng(3)
    x = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    assert rel(e.fft3d(x, -1), np.fft.fftn(x)) < 1e-13
    assert rel(e.fft3d(x, +1), np.fft.ifftn(x) * x.size) < 1e-13


@pytest.mark.parametrize("m", [(32, 24, 48), (128, 128, 128), (16, 192, 8), (256, 16, 32)])
def test_fft_line_path_matches_numpy(m):
    """Grids the plane pass cannot hold run every FFT as line passes (and their bead
    steps on the separable line schedule)."""
    print(f"fft line path {m}", flush=True)
    e = hip_engine(bare_system(m))
    assert e.info.fft_schedule == SCHED_SEPARABLE
    rng = np.random.default_rng(3)
    x = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    assert rel(e.fft3d(x, -1), np.fft.fftn(x)) < 1e-13
    assert rel(e.fft3d(x, +1), np.fft.ifftn(x) * x.size) < 1e-13


@pytest.mark.parametrize("sched", [SCHED_PLANE, SCHED_LINE, SCHED_CONV, SCHED_SEPARABLE])
@pytest.mark.parametrize("m,bond", [((32, 32, 32), 1.3), ((96, 96, 96), 1.0), ((16, 24, 24), 0.7)])
def test_bead_step_matches_oracle(m, bond, sched):
    print(f"bead step {m} b={bond} sched={sched}", flush=True)
    s = bare_system(m, fft_schedule=sched)
    g, o = hip_engine(s), oracle_engine(s)
    assert g.info.fft_schedule == sched
    rng = np.random.default_rng(5)
    q = 1 + 0.1 * (rng.standard_normal(m) + 1j * rng.standard_normal(m))
    w = 0.2 * (rng.standard_normal(m) + 1j * rng.standard_normal(m))
    emw = np.exp(-w)
    ref = o.bead_step(q, emw, bond)
    assert rel(g.bead_step(q, emw, bond), ref) < 1e-13
    s32 = bare_system(m, precision=PREC_FP32, fft_schedule=sched)
    assert rel(hip_engine(s32).bead_step(q, emw, bond), ref) < 2e-5


@pytest.mark.parametrize("sched", [SCHED_LINE, SCHED_SEPARABLE])
@pytest.mark.parametrize("m,bond", [((32, 24, 48), 1.1), ((16, 192, 8), 0.9), ((128, 128, 128), 1.2), ((256, 16, 32), 0.8)])
def test_bead_step_line_schedules_match_oracle(m, bond, sched):
    """The line schedules on grids the plane pass cannot hold."""
    print(f"bead step {m} b={bond} sched={sched}", flush=True)
    s = bare_system(m, fft_schedule=sched)
    g, o = hip_engine(s), oracle_engine(s)
    assert g.info.fft_schedule == sched
    rng = np.random.default_rng(6)
    q = 1 + 0.1 * (rng.standard_normal(m) + 1j * rng.standard_normal(m))
    w = 0.2 * (rng.standard_normal(m) + 1j * rng.standard_normal(m))
    emw = np.exp(-w)
    ref = o.bead_step(q, emw, bond)
    assert rel(g.bead_step(q, emw, bond), ref) < 1e-13
    s32 = bare_system(m, precision=PREC_FP32, fft_schedule=sched)
    assert rel(hip_engine(s32).bead_step(q, emw, bond), ref) < 2e-5


@pytest.mark.parametrize("m,bond", [((32, 32, 32), 1.3), ((96, 96, 96), 1.0), ((48, 24, 24), 0.7), ((64, 32, 32), 1.1)])
def test_bead_step_conv_mfma_matches_oracle(m, bond):
    """The conv schedule with its x pass as DFT-as-GEMM on the matrix cores (FP32 only),
    against the FP64 oracle at the FP32 bar; FP64 asks for it are refused."""
    print(f"bead step conv_mfma {m} b={bond}", flush=True)
    s = bare_system(m, fft_schedule=SCHED_CONV)
    o = oracle_engine(s)
    rng = np.random.default_rng(7)
    q = 1 + 0.1 * (rng.standard_normal(m) + 1j * rng.standard_normal(m))
    w = 0.2 * (rng.standard_normal(m) + 1j * rng.standard_normal(m))
    emw = np.exp(-w)
    ref = o.bead_step(q, emw, bond)
    g32 = hip_engine(bare_system(m, precision=PREC_FP32, fft_schedule=SCHED_CONV_MFMA))
    assert g32.info.fft_schedule == SCHED_CONV_MFMA
    assert rel(g32.bead_step(q, emw, bond), ref) < 2e-5
    with pytest.raises(FtsError):
        hip_engine(bare_system(m, fft_schedule=SCHED_CONV_MFMA))


def test_autotuner_times_both_schedules():
    e = hip_engine(make_system(96, 100, precision=PREC_FP32))
    print("tune ms (conv, separable):", list(e.info.tune_ms), "chosen", e.info.fft_schedule, flush=True)
    assert e.info.fft_schedule in (SCHED_CONV, SCHED_SEPARABLE)
    tune = list(e.info.tune_ms)
    assert min(tune) > 0 and tune[0 if e.info.fft_schedule == SCHED_CONV else 1] == min(tune)


def c1_like(n=32, N=20, **kw):
    return make_system(n, N, rho=0.2, salt=0.05, B=0.5, lB=1.0, dt=0.01, seed=1234, **kw)


def copolymer_system(**kw):
    """Two species, two monomer types with chi, non-uniform charges and bonds."""
    N1, N2 = 12, 8
    sp1 = dict(num_chains=6, types=[0] * 6 + [1] * 6, charges=[-1.0, 0.0] * 6, bonds=np.linspace(0.8, 1.2, N1 - 1))
    sp2 = dict(num_chains=9, types=[1, 0] * 4, charges=[0.5] * N2)
    rho0 = (6 * N1 + 9 * N2) / (8.0 ** 3)
    return System(m=(16, 16, 16), dL=0.5, a=1.0, species=[sp1, sp2], monomer_types=2, xi=0.5 * rho0,
                  chi=[[0.0, 0.3], [0.3, 0.0]], lB=0.8, counterion=dict(auto=True, valence=1.0),
                  salt=dict(num_salt=10, smearing_length=0.7), dt=0.005, lambda_omega=(1.0,), seed=99, **kw)


def set_random_fields(engines, n_omega, m, sigma, seed):
    rng = np.random.default_rng(seed)
    om = sigma * (rng.standard_normal((n_omega,) + m) + 1j * rng.standard_normal((n_omega,) + m))
    ps = sigma * (rng.standard_normal(m) + 1j * rng.standard_normal(m))
    for e in engines:
        e.set_fields(om, ps)


FIELDS = [FIELD_OMEGA, FIELD_PSI, FIELD_RHO_TYPE, FIELD_RHO_P, FIELD_RHO_C, FIELD_RHO_C_POLY, FIELD_RHO_C_ION, FIELD_ION_RHO]


@pytest.mark.parametrize("which", ["c1", "copolymer"])
def test_densities_match_oracle(which):
    print(f"densities {which}", flush=True)
    s = c1_like() if which == "c1" else copolymer_system()
    g, o = hip_engine(s), oracle_engine(s)
    set_random_fields([g, o], g.info.n_omega, s.m, 0.1, 17)
    Qg, Qo = g.compute_densities(), o.compute_densities()
    assert rel(Qg, Qo) < 1e-11
    for fid in FIELDS:
        assert rel(g.get_field(fid), o.get_field(fid)) < 1e-10, fid
    if which == "c1":
        assert rel(g.get_field(FIELD_QF, 20), o.get_field(FIELD_QF, 20)) < 1e-11
        assert rel(g.get_field(FIELD_QB, 20), o.get_field(FIELD_QB, 20)) < 1e-11


def obs_rel(a, b):
    """Relative differences of the observables.  mu_ex = -ln Q is compared with the
    error scale of Q itself: an absolute error dQ/Q in mu is all a last-bit error
    in Q can cause, and mu near 0 (Q near 1, the ions) makes a relative measure of
    mu meaningless, so mu uses |d mu| / max(|mu|, 1)."""
    out = {}
    for key in ("H", "p_iso"):
        out[key] = abs(a[key] - b[key]) / abs(b[key])
    for key in ("stress", "Q_poly", "Q_ion"):
        out[key] = rel(a[key], b[key]) if len(b[key]) else 0.0
    for key in ("mu_ex_poly", "mu_ex_ion"):
        if len(b[key]):
            out[key] = float(np.abs(a[key] - b[key]).max() / max(np.abs(b[key]).max(), 1.0))
    return out


@pytest.mark.parametrize("which", ["c1", "copolymer"])
def test_measure_matches_oracle(which):
    print(f"measure {which}", flush=True)
    s = c1_like(stress_mask=0x7F) if which == "c1" else copolymer_system(stress_mask=0x7F)
    g, o = hip_engine(s), oracle_engine(s)
    set_random_fields([g, o], g.info.n_omega, s.m, 0.1, 23)
    r = obs_rel(g.measure(), o.measure())
    assert max(r.values()) < 1e-10, r


@pytest.mark.parametrize("scheme,sched", [(SCHEME_PC_REFERENCE, SCHED_PLANE), (SCHEME_PC_REFERENCE, SCHED_LINE),
                                          (SCHEME_EM, SCHED_PLANE), (SCHEME_SEMI_IMPLICIT, SCHED_LINE),
                                          (SCHEME_PC_REFERENCE, SCHED_CONV), (SCHEME_PC_REFERENCE, SCHED_SEPARABLE),
                                          (SCHEME_SEMI_IMPLICIT, SCHED_SEPARABLE), (SCHEME_SIPC, SCHED_PLANE),
                                          (SCHEME_SIPC, SCHED_LINE)])
def test_trajectory_matches_oracle(scheme, sched):
    print(f"trajectory scheme={scheme} sched={sched}", flush=True)
    s = c1_like(stress_mask=0x7F, scheme=scheme, fft_schedule=sched)
    g, o = hip_engine(s), oracle_engine(s)
    om, ps = initial_fields(s, sigma=1e-3, cfg=1)
    for e in (g, o):
        e.set_fields(om, ps)
    rg, ro = g.step(6, record_every=2), o.step(6, record_every=2)
    assert len(rg) == len(ro) == 3
    for a, b in zip(rg, ro):
        assert a["step"] == b["step"]
        r = obs_rel(a, b)
        assert max(r.values()) < 1e-10, r
    assert rel(g.get_field(FIELD_OMEGA), o.get_field(FIELD_OMEGA)) < 1e-10
    assert rel(g.get_field(FIELD_PSI), o.get_field(FIELD_PSI)) < 1e-10


@pytest.mark.parametrize("sched", [SCHED_PLANE, SCHED_CONV])
def test_copolymer_pc_trajectory(sched):
    print(f"copolymer pc sched={sched}", flush=True)
    s = copolymer_system(stress_mask=0x7F, fft_schedule=sched)
    g, o = hip_engine(s), oracle_engine(s)
    set_random_fields([g, o], g.info.n_omega, s.m, 0.05, 31)
    rg, ro = g.step(3, record_every=1), o.step(3, record_every=1)
    for a, b in zip(rg, ro):
        r = obs_rel(a, b)
        assert max(r.values()) < 1e-10, r


def triblock_system(**kw):
    """A palindromic chain with three bead classes, an odd N (a middle bead) and
    symmetric bonds: A4 B5 A4 with charges -1 on A, +0.5 on B."""
    types = [0] * 4 + [1] * 5 + [0] * 4
    charges = [-1.0] * 4 + [0.5] * 5 + [-1.0] * 4
    bonds = [0.9, 1.1, 1.0, 1.2, 0.8, 1.0, 1.0, 0.8, 1.2, 1.0, 1.1, 0.9]
    sp = dict(num_chains=8, types=types, charges=charges, bonds=bonds)
    rho0 = 8 * 13 / (8.0 ** 3)
    return System(m=(16, 16, 16), dL=0.5, a=1.0, species=[sp], monomer_types=2, xi=0.5 * rho0,
                  chi=[[0.0, 0.4], [0.4, 0.0]], lB=0.9, counterion=dict(auto=True, valence=1.0),
                  salt=dict(num_salt=12), dt=0.005, seed=41, **kw)


@pytest.mark.parametrize("which", ["c1", "triblock"])
def test_mirrored_sweep_matches_both_sweeps_and_oracle(which):
    """A palindromic species runs the forward sweep only (qB_s = qF_{N-1-s}); the
    result must match the engine with both sweeps and the oracle."""
    print(f"mirrored sweep {which}", flush=True)
    mk = (lambda **kw: c1_like(stress_mask=0x7F, **kw)) if which == "c1" else (lambda **kw: triblock_system(stress_mask=0x7F, **kw))
    gm, gb, o = hip_engine(mk()), hip_engine(mk(chain_symmetry=SYMMETRY_OFF)), oracle_engine(mk())
    assert gm.info.mirrored_species == 1 and gb.info.mirrored_species == 0
    set_random_fields([gm, gb, o], gm.info.n_omega, mk().m, 0.1, 29)
    Qm, Qb, Qo = gm.compute_densities(), gb.compute_densities(), o.compute_densities()
    assert rel(Qm, Qo) < 1e-11 and rel(Qm, Qb) < 1e-12
    for fid in FIELDS:
        assert rel(gm.get_field(fid), o.get_field(fid)) < 1e-10, fid
    N = mk().desc.species[0].N
    assert rel(gm.get_field(FIELD_QB, N), o.get_field(FIELD_QB, N)) < 1e-11
    assert rel(gm.get_field(FIELD_QB, N), gb.get_field(FIELD_QB, N)) < 1e-12
    r = obs_rel(gm.measure(), o.measure())
    assert max(r.values()) < 1e-10, r
    rg, ro = gm.step(3, record_every=1), o.step(3, record_every=1)
    for a, b in zip(rg, ro):
        r = obs_rel(a, b)
        assert max(r.values()) < 1e-10, r
    # FP32 with the mirror against FP32 with both sweeps: the same propagators
    f32m, f32b = hip_engine(mk(precision=PREC_FP32)), hip_engine(mk(precision=PREC_FP32, chain_symmetry=SYMMETRY_OFF))
    set_random_fields([f32m, f32b], f32m.info.n_omega, mk().m, 0.1, 29)
    a, b = f32m.measure(), f32b.measure()
    assert abs(a["H"] - b["H"]) / abs(b["H"]) < 1e-6
    assert rel(a["stress"], b["stress"]) < 1e-5


def test_fp32_engine_tracks_fp64_at_96():
    """FP32 propagators (per-bead renormalised) vs the FP64 engine at the headline size,
    on equilibrated fields (100 plain EM steps), where |Q| spans many decades."""
    print("fp32 vs fp64 96^3 N=100", flush=True)
    s64 = make_system(96, 100, stress_mask=0x7F, seed=5, scheme=SCHEME_EM)
    g64 = hip_engine(s64)
    om, ps = initial_fields(s64, cfg=4)
    g64.set_fields(om, ps)
    g64.step(100)
    om, ps = g64.get_field(FIELD_OMEGA), g64.get_field(FIELD_PSI)[0]
    s32 = make_system(96, 100, stress_mask=0x7F, seed=5, scheme=SCHEME_EM, precision=PREC_FP32)
    g32 = hip_engine(s32)
    g32.set_fields(om, ps)
    g64.set_fields(om, ps)
    a, b = g32.measure(), g64.measure()
    assert abs(a["H"] - b["H"]) / abs(b["H"]) < 1e-4, (a["H"], b["H"])
    assert abs(a["p_iso"] - b["p_iso"]) / abs(b["p_iso"]) < 1e-4
    assert rel(a["stress"], b["stress"]) < 1e-4
    # mu = -ln Q compared absolutely: |Q| itself varies over decades between states
    assert np.abs(a["mu_ex_poly"] - b["mu_ex_poly"]).max() < 1e-4 * max(1.0, np.abs(b["mu_ex_poly"]).max())
    # density means are exact: <rho_p> = n N / V and <rho_c> = 0
    rho_p = g64.get_field(FIELD_RHO_P)
    nN_V = s64.desc.species[0].num_chains * 100 / g64.info.volume
    assert abs(rho_p.mean() - nN_V) < 1e-10 * nN_V
    assert abs(g64.get_field(FIELD_RHO_C).mean()) < 1e-10
    r32 = g32.get_field(FIELD_RHO_P)
    assert rel(r32, rho_p) < 1e-3
    # the fields the FP32 form transforms in FP32 (smoothing, ion potentials and charges)
    for fid in (FIELD_RHO_C, FIELD_RHO_C_ION, FIELD_ION_RHO):
        assert rel(g32.get_field(fid), g64.get_field(fid)) < 1e-3, fid


def test_noncubic_line_schedule_measure():
    """m2 != m3 only runs with line passes; densities and observables still match."""
    print("non-cubic 32x24x16 measure", flush=True)
    sp = [dict(num_chains=6, types=[0] * 10, charges=[-0.5] * 10)]
    s = System(m=(32, 24, 16), dL=0.5, species=sp, xi=0.05, lB=1.0, counterion=dict(auto=True), salt=dict(num_salt=12),
               stress_mask=0x7F, seed=3)
    g, o = hip_engine(s), oracle_engine(s)
    set_random_fields([g, o], 1, s.m, 0.1, 45)
    r = obs_rel(g.measure(), o.measure())
    assert max(r.values()) < 1e-10, r


def test_headline_measure_matches_oracle_64():
    print("measure 64^3 N=50 vs oracle", flush=True)
    s = make_system(64, 50, salt=0.1, stress_mask=0x7F, seed=8)
    g, o = hip_engine(s), oracle_engine(s)
    set_random_fields([g, o], 1, s.m, 0.05, 43)
    r = obs_rel(g.measure(), o.measure())
    assert max(r.values()) < 1e-10, r


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_FP32])
def test_ion_hot_spot_matches_oracle(precision):
    """A psi hot spot that drives an ion partition function to 4e198 (|Q|^2 beyond FP64):
    mu = -log Q and H stay finite and match the oracle; a higher spot whose ion Q
    overflows fails the step as a non-finite Q does."""
    from test_oracle import ion_hot_spot_state
    print(f"ion hot spot precision={precision}", flush=True)
    s, om, ps = ion_hot_spot_state(SCHEME_PC_REFERENCE, 800.0, precision=precision)
    g, o = hip_engine(s), oracle_engine(s)
    for e in (g, o):
        e.set_fields(om, ps)
    rg, ro = g.measure(), o.measure()
    tol = 1e-10 if precision == PREC_FP64 else 1e-5
    assert np.isfinite(rg["H"])
    for j in range(2):
        assert abs(rg["mu_ex_ion"][j] - ro["mu_ex_ion"][j]) <= tol * abs(ro["mu_ex_ion"][j]), j
    assert abs(rg["H"] - ro["H"]) <= tol * abs(ro["H"])
    s2, om2, ps2 = ion_hot_spot_state(SCHEME_PC_REFERENCE, 3000.0, precision=precision)
    g2 = hip_engine(s2)
    g2.set_fields(om2, ps2)
    with pytest.raises(FtsError) as err:
        g2.step(1, record_every=1)
    assert "non-finite Q" in str(err.value)

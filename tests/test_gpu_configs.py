"""GPU: the configurations the other tests do not reach at full size.

Precision form (64^3, L = 32, N = 50, salt 0.1, semi-implicit at dt = 0.05): FP32
against FP64 on the same noise stream.  At identical fields (states along an FP64
trajectory) H, mu, the stress tensor and p_iso agree within the precision form's bar
of 1e-4; the same-noise trajectories agree step by step (H within 1e-5) for the
first steps, before the CL dynamics amplify the rounding differences.

Large single system (256^3, L = 128, N = 200, charge -1 on every second bead, salt
0.02, B = 1, lB = 0.7): FP32 propagators (both sweeps: the chain is not
palindromic), FP64 reductions and the learned preconditioner, three steps.  Checked
by what holds at any size (the density means nN/V and 0, exact up to the FP32
propagators' rounding; p_iso = tr sigma/3; finite Q), and by the precision form's own bar: H, mu and the stress tensor of the
FP32 engine within 1e-4 of the FP64 engine on the same fields.

Batched form (64 replicas, n in {32, 48, 64, 96}, N in [20, 100], salt and lB drawn
per replica, own seeds): one batch, three steps; every replica keeps the density
means (to FP32 rounding), and one replica of each mesh reproduces its single-engine run to
1e-12.
"""
import numpy as np
import pytest

from ftspolysol_amd.api import (FIELD_OMEGA, FIELD_PSI, FIELD_RHO_C, FIELD_RHO_P, PREC_FP32, PREC_FP64, SCHED_PLANE, SCHED_SEPARABLE,
                                SCHEME_PC_REFERENCE, SCHEME_SEMI_IMPLICIT, SCHEME_SI_LEARNED, hip_batch, hip_engine)
from ftspolysol_amd.lrk import load_params
from ftspolysol_amd.system import config, initial_fields, make_system, replica_params

pytestmark = pytest.mark.gpu


def exact_means(e, system):
    """<rho_p> = nN/V and <rho_c> = 0: exact in FP64, to the rounding of the
    propagators (measured ~2e-7 relative) with FP32 ones."""
    d = system.desc
    tol = 1e-5 if d.precision == PREC_FP32 else 1e-10
    nN = sum(d.species[p].num_chains * d.species[p].N for p in range(d.n_species))
    rho_p = e.get_field(FIELD_RHO_P)
    target = nN / e.info.volume
    assert abs(rho_p.mean() - target) < tol * target
    assert abs(e.get_field(FIELD_RHO_C).mean()) < tol * target


def test_large_system_all_forms():
    print("256^3 N=200 FP32 + learned", flush=True)
    s32 = config(5, precision=PREC_FP32, scheme=SCHEME_SI_LEARNED, stress_mask=0x7F, seed=0xC5)
    g = hip_engine(s32)
    assert g.info.fft_schedule == SCHED_SEPARABLE and g.info.mirrored_species == 0
    g.lrk_set_network(load_params("ftspolysol_amd/lrk_c4_48cubed.npy"), 1e-3)
    g.set_fields(*initial_fields(s32, sigma=1e-3, cfg=5))
    g.set_timing(True)
    recs = g.step(3, record_every=1)
    t = g.get_timing()
    print(f"bead steps {t['bead_steps']}, {t['algorithmic_bytes'] / (t['sweep_ms'] * 1e-3) / 1e9:.0f} GB/s algorithmic", flush=True)
    assert len(recs) == 3 and all(np.isfinite(r["H"]) for r in recs)
    st = g.lrk_get_stats()
    assert st["judged"] == 2 and st["learned_steps"] + st["fallback_steps"] == 3
    a = g.measure()
    for r in recs + [a]:
        sig = r["stress"]
        assert abs(r["p_iso"] - (sig[0] + sig[1] + sig[2]) / 3) <= 1e-12 * abs(r["p_iso"])
    exact_means(g, s32)
    om, ps = g.get_field(FIELD_OMEGA), g.get_field(FIELD_PSI)[0]
    g.close()
    s64 = config(5, precision=PREC_FP64, scheme=SCHEME_PC_REFERENCE, stress_mask=0x7F, seed=0xC5)
    h = hip_engine(s64)
    h.set_fields(om, ps)
    b = h.measure()
    assert abs(a["H"] - b["H"]) <= 1e-4 * abs(b["H"])
    assert np.abs(a["stress"] - b["stress"]).max() <= 1e-4 * np.abs(b["stress"]).max()
    assert np.abs(a["mu_ex_poly"] - b["mu_ex_poly"]).max() <= 1e-4 * max(1.0, np.abs(b["mu_ex_poly"]).max())
    h.close()


def test_batched_form_64_replicas():
    print("64 replicas", flush=True)
    systems = [make_system(**replica_params(r), dt=0.01, precision=PREC_FP32, fft_schedule=SCHED_PLANE) for r in range(64)]
    b = hip_batch(systems)
    assert sorted(b.info.group_mesh[k][0] for k in range(b.info.n_groups)) == [32, 48, 64, 96]
    for r, (e, s) in enumerate(zip(b.replicas, systems)):
        e.set_fields(*initial_fields(s, sigma=1e-3, seed=900 + r))
    recs = b.step(3, record_every=1)
    assert all(len(x) == 3 for x in recs)
    for e, s in zip(b.replicas, systems):
        exact_means(e, s)
    first = {}
    for r, s in enumerate(systems):
        first.setdefault(s.m[0], r)
    for n, r in sorted(first.items()):
        e = hip_engine(systems[r])
        e.set_fields(*initial_fields(systems[r], sigma=1e-3, seed=900 + r))
        single = e.step(3, record_every=1)
        for x, y in zip(single, recs[r]):
            assert abs(x["H"] - y["H"]) <= 1e-12 * abs(x["H"]), (n, r)
        w = b.replicas[r].get_field(FIELD_OMEGA)
        assert np.abs(e.get_field(FIELD_OMEGA) - w).max() <= 1e-12 * np.abs(w).max(), (n, r)
        e.close()
    b.close()


def test_precision_form_64():
    print("64^3 N=50 FP32 vs FP64, semi-implicit dt=0.05", flush=True)
    mk = lambda prec: config(2, precision=prec, scheme=SCHEME_SEMI_IMPLICIT, stress_mask=0x7F, seed=0xC2)
    s64, s32 = mk(PREC_FP64), mk(PREC_FP32)
    assert s64.m == (64, 64, 64) and s64.desc.species[0].N == 50 and s64.desc.dt == 0.05
    g64, g32 = hip_engine(s64), hip_engine(s32)
    init = initial_fields(s64, sigma=1e-3, cfg=2)
    for e in (g64, g32):
        e.set_fields(*init)
    r64, r32 = g64.step(8, record_every=1), g32.step(8, record_every=1)
    # the same noise: the trajectories agree step by step at first (record 0 is the
    # initial state, where H is near 0)
    for a, b in zip(r32[1:], r64[1:]):
        assert abs(a["H"] - b["H"]) <= 1e-5 * abs(b["H"]), (a["step"], a["H"], b["H"])
    # identical fields: the precision error of every observable
    probe = hip_engine(mk(PREC_FP32))
    g64.set_restart(20, 10)   # dt = 0.05 sees rare non-finite excursions (restarted from a checkpoint)
    for k in range(3):
        g64.step(40)
        om, ps = g64.get_field(FIELD_OMEGA), g64.get_field(FIELD_PSI)[0]
        probe.set_fields(om, ps)
        a, b = probe.measure(), g64.measure()
        assert abs(a["H"] - b["H"]) <= 1e-4 * abs(b["H"]), (k, a["H"], b["H"])
        assert abs(a["p_iso"] - b["p_iso"]) <= 1e-4 * abs(b["p_iso"])
        assert np.abs(a["stress"] - b["stress"]).max() <= 1e-4 * np.abs(b["stress"]).max()
        assert np.abs(a["mu_ex_poly"] - b["mu_ex_poly"]).max() <= 1e-4 * max(1.0, np.abs(b["mu_ex_poly"]).max())
    for e in (g64, g32, probe):
        e.close()


def test_fp32_stress_finite_where_ion_boltzmann_factors_leave_fp32_range():
    """A strong imaginary psi mode (amplitude 150) makes the ions' Boltzmann factors
    b_j = e^{-i z_j psi~_j} reach about e^150, beyond FP32 (about e^88), as in rare CL
    excursions of C2.  The FP32 form must still give a finite stress tensor, within its
    precision bar of the FP64 engine on the same fields (the ion rows contract the
    normalised ion density, which stays of the order of the density)."""
    print("strong imaginary psi mode, FP32 vs FP64 measure", flush=True)
    mk = lambda prec: config(1, precision=prec, scheme=SCHEME_SEMI_IMPLICIT, stress_mask=0x7F, seed=0xC1)
    s64 = mk(PREC_FP64)
    n = s64.m[0]
    x = np.arange(n)
    psi = 150j * np.cos(2 * np.pi * x / n)[:, None, None] * np.ones(s64.m)
    om = initial_fields(s64, sigma=0.05, cfg=1)[0]
    out = []
    for prec in (PREC_FP64, PREC_FP32):
        e = hip_engine(mk(prec))
        e.set_fields(om, psi)
        out.append(e.measure())
        e.close()
    b, a = out
    assert np.all(np.isfinite(b["stress"])) and np.isfinite(b["p_iso"])
    assert np.all(np.isfinite(a["stress"])) and np.isfinite(a["p_iso"])
    assert np.abs(a["stress"] - b["stress"]).max() <= 1e-4 * np.abs(b["stress"]).max()
    assert abs(a["H"] - b["H"]) <= 1e-4 * abs(b["H"])

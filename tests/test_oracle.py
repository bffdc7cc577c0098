"""CPU pins of the FP64 oracle (oracle/libfts_oracle.so).

The reference ships no tests, fixtures or golden outputs, so the oracle is pinned by
known answers and mathematics that a plausible mistake would break:
  - FFT: brute-force DFT, numpy, round trip;
  - Philox4x32-10: the known-answer fixture from PyTorch's philox_engine, and the
    moments of the Box-Muller normals;
  - zero fields: Q = 1, uniform densities, H = 0, the closed-form stress tensor
    with the brute-force mode term;
  - uniform potential W = w0: Q = exp(-N w0);
  - propagator invariant <qF_s qB_s e^{W_s}> = Q for every s;
  - weak single mode: the discrete-chain (Debye) linear response of the density;
  - interaction modes: M = 1 and M = 2 closed forms;
  - CL updates: Heun (PC) second order / EM first order with the noise off, and the
    semi-implicit step's exact damping 1/(1 + dt lam A(k)) of a linear mode;
  - p_iso = tr(sigma)/3 on random fields.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from ftspolysol_amd.api import (FTS_NONFINITE, FtsError, FIELD_OMEGA, FIELD_PSI, FIELD_QB, FIELD_QF, FIELD_RHO_C, FIELD_RHO_P, FIELD_ION_RHO, SCHEME_EM,
                                SCHEME_PC_REFERENCE, SCHEME_SEMI_IMPLICIT, SCHEME_SIPC)
from ftspolysol_amd.system import System, config, make_system
from oracle_lib import oracle_engine, oracle_library

HERE = os.path.dirname(os.path.abspath(__file__))


def homopolymer(m, N, n_chains, lB=0.0, charge=0.0, xi=0.1, dL=0.5, a=1.0, ions=False, salt=0, **kw):
    sp = [dict(num_chains=n_chains, types=[0] * N, charges=[charge] * N)]
    return System(m=m, dL=dL, a=a, species=sp, xi=xi, lB=lB,
                  counterion=dict(auto=True, valence=1.0) if ions else None,
                  salt=dict(num_salt=salt) if salt else None, **kw)


def kvecs(m, dL):
    out = []
    for n in m:
        idx = np.arange(n)
        mode = np.where(idx > n // 2, idx - n, idx)  # Nyquist -> +n/2 (common.h:219-222)
        out.append(2 * np.pi * mode / (dL * n))
    return np.meshgrid(*out, indexing="ij")


# ---------------------------------------------------------------------------- FFT
def brute_dft(x, sign):
    out = x.astype(np.complex128)
    for ax in range(3):
        n = x.shape[ax]
        F = np.exp(sign * 2j * np.pi * np.outer(np.arange(n), np.arange(n)) / n)
        out = np.moveaxis(np.tensordot(F, np.moveaxis(out, ax, 0), axes=(1, 0)), 0, ax)
    return out


@pytest.mark.parametrize("m", [(4, 6, 5), (3, 8, 7), (2, 9, 10)])
def test_fft_brute_force(m):
    e = oracle_engine(homopolymer(m, 2, 1))
    rng = np.random.default_rng(0)
    x = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    for sign in (-1, +1):
        ref = brute_dft(x, sign)
        assert np.abs(e.fft3d(x, sign) - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("m", [(8, 12, 16), (96, 8, 6), (32, 32, 32)])
def test_fft_numpy_and_round_trip(m):
    e = oracle_engine(homopolymer(m, 2, 1))
    rng = np.random.default_rng(1)
    x = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    f = e.fft3d(x, -1)
    assert np.abs(f - np.fft.fftn(x)).max() < 1e-13 * np.abs(f).max()
    back = e.fft3d(f, +1) / x.size
    assert np.abs(back - x).max() < 1e-13 * np.abs(x).max()
    assert abs(np.vdot(f, f).real / x.size - np.vdot(x, x).real) < 1e-12 * np.vdot(x, x).real  # Parseval


# --------------------------------------------------------------------------- noise
def test_philox_known_answers():
    lib = oracle_library()
    f = lib.fto_philox4x32_10
    f.argtypes = [ctypes.POINTER(ctypes.c_uint)] * 3
    for case in json.load(open(os.path.join(HERE, "golden", "philox_kat.json"))):
        ctr = (ctypes.c_uint * 4)(*case["counter"])
        key = (ctypes.c_uint * 2)(*case["key"])
        out = (ctypes.c_uint * 4)()
        f(ctr, key, out)
        assert list(out) == case["out"]


def test_noise_moments_and_independence():
    lib = oracle_library()
    g = lib.fto_cl_normal
    g.restype = ctypes.c_double
    g.argtypes = [ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_uint, ctypes.c_uint]
    n = 40000
    a = np.array([g(7, 3, 0, i) for i in range(n)])
    b = np.array([g(7, 4, 0, i) for i in range(n)])
    c = np.array([g(7, 3, 0x100, i) for i in range(n)])
    for x in (a, b, c):
        assert abs(x.mean()) < 4 / np.sqrt(n)
        assert abs(x.var() - 1) < 5 * np.sqrt(2 / n)
        assert abs(np.mean(x ** 4) - 3) < 0.1
    for x, y in ((a, b), (a, c)):
        assert abs(np.corrcoef(x, y)[0, 1]) < 4 / np.sqrt(n)


# ------------------------------------------------------------------ closed forms
def test_zero_fields_closed_forms():
    s = config(1, n=16, N=8, stress_mask=0x7F)
    e = oracle_engine(s)
    info = e.info
    assert info.n_ions == 2
    Q = e.compute_densities()
    assert abs(Q[0] - 1) < 1e-14
    nN_V = s.desc.species[0].num_chains * 8 / info.volume
    assert np.abs(e.get_field(FIELD_RHO_P) - nN_V).max() < 1e-13
    assert np.abs(e.get_field(FIELD_RHO_C)).max() < 1e-13
    rho_ion = e.get_field(FIELD_ION_RHO)
    for j in range(info.n_ions):
        assert np.abs(rho_ion[j] - info.ion_count[j] / info.volume).max() < 1e-13
    rec = e.measure()
    assert abs(rec["H"]) < 1e-14
    # sigma_ab = (ideal + P_sr) delta_ab + T_ab, T by brute force (stress.cu:106-149, 198-207)
    kx, ky, kz = kvecs(s.m, s.dL)
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    nz = k2 > 0
    comps = [kx * kx, ky * ky, kz * kz, kx * ky, kx * kz, ky * kz]
    ideal = (s.desc.species[0].num_chains + sum(info.ion_count[j] for j in range(info.n_ions))) / info.volume
    P_sr = s.desc.xi / (16 * np.pi ** 1.5 * s.desc.a ** 3)
    assert abs(info.short_range_self_pressure - P_sr) < 1e-15
    for c in range(6):
        d = 1.0 if c < 3 else 0.0
        T = (nz.sum() * d / 3 - (comps[c][nz] / k2[nz]).sum()) / info.volume
        assert abs(rec["stress"][c] - ((ideal + P_sr) * d + T)) < 1e-12
    assert abs(rec["p_iso"] - (ideal + P_sr)) < 1e-12
    # off-diagonal mode term is non-zero on an even grid (Nyquist at +m/2)
    assert abs(rec["stress"][3]) > 1e-6


def test_ion_counts_config1():
    e = oracle_engine(config(1))
    info = e.info
    # 41 chains x 20 beads x 0.5 = 410 counterions, merged with 205 salt cations
    assert [info.ion_valence[j] for j in range(info.n_ions)] == [1.0, -1.0]
    assert [info.ion_count[j] for j in range(info.n_ions)] == [615, 205]


@pytest.mark.parametrize("w0", [0.03, -0.05 + 0.02j])
def test_uniform_potential(w0):
    N = 7
    s = homopolymer((8, 8, 8), N, 3)
    e = oracle_engine(s)
    # M = 1: gamma = i, W~ = i omega~; omega = -i w0 gives W = w0
    e.set_fields(np.full((1, 8, 8, 8), -1j * w0))
    Q = e.compute_densities()[0]
    assert abs(Q - np.exp(-N * w0)) < 1e-13 * abs(Q)


def smoothed(f, m, dL, a):
    kx, ky, kz = kvecs(m, dL)
    G = np.exp(-0.5 * a * a * (kx ** 2 + ky ** 2 + kz ** 2))
    return np.fft.ifftn(G * np.fft.fftn(f))


def test_propagator_invariant():
    m, N = (8, 12, 12), 9
    s = homopolymer(m, N, 2)
    e = oracle_engine(s)
    rng = np.random.default_rng(4)
    om = 0.2 * (rng.standard_normal((1,) + m) + 1j * rng.standard_normal((1,) + m))
    e.set_fields(om)
    Q = e.compute_densities()[0]
    W = 1j * smoothed(om[0], m, s.dL, s.desc.a)
    qF, qB = e.get_field(FIELD_QF, N), e.get_field(FIELD_QB, N)
    for sb in range(N):
        val = np.mean(qF[sb] * qB[sb] * np.exp(W))
        assert abs(val - Q) < 1e-12 * abs(Q)
    # forward and backward sweeps agree on Q
    assert abs(np.mean(qB[0]) - Q) < 1e-12 * abs(Q)


def debye_pair_sum(N, phi):
    j = np.arange(1, N)
    return N + 2 * np.sum((N - j) * phi ** j)


def test_debye_linear_response():
    m, N, n, b = (16, 16, 16), 12, 5, 1.0
    s = homopolymer(m, N, n, dL=0.5, a=1.0)
    e = oracle_engine(s)
    L = 0.5 * 16
    k0 = 2 * np.pi / L
    x = np.arange(16) * 0.5
    eps = 1e-6
    om = np.zeros((1,) + m, dtype=complex)
    om[0] = -1j * eps * np.cos(k0 * x)[:, None, None]
    e.set_fields(om)
    e.compute_densities()
    rho = e.get_field(FIELD_RHO_P)[0]
    G = np.exp(-0.5 * k0 ** 2)
    phi = np.exp(-b * b * k0 ** 2 / 6)
    expected = -(n / e.info.volume) * debye_pair_sum(N, phi) * eps * G * G
    amp = 2 * np.mean(rho * np.cos(k0 * x)[:, None, None])
    assert abs(amp - expected) < 1e-4 * abs(expected)


def test_modes_closed_forms():
    e1 = oracle_engine(homopolymer((8, 8, 8), 2, 1, xi=0.3))
    assert e1.info.n_omega == 1 and abs(e1.info.mode_eigenvalue[0] - 0.3) < 1e-15
    assert e1.info.gamma_is_imag[0] == 1.0 and e1.info.mode_O[0][0] == 1.0
    xi, chi = 0.1, 0.3
    sp = [dict(num_chains=2, types=[0, 1, 0, 1], charges=[0.0] * 4)]
    s2 = System(m=(8, 8, 8), dL=0.5, species=sp, monomer_types=2, xi=xi, chi=[[0, chi], [chi, 0]])
    i2 = oracle_engine(s2).info
    assert i2.n_omega == 2
    assert abs(i2.mode_eigenvalue[0] - (2 * xi + chi)) < 1e-14 and abs(i2.mode_eigenvalue[1] + chi) < 1e-14
    r = 1 / np.sqrt(2)
    assert np.allclose([i2.mode_O[0][0], i2.mode_O[0][1], i2.mode_O[1][0], i2.mode_O[1][1]], [r, r, r, -r], atol=1e-14)
    assert i2.gamma_is_imag[0] == 1.0 and i2.gamma_is_imag[1] == 0.0
    rho0 = 2 * 4 / (4.0 ** 3)
    assert abs(i2.omega_coefficient[1] - rho0 / chi) < 1e-14


def test_p_iso_is_trace_over_three():
    s = config(1, n=16, N=8, stress_mask=0x7F)
    e = oracle_engine(s)
    rng = np.random.default_rng(9)
    m = s.m
    e.set_fields(0.1 * (rng.standard_normal((1,) + m) + 1j * rng.standard_normal((1,) + m)),
                 0.1 * (rng.standard_normal(m) + 1j * rng.standard_normal(m)))
    rec = e.measure()
    tr = (rec["stress"][0] + rec["stress"][1] + rec["stress"][2]) / 3
    assert abs(rec["p_iso"] - tr) < 1e-12 * abs(tr)


# ------------------------------------------------------------------- CL updates
def run_deterministic(scheme, dt, T, om, ps):
    s = make_system(8, 4, rho=0.2, salt=0.05, dt=dt, scheme=scheme, noise=False)
    e = oracle_engine(s)
    e.set_fields(om, ps)
    e.step(int(round(T / dt)))
    return e.get_field(FIELD_OMEGA)


@pytest.mark.parametrize("scheme,order", [(SCHEME_PC_REFERENCE, 2), (SCHEME_EM, 1), (SCHEME_SIPC, 2)])
def test_time_step_convergence_order(scheme, order):
    rng = np.random.default_rng(12)
    om = 0.3 * (rng.standard_normal((1, 8, 8, 8)) + 1j * rng.standard_normal((1, 8, 8, 8)))
    ps = 0.3 * (rng.standard_normal((8, 8, 8)) + 1j * rng.standard_normal((8, 8, 8)))
    T = 0.08
    ref = run_deterministic(SCHEME_PC_REFERENCE, T / 64, T, om, ps)
    e1 = np.abs(run_deterministic(scheme, T / 4, T, om, ps) - ref).max()
    e2 = np.abs(run_deterministic(scheme, T / 8, T, om, ps) - ref).max()
    ratio = e1 / e2
    assert abs(np.log2(ratio) - order) < 0.25, ratio


@pytest.mark.parametrize("scheme", [SCHEME_EM, SCHEME_SEMI_IMPLICIT, SCHEME_SIPC])
def test_linear_mode_damping(scheme):
    """Around the homogeneous fixed point omega0 = -i xi, one deterministic step
    multiplies a weak mode k0 by 1 - dt lam A(k0) (EM), 1/(1 + dt lam A(k0)) (SI) or
    (1 - a/2)/(1 + a/2), a = dt lam A(k0) (SIPC: the Crank-Nicolson factor),
    A(k) = rho0/xi + Gamma(k)^2 (n/V) sum_{ss'} phi^{|s-s'|} from the chain's linear response."""
    m, N, n, xi, dt = (16, 16, 16), 10, 6, 0.05, 0.01
    s = homopolymer(m, N, n, xi=xi, dt=dt, scheme=scheme, noise=False, lambda_omega=(1.0,))
    e = oracle_engine(s)
    k0 = 2 * np.pi / 8.0
    x = np.arange(16) * 0.5
    eps = 1e-6
    om = np.full((1,) + m, -1j * xi, dtype=complex)
    om[0] += -1j * eps * np.cos(k0 * x)[:, None, None]
    e.set_fields(om)
    e.step(1)
    w = e.get_field(FIELD_OMEGA)[0] + 1j * xi
    amp = 2 * np.mean(w * np.cos(k0 * x)[:, None, None]) / (-1j * eps)
    V = e.info.volume
    rho0 = n * N / V
    A = rho0 / xi + np.exp(-k0 ** 2) * (n / V) * debye_pair_sum(N, np.exp(-k0 ** 2 / 6))
    a = dt * A
    expected = {SCHEME_EM: 1 - a, SCHEME_SEMI_IMPLICIT: 1 / (1 + a), SCHEME_SIPC: (1 - a / 2) / (1 + a / 2)}[scheme]
    assert abs(amp - expected) < 1e-4, (amp, expected)


def si_noise_spectrum(seed, step, field, m, P, drop_zero, corrector=False):
    """g(k) Pn(k) of the semi-implicit noise, restated from its definition (g(k) P2(k),
    P2 = 2 P / (1 + P), for the predictor-corrector's corrector): the pair
    {k, -k} takes the two Box-Muller normals a, b of one Philox draw (field `field`,
    site = the pair's smaller grid index), g = (a + i b) / sqrt 2 at the smaller index and
    its conjugate at the other; a self-conjugate k takes a alone; Pn = sqrt(P (1 + P) / 2)."""
    lib = oracle_library()
    g = lib.fto_cl_normal
    g.restype = ctypes.c_double
    g.argtypes = [ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_uint, ctypes.c_uint]
    g2 = lib.fto_cl_normal_pair
    g2.restype = None
    g2.argtypes = [ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(ctypes.c_double)]
    ab = (ctypes.c_double * 2)()
    m1, m2, m3 = m
    out = np.zeros(m, dtype=complex)
    for x in range(m1):
        for y in range(m2):
            for z in range(m3):
                gi = (x * m2 + y) * m3 + z
                px, py, pz = (-x) % m1, (-y) % m2, (-z) % m3
                pi = (px * m2 + py) * m3 + pz
                if drop_zero and gi == 0:
                    continue
                if pi == gi:
                    v = g(seed, step, field, gi)
                else:
                    g2(seed, step, field, min(gi, pi), ab)
                    a, b = ab[0], ab[1]
                    assert abs(a - g(seed, step, field, min(gi, pi))) <= 1e-15 * max(1.0, abs(a))   # cl_normal's value
                    v = (a + 1j * (b if gi < pi else -b)) / np.sqrt(2)
                out[x, y, z] = v
    return out * (2 * P / (1 + P) if corrector else np.sqrt(P * (1 + P) / 2))


def test_semi_implicit_noise_factor():
    """The noise of a semi-implicit step is drawn in k-space and transformed with
    Pn = sqrt(P (1 + P) / 2), P = 1 / (1 + dt lam A(k)) the drift's factor.  One step
    from the same state with the noise on and off gives the step's noise increment,
    which must equal sigma sqrt(M) IFFT[Pn g], g restated here from its definition and
    the exported Philox generator, with A(k) from the closed forms: for omega the
    chain's linear response (as above), for psi of a neutral chain without ions
    k^2 / (4 pi lB) (k = 0 dropped: mean-free noise).  g is the spectrum of real white
    noise / sqrt(M): the increment is real, and E|g|^2 = 1 at every k.  For a linear
    mode Pn keeps the stationary variance at 1 / A: Var = Pn^2 2 lam dt / (1 - (1 - P a)^2)
    with a = dt lam A and 1 - P a = P gives lam dt / (1 - P) = 1 / A."""
    m, N, n, xi, dt, lB, lam_psi, seed = (16, 16, 16), 10, 6, 0.05, 0.2, 1.0, 0.7, 77
    rng = np.random.default_rng(12)
    om0 = np.full((1,) + m, -1j * xi) + 1e-3 * (rng.standard_normal((1,) + m) + 1j * rng.standard_normal((1,) + m))
    ps0 = 1e-3 * (rng.standard_normal(m) + 1j * rng.standard_normal(m))

    def step(noise):
        e = oracle_engine(homopolymer(m, N, n, xi=xi, lB=lB, dt=dt, scheme=SCHEME_SEMI_IMPLICIT, noise=noise, seed=seed,
                                      lambda_omega=(1.0,), lambda_psi=lam_psi))
        e.set_fields(om0, ps0)
        e.step(1)
        return e.get_field(FIELD_OMEGA)[0], e.get_field(FIELD_PSI)[0], e.info.volume

    (w1, p1, V), (w0, p0, _) = step(True), step(False)
    kx, ky, kz = kvecs(m, 0.5)
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    M = np.prod(m)
    dV = V / M
    j = np.arange(1, N)
    phi = np.exp(-k2 / 6)[..., None]
    pair = N + 2 * np.sum((N - j) * phi ** j, axis=-1)
    A_w = n * N / V / xi + np.exp(-k2) * (n / V) * pair
    A_p = k2 / (4 * np.pi * lB)
    for d, A, lam, field, drop in ((w1 - w0, A_w, 1.0, 0x400, False), (p1 - p0, A_p, lam_psi, 0x500, True)):
        P = 1 / (1 + dt * lam * A)
        sigma = np.sqrt(2 * lam * dt / dV)
        expected = sigma * np.sqrt(M) * np.fft.ifftn(si_noise_spectrum(seed, 0, field, m, P, drop))
        scale = np.abs(expected).max()
        assert np.abs(expected.imag).max() < 1e-12 * scale, field   # Hermitian: a real increment
        assert np.abs(d - expected).max() < 1e-10 * scale, field
        # unit spectrum: |g|^2 averaged over the grid is 1 (within sampling error)
        gk = np.fft.fftn(d) / (sigma * np.sqrt(M)) / np.sqrt(P * (1 + P) / 2)
        if drop:
            gk.flat[0] = 1.0
        assert abs(np.mean(np.abs(gk) ** 2) - 1) < 5 * np.sqrt(2 / M), field


def test_sipc_linear_field_is_crank_nicolson():
    """psi of a neutral chain without ions has the exactly linear drift A_p psi,
    A_p(k) = k^2 / (4 pi lB).  For it one step of the semi-implicit predictor-corrector
    is, mode by mode, psi <- (1 - a/2)/(1 + a/2) psi + P2 eta, a = dt lam A_p,
    P2 = 2 P / (1 + P) = 1 / (1 + a/2): the predictor drops out (its coefficient
    1 - P2 - P2 a/2 vanishes), so the predictor's own noise leaves no trace, and the
    stationary variance P2^2 2 lam dt / (1 - c^2) is exactly 1 / A_p.  Checked to round-off
    on both parts: the noise-off step against the closed-form factor, and the noise-on
    minus noise-off increment against sigma sqrt(M) IFFT[P2 g] with g restated from its
    definition and the exported Philox generator (k = 0 dropped).  A step that reused the
    predictor's noise factor, forgot the trapezoid's half, or kept a share of the predictor
    fails one of the two."""
    m, N, n, xi, dt, lB, lam_psi, seed = (16, 16, 16), 10, 6, 0.05, 0.2, 1.0, 0.7, 77
    rng = np.random.default_rng(13)
    om0 = np.full((1,) + m, -1j * xi) + 1e-3 * (rng.standard_normal((1,) + m) + 1j * rng.standard_normal((1,) + m))
    ps0 = 0.05 * (rng.standard_normal(m) + 1j * rng.standard_normal(m))
    ps0 -= ps0.mean()

    def step(noise):
        e = oracle_engine(homopolymer(m, N, n, xi=xi, lB=lB, dt=dt, scheme=SCHEME_SIPC, noise=noise, seed=seed,
                                      lambda_omega=(1.0,), lambda_psi=lam_psi))
        e.set_fields(om0, ps0)
        e.step(1)
        return e.get_field(FIELD_PSI)[0], e.info.volume

    (p1, V), (p0, _) = step(True), step(False)
    kx, ky, kz = kvecs(m, 0.5)
    A = (kx ** 2 + ky ** 2 + kz ** 2) / (4 * np.pi * lB)
    a = dt * lam_psi * A
    M = np.prod(m)
    expected0 = np.fft.ifftn((1 - a / 2) / (1 + a / 2) * np.fft.fftn(ps0))
    assert np.abs(p0 - expected0).max() < 1e-12 * np.abs(ps0).max()
    P = 1 / (1 + a)
    sigma = np.sqrt(2 * lam_psi * dt / (V / M))
    expected = sigma * np.sqrt(M) * np.fft.ifftn(si_noise_spectrum(seed, 0, 0x500, m, P, True, corrector=True))
    scale = np.abs(expected).max()
    assert np.abs(p1 - p0 - expected).max() < 1e-10 * scale


def ion_hot_spot_state(scheme, height, **kw):
    """A neutral chain with salt ions and psi = i height on a 6^3 block: the ion
    partition functions grow as exp(z Im psi_smeared) there while the polymer's Q (blind
    to psi) stays 1."""
    m = (16, 16, 16)
    s = homopolymer(m, 8, 20, lB=1.0, charge=0.0, salt=40, dt=0.01, scheme=scheme, noise=False, **kw)
    om = np.zeros((1,) + m, dtype=complex)
    ps = np.zeros(m, dtype=complex)
    ps[5:11, 5:11, 5:11] = 1j * height
    return s, om, ps


def test_chemical_potential_of_a_large_partition_function():
    """mu = -log Q with |Q| = 4e198: |Q|^2 is above the FP64 range, so |Q| must come
    from hypot, not from the square root of the sum of squares (which gave mu = -inf and
    H = -inf for a finite state)."""
    s, om, ps = ion_hot_spot_state(SCHEME_PC_REFERENCE, 800.0)
    e = oracle_engine(s)
    e.set_fields(om, ps)
    r = e.measure()
    Q = r["Q_ion"][0]
    assert np.isfinite(Q) and abs(Q) > 1e190
    assert abs(r["mu_ex_ion"][0] - (-np.log(abs(Q)) - 1j * np.angle(Q))) <= 1e-12 * abs(np.log(abs(Q)))
    assert np.isfinite(r["H"])


@pytest.mark.parametrize("scheme", [SCHEME_PC_REFERENCE, SCHEME_EM, SCHEME_SIPC])
def test_ion_partition_overflow_fails_the_step(scheme):
    """An ion Q that is not finite fails the step (as a polymer Q does): without the
    check the step would go on with zero ion densities."""
    s, om, ps = ion_hot_spot_state(scheme, 3000.0)
    e = oracle_engine(s)
    e.set_fields(om, ps)
    with pytest.raises(FtsError) as err:
        e.step(1, record_every=1)
    assert err.value.code == FTS_NONFINITE and "non-finite Q" in str(err.value)
    e.set_fields(om, ps / 10)
    r = e.step(1, record_every=1)
    assert np.isfinite(r[0]["H"])


def test_sipc_record_is_the_predictor_state():
    """The semi-implicit predictor-corrector records at its predictor state
    psi_bar = S_P[psi] + sigma sqrt(M) IFFT[Pn g] (the semi-implicit step, noise factor
    Pn = sqrt(P (1 + P) / 2)), as the reference's predictor-corrector records at its
    predictor.  For a neutral chain without ions psi decouples from omega, so two runs
    that differ only in psi's initial field record H values whose difference is the
    electrostatic term (1/8 pi lB) <psi_bar S_{k^2} psi_bar> of the two predicted fields,
    restated here in numpy with g from the exported Philox generator.  The cross term
    between S_P psi and the noise fixes the predictor's noise factor, which the linear
    update itself does not see."""
    m, N, n, xi, dt, lB, lam_psi, seed = (16, 16, 16), 10, 6, 0.05, 0.2, 1.0, 0.7, 91
    rng = np.random.default_rng(14)
    om0 = np.full((1,) + m, -1j * xi) + 1e-3 * (rng.standard_normal((1,) + m) + 1j * rng.standard_normal((1,) + m))
    psis = []
    for _ in range(2):
        p = 0.05 * (rng.standard_normal(m) + 1j * rng.standard_normal(m))
        psis.append(p - p.mean())
    H = []
    for ps in psis:
        e = oracle_engine(homopolymer(m, N, n, xi=xi, lB=lB, dt=dt, scheme=SCHEME_SIPC, noise=True, seed=seed,
                                      lambda_omega=(1.0,), lambda_psi=lam_psi))
        e.set_fields(om0, ps)
        H.append(e.step(1, record_every=1)[0]["H"])
        V = e.info.volume
    kx, ky, kz = kvecs(m, 0.5)
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    P = 1 / (1 + dt * lam_psi * k2 / (4 * np.pi * lB))
    M = np.prod(m)
    sigma = np.sqrt(2 * lam_psi * dt / (V / M))
    noise = sigma * np.sqrt(M) * np.fft.ifftn(si_noise_spectrum(seed, 0, 0x500, m, P, True))

    def h_psi(ps):
        pb = np.fft.ifftn(P * np.fft.fftn(ps)) + noise
        return np.mean(pb * np.fft.ifftn(k2 * np.fft.fftn(pb))) / (8 * np.pi * lB)

    expected = h_psi(psis[1]) - h_psi(psis[0])
    assert abs((H[1] - H[0]) - expected) < 1e-10 * max(abs(expected), 1e-3), (H[1] - H[0], expected)
    # the same with the corrector's factor P2 on the predictor's noise differs by far more
    # than the tolerance: the check sees the factor
    noise_p2 = sigma * np.sqrt(M) * np.fft.ifftn(si_noise_spectrum(seed, 0, 0x500, m, P, True, corrector=True))

    def h_psi_p2(ps):
        pb = np.fft.ifftn(P * np.fft.fftn(ps)) + noise_p2
        return np.mean(pb * np.fft.ifftn(k2 * np.fft.fftn(pb))) / (8 * np.pi * lB)

    assert abs((h_psi_p2(psis[1]) - h_psi_p2(psis[0])) - expected) > 1e-4 * abs(expected)

"""System descriptions for the engine's C ABI (fts_system), and the benchmark
configurations.

`System` holds the ctypes arrays a `fts_system` points into.  `config()` builds
the five workload configurations:

  C1  32^3, L=16, N=20,  z=-0.5/bead, salt 0.05, B=0.5, lB=1    (plain method)
  C2  64^3, L=32, N=50,  z=-0.5/bead, salt 0.10, B=0.5, lB=1    (precision form)
  C3  one replica of the batched form (n, N, salt, lB drawn per replica)
  C4  96^3, L=48, N=100, z=-0.5/bead, salt 0.05, B=0.5, lB=1    (headline)
  C5  256^3, L=128, N=200, z=-1 on odd beads, salt 0.02, B=1, lB=0.7

Common choices: grid spacing dL = 0.5 b, a = 1 b, bonds b = 1, one monomer type
(one active mode, gamma = i), excluded volume B mapped to xi = B * rho0 (the
Hamiltonian's (rho0 / 2 xi) <w^2> equals (1 / 2B) <w^2> only then), counterions
neutralising (z = +1), 1:1 salt merged with equal counterions
(ion.cu:117-127).  The reference only accepts integer counts (param.h:10,
ion.cu:13-15), so the chain and salt numbers are rounded:
n_chains = round(rho * V / N), num_salt = round(c_s * V).  The configurations do
not give mobilities; lambda_omega = lambda_psi = 1.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from .api import (CounterionDesc, PREC_FP64, SaltDesc, SCHEME_PC_REFERENCE, SpeciesDesc, SystemDesc)


class System:
    """An fts_system plus the arrays it points to."""

    def __init__(self, *, m, dL, a=1.0, species, monomer_types=1, xi=0.0, chi=None, lB=0.0,
                 counterion=None, salt=None, ions_enabled=None, default_smearing_length=0.0,
                 dt=0.01, lambda_omega=(1.0,), lambda_psi=1.0, noise=True,
                 scheme=SCHEME_PC_REFERENCE, precision=PREC_FP64, seed=0, stress_mask=0,
                 eigen_tol_rel=-1.0, fft_schedule=0, chain_symmetry=0):
        self._arrays = []
        d = SystemDesc()
        d.m1, d.m2, d.m3 = m
        d.dL, d.a = dL, a
        d.monomer_types = monomer_types
        d.xi = xi
        if chi is not None:
            arr = np.ascontiguousarray(np.asarray(chi, dtype=np.float64).reshape(-1))
            self._arrays.append(arr)
            d.chi = arr.ctypes.data_as(C.POINTER(C.c_double))
        d.eigen_tol_rel = eigen_tol_rel
        d.lB = lB
        sp_arr = (SpeciesDesc * len(species))()
        for i, sp in enumerate(species):
            types = np.ascontiguousarray(np.asarray(sp["types"], dtype=np.int32))
            charges = np.ascontiguousarray(np.asarray(sp["charges"], dtype=np.float64))
            self._arrays += [types, charges]
            sp_arr[i].num_chains = int(sp["num_chains"])
            sp_arr[i].N = len(types)
            sp_arr[i].types = types.ctypes.data_as(C.POINTER(C.c_int))
            sp_arr[i].charges = charges.ctypes.data_as(C.POINTER(C.c_double))
            if sp.get("bonds") is not None:
                bonds = np.ascontiguousarray(np.asarray(sp["bonds"], dtype=np.float64))
                self._arrays.append(bonds)
                sp_arr[i].bonds = bonds.ctypes.data_as(C.POINTER(C.c_double))
        self._arrays.append(sp_arr)
        d.n_species = len(species)
        d.species = C.cast(sp_arr, C.POINTER(SpeciesDesc))
        if ions_enabled is None:
            ions_enabled = counterion is not None or salt is not None
        d.ions_enabled = int(ions_enabled)
        d.default_smearing_length = default_smearing_length
        ci = CounterionDesc()
        if counterion is not None:
            ci.enabled = 1
            ci.auto_number = int(counterion.get("auto", True))
            ci.number = int(counterion.get("number", 0))
            ci.valence = float(counterion.get("valence", 1.0))
            ci.smearing_length = float(counterion.get("smearing_length", 0.0))
        d.counterions = ci
        sl = SaltDesc()
        if salt is not None:
            sl.enabled = 1
            sl.num_salt = int(salt["num_salt"])
            sl.cation_valence = float(salt.get("cation_valence", 1.0))
            sl.anion_valence = float(salt.get("anion_valence", -1.0))
            sl.cation_stoich = int(salt.get("cation_stoich", 1))
            sl.anion_stoich = int(salt.get("anion_stoich", 1))
            sl.smearing_length = float(salt.get("smearing_length", 0.0))
            sl.combine_identical_species = int(salt.get("combine", True))
        else:
            sl.combine_identical_species = 1
        d.salt = sl
        d.dt = dt
        lam = np.ascontiguousarray(np.asarray(lambda_omega, dtype=np.float64))
        self._arrays.append(lam)
        d.n_lambda_omega = lam.size
        d.lambda_omega = lam.ctypes.data_as(C.POINTER(C.c_double))
        d.lambda_psi = lambda_psi
        d.noise = int(noise)
        d.scheme = scheme
        d.precision = precision
        d.seed = seed
        d.stress_mask = stress_mask
        d.fft_schedule = fft_schedule
        d.chain_symmetry = chain_symmetry
        self.desc = d
        self.m = tuple(m)
        self.dL = dL

    @property
    def M(self):
        return self.m[0] * self.m[1] * self.m[2]


CONFIGS = {
    1: dict(n=32, N=20, rho=0.2, salt=0.05, B=0.5, lB=1.0, charge="uniform", z=-0.5, dt=0.01),
    2: dict(n=64, N=50, rho=0.2, salt=0.10, B=0.5, lB=1.0, charge="uniform", z=-0.5, dt=0.05),
    4: dict(n=96, N=100, rho=0.2, salt=0.05, B=0.5, lB=1.0, charge="uniform", z=-0.5, dt=0.01),
    5: dict(n=256, N=200, rho=0.2, salt=0.02, B=1.0, lB=0.7, charge="odd", z=-1.0, dt=0.01),
}


def chemistry(n, N, rho, salt, B, lB, charge, z, dL=0.5):
    """Derived counts for an n^3 box: returns (n_chains, num_salt, rho0, xi, charges).
    n_chains = round(rho V / N), raised by one where the chains' total charge would
    not be a whole number (auto-neutralising counterions need an integer count,
    ion.cu:11-19): half charges on an odd number of beads in an odd number of chains."""
    V = (dL * n) ** 3
    n_chains = int(round(rho * V / N))
    num_salt = int(round(salt * V))
    if charge == "uniform":
        charges = np.full(N, z)
    else:  # charge z on every second bead (odd index)
        charges = np.where(np.arange(N) % 2 == 1, z, 0.0)
    if lB > 0 and abs(n_chains * charges.sum() - round(n_chains * charges.sum())) > 1e-9:
        n_chains += 1
    rho0 = n_chains * N / V
    return n_chains, num_salt, rho0, B * rho0, charges


def make_system(n, N, rho=0.2, salt=0.05, B=0.5, lB=1.0, charge="uniform", z=-0.5, dL=0.5, **kw):
    n_chains, num_salt, rho0, xi, charges = chemistry(n, N, rho, salt, B, lB, charge, z, dL)
    species = [dict(num_chains=n_chains, types=np.zeros(N, dtype=np.int32), charges=charges)]
    return System(m=(n, n, n), dL=dL, a=kw.pop("a", 1.0), species=species, monomer_types=1, xi=xi, lB=lB,
                  counterion=dict(auto=True, valence=1.0) if lB > 0 else None,
                  salt=dict(num_salt=num_salt) if num_salt > 0 else None, **kw)


def config(cfg, **kw):
    """The system of configuration `cfg` (1, 2, 4 or 5); keyword overrides pass to System."""
    c = dict(CONFIGS[cfg])
    for key in ("n", "N", "rho", "salt", "B", "lB", "charge", "z", "dt"):
        if key in kw:
            c[key] = kw.pop(key)
    dt = c.pop("dt")
    return make_system(dt=dt, **c, **kw)


def initial_fields(system, n_omega=1, sigma=1e-3, seed=None, cfg=0):
    """w, psi = 0 + complex Gaussian noise (Re, Im iid N(0, sigma^2)), seeded."""
    rng = np.random.default_rng(0x5EED0000 + cfg if seed is None else seed)
    shape = (n_omega,) + tuple(system.m)
    om = sigma * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    ps = sigma * (rng.standard_normal(system.m) + 1j * rng.standard_normal(system.m))
    return om, ps


def replica_params(r, base_seed=0xBA7C4000):
    """Batched-form replica r: n in {32,48,64,96}, L = 0.5 n, N ~ U[20,100],
    salt log-uniform in [1e-3, 0.5], lB ~ U[0.5, 2], seed base + r."""
    rng = np.random.default_rng(base_seed + r)
    n = int(rng.choice([32, 48, 64, 96]))
    N = int(rng.integers(20, 101))
    salt = float(math.exp(rng.uniform(math.log(1e-3), math.log(0.5))))
    lB = float(rng.uniform(0.5, 2.0))
    return dict(n=n, N=N, rho=0.2, salt=salt, B=0.5, lB=lB, charge="uniform", z=-0.5, seed=base_seed + r)

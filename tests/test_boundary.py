"""CPU checks of the C-ABI boundary and host logic (no GPU needed).

- libfts_hip.so loads and exports every entry point include/fts.h declares, and the
  oracle exports the same set under the fto_ prefix;
- fts_create rejects invalid systems with FTS_INVALID and a message before it
  touches a device (the reference returns false with "error: ..." on stderr,
  errHandle.cu:5-9);
- the ctypes structures have the C layout of the header;
- bench.py's integrated autocorrelation time on AR(1) series with known tau, its
  pooled and jackknife forms on a rare-excursion process, and the shape of its line.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from ftspolysol_amd.api import FTS_INVALID, FtsError, Engine, Info, Observables, SystemDesc, load_hip_library
from ftspolysol_amd.system import System, make_system
from oracle_lib import oracle_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fts.h")).read()
    return sorted(set(re.findall(r"\b(fts_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("fts_create", "fts_destroy", "fts_last_error", "fts_set_fields", "fts_get_field",
                     "fts_compute_densities", "fts_measure", "fts_step"):
        assert required in names


def test_hip_library_exports_every_declared_symbol():
    lib = load_hip_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_oracle_exports_the_same_interface():
    lib = oracle_library()
    # timing and replica batches are device-scheduling concepts: the oracle checks a
    # batch replica by replica through the single-system interface.  The front end
    # (decks, .rf files, resize) is host I/O that every caller shares; its own tests
    # pin it, and there is no device arithmetic for the oracle to check.  The no-alloc
    # bead step takes device arrays; its test compares it with fts_bead_step.
    engine_only = {"fts_set_timing", "fts_get_timing", "fts_field_resize", "fts_bead_step_noalloc"}
    missing = [n for n in declared_functions()
               if n not in engine_only and not n.startswith(("fts_batch_", "fts_deck_", "fts_rf_"))
               and not hasattr(lib, "fto_" + n[4:])]
    assert not missing, missing


def test_struct_layout(tmp_path):
    """The ctypes mirrors have the C compiler's sizes and offsets for include/fts.h."""
    from ftspolysol_amd.api import BatchInfo, CounterionDesc, LrkStats, SaltDesc, SpeciesDesc, Timing
    structs = {"fts_species_desc": SpeciesDesc, "fts_counterion_desc": CounterionDesc, "fts_salt_desc": SaltDesc,
               "fts_system": SystemDesc, "fts_observables": Observables, "fts_info": Info, "fts_timing": Timing,
               "fts_batch_info": BatchInfo, "fts_lrk_stats": LrkStats}
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "fts.h"', "int main(void) {"]
    for cname, py in structs.items():
        src.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            src.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    src.append("return 0; }")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in out:
        if not line:
            continue
        cname, what, val = line.split()
        py = structs[cname]
        got = ctypes.sizeof(py) if what == "size" else getattr(py, what).offset
        assert got == int(val), (cname, what, got, val)


def bad_create(lib_prefix, system):
    lib = load_hip_library() if lib_prefix == "fts_" else oracle_library()
    with pytest.raises(FtsError) as ei:
        Engine(lib, lib_prefix, system)
    return ei.value


def test_hip_rejects_unsupported_mesh():
    err = bad_create("fts_", make_system(10, 4))
    assert err.code == FTS_INVALID and "mesh" in str(err)
    err = bad_create("fts_", System(m=(32, 32, 20), dL=0.5, species=[dict(num_chains=1, types=[0], charges=[0.0])], xi=0.1))
    assert err.code == FTS_INVALID
    # the plane schedule needs m2 == m3 and the plane in LDS: FP64 128^2 does not fit
    from ftspolysol_amd.api import SCHED_PLANE
    err = bad_create("fts_", System(m=(128, 128, 128), dL=0.5, species=[dict(num_chains=1, types=[0], charges=[0.0])], xi=0.1,
                                    fft_schedule=SCHED_PLANE))
    assert err.code == FTS_INVALID and "plane" in str(err)


@pytest.mark.parametrize("prefix", ["fts_", "fto_"])
def test_rejects_invalid_physics(prefix):
    # salt that is not neutral (ion.cu:26-30)
    s = System(m=(16, 16, 16), dL=0.5, species=[dict(num_chains=2, types=[0] * 4, charges=[0.0] * 4)], xi=0.1, lB=1.0,
               salt=dict(num_salt=3, cation_valence=2.0, anion_valence=-1.0))
    assert bad_create(prefix, s).code == FTS_INVALID
    # counterions cannot neutralise a fractional total charge (ion.cu:13-15)
    s = System(m=(16, 16, 16), dL=0.5, species=[dict(num_chains=1, types=[0] * 3, charges=[-0.5] * 3)], xi=0.1, lB=1.0,
               counterion=dict(auto=True, valence=1.0))
    assert "counterion" in str(bad_create(prefix, s))
    # chi must be symmetric with a zero diagonal (interactionMode.cu:36-48)
    s = System(m=(16, 16, 16), dL=0.5, species=[dict(num_chains=1, types=[0, 1], charges=[0.0, 0.0])], monomer_types=2,
               xi=0.1, chi=[[0.0, 0.2], [0.3, 0.0]])
    assert "symmetric" in str(bad_create(prefix, s))
    # bead type out of range
    s = System(m=(16, 16, 16), dL=0.5, species=[dict(num_chains=1, types=[0, 3], charges=[0.0, 0.0])], xi=0.1)
    assert bad_create(prefix, s).code == FTS_INVALID


def test_tau_int_ar1():
    import bench
    rng = np.random.default_rng(0)
    for rho in (0.0, 0.5, 0.8):
        n = 200000
        x = np.empty(n)
        x[0] = 0.0
        e = rng.standard_normal(n)
        for i in range(1, n):
            x[i] = rho * x[i - 1] + e[i]
        tau, _ = bench.tau_int(x)
        exact = 0.5 * (1 + rho) / (1 - rho)
        assert abs(tau - exact) < 0.08 * exact + 0.02, (rho, tau, exact)


def test_tau_int_pooled_over_replicas_with_rare_excursions():
    """Replicas of white noise plus rare large excursions that decay as 0.9^t (events at
    rate 1e-4 per step, height 134: a shot-noise process holding 90 % of the variance).
    Its tau_int is exact: (1 * 0.5 + v_s tau_s) / (1 + v_s), v_s = p A^2 / (1 - phi^2),
    tau_s = (1 + phi) / (2 (1 - phi)) = 8.64.  Most replicas hold no event, so a tau that
    averages per-replica normalised autocorrelations comes out near 2.5; the pooled one
    and block averaging must find the exact value."""
    import bench
    rng = np.random.default_rng(4)
    R, N, p, A, phi = 64, 4000, 1e-4, 134.0, 0.9
    x = rng.standard_normal((R, N))
    jumps = (rng.random((R, N)) < p) * A
    s = np.zeros((R, N))
    for i in range(1, N):
        s[:, i] = phi * s[:, i - 1] + jumps[:, i]
    x += s
    vs, ts = p * A * A / (1 - phi ** 2), 0.5 * (1 + phi) / (1 - phi)
    exact = (0.5 + vs * ts) / (1 + vs)
    tau, _ = bench.tau_int(list(x))
    assert abs(tau - exact) < 0.1 * exact, (tau, exact)
    assert abs(bench.tau_block(x, 200) - exact) < 0.15 * exact


def test_tau_jackknife_matches_the_spread_over_realisations():
    """The replica-jackknife error of tau_int against the spread of tau_int over 30
    independent realisations of the rare-excursion process above (32 replicas x 2000
    samples, about 6 excursions per realisation): the mean jackknife error must agree
    with the empirical standard deviation within 30 %."""
    import bench
    taus, errs = [], []
    for seed in range(100, 130):
        rng = np.random.default_rng(seed)
        R, N, p, A, phi = 32, 2000, 1e-4, 134.0, 0.9
        x = rng.standard_normal((R, N))
        jumps = (rng.random((R, N)) < p) * A
        s = np.zeros((R, N))
        for i in range(1, N):
            s[:, i] = phi * s[:, i - 1] + jumps[:, i]
        x += s
        taus.append(bench.tau_int(list(x))[0])
        errs.append(bench.tau_jackknife(x))
    spread = np.std(taus)
    assert abs(np.mean(errs) - spread) < 0.3 * spread, (np.mean(errs), spread)


def test_long_record_estimate_reads_the_committed_records():
    """The bench line's long_record_estimate: this run's rates over twice the tau_int of
    the longest committed record of each leg."""
    import bench
    est = bench.long_record_estimate({"replica_steps_per_s": 400.0}, {"replica_steps_per_s": 100.0},
                                     {"replica_steps_per_s": 200.0}, "sipc_learned", 0.01, 0.01)
    assert est is not None and est["record_steps"]["fast"] >= 10000
    tf, tp = est["tau_int"]["fast"], est["tau_int"]["plain_fp64"]
    assert abs(est["samples_per_s"]["fast"] - 400.0 / (2 * tf)) < 1e-12
    assert abs(est["speedup_vs_plain_fp64"]["ratio"] - 4.0 * tp / tf) < 1e-12
    assert abs(est["speedup_vs_plain_fp64_batched"]["ratio"] - 2.0 * tp / tf) < 1e-12
    pc = est["fp32_pc_committed"]
    assert pc is not None and abs(pc["speedup_vs_plain_fp64"]["ratio"] - pc["samples_per_s"] / est["samples_per_s"]["plain_fp64"]) < 1e-12
    assert bench.long_record_estimate({"replica_steps_per_s": 1.0}, {"replica_steps_per_s": 1.0}, {}, "sipc_learned", 0.5, 0.01) is None


def test_bench_refuses_more_than_one_gpu():
    # the engine runs a system (or a replica batch) on one GPU; bench.py says so for --gpus 2
    # and exits before it touches a device
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and not r.stdout.strip()


def test_bench_reports_a_failed_leg(monkeypatch):
    """A leg whose replica exhausts its non-finite restarts (the library's FTS_NONFINITE
    error out of fts_batch_step) is reported with the reason and no statistics, and the
    batch is released; the run goes on.  A stand-in batch raises the error the binding
    raises, so no device is touched."""
    import bench
    from ftspolysol_amd.api import FTS_NONFINITE, FtsError
    from ftspolysol_amd.system import make_system

    closed = []

    class Replica:
        def set_fields(self, *a):
            pass

        def set_restart(self, *a):
            pass

    class FailingBatch:
        def __init__(self, systems):
            self.replicas = [Replica() for _ in systems]

        def set_graph(self, level):
            pass

        def step(self, n, record_every=0):
            raise FtsError(FTS_NONFINITE, "replica 1: updated fields have non-finite values at step 252; "
                                          "maximum non-finite restarts (50) exceeded")

        def close(self):
            closed.append(True)

    monkeypatch.setattr(bench, "hip_batch", FailingBatch)
    systems = [make_system(16, 4, dt=0.04, seed=r) for r in range(2)]
    leg = bench.run_leg(systems, "stand-in", 5, 20, 100)
    assert "maximum non-finite restarts" in leg["failed"]
    assert leg["samples_per_s"] is None and leg["tau_int"] is None and leg["H_mean"] is None
    assert leg["dt"] == 0.04 and leg["replicas"] == 2
    assert closed == [True]
    assert bench.speedup(leg, {"samples_per_s": 1.0, "tau_int": 1.0, "tau_int_err": 0.1}) is None
    assert set(bench.leg_summary(leg)) >= {"failed", "dt", "replicas"}


def test_bench_line_contract(monkeypatch, capsys):
    """The headline line's shape, on the CPU: stand-in batches return records of H from
    an AR(1) process (so every leg has a finite tau_int) and timing counters; the FP32/FP64
    same-state comparison and the CPU baseline are stubbed.  The one JSON line must carry
    the keys a reader of the line relies on, with the types they need."""
    import json
    import sys
    import bench

    class EInfo:
        restarts = 0

    class Replica:
        def __init__(self, r):
            self.rng = np.random.default_rng(100 + r)
            self.x = 0.0
            self.judged = 0

        def set_fields(self, *a):
            pass

        def set_restart(self, *a):
            pass

        def lrk_set_network(self, *a):
            pass

        def get_info(self):
            return EInfo()

        def lrk_get_stats(self):
            j = self.judged
            return {"judged": j, "learned_steps": j, "fallback_steps": 0, "sum_rel_residual_learned": [0.8 * j, 0.8 * j],
                    "sum_rel_residual_debye": [0.9 * j, 0.9 * j]}

        def get_field(self, fid):
            return np.zeros((1, 4, 4, 4), dtype=complex)

        def record(self):
            self.x = 0.6 * self.x + self.rng.standard_normal()
            self.judged += 1
            return {"H": complex(7.9 + 0.3 * self.x, 0.0)}

    class BInfo:
        group_schedule = [3]
        group_tune_ms = [[1.0, 2.0]]

    class Batch:
        def __init__(self, systems):
            self.replicas = [Replica(r) for r in range(len(systems))]
            self.info = BInfo()
            self.steps = 0

        def set_graph(self, level):
            pass

        def set_streams(self, n):
            pass

        def set_timing(self, on):
            self.steps = 0

        def get_timing(self, g):
            return {"sweep_ms": 0.1 * self.steps, "step_launches": 99 * self.steps, "algorithmic_bytes": 2.2e8 * 99 * self.steps,
                    "sweeps": self.steps, "bead_steps": 99 * self.steps}

        def step(self, n, record_every=0):
            self.steps += n
            return [[rep.record() for _ in range(n)] if record_every else [] for rep in self.replicas]

        def close(self):
            pass

    monkeypatch.setattr(bench, "hip_batch", Batch)
    monkeypatch.setattr(bench, "same_state_errors", lambda wl, f: {"H": 1e-8, "mu_ex_poly": 1e-7, "stress": 1e-7, "p_iso": 1e-7,
                                                                     "H_fp64": 7.9})
    monkeypatch.setattr(bench, "cpu_baseline", lambda *a: {"value": 0.1, "unit": "decorrelated CL samples/s", "cores": 1,
                                                           "kind": "port", "sample": "stand-in"})
    monkeypatch.setattr(sys, "argv", ["bench.py", "--steps", "20", "--warmup", "5", "--equilibration", "10", "--record-steps", "2000",
                                      "--plain-record-steps", "2000", "--replicas", "4", "--plain-replicas", "4"])
    bench.main()
    line = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
                "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in line, key
    assert line["n_gpus"] == 1 and line["steps"] == 20 and line["warmup"] == 5 and line["higher_is_better"] is True
    assert isinstance(line["value"], float) and line["value"] > 0 and line["vs_baseline"] is None
    assert "workload" in line["config"]
    r = line["roofline"]
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-12
    assert set(line["cpu_baseline"]) >= {"value", "unit", "cores", "kind", "sample"}
    for leg in ("fast", "si_debye", "plain_fp64_batched", "plain_fp64"):
        assert line[leg]["tau_int"] > 0 and line[leg]["tau_int_err"] > 0, leg

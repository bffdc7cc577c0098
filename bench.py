"""Headline benchmark: decorrelated complex-Langevin samples/s at 96^3, N = 100.

Workload (one GPU): 96^3 grid, L = 48 b, discrete Gaussian chains N = 100,
b = 1, charge -0.5 on every bead, counterions, 1:1 salt 0.05 b^-3, bead density
0.2 b^-3 (221 chains), excluded volume B = 0.5 b^3, lB = 1 b, a = 1 b,
lambda = 1.  Synthetic initial fields (w, psi = 1e-3 complex Gaussian).

Legs (engines stepping the workload; every batch carries all its replicas in each
bead-step launch, each replica with its own seed and initial fields; replica 0 of
every leg has the same seed and fields):
  fast        FP32 + learned: R replicas with complex-FP32 propagators and FFTs,
              per-bead renormalisation, FP64 reductions and field accumulation,
              stepping the semi-implicit predictor-corrector with the learned
              linear-response kernel (FTS_SCHEME_SIPC_LEARNED; --fast-scheme si_learned
              for the one-stage semi-implicit step; network trained by
              ftspolysol_amd.train_lrk on 48^3 data, lrk_c4_48cubed.npy), judged
              every step by the exact force, falling back to the Debye kernel when
              its excess force residual passes 1e-3, at --dt-fast (0.01, the
              plain dt).  `value` is this leg.
  si_debye    the same batch, scheme and seeds with the Debye kernel only
              (FTS_SCHEME_SIPC, or FTS_SCHEME_SEMI_IMPLICIT) at --dt-debye (0.01): what
              the learned kernel gains or loses.  --learned-vs-debye runs these two legs
              alone; a leg whose replica exhausts its non-finite restarts is reported as
              failed.
  plain_fp64_batched  FP64, the reference's predictor-corrector
              (complexLangevin.cu:669-698), a batch of --plain-replicas: the plain
              engine at its own best throughput, and the plain scheme's tau_int and
              <H>.
  plain_fp64  one FP64 predictor-corrector system, as the reference runs it (its
              rate; its tau_int and <H> are the batched leg's, the same dynamics).
  fp32_pc     (--with-fp32-pc) FP32 predictor-corrector batch.

Each leg: max(warmup, 200) untimed equilibration steps; then exactly `steps` timed
steps (bracketed by device synchronisation, recording H every step as the
reference measures its observables every step, complexLangevin.cu:684), which set
the leg's rate; then untimed steps that extend the record of H to --record-steps
(2000 steps for the FP32 legs and --plain-record-steps 2000 for the FP64 batch) whatever
`steps` is.  tau_int and <H> come from that record, so they do not depend on
--steps.  A leg whose record is shorter than 20 Sokal windows, or whose
tau has a Sokal error above 10 %, reports samples_per_s null.

Decorrelated samples/s = (replica CL steps/s) / (2 tau_int), tau_int by Sokal's
automatic window (c = 6) on the pooled autocorrelation of Re H over the replicas,
with Sokal's error tau sqrt(2 (2W + 1) / n).  Every replica keeps a device checkpoint
every 50 steps; a replica whose step turns non-finite is restored from it with a new
noise key (the repeated steps count in the elapsed time); restarts are reported.

<H> error: `same_state_rel_err` (the precision error) measures H, mu and the
stress tensor with an FP32 engine and an FP64 engine on the same fields (fast
replica 0's final state); `H_err_vs_plain` compares the time averages of the fast
batch and the plain FP64 batch (statistical once trajectories decorrelate, plus
the time-step bias of the semi-implicit scheme against the predictor-corrector;
`H_err_sigma` is the combined standard error).

roofline: the dominant kernels are the propagator bead-step passes (one pencil
pass and one plane pass per bead index, or the four line passes, every replica's
job in each launch).  Algorithmic bytes per bead step = read q_s + write q_{s+1}
= 2 c M (c = 8 B for complex FP32), summed over the jobs of a launch group and
divided by the group's average duration from HIP events on the batch stream over
the timed steps.  `copy_achievable` is a 1 GiB device-to-device hipMemcpy timed on
the same GPU; `traffic_rate` the measured HBM traffic of a launch group (rocprofv3
FETCH_SIZE + WRITE_SIZE, committed under profiles/) over its duration.

cpu_baseline: the FP64 CPU restatement (oracle/) running the plain scheme on the
host cores for a bounded sample of steps of the same workload.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from ftspolysol_amd.api import (FIELD_OMEGA, FIELD_PSI, Engine, FtsError, PREC_FP32, PREC_FP64,  # noqa: E402
                                SCHEME_PC_REFERENCE, SCHEME_SEMI_IMPLICIT, SCHEME_SI_LEARNED, SCHEME_SIPC, SCHEME_SIPC_LEARNED,
                                hip_batch, hip_engine)
from ftspolysol_amd.lrk import load_params  # noqa: E402
from ftspolysol_amd.system import initial_fields, make_system  # noqa: E402

HBM_PEAK_GBS = 8000.0
LRK_PARAMS = os.path.join(ROOT, "ftspolysol_amd", "lrk_c4_48cubed.npy")
LRK_PARAMS_C5 = os.path.join(ROOT, "ftspolysol_amd", "lrk_c5_48cubed.npy")


def tau_int(x, c=6.0):
    """Integrated autocorrelation time (in samples) with Sokal's automatic window.
    x: one series, or a list of equally long series of the same stationary process
    (the replicas of a leg).  The autocorrelation is pooled: covariances about the grand
    mean summed over the series and normalised by the pooled variance.  (Normalising each
    series by its own variance and averaging lets a series that holds one large excursion
    of H count as nearly white, so that rare slow excursions, which carry most of the
    variance of H here, drop out of tau: 2-3x too small on these records.)  The pooled
    estimate agrees with block averaging (tau_block)."""
    series = [np.asarray(x, dtype=np.float64)] if np.ndim(x) == 1 else [np.asarray(v, dtype=np.float64) for v in x]
    n = series[0].size
    if n < 8:
        return float("nan"), 0
    X = np.stack(series)
    Y = X - X.mean()
    f = np.fft.rfft(Y, 2 * n, axis=1)
    cov = np.fft.irfft(f * np.conj(f), axis=1)[:, :n].sum(axis=0) / (X.shape[0] * (n - np.arange(n)))
    if not cov[0] > 0:
        return 0.5, 0
    rho = cov / cov[0]
    tau = 0.5
    for w in range(1, n):
        tau += rho[w]
        if w >= c * tau:
            return float(tau), w
    return float(tau), n - 1


def tau_jackknife(x):
    """Standard error of tau_int (pooled) from the delete-one-replica jackknife: it sees
    how much of tau rests on the excursions of a few replicas, which Sokal's error (a
    property of the window alone) does not.  None for fewer than 4 replicas."""
    X = np.atleast_2d(np.asarray(x, dtype=np.float64))
    R = X.shape[0]
    if R < 4:
        return None
    t = np.array([tau_int(list(np.delete(X, r, axis=0)))[0] for r in range(R)])
    return float(np.sqrt((R - 1) / R * np.sum((t - t.mean()) ** 2)))


def tau_block(x, block=200):
    """tau_int from block means: Var(block mean) B / (2 Var), blocks of `block` samples of
    every series (a check of tau_int that assumes no window)."""
    X = np.atleast_2d(np.asarray(x, dtype=np.float64))
    nb = X.shape[1] // block
    if nb < 2 or not X.var() > 0:
        return None
    bm = X[:, :nb * block].reshape(X.shape[0], nb, block).mean(axis=2)
    return float(bm.var(ddof=1) * block / X.var() / 2)


MIN_EQUILIBRATION = 200


def lrk_totals(stats):
    return {"judged": sum(x["judged"] for x in stats), "learned": sum(x["learned_steps"] for x in stats),
            "fallback": sum(x["fallback_steps"] for x in stats),
            "res_learned": [sum(x["sum_rel_residual_learned"][c] for x in stats) for c in range(2)],
            "res_debye": [sum(x["sum_rel_residual_debye"][c] for x in stats) for c in range(2)]}


def run_leg(systems, label, *args, **kw):
    """One leg (run_leg_steps); a leg whose replica exhausts its non-finite restarts is
    reported as failed, with the reason, instead of ending the run."""
    try:
        out = run_leg_steps(systems, label, *args, **kw)
    except FtsError as exc:
        print(f"[bench] {label}: failed: {exc}", file=sys.stderr, flush=True)
        out = {"replicas": len(systems), "failed": str(exc), "replica_steps_per_s": None, "samples_per_s": None,
               "tau_int": None, "tau_int_err": None, "H_mean": None, "H_stderr": None, "lrk": None}
    out["dt"] = systems[0].desc.dt
    return out


def run_leg_steps(systems, label, warmup, steps, record_steps, ckpt_every=50, max_restarts=50, lrk=None,
                  lrk_threshold=1e-3, stats=True, min_equilibration=MIN_EQUILIBRATION, graph_level=2, streams=-1):
    """One leg: equilibrate max(warmup, MIN_EQUILIBRATION) untimed steps, time exactly
    `steps` steps (recording H every step), then extend the record of H untimed to
    `record_steps` steps when it is shorter.  Returns the rate of the timed steps and the
    statistics of the record."""
    b = hip_batch(systems)
    try:
        return leg_steps(b, systems, label, warmup, steps, record_steps, ckpt_every, max_restarts, lrk, lrk_threshold,
                         stats, min_equilibration, graph_level, streams)
    finally:
        b.close()


def leg_steps(b, systems, label, warmup, steps, record_steps, ckpt_every, max_restarts, lrk, lrk_threshold, stats,
              min_equilibration, graph_level, streams=-1):
    b.set_graph(graph_level)
    if streams >= 0:
        b.set_streams(streams)
    for r, (e, s) in enumerate(zip(b.replicas, systems)):
        om, ps = initial_fields(s, sigma=1e-3, seed=0x5EED0004 + r)
        e.set_fields(om, ps)
        e.set_restart(ckpt_every, max_restarts)
        if lrk is not None:
            e.lrk_set_network(lrk, lrk_threshold)
    R = len(systems)
    equil = max(warmup, min_equilibration)
    t0 = time.perf_counter()
    done = 0
    while done < equil:
        k = min(100, equil - done)
        b.step(k)
        done += k
    print(f"[bench] {label}: {equil} equilibration steps in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    lrk0 = lrk_totals([e.lrk_get_stats() for e in b.replicas]) if lrk is not None else None
    restarts0 = sum(e.get_info().restarts for e in b.replicas)
    H = [[] for _ in range(R)]
    # timed region: exactly `steps` steps; fts_batch_step returns after its last device
    # synchronisation, so the wall clock brackets the device work
    b.set_timing(True)
    t0 = time.perf_counter()
    recs = b.step(steps, record_every=1)
    elapsed = time.perf_counter() - t0
    timing = b.get_timing(0)   # before set_timing, which clears the accumulated events
    b.set_timing(False)
    for r in range(R):
        H[r].extend(x["H"] for x in recs[r])
    print(f"[bench] {label}: {steps} timed steps in {elapsed:.2f} s", file=sys.stderr, flush=True)
    done = steps
    t1 = time.perf_counter()
    while stats and done < record_steps:
        k = min(250, record_steps - done)
        recs = b.step(k, record_every=1)
        for r in range(R):
            H[r].extend(x["H"] for x in recs[r])
        done += k
        print(f"[bench] {label}: record {done}/{record_steps} steps", file=sys.stderr, flush=True)
    record_s = time.perf_counter() - t1
    restarts = sum(e.get_info().restarts for e in b.replicas) - restarts0
    lrk_stats = None
    if lrk is not None:
        t = lrk_totals([e.lrk_get_stats() for e in b.replicas])
        judged = max(1, t["judged"] - lrk0["judged"])
        lrk_stats = {"judged_steps": int(t["judged"] - lrk0["judged"]),
                     "learned_steps": int(t["learned"] - lrk0["learned"]),
                     "fallback_steps": int(t["fallback"] - lrk0["fallback"]),
                     "learned_fraction": (t["learned"] - lrk0["learned"]) / judged,
                     "rel_force_residual_learned": [(t["res_learned"][c] - lrk0["res_learned"][c]) / judged for c in range(2)],
                     "rel_force_residual_debye": [(t["res_debye"][c] - lrk0["res_debye"][c]) / judged for c in range(2)]}
    final = (b.replicas[0].get_field(FIELD_OMEGA), b.replicas[0].get_field(FIELD_PSI)[0])
    sched = b.info.group_schedule[0]
    tune = list(b.info.group_tune_ms[0])
    out = dict(replicas=R, replica_steps_per_s=R * steps / elapsed, ms_per_step=1e3 * elapsed / steps,
               equilibration_steps=equil, timed_steps=steps, timing=timing, elapsed=elapsed, schedule=sched,
               tune_ms=tune, restarts=int(restarts), final=final, lrk=lrk_stats)
    if not stats:
        return out
    H = np.array(H)
    tau, window = tau_int(H.real)
    n = H.shape[1]
    # Sokal's estimate of the statistical error of tau_int: var = 2 (2W + 1) / n tau^2,
    # n the number of samples pooled
    tau_err_sokal = float(tau * np.sqrt(2.0 * (2 * window + 1) / H.size))
    jk = tau_jackknife(H.real)
    tau_err = max(tau_err_sokal, jk) if jk is not None else tau_err_sokal
    # samples/s needs a record of at least 20 Sokal windows and Sokal's relative error of
    # tau under 10 %; tau_int_err is the larger of Sokal's and the replica jackknife's
    valid = n >= 20 * max(window, 1) and tau_err_sokal <= 0.1 * tau
    out.update(record_steps=n, tau_record_samples=int(H.size), record_replica_steps_per_s=R * (n - steps) / record_s
               if n > steps and record_s > 0 else None,
               tau_int=tau, window=window, tau_int_err=tau_err, tau_int_err_sokal=tau_err_sokal, tau_int_err_jackknife=jk,
               tau_int_each=[round(float(tau_int(H.real[r:r + 1])[0]), 3) for r in range(R)],
               tau_block=tau_block(H.real, 200),
               samples_per_s=(out["replica_steps_per_s"] / (2 * tau)) if valid else None,
               H_mean=float(H.real.mean()), H_imag_mean=float(H.imag.mean()), H_std=float(H.real.std()),
               H_median=float(np.median(H.real)),
               H_stderr=float(H.real.std(ddof=1) * np.sqrt(2 * tau / H.size)))
    return out


def copy_bandwidth(gib=1.0, reps=10):
    """Achievable HBM bandwidth on this GPU: a device-to-device hipMemcpy of `gib` GiB
    (read + write bytes over the copy time, HIP events), through the HIP runtime the
    engine library already loaded (a second runtime in the process does not see the GPU)."""
    hip = ctypes.CDLL("libamdhip64.so.7" if os.path.exists("/opt/rocm/lib/libamdhip64.so.7") else "libamdhip64.so")
    nbytes = int(gib * (1 << 30))
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    ms = ctypes.c_float(0.0)

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc})")
    ok(hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)), "hipMalloc")
    try:
        ok(hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)), "hipMalloc")
        ok(hip.hipMemset(a, 0, ctypes.c_size_t(nbytes)), "hipMemset")
        ok(hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3), "hipMemcpy")  # 3 = device to device
        ok(hip.hipEventCreate(ctypes.byref(e0)), "hipEventCreate")
        ok(hip.hipEventCreate(ctypes.byref(e1)), "hipEventCreate")
        ok(hip.hipEventRecord(e0, None), "hipEventRecord")
        for _ in range(reps):
            ok(hip.hipMemcpyAsync(b, a, ctypes.c_size_t(nbytes), 3, None), "hipMemcpyAsync")
        ok(hip.hipEventRecord(e1, None), "hipEventRecord")
        ok(hip.hipEventSynchronize(e1), "hipEventSynchronize")
        ok(hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1), "hipEventElapsedTime")
    finally:
        for ev in (e0, e1):
            if ev.value:
                hip.hipEventDestroy(ev)
        for ptr in (a, b):
            if ptr.value:
                hip.hipFree(ptr)
    return 2.0 * nbytes / (ms.value / reps * 1e-3) / 1e9


def same_state_errors(wl, fields):
    """FP32 engine vs FP64 engine observables on the same fields."""
    out = {}
    e32 = hip_engine(make_system(**wl, precision=PREC_FP32, stress_mask=0x7F))
    e64 = hip_engine(make_system(**wl, precision=PREC_FP64, stress_mask=0x7F))
    for e in (e32, e64):
        e.set_fields(*fields)
    a, b = e32.measure(), e64.measure()
    out["H"] = abs(a["H"] - b["H"]) / abs(b["H"])
    out["mu_ex_poly"] = float(np.abs(a["mu_ex_poly"] - b["mu_ex_poly"]).max() / max(1.0, np.abs(b["mu_ex_poly"]).max()))
    out["stress"] = float(np.abs(a["stress"] - b["stress"]).max() / np.abs(b["stress"]).max())
    out["p_iso"] = abs(a["p_iso"] - b["p_iso"]) / abs(b["p_iso"])
    out["H_fp64"] = b["H"].real
    e32.close()
    e64.close()
    return out


def cpu_baseline(system, n_steps, tau_plain, scheme_name):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "libfts_oracle.so"))
    eng = Engine(lib, "fto_", system)
    om, ps = initial_fields(system, sigma=1e-3, seed=0x5EED0004)
    eng.set_fields(om, ps)
    t0 = time.perf_counter()
    eng.step(n_steps)
    el = time.perf_counter() - t0
    sps = n_steps / el
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": sps / (2 * tau_plain), "unit": "decorrelated CL samples/s", "cores": cores, "kind": "port",
            "steps_per_s": sps,
            "sample": f"{n_steps} plain FP64 {scheme_name} steps of the 96^3 N=100 workload on the host "
                      f"({el:.1f} s, own mixed-radix FFT, OpenMP); samples/s uses the GPU plain engine's tau_int"}


def speedup(x, y):
    """Ratio of two legs' decorrelated samples/s with its relative error."""
    if x is None or y is None or x.get("samples_per_s") is None or y.get("samples_per_s") is None:
        return None
    return {"ratio": x["samples_per_s"] / y["samples_per_s"], "rel_err": ratio_err(x, y)}


def ratio_err(x, y):
    """Relative error of x/y's decorrelated-sample ratio from the two tau_int errors."""
    return float(np.hypot(x["tau_int_err"] / x["tau_int"], y["tau_int_err"] / y["tau_int"]))


LEG_KEYS = ("failed", "dt", "replicas", "replica_steps_per_s", "ms_per_step", "equilibration_steps", "timed_steps", "record_steps",
            "tau_record_samples", "record_replica_steps_per_s", "tau_int", "tau_int_err", "tau_int_err_sokal",
            "tau_int_err_jackknife", "window", "tau_int_each",
            "tau_block", "samples_per_s", "H_mean", "H_stderr", "H_std", "H_median", "schedule", "tune_ms", "restarts")


def leg_summary(leg):
    return {k: leg[k] for k in LEG_KEYS if k in leg}


def obs_vector(rec):
    """The observables a record reports (real parts, as observables.dat prints them,
    complexLangevin.cu:72-91): H, mu_ex of the polymer, the six stress components, p_iso."""
    return np.concatenate([[rec["H"].real], rec["mu_ex_poly"].real, rec["stress"].real, [rec["p_iso"].real]])


OBS_NAMES = ["H", "mu_ex_poly", "sigma_xx", "sigma_yy", "sigma_zz", "sigma_xy", "sigma_xz", "sigma_yz", "p_iso"]


def same_state(sys64, sys32, om, ps):
    """Observables of an FP32 and an FP64 engine on the same fields."""
    e64, e32 = hip_engine(sys64), hip_engine(sys32)
    for e in (e64, e32):
        e.set_fields(om, ps)
    a, b = obs_vector(e32.measure()), obs_vector(e64.measure())
    e64.close()
    e32.close()
    return a, b


def run_c2(args):
    """Precision form, single system: 64^3, L = 32, N = 50, salt 0.1, semi-implicit (Debye)
    at dt = 0.05; FP32 (propagators, FFTs; per-bead renormalisation; FP64 reductions and
    field updates) against FP64 on the same noise stream.  Three measurements:
      same_state  the precision error: FP32 and FP64 observables on identical fields, at
                  --c2-states states along an FP64 trajectory (max relative error, per
                  observable; mu compared absolutely where |mu| < 1, the stress components
                  relative to the largest |sigma_ab|);
      divergence  the same-noise FP32 and FP64 trajectories step by step from identical
                  initial fields: relative difference of H at every step (they separate
                  as the CL dynamics amplify the rounding differences);
      ensemble    time averages over R replicas x --c2-steps steps (records every
                  --c2-every, the first tenth dropped) in FP32 and in FP64 batches with the
                  same seeds and initial fields: relative difference of the averages with
                  the combined standard error (Sokal tau_int of each record)."""
    from ftspolysol_amd.system import config
    R, steps, every = args.c2_replicas, args.c2_steps, args.c2_every
    mk = lambda prec, seed, mask=0x7F: config(2, precision=prec, scheme=SCHEME_SEMI_IMPLICIT, stress_mask=mask, seed=seed)
    # divergence of one same-noise pair
    pair = {}
    for name, prec in (("fp64", PREC_FP64), ("fp32", PREC_FP32)):
        e = hip_engine(mk(prec, 0xC2))
        e.set_fields(*initial_fields(mk(prec, 0xC2), sigma=1e-3, cfg=2))
        pair[name] = np.array([r["H"].real for r in e.step(args.c2_div_steps, record_every=1)])
        e.close()
    div = np.abs(pair["fp32"] - pair["fp64"]) / np.abs(pair["fp64"])
    print(f"[bench c2] divergence over {args.c2_div_steps} steps measured", file=sys.stderr, flush=True)
    # ensemble batches, with same-state probes of replica 0 along the FP64 run
    recs, rate, restarts = {}, {}, {}
    probes = []
    chunk = max(every, steps // max(1, args.c2_states))
    for name, prec in (("fp64", PREC_FP64), ("fp32", PREC_FP32)):
        systems = [mk(prec, 0xC2 + r) for r in range(R)]
        b = hip_batch(systems)
        for r, (e, sy) in enumerate(zip(b.replicas, systems)):
            e.set_fields(*initial_fields(sy, sigma=1e-3, seed=0x5EED0002 + r))
            e.set_restart(50, 50)
        rr = [[] for _ in range(R)]
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            k = min(chunk, steps - done)
            out = b.step(k, record_every=every)
            for r in range(R):
                rr[r] += out[r]
            done += k
            if name == "fp64" and len(probes) < args.c2_states:
                om = b.replicas[0].get_field(FIELD_OMEGA)
                ps = b.replicas[0].get_field(FIELD_PSI)[0]
                probes.append((om, ps))
            print(f"[bench c2] {name} x{R}: {done}/{steps} steps", file=sys.stderr, flush=True)
        rate[name] = R * steps / (time.perf_counter() - t0)
        restarts[name] = sum(e.get_info().restarts for e in b.replicas)
        recs[name] = rr
        b.close()
    errs = []
    for om, ps in probes:
        a, b64 = same_state(mk(PREC_FP64, 0xC2), mk(PREC_FP32, 0xC2), om, ps)
        e = np.abs(a - b64) / np.maximum(np.abs(b64), 1e-300)
        e[1] = abs(a[1] - b64[1]) / max(1.0, abs(b64[1]))   # mu = -ln Q: absolute where |mu| < 1
        # stress components relative to the tensor's scale (off-diagonal ones are near 0)
        e[2:8] = np.abs(a[2:8] - b64[2:8]) / np.abs(b64[2:8]).max()
        errs.append(e)
    errs = np.array(errs)
    per = {}
    for i, nm in enumerate(OBS_NAMES):
        stats = {}
        for name in ("fp64", "fp32"):
            X = np.array([[obs_vector(x)[i] for x in rr] for rr in recs[name]])
            X = X[:, X.shape[1] // 10:]
            tau, _ = tau_int(X)
            stats[name] = (float(X.mean()), float(X.std(ddof=1) * np.sqrt(2 * tau / X.size)), float(tau))
        m64, se64, _ = stats["fp64"]
        m32, se32, _ = stats["fp32"]
        per[nm] = {"same_state_max_rel_err": float(errs[:, i].max()),
                   "fp64_mean": m64, "fp64_stderr": se64, "fp32_mean": m32, "fp32_stderr": se32,
                   "tau_int_records": stats["fp64"][2],
                   "ensemble_rel_diff": abs(m32 - m64) / abs(m64), "ensemble_rel_sigma": float(np.hypot(se32, se64) / abs(m64))}
    first = lambda thr: next((int(i) + 1 for i, v in enumerate(div) if v > thr), None)
    line = {"metric": "C2 precision form: FP32 vs FP64 per-observable relative error on the same states",
            "value": float(errs.max()), "unit": "max relative error, same fields", "higher_is_better": False, "n_gpus": 1,
            "steps": steps,
            "config": {"workload": "64^3 L=32 N=50 z=-0.5 salt=0.1 B=0.5 lB=1, semi-implicit (Debye) dt=0.05",
                       "ensemble": f"{R} replicas per precision, {steps} steps, records every {every}, first tenth dropped",
                       "same_state": f"{len(probes)} states of FP64 replica 0, one every {chunk} steps"},
            "replica_steps_per_s": rate, "restarts": restarts, "observables": per,
            "divergence": {"steps_to_1e-6": first(1e-6), "steps_to_1e-4": first(1e-4), "steps_to_1e-2": first(1e-2),
                           "H_rel_diff": [float(f"{v:.3g}") for v in div]}}
    print(json.dumps(line))


def run_c3(args):
    """Batched form: 64 replicas (n in {32, 48, 64, 96}, L = n/2, N in [20, 100], salt
    log-uniform in [1e-3, 0.5], lB in [0.5, 2], own seeds), FP32 propagators, FP64
    reductions, semi-implicit (Debye) steps at dt = 0.01 in one replica batch: replica
    CL steps/s over exactly `steps` timed steps after max(warmup, 50) untimed ones."""
    from ftspolysol_amd.system import replica_params
    systems = [make_system(**replica_params(r), dt=0.01, precision=PREC_FP32, scheme=SCHEME_SEMI_IMPLICIT)
               for r in range(64)]
    b = hip_batch(systems)
    for r, (e, s) in enumerate(zip(b.replicas, systems)):
        e.set_fields(*initial_fields(s, sigma=1e-3, seed=900 + r))
        e.set_restart(50, 50)
    t0 = time.perf_counter()
    b.step(max(args.warmup, 50))
    print(f"[bench c3] warmup in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    b.set_timing(True)
    t0 = time.perf_counter()
    recs = b.step(args.steps, record_every=1)
    el = time.perf_counter() - t0
    groups = []
    for g in range(b.info.n_groups):
        t = b.get_timing(g)
        us = 1e3 * t["sweep_ms"] / max(1, t["step_launches"])
        groups.append({"mesh": b.info.group_mesh[g][0], "replicas": b.info.group_size[g],
                       "schedule": b.info.group_schedule[g], "us_per_launch_group": us,
                       "alg_GBps": t["algorithmic_bytes"] / max(1e-30, t["sweep_ms"] * 1e-3) / 1e9})
    restarts = sum(e.get_info().restarts for e in b.replicas)
    finite = all(np.isfinite(x["H"]) for rr in recs for x in rr)
    b.close()
    print(json.dumps({"metric": "C3 batched form: replica CL steps/s, 64 mixed replicas", "value": 64 * args.steps / el,
                      "unit": "replica CL steps/s", "higher_is_better": True, "n_gpus": 1, "steps": args.steps,
                      "ms_per_step": 1e3 * el / args.steps, "config": {"workload": "64 replicas, n in {32,48,64,96}, "
                      "N in [20,100], salt log-uniform [1e-3,0.5], lB in [0.5,2]; FP32 semi-implicit dt=0.01"},
                      "groups": groups, "restarts": restarts, "records_finite": finite}))


def run_c5(args):
    """Large single system: 256^3, L = 128, N = 200, charge -1 on every second bead, salt
    0.02, B = 1, lB = 0.7; FP32 propagators (both sweeps), FP64 reductions, the learned
    preconditioner; `steps` timed steps with the full stress tensor and p_iso in a record
    every --c5-stress-every steps, then one fts_measure (densities + all observables) timed
    on its own."""
    from ftspolysol_amd.system import config
    s = config(5, precision=PREC_FP32, scheme=SCHEME_SI_LEARNED, stress_mask=0x7F, seed=0xC5)
    e = hip_engine(s)
    e.lrk_set_network(load_params(args.lrk_c5), args.lrk_threshold)
    e.set_fields(*initial_fields(s, sigma=1e-3, cfg=5))
    e.set_restart(50, 20)
    e.step(max(args.warmup, 5))
    print("[bench c5] warmed up", file=sys.stderr, flush=True)
    e.set_timing(True)
    t0 = time.perf_counter()
    recs = []
    done = 0
    chunk = max(args.c5_stress_every, 250 // args.c5_stress_every * args.c5_stress_every)
    while done < args.steps:
        k = min(chunk, args.steps - done)
        recs += e.step(k, record_every=args.c5_stress_every)
        done += k
        if args.steps > chunk:
            print(f"[bench c5] {done}/{args.steps} steps, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    el = time.perf_counter() - t0
    t = e.get_timing()
    e.set_timing(False)
    info = e.get_info()
    lrk_stats = e.lrk_get_stats()
    print(f"[bench c5] {args.steps} timed steps in {el:.2f} s", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    rec = e.measure()
    measure_s = time.perf_counter() - t1
    e.close()
    us = 1e3 * t["sweep_ms"] / max(1, t["step_launches"])
    print(json.dumps({"metric": "C5 large system: CL steps/s at 256^3 N=200", "value": args.steps / el, "unit": "CL steps/s",
                      "higher_is_better": True, "n_gpus": 1, "steps": args.steps, "ms_per_step": 1e3 * el / args.steps,
                      "config": {"workload": "256^3 L=128 N=200 z=-1 on odd beads salt=0.02 B=1 lB=0.7; FP32 + learned SI",
                                 "stress_every": args.c5_stress_every, "learned_kernel": os.path.basename(args.lrk_c5)},
                      "bead_step_us": us, "bead_step_alg_GBps": t["algorithmic_bytes"] / max(1e-30, t["sweep_ms"] * 1e-3) / 1e9,
                      "stress_records": len(recs), "measure_with_stress_s": measure_s,
                      "stress_finite": bool(np.all(np.isfinite(rec["stress"]))),
                      "records_finite": bool(all(np.isfinite(x["H"]) and np.all(np.isfinite(x["stress"])) for x in recs)),
                      "restarts": int(info.restarts), "learned_steps": int(lrk_stats["learned_steps"]),
                      "judged_steps": int(lrk_stats["judged"]),
                      "H_records": [float(x["H"].real) for x in recs[:: max(1, len(recs) // 20)]]}))


def long_record_estimate(fast, plain_1, plain_b, fast_scheme, dt_fast, dt_plain):
    """This run's rates with the tau_int of the longest committed records of the same
    legs (profiles/long_record_96cubed_N100.jsonl, bench.py --config dtstudy): rare long
    excursions of H set tau_int, and the run's own 2000-step records bound it from below.
    None when no committed record matches the legs."""
    path = os.path.join(ROOT, "profiles", "long_record_96cubed_N100.jsonl")
    if not os.path.exists(path):
        return None
    best = {}
    for line in open(path):
        d = json.loads(line)
        key = (d.get("scheme"), d.get("precision"), d.get("dt"))
        if d.get("tau_int") and (key not in best or d["record_steps"] > best[key]["record_steps"]):
            best[key] = d
    f, p = best.get((fast_scheme, "fp32", dt_fast)), best.get(("pc", "fp64", dt_plain))
    if f is None or p is None or not fast.get("replica_steps_per_s") or not plain_1.get("replica_steps_per_s"):
        return None
    fast_s = fast["replica_steps_per_s"] / (2 * f["tau_int"])
    plain_s = plain_1["replica_steps_per_s"] / (2 * p["tau_int"])
    plain_bs = plain_b["replica_steps_per_s"] / (2 * p["tau_int"]) if plain_b.get("replica_steps_per_s") else None
    rel = float(np.hypot(f["tau_int_err"] / f["tau_int"], p["tau_int_err"] / p["tau_int"]))
    pc32 = best.get(("pc", "fp32", dt_plain))
    fp32_pc = None
    if pc32 is not None:
        # the FP32 predictor-corrector batch is not a leg of this run: its committed rate
        s32 = pc32["replica_steps_per_s"] / (2 * pc32["tau_int"])
        fp32_pc = {"record_steps": pc32["record_steps"], "replica_steps_per_s": pc32["replica_steps_per_s"],
                   "tau_int": pc32["tau_int"], "samples_per_s": s32,
                   "speedup_vs_plain_fp64": {"ratio": s32 / plain_s, "rel_err": float(
                       np.hypot(pc32["tau_int_err"] / pc32["tau_int"], p["tau_int_err"] / p["tau_int"]))}}
    return {"source": "profiles/long_record_96cubed_N100.jsonl",
            "fp32_pc_committed": fp32_pc,
            "record_steps": {"fast": f["record_steps"], "plain_fp64": p["record_steps"]},
            "tau_int": {"fast": f["tau_int"], "plain_fp64": p["tau_int"]},
            "samples_per_s": {"fast": fast_s, "plain_fp64": plain_s, "plain_fp64_batched": plain_bs},
            "speedup_vs_plain_fp64": {"ratio": fast_s / plain_s, "rel_err": rel},
            "speedup_vs_plain_fp64_batched": None if plain_bs is None else {"ratio": fast_s / plain_bs, "rel_err": rel}}


STUDY_SCHEMES = {"pc": SCHEME_PC_REFERENCE, "si": SCHEME_SEMI_IMPLICIT, "si_learned": SCHEME_SI_LEARNED,
                 "sipc": SCHEME_SIPC, "sipc_learned": SCHEME_SIPC_LEARNED}


def run_dtstudy(args):
    """Time-step study of the headline workload: for each leg scheme:precision:dt of
    --legs, a batch of --replicas replicas (the headline's seeds and initial fields)
    equilibrated max(warmup, --equilibration) steps, timed over `steps` steps and
    recorded over --record-steps steps.  One JSON line per leg: the rate, tau_int (pooled
    over the replicas, with the block-average check), <H> with its standard error, the
    median and spread of H (the spread holds the excursions of H), samples/s, restarts
    and the learned kernel's use."""
    wl = dict(n=96, N=100, rho=0.2, salt=0.05, B=0.5, lB=1.0, charge="uniform", z=-0.5)
    lrk = load_params(args.lrk)
    EQ = dict(min_equilibration=args.equilibration, graph_level=args.graph_level, streams=args.streams)
    for spec in args.legs.split(","):
        name, prec, dt = spec.split(":")
        scheme = STUDY_SCHEMES[name]
        learned = scheme in (SCHEME_SI_LEARNED, SCHEME_SIPC_LEARNED)
        systems = [make_system(**wl, dt=float(dt), precision=PREC_FP32 if prec == "fp32" else PREC_FP64, scheme=scheme,
                               seed=0xC4 + r) for r in range(args.replicas)]
        leg = run_leg(systems, f"{name} {prec} dt={dt} x{args.replicas}", args.warmup, args.steps,
                      max(args.record_steps, args.steps), lrk=lrk if learned else None, lrk_threshold=args.lrk_threshold, **EQ)
        print(json.dumps({"workload": "96^3 L=48 N=100 z=-0.5 salt=0.05 B=0.5 lB=1 (221 chains)", "scheme": name,
                          "precision": prec, **leg_summary(leg), "learned_fraction": leg["lrk"]["learned_fraction"]
                          if leg.get("lrk") else None}), flush=True)


def run_beadstep(args):
    """The bead-step launch groups alone, at 96^3 N=100 (the headline workload): for each
    case prec:replicas:schedule of --cases, the average duration of one bead-step launch
    group (HIP events on the batch stream over the sweeps of `steps` CL steps after
    `warmup`), its algorithmic bytes (jobs x read q_s + write q_{s+1}) and their rate
    against the 8 TB/s peak.  FP64 cases step the predictor-corrector, FP32 ones the
    semi-implicit scheme; schedules 3 = conv, 4 = separable, 0 = timed at creation."""
    wl = dict(n=96, N=100, rho=0.2, salt=0.05, B=0.5, lB=1.0, charge="uniform", z=-0.5)
    rows = []
    for case in args.cases.split(","):
        prec, R, sched = case.split(":")
        R, sched = int(R), int(sched)
        p = PREC_FP64 if prec == "fp64" else PREC_FP32
        scheme = SCHEME_PC_REFERENCE if prec == "fp64" else SCHEME_SEMI_IMPLICIT
        systems = [make_system(**wl, dt=0.01, precision=p, scheme=scheme, seed=0xC4 + r, fft_schedule=sched)
                   for r in range(R)]
        b = hip_batch(systems)
        b.set_graph(args.graph_level)
        for r, (e, s) in enumerate(zip(b.replicas, systems)):
            e.set_fields(*initial_fields(s, sigma=1e-3, seed=0x5EED0004 + r))
        b.step(max(args.warmup, 1))
        b.set_timing(True)
        t0 = time.perf_counter()
        b.step(args.steps)
        el = time.perf_counter() - t0
        t = b.get_timing(0)
        used = b.info.group_schedule[0]
        b.close()
        us = 1e3 * t["sweep_ms"] / max(1, t["step_launches"])
        gbs = t["algorithmic_bytes"] / max(1e-30, t["sweep_ms"] * 1e-3) / 1e9
        rows.append({"precision": prec, "replicas": R, "schedule": used, "us_per_launch_group": us,
                     "alg_bytes_per_launch_group": t["algorithmic_bytes"] / max(1, t["step_launches"]),
                     "alg_GBps": gbs, "frac": gbs / HBM_PEAK_GBS, "ms_per_step": 1e3 * el / args.steps})
        print(f"[bench beadstep] {case}: {us:.2f} us per launch group, {gbs:.0f} GB/s", file=sys.stderr, flush=True)
    # a device-to-device copy of 256 MiB (twice): the rate, and under rocprofv3 --pmc the
    # calibration of FETCH_SIZE / WRITE_SIZE on a copy of known size
    copy_gbs = copy_bandwidth(gib=0.25, reps=2)
    print(json.dumps({"metric": "bead-step launch group at 96^3 N=100", "unit": "us per launch group",
                      "steps": args.steps, "warmup": args.warmup, "graph_level": args.graph_level, "cases": rows,
                      "copy_GBps": copy_gbs, "copy_bytes": 1 << 28}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed steps of every leg (their rate)")
    ap.add_argument("--warmup", type=int, default=200,
                    help=f"untimed steps before the timed ones (at least {MIN_EQUILIBRATION} are taken)")
    ap.add_argument("--record-steps", type=int, default=2000,
                    help="steps of the H record of the FP32 legs (tau_int, <H>), timed steps included")
    ap.add_argument("--plain-record-steps", type=int, default=2000, help="the same for the plain FP64 batch")
    ap.add_argument("--replicas", type=int, default=16, help="replicas in the FP32 batches")
    ap.add_argument("--plain-replicas", type=int, default=16, help="replicas in the plain FP64 batch")
    ap.add_argument("--dt-fast", type=float, default=0.01,
                    help="time step of the learned leg (the learned kernel keeps dt 0.04 finite where the Debye scheme "
                         "diverges, and lowers the <H> bias at dt >= 0.02; see --learned-vs-debye)")
    ap.add_argument("--dt-debye", type=float, default=0.01, help="time step of the Debye semi-implicit leg")
    ap.add_argument("--dt-plain", type=float, default=0.01)
    ap.add_argument("--lrk", default=LRK_PARAMS, help="parameters of the learned kernel (.npy)")
    ap.add_argument("--lrk-c5", default=LRK_PARAMS_C5 if os.path.exists(LRK_PARAMS_C5) else LRK_PARAMS,
                    help="parameters of the learned kernel for --config c5 (trained on C5's chemistry)")
    ap.add_argument("--lrk-threshold", type=float, default=1e-3)
    ap.add_argument("--with-fp32-pc", action="store_true", help="add the FP32 predictor-corrector batch")
    ap.add_argument("--no-debye", action="store_true", help="skip the FP32 Debye semi-implicit batch")
    ap.add_argument("--fast-only", action="store_true", help="run the fast leg alone and print its summary (profiling runs)")
    ap.add_argument("--fast-scheme", choices=["sipc_learned", "si_learned"], default="sipc_learned",
                    help="scheme of the fast leg (its Debye leg takes the same scheme with the Debye kernel)")
    ap.add_argument("--learned-vs-debye", action="store_true",
                    help="run the fast (learned) and Debye semi-implicit legs only and print their comparison")
    ap.add_argument("--streams", type=int, default=-1,
                    help="side streams per replica group for the replicas' own work (-1: the library's default)")
    ap.add_argument("--graph-level", type=int, choices=[0, 1, 2], default=2,
                    help="batch launches: 2 replays each whole CL step as a captured graph, 1 only the sweeps, 0 none")
    ap.add_argument("--equilibration", type=int, default=MIN_EQUILIBRATION,
                    help="least untimed steps before the timed ones (lower only for short profiling runs)")
    ap.add_argument("--schedule", choices=["auto", "plane", "line", "conv", "separable", "conv_mfma"], default="auto",
                    help="bead-step schedule of the FP32 batches (auto: timed at creation)")
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config", choices=["c4", "c2", "c3", "c5", "beadstep", "dtstudy"], default="c4",
                    help="c4: the headline (default); c2: FP32 vs FP64 errors at 64^3; c3: 64 mixed replicas; c5: 256^3; "
                         "beadstep: the bead-step launch groups alone at 96^3 (--cases); dtstudy: the headline workload "
                         "per scheme, precision and dt (--legs)")
    ap.add_argument("--legs", default="pc:fp64:0.01,sipc_learned:fp32:0.01",
                    help="dtstudy legs scheme:precision:dt (schemes pc, si, si_learned, sipc, sipc_learned)")
    ap.add_argument("--cases", default="fp64:1:3,fp64:1:4,fp64:8:3,fp32:16:3",
                    help="beadstep cases precision:replicas:schedule (3 conv, 4 separable, 5 conv on the matrix cores, 0 timed)")
    ap.add_argument("--c2-steps", type=int, default=10000)
    ap.add_argument("--c2-every", type=int, default=10)
    ap.add_argument("--c2-replicas", type=int, default=8)
    ap.add_argument("--c2-states", type=int, default=10)
    ap.add_argument("--c2-div-steps", type=int, default=100)
    ap.add_argument("--c5-stress-every", type=int, default=50)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine runs each system (or replica batch) on a single GPU, "
              "and runs on several GPUs are not supported", file=sys.stderr)
        sys.exit(2)
    if args.config != "c4":
        {"c2": run_c2, "c3": run_c3, "c5": run_c5, "beadstep": run_beadstep, "dtstudy": run_dtstudy}[args.config](args)
        return
    if args.steps < 1:
        print("--steps must be at least 1", file=sys.stderr)
        sys.exit(2)

    wl = dict(n=96, N=100, rho=0.2, salt=0.05, B=0.5, lB=1.0, charge="uniform", z=-0.5)
    lrk = load_params(args.lrk)
    sched = {"auto": 0, "plane": 1, "line": 2, "conv": 3, "separable": 4, "conv_mfma": 5}[args.schedule]
    fp32 = lambda scheme, dt: [make_system(**wl, dt=dt, precision=PREC_FP32, scheme=scheme, seed=0xC4 + r,
                                           fft_schedule=sched) for r in range(args.replicas)]
    plain = lambda n: [make_system(**wl, dt=args.dt_plain, precision=PREC_FP64, scheme=SCHEME_PC_REFERENCE,
                                   seed=0xC4 + r) for r in range(n)]
    rec = max(args.record_steps, args.steps)
    prec = max(args.plain_record_steps, args.steps)
    W, K = args.warmup, args.steps

    EQ = dict(min_equilibration=args.equilibration, graph_level=args.graph_level, streams=args.streams)
    sipc = args.fast_scheme == "sipc_learned"
    fast_scheme, debye_scheme = (SCHEME_SIPC_LEARNED, SCHEME_SIPC) if sipc else (SCHEME_SI_LEARNED, SCHEME_SEMI_IMPLICIT)
    debye_name = "sipc" if sipc else "si"
    fast = run_leg(fp32(fast_scheme, args.dt_fast), f"fast fp32 {args.fast_scheme} x{args.replicas}", W, K, rec, lrk=lrk,
                   lrk_threshold=args.lrk_threshold, **EQ)
    debye = None if (args.no_debye or args.fast_only) else run_leg(fp32(debye_scheme, args.dt_debye),
                                               f"{debye_name} (Debye) fp32 x{args.replicas}", W, K, rec, **EQ)
    fp32_pc = run_leg(fp32(SCHEME_PC_REFERENCE, args.dt_plain), f"fp32 pc x{args.replicas}", W, K, rec, **EQ) \
        if args.with_fp32_pc and not args.fast_only else None
    if args.learned_vs_debye:
        print(json.dumps({"metric": "learned vs Debye kernel at 96^3 N=100", "scheme": args.fast_scheme, "dt": args.dt_fast,
                          "dt_debye": args.dt_debye,
                          "replicas": args.replicas, "fast": leg_summary(fast), "si_debye": leg_summary(debye),
                          "learned_kernel": fast["lrk"], "learned_vs_debye": speedup(fast, debye)}))
        return
    if args.fast_only:
        # a short run for profilers: the fast leg alone (no statistics beyond its record)
        print(json.dumps({"fast": leg_summary(fast), "roofline_us_per_launch_group":
                          1e3 * fast["timing"]["sweep_ms"] / max(1, fast["timing"]["step_launches"])}))
        return
    plain_b = run_leg(plain(args.plain_replicas), f"plain fp64 pc x{args.plain_replicas}", W, K, prec, **EQ)
    plain_1 = run_leg(plain(1), "plain fp64 pc x1", W, K, 0, stats=False, **EQ)
    # one system's rate; its tau_int and <H> are those of the same dynamics in the batch
    for key in ("tau_int", "tau_int_err", "window", "H_mean", "H_stderr", "record_steps", "tau_record_samples"):
        plain_1[key] = plain_b.get(key)
    plain_1["samples_per_s"] = plain_1["replica_steps_per_s"] / (2 * plain_b["tau_int"]) \
        if plain_b["samples_per_s"] is not None and plain_1["replica_steps_per_s"] is not None else None
    plain_1["tau_source"] = f"the plain FP64 batch's record ({args.plain_replicas} replicas)"
    failed = {name: leg["failed"] for name, leg in (("fast", fast), ("plain_fp64_batched", plain_b), ("plain_fp64", plain_1))
              if leg.get("failed")}
    if failed:
        print(json.dumps({"metric": "decorrelated CL samples/s at 96^3 N=100, FP32+learned vs plain FP64, <H> err",
                          "value": None, "unit": "decorrelated CL samples/s", "n_gpus": 1, "steps": K, "warmup": W,
                          "higher_is_better": True, "failed": failed}))
        sys.exit(1)

    same = same_state_errors(wl, fast["final"])

    def roof(t):
        if not t["step_launches"]:
            return float("nan"), float("nan"), float("nan")
        us = 1e3 * t["sweep_ms"] / t["step_launches"]
        byt = t["algorithmic_bytes"] / t["step_launches"]
        return us, byt, byt / (us * 1e-6) / 1e9

    us, byt, achieved = roof(fast["timing"])
    # measured HBM traffic per launch group (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, calibrated
    # on a copy kernel) when the committed measurement is for this configuration
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic_96cubed_N100.json")
    if os.path.exists(tpath):
        for t in json.load(open(tpath)).get("entries", []):
            if t.get("replicas") == args.replicas and t.get("schedule") == fast["schedule"] and t.get("precision") == "fp32":
                traffic = t["traffic_bytes"]
    _, _, plain_achieved = roof(plain_1["timing"])
    _, _, plain_b_achieved = roof(plain_b["timing"])
    try:
        copy_gbs = copy_bandwidth()
    except Exception as exc:  # report the peak fraction only
        print(f"[bench] copy bandwidth not measured: {exc}", file=sys.stderr)
        copy_gbs = None
    def h_err(x):
        """|<H>_x - <H>_plain| / <H>_plain and its combined standard error"""
        if x is None or x.get("H_mean") is None or plain_b.get("H_mean") is None:
            return None, None
        return (abs(x["H_mean"] - plain_b["H_mean"]) / abs(plain_b["H_mean"]),
                float(np.hypot(x["H_stderr"], plain_b["H_stderr"]) / abs(plain_b["H_mean"])))
    H_err_vs_plain, H_err_sigma = h_err(fast)
    H_err_debye, H_err_debye_sigma = h_err(debye)

    sched_name = {1: "plane", 2: "line", 3: "conv", 4: "separable", 5: "conv_mfma"}
    out = {
        "metric": "decorrelated CL samples/s at 96^3 N=100, FP32+learned vs plain FP64, <H> err",
        "value": fast["samples_per_s"],
        "unit": "decorrelated CL samples/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": fast["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 propagators/FFT, f64 reductions",
        "data": "synthetic",
        "config": {"workload": "96^3 L=48 N=100 z=-0.5 salt=0.05 B=0.5 lB=1 (221 chains)",
                   "fast": f"FP32 {args.fast_scheme} dt={args.dt_fast}, {args.replicas} replicas per launch, learned kernel "
                           f"{os.path.basename(args.lrk)}, fallback threshold {args.lrk_threshold}",
                   "si_debye": None if debye is None else f"FP32 {debye_name} (Debye kernel) dt={args.dt_debye}, {args.replicas} replicas",
                   "plain_fp64_batched": f"FP64 pc dt={args.dt_plain}, {args.plain_replicas} replicas per launch",
                   "plain_fp64": f"FP64 pc dt={args.dt_plain}, 1 system",
                   "statistics": f"tau_int and <H> from an H record of {rec} steps (FP32 legs) / {prec} (FP64 batch) "
                                 f"after {max(W, args.equilibration)} equilibration steps; rate from the {K} timed steps; "
                                 "tau_int pooled over the replicas, its error the larger of Sokal's and the replica "
                                 "jackknife's; rare long excursions of H set tau_int, so records this short give lower "
                                 "bounds (bench.py --config dtstudy --record-steps 10000 measures longer records)",
                   "parallelism": "single GPU"},
        "equilibration_steps": fast["equilibration_steps"],
        "tau_record_samples": fast["tau_record_samples"],
        "learned_fraction": fast["lrk"]["learned_fraction"] if fast.get("lrk") else None,
        "fast": leg_summary(fast),
        "si_debye": None if debye is None else leg_summary(debye),
        "fp32_pc": None if fp32_pc is None else leg_summary(fp32_pc),
        "plain_fp64_batched": leg_summary(plain_b),
        "plain_fp64": dict(leg_summary(plain_1), tau_source=plain_1["tau_source"]),
        "learned_kernel": fast["lrk"],
        "speedup_vs_plain_fp64": speedup(fast, plain_1),
        "speedup_vs_plain_fp64_batched": speedup(fast, plain_b),
        "learned_vs_debye": speedup(fast, debye),
        "debye_speedup_vs_plain_fp64": speedup(debye, plain_1),
        "fp32_pc_speedup_vs_plain_fp64": speedup(fp32_pc, plain_1),
        "H_err_same_state": float(same["H"]),
        "same_state_rel_err": {k: float(v) for k, v in same.items() if k != "H_fp64"},
        "H_err_vs_plain": H_err_vs_plain,
        "H_err_sigma": H_err_sigma,
        "H_err_vs_plain_si_debye": H_err_debye,
        "H_err_sigma_si_debye": H_err_debye_sigma,
        "H_mean": {"fast": fast["H_mean"], "si_debye": None if debye is None else debye["H_mean"],
                   "plain_fp64_batched": plain_b["H_mean"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"bead-step launch group ({sched_name.get(fast['schedule'], '?')} schedule), FP32, "
                               f"{args.replicas} replica jobs per launch",
                     "us_per_launch_group": us, "alg_bytes_per_launch_group": byt,
                     "plain_fp64_achieved": plain_achieved,
                     "plain_fp64_batched_achieved": plain_b_achieved,
                     "copy_achievable": copy_gbs,
                     "frac_of_copy": None if not copy_gbs else achieved / copy_gbs,
                     "traffic_rate": None if traffic is None else traffic / (us * 1e-6) / 1e9},
    }
    out["long_record_estimate"] = long_record_estimate(fast, plain_1, plain_b, args.fast_scheme, args.dt_fast, args.dt_plain)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(plain(1)[0], args.cpu_steps, plain_b["tau_int"], "PC")
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Headline benchmark: decorrelated complex-Langevin samples/s at 96^3, N = 100.

Workload (one GPU): 96^3 grid, L = 48 b, discrete Gaussian chains N = 100,
b = 1, charge -0.5 on every bead, counterions, 1:1 salt 0.05 b^-3, bead density
0.2 b^-3 (221 chains), excluded volume B = 0.5 b^3, lB = 1 b, a = 1 b,
lambda = 1.  Synthetic initial fields (w, psi = 1e-3 complex Gaussian).

Three engines run the workload:
  fast    FP32 + learned: a batch of R independent replicas (own seeds and
          initial fields) with complex-FP32 propagators and FFTs, per-bead
          renormalisation, FP64 reductions and field accumulation, stepping the
          semi-implicit scheme with the learned linear-response kernel
          (FTS_SCHEME_SI_LEARNED; network trained by ftspolysol_amd.train_lrk on
          48^3 data, ftspolysol_amd/lrk_c4_48cubed.npy), judged every step by the
          exact force, falling back to the Debye kernel when its excess force
          residual passes 1e-3.  Every bead-step launch carries all replicas.
  fp32_pc the same batch without the learned part: FP32, the reference's
          predictor-corrector step (complexLangevin.cu:669-698).
  plain   one complex-FP64 system, predictor-corrector, as the reference runs it.
Replica 0 of each batch has the plain engine's seed and initial fields.  Each
takes `warmup` untimed steps, then timed steps recording Re H every step: `steps`
(2000) for the fast leg, whose tau_int sets `value`, `--aux-steps` (1000) for the
others.
tau_int of one 1000-step trajectory scatters by a factor of ten between seeds
(0.5 .. 5.8 measured), so the plain scheme's tau_int and <H> come from the pooled
statistics of `--tau-replicas` independent FP64 replicas of the same dynamics
(one batch, untimed for the rate); its throughput is the timed plain engine's.
Every replica keeps a device checkpoint every 50 steps; a replica whose step
turns non-finite is restored from it with a new noise key and the repeated steps
count in the elapsed time; the restarts are reported.

<H> error: `H_err_same_state` (the precision error) measures H, mu and the
stress tensor with an FP32 engine and an FP64 engine on the same fields (fast
replica 0's final state); `H_err_vs_plain` compares the time averages of the fast
batch and the plain engine (statistical once trajectories decorrelate, plus the
time-step bias of the fast scheme against the predictor-corrector; `H_err_sigma`
is the standard error).  Decorrelated samples/s = (replica CL steps/s) /
(2 tau_int), tau_int by Sokal's automatic window (c = 6) on the replica-averaged
normalised autocorrelation of Re H.  `value` is the fast batch's whole-GPU rate.

roofline: `copy_achievable` is a 1 GiB device-to-device hipMemcpy timed on the same GPU
(the bandwidth a streaming kernel can reach); `traffic_rate` the measured HBM traffic
of a launch group over its duration.  The dominant kernels are the propagator bead-step passes (one pencil
pass and one plane pass per bead index, or the four line passes, every replica's
job in each launch).  Algorithmic bytes per bead step = read q_s + write q_{s+1}
= 2 c M (c = 8 B for complex FP32), summed over the jobs of a launch group and
divided by the group's average duration from HIP events on the batch stream.

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

from ftspolysol_amd.api import (FIELD_OMEGA, FIELD_PSI, Engine, PREC_FP32, PREC_FP64, SCHEME_EM,  # noqa: E402
                                SCHEME_PC_REFERENCE, SCHEME_SEMI_IMPLICIT, SCHEME_SI_LEARNED, hip_batch, hip_engine)
from ftspolysol_amd.lrk import load_params  # noqa: E402
from ftspolysol_amd.system import initial_fields, make_system  # noqa: E402

HBM_PEAK_GBS = 8000.0
LRK_PARAMS = os.path.join(ROOT, "ftspolysol_amd", "lrk_c4_48cubed.npy")


def autocorr(x):
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    y = x - x.mean()
    f = np.fft.rfft(y, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n]
    return acf / acf[0] if acf[0] > 0 else None


def tau_int(x, c=6.0):
    """Integrated autocorrelation time (in samples) with Sokal's automatic window.
    x: one series, or a list of equally long series whose normalised
    autocorrelations are averaged first."""
    series = [np.asarray(x, dtype=np.float64)] if np.ndim(x) == 1 else [np.asarray(v, dtype=np.float64) for v in x]
    n = series[0].size
    if n < 8:
        return float("nan"), 0
    rhos = [r for r in (autocorr(v) for v in series) if r is not None]
    if not rhos:
        return 0.5, 0
    rho = np.mean(rhos, axis=0)
    tau = 0.5
    for w in range(1, n):
        tau += rho[w]
        if w >= c * tau:
            return float(tau), w
    return float(tau), n - 1


def run_batch(systems, warmup, steps, label, ckpt_every=50, max_restarts=50, lrk=None, lrk_threshold=1e-3):
    b = hip_batch(systems)
    for r, (e, s) in enumerate(zip(b.replicas, systems)):
        om, ps = initial_fields(s, sigma=1e-3, seed=0x5EED0004 + r)
        e.set_fields(om, ps)
        e.set_restart(ckpt_every, max_restarts)
        if lrk is not None:
            e.lrk_set_network(lrk, lrk_threshold)
    R = len(systems)
    t0 = time.perf_counter()
    done = 0
    while done < warmup:
        k = min(100, warmup - done)
        b.step(k)
        done += k
    print(f"[bench] {label}: warmup {warmup} steps in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    b.set_timing(True)
    H = [[] for _ in range(R)]
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(200, steps - done)
        recs = b.step(k, record_every=1)
        for r in range(R):
            H[r].extend(x["H"] for x in recs[r])
        done += k
        print(f"[bench] {label}: {done}/{steps} timed steps", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    restarts = [e.get_info().restarts for e in b.replicas]
    lrk_stats = None
    if lrk is not None:
        st = [e.lrk_get_stats() for e in b.replicas]
        judged = max(1, sum(x["judged"] for x in st))
        lrk_stats = {"learned_steps": int(sum(x["learned_steps"] for x in st)),
                     "fallback_steps": int(sum(x["fallback_steps"] for x in st)),
                     "rel_force_residual_learned": [float(sum(x["sum_rel_residual_learned"][c] for x in st) / judged) for c in range(2)],
                     "rel_force_residual_debye": [float(sum(x["sum_rel_residual_debye"][c] for x in st) / judged) for c in range(2)]}
    final = (b.replicas[0].get_field(FIELD_OMEGA), b.replicas[0].get_field(FIELD_PSI)[0])
    timing = b.get_timing(0)
    sched = b.info.group_schedule[0]
    tune = list(b.info.group_tune_ms[0])
    b.close()
    H = np.array(H)
    tau, window = tau_int(H.real)
    tau_each = [tau_int(H.real[r:r + 1])[0] for r in range(R)]
    # Sokal's estimate of the statistical error of tau_int: var = 2 (2W + 1) / n tau^2,
    # n the number of samples pooled
    tau_err = float(tau * np.sqrt(2.0 * (2 * window + 1) / H.size))
    sps = R * steps / elapsed
    stderr_H = float(H.real.std(ddof=1) * np.sqrt(2 * tau / H.size))
    return dict(replicas=R, replica_steps_per_s=sps, ms_per_step=1e3 * elapsed / steps, tau_int=tau, window=window,
                tau_int_each=[round(float(x), 3) for x in tau_each], tau_int_err=tau_err,
                samples_per_s=sps / (2 * tau), H_mean=float(H.real.mean()), H_imag_mean=float(H.imag.mean()),
                H_stderr=stderr_H, H0=H[0].real, timing=timing, elapsed=elapsed, schedule=sched, tune_ms=tune,
                restarts=int(sum(restarts)), final=final, lrk=lrk_stats)


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000, help="timed steps of the fast leg (the measured one)")
    ap.add_argument("--aux-steps", type=int, default=1000,
                    help="timed steps of the comparison legs (FP32 PC, plain FP64 and its tau_int replicas), at most --steps")
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--replicas", type=int, default=16, help="replicas in the fast batch")
    ap.add_argument("--dt-fast", type=float, default=0.01)
    ap.add_argument("--fast-scheme", choices=["si_learned", "si", "em", "pc"], default="si_learned")
    ap.add_argument("--lrk", default=LRK_PARAMS, help="parameters of the learned kernel (.npy)")
    ap.add_argument("--lrk-threshold", type=float, default=1e-3)
    ap.add_argument("--dt-plain", type=float, default=0.01)
    ap.add_argument("--plain-replicas", type=int, default=1)
    ap.add_argument("--tau-replicas", type=int, default=8,
                    help="independent FP64 replicas (one batch) whose pooled autocorrelation gives the plain scheme's "
                         "tau_int; one trajectory of `steps` steps estimates it too poorly")
    ap.add_argument("--no-fp32-pc", action="store_true", help="skip the FP32 predictor-corrector comparison batch")
    ap.add_argument("--schedule", choices=["auto", "plane", "line", "conv"], default="auto",
                    help="bead-step schedule of the FP32 batches (auto: timed at creation)")
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine runs each system (or replica batch) on a single GPU, "
              "and runs on several GPUs are not supported", file=sys.stderr)
        sys.exit(2)

    wl = dict(n=96, N=100, rho=0.2, salt=0.05, B=0.5, lB=1.0, charge="uniform", z=-0.5)
    schemes = {"em": SCHEME_EM, "si": SCHEME_SEMI_IMPLICIT, "pc": SCHEME_PC_REFERENCE, "si_learned": SCHEME_SI_LEARNED}
    lrk = load_params(args.lrk) if args.fast_scheme == "si_learned" else None
    sched = {"auto": 0, "plane": 1, "line": 2, "conv": 3}[args.schedule]
    fast_sys = [make_system(**wl, dt=args.dt_fast, precision=PREC_FP32, scheme=schemes[args.fast_scheme], seed=0xC4 + r,
                            fft_schedule=sched) for r in range(args.replicas)]
    pc_sys = [make_system(**wl, dt=args.dt_plain, precision=PREC_FP32, scheme=SCHEME_PC_REFERENCE, seed=0xC4 + r,
                          fft_schedule=sched) for r in range(args.replicas)]
    plain_sys = [make_system(**wl, dt=args.dt_plain, precision=PREC_FP64, scheme=SCHEME_PC_REFERENCE, seed=0xC4 + r)
                 for r in range(args.plain_replicas)]

    fast = run_batch(fast_sys, args.warmup, args.steps, f"fast fp32 {args.fast_scheme} x{args.replicas}", lrk=lrk,
                     lrk_threshold=args.lrk_threshold)
    aux = min(args.aux_steps, args.steps)
    fp32_pc = None if args.no_fp32_pc else run_batch(pc_sys, args.warmup, aux, f"fp32 pc x{args.replicas}")
    plain = run_batch(plain_sys, args.warmup, aux, f"plain fp64 pc x{args.plain_replicas}")
    plain["tau_source"] = f"{args.plain_replicas} replica(s) timed"
    if args.tau_replicas > args.plain_replicas:
        # the same dynamics in more independent replicas: tau_int (and <H>) from their pooled
        # statistics; the throughput stays that of the timed plain engine
        tau_sys = [make_system(**wl, dt=args.dt_plain, precision=PREC_FP64, scheme=SCHEME_PC_REFERENCE, seed=0xC4 + r)
                   for r in range(args.tau_replicas)]
        st = run_batch(tau_sys, args.warmup, aux, f"plain fp64 pc statistics x{args.tau_replicas}")
        plain["tau_int_timed_replicas"] = plain["tau_int"]
        for k in ("tau_int", "window", "tau_int_each", "tau_int_err", "H_mean", "H_stderr"):
            plain[k] = st[k]
        plain["samples_per_s"] = plain["replica_steps_per_s"] / (2 * plain["tau_int"])
        plain["tau_source"] = f"pooled autocorrelation of {args.tau_replicas} FP64 replicas"
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
    _, _, plain_achieved = roof(plain["timing"])
    try:
        copy_gbs = copy_bandwidth()
    except Exception as exc:  # report the peak fraction only
        print(f"[bench] copy bandwidth not measured: {exc}", file=sys.stderr)
        copy_gbs = None
    H_err_vs_plain = abs(fast["H_mean"] - plain["H_mean"]) / abs(plain["H_mean"])
    H_err_sigma = np.hypot(fast["H_stderr"], plain["H_stderr"]) / abs(plain["H_mean"])
    keys = ("replicas", "replica_steps_per_s", "ms_per_step", "tau_int", "tau_int_err", "tau_int_each", "samples_per_s",
            "H_mean", "H_stderr", "schedule", "tune_ms", "restarts")
    sched_name = {1: "plane", 2: "line", 3: "conv"}
    out = {
        "metric": "decorrelated CL samples/s at 96^3 N=100, FP32+learned vs plain FP64, <H> err",
        "value": fast["samples_per_s"],
        "unit": "decorrelated CL samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fast["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 propagators/FFT, f64 reductions",
        "data": "synthetic",
        "config": {"workload": "96^3 L=48 N=100 z=-0.5 salt=0.05 B=0.5 lB=1 (221 chains)",
                   "fast": f"FP32 {args.fast_scheme} dt={args.dt_fast}, {args.replicas} replicas per launch"
                           + (f", learned kernel {os.path.basename(args.lrk)}, fallback threshold {args.lrk_threshold}"
                              if lrk is not None else ""),
                   "fp32_pc": None if fp32_pc is None else f"FP32 pc dt={args.dt_plain}, {args.replicas} replicas per launch",
                   "plain": f"FP64 pc dt={args.dt_plain}, {args.plain_replicas} replica(s)",
                   "parallelism": "single GPU"},
        "fast": {k: fast[k] for k in keys},
        "fp32_pc": None if fp32_pc is None else {k: fp32_pc[k] for k in keys},
        "plain_fp64": dict({k: plain[k] for k in keys}, tau_source=plain["tau_source"],
                           **({"tau_int_timed_replicas": plain["tau_int_timed_replicas"]} if "tau_int_timed_replicas" in plain else {})),
        "learned_kernel": fast["lrk"],
        "speedup_vs_plain_fp64": fast["samples_per_s"] / plain["samples_per_s"],
        "speedup_rel_err": float(np.hypot(fast["tau_int_err"] / fast["tau_int"], plain["tau_int_err"] / plain["tau_int"])),
        "fp32_pc_speedup_vs_plain_fp64": None if fp32_pc is None else fp32_pc["samples_per_s"] / plain["samples_per_s"],
        "H_err_same_state": float(same["H"]),
        "same_state_rel_err": {k: float(v) for k, v in same.items() if k != "H_fp64"},
        "H_err_vs_plain": float(H_err_vs_plain),
        "H_err_sigma": float(H_err_sigma),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"bead-step launch group ({sched_name.get(fast['schedule'], '?')} schedule), FP32, "
                               f"{args.replicas} replica jobs per launch",
                     "us_per_launch_group": us, "alg_bytes_per_launch_group": byt,
                     "plain_fp64_achieved": plain_achieved,
                     "copy_achievable": copy_gbs,
                     "frac_of_copy": None if not copy_gbs else achieved / copy_gbs,
                     "traffic_rate": None if traffic is None else traffic / (us * 1e-6) / 1e9},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(plain_sys[0], args.cpu_steps, plain["tau_int"], "PC")
    print(json.dumps(out))


if __name__ == "__main__":
    main()

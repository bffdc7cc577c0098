"""Headline benchmark: decorrelated complex-Langevin samples/s at 96^3, N = 100.

Workload (one system, one GPU): 96^3 grid, L = 48 b, discrete Gaussian chains
N = 100, b = 1, charge -0.5 on every bead, counterions, 1:1 salt 0.05 b^-3,
bead density 0.2 b^-3 (221 chains), excluded volume B = 0.5 b^3, lB = 1 b,
a = 1 b, lambda = 1.  Synthetic initial fields (w, psi = 1e-3 complex Gaussian).

Two engines run the same workload:
  fast   complex-FP32 propagators and FFTs with per-bead renormalisation, FP64
         reductions and field accumulation, scheme and dt_fast selectable
         (default: Euler-Maruyama, dt = 0.01)
  plain  complex-FP64 engine, plain explicit Euler-Maruyama at dt_plain = 0.01
Each takes `warmup` untimed steps, then `steps` timed steps recording Re H every
step.  Decorrelated samples/s = (CL steps/s) / (2 tau_int), tau_int of the Re H
series by Sokal's automatic window (c = 6).  `value` is the fast engine's.

roofline: the dominant kernel pair is the propagator bead step (one pencil pass
plus one plane pass advancing the forward and backward sweeps together).
Algorithmic bytes per bead step = read q_s + write q_{s+1} = 2 c M (c = 8 B for
complex FP32), timed with HIP events on the engine's stream around every sweep.

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

from ftspolysol_amd.api import (Engine, PREC_FP32, PREC_FP64, SCHEME_EM, SCHEME_PC_REFERENCE,  # noqa: E402
                                SCHEME_SEMI_IMPLICIT, hip_engine)
from ftspolysol_amd.system import initial_fields, make_system  # noqa: E402

HBM_PEAK_GBS = 8000.0


def tau_int(x, c=6.0):
    """Integrated autocorrelation time (in samples) with Sokal's automatic window."""
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    if n < 8:
        return float("nan"), 0
    y = x - x.mean()
    f = np.fft.rfft(y, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n]
    if acf[0] <= 0:
        return 0.5, 0
    rho = acf / acf[0]
    tau = 0.5
    for w in range(1, n):
        tau += rho[w]
        if w >= c * tau:
            return float(tau), w
    return float(tau), n - 1


def run_engine(system, warmup, steps, label):
    eng = hip_engine(system)
    om, ps = initial_fields(system, sigma=1e-3, cfg=4)
    eng.set_fields(om, ps)
    t0 = time.perf_counter()
    done = 0
    while done < warmup:
        k = min(100, warmup - done)
        eng.step(k)
        done += k
    print(f"[bench] {label}: warmup {warmup} steps in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    eng.set_timing(True)
    H = []
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(200, steps - done)
        recs = eng.step(k, record_every=1)
        H.extend(r["H"] for r in recs)
        done += k
        print(f"[bench] {label}: {done}/{steps} timed steps", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    timing = eng.get_timing()
    eng.close()
    H = np.array(H)
    tau, window = tau_int(H.real)
    sps = steps / elapsed
    stderr_H = float(H.real.std(ddof=1) * np.sqrt(2 * tau / len(H))) if len(H) > 1 else float("nan")
    return dict(steps_per_s=sps, ms_per_step=1e3 * elapsed / steps, tau_int=tau, window=window,
                samples_per_s=sps / (2 * tau), H_mean=float(H.real.mean()), H_imag_mean=float(H.imag.mean()),
                H_stderr=stderr_H, timing=timing, elapsed=elapsed)


def cpu_baseline(system, n_steps, tau_plain):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "libfts_oracle.so"))
    eng = Engine(lib, "fto_", system)
    om, ps = initial_fields(system, sigma=1e-3, cfg=4)
    eng.set_fields(om, ps)
    t0 = time.perf_counter()
    eng.step(n_steps)
    el = time.perf_counter() - t0
    sps = n_steps / el
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": sps / (2 * tau_plain), "unit": "decorrelated CL samples/s", "cores": cores, "kind": "port",
            "steps_per_s": sps,
            "sample": f"{n_steps} plain FP64 Euler-Maruyama steps of the 96^3 N=100 workload on the host "
                      f"({el:.1f} s, own mixed-radix FFT, OpenMP); samples/s uses the GPU plain engine's tau_int"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--dt-fast", type=float, default=0.01)
    ap.add_argument("--fast-scheme", choices=["em", "si", "pc"], default="em")
    ap.add_argument("--dt-plain", type=float, default=0.01)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs one system on one GPU; multi-GPU runs are not part of this build "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    wl = dict(n=96, N=100, rho=0.2, salt=0.05, B=0.5, lB=1.0, charge="uniform", z=-0.5)
    fast_scheme = {"em": SCHEME_EM, "si": SCHEME_SEMI_IMPLICIT, "pc": SCHEME_PC_REFERENCE}[args.fast_scheme]
    fast_sys = make_system(**wl, dt=args.dt_fast, precision=PREC_FP32, scheme=fast_scheme, seed=0xC4)
    plain_sys = make_system(**wl, dt=args.dt_plain, precision=PREC_FP64, scheme=SCHEME_EM, seed=0xC4)

    fast = run_engine(fast_sys, args.warmup, args.steps, f"fast fp32 {args.fast_scheme}")
    plain = run_engine(plain_sys, args.warmup, args.steps, "plain fp64 EM")

    t = fast["timing"]
    per_pair_s = (t["sweep_ms"] * 1e-3) / (t["bead_steps"] / 2) if t["bead_steps"] else float("nan")
    bytes_per_pair = t["algorithmic_bytes"] / (t["bead_steps"] / 2) if t["bead_steps"] else float("nan")
    achieved = bytes_per_pair / per_pair_s / 1e9
    tp = plain["timing"]
    plain_pair_s = (tp["sweep_ms"] * 1e-3) / (tp["bead_steps"] / 2) if tp["bead_steps"] else float("nan")
    plain_achieved = (tp["algorithmic_bytes"] / (tp["bead_steps"] / 2)) / plain_pair_s / 1e9

    H_err = abs(fast["H_mean"] - plain["H_mean"]) / abs(plain["H_mean"])
    H_err_sigma = np.hypot(fast["H_stderr"], plain["H_stderr"]) / abs(plain["H_mean"])

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
                   "fast": f"FP32 {args.fast_scheme} dt={args.dt_fast} (learned preconditioner not yet in this engine)",
                   "plain": f"FP64 Euler-Maruyama dt={args.dt_plain}", "parallelism": "single GPU"},
        "fast": {k: fast[k] for k in ("steps_per_s", "tau_int", "samples_per_s", "H_mean", "H_stderr")},
        "plain_fp64": {k: plain[k] for k in ("steps_per_s", "tau_int", "samples_per_s", "H_mean", "H_stderr")},
        "speedup_vs_plain_fp64": fast["samples_per_s"] / plain["samples_per_s"],
        "H_err": H_err,
        "H_err_sigma": H_err_sigma,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "bead step (pencil + plane pass, fwd and bwd sweeps), FP32",
                     "us_per_launch_pair": per_pair_s * 1e6, "alg_bytes_per_pair": bytes_per_pair,
                     "plain_fp64_achieved": plain_achieved},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(plain_sys, args.cpu_steps, plain["tau_int"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()

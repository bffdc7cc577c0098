"""Training data and training of the learned linear-response kernel.

  python -m ftspolysol_amd.train_lrk --out ftspolysol_amd/lrk_c4.npy

Data: `trajectories` seeded plain semi-implicit (Debye-kernel) CL runs of the
given system, stepped together as one replica batch on the GPU with training
records on.  After `burn` steps every judged step is one record: the field
moments at its start and the |k|-binned statistics of its (field change, exact
force change) pair (FTS_LRK_RECORD in include/fts.h).  Training: Adam on the
device (fts_lrk_trainer_run) over a random 90 % of the records; the other 10 %
measure the relative force residual of the learned kernel against the Debye
kernel.  The parameters go to `out` (.npy, FP32) with a JSON of the run beside it.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .api import LRK_RECORD, SCHEME_SEMI_IMPLICIT, PREC_FP32, hip_batch, hip_lrk
from .lrk import init_params, network_r, relative_residuals, save_params
from .system import CONFIGS, initial_fields, make_system


def generate_records(n=48, N=100, rho=0.2, salt=0.05, B=0.5, lB=1.0, z=-0.5, dt=0.01, trajectories=200, steps=1000,
                     burn=100, precision=PREC_FP32, seed0=0x7A1E0000, group=50, chunk=100, log=sys.stderr, charge="uniform"):
    """Records of `trajectories` plain semi-implicit runs (groups of `group` replicas per batch)."""
    out = []
    t0 = time.perf_counter()
    for g0 in range(0, trajectories, group):
        R = min(group, trajectories - g0)
        systems = [make_system(n, N, rho=rho, salt=salt, B=B, lB=lB, charge=charge, z=z, dt=dt, scheme=SCHEME_SEMI_IMPLICIT,
                               precision=precision, seed=seed0 + g0 + r) for r in range(R)]
        b = hip_batch(systems)
        try:
            for r, (e, s) in enumerate(zip(b.replicas, systems)):
                om, ps = initial_fields(s, sigma=1e-3, seed=seed0 + 0x1000 + g0 + r)
                e.set_fields(om, ps)
                e.set_restart(50, 20)
            done = 0
            while done < burn:
                k = min(chunk, burn - done)
                b.step(k)
                done += k
            for e in b.replicas:
                e.lrk_set_recording(chunk)
            done = 0
            while done < steps:
                k = min(chunk, steps - done)
                b.step(k)
                for e in b.replicas:
                    rec = e.lrk_get_records(chunk)
                    out.append(rec[np.all(np.isfinite(rec), axis=1)])
                done += k
            restarts = sum(e.get_info().restarts for e in b.replicas)
        finally:
            b.close()
        print(f"[train_lrk] replicas {g0}..{g0 + R - 1}: {sum(len(x) for x in out)} records, {restarts} restarts, "
              f"{time.perf_counter() - t0:.1f} s", file=log, flush=True)
    return np.concatenate(out, axis=0) if out else np.zeros((0, LRK_RECORD))


def train(records, epochs=20, batch_records=64, lr0=3e-3, lr1=1e-4, seed=1, val_fraction=0.1, log=sys.stderr):
    """Train on a random (1 - val_fraction) of the records; returns (params, report)."""
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(records))
    n_val = max(1, int(round(val_fraction * len(records))))
    val, tr = records[order[:n_val]], records[order[n_val:]]
    net = hip_lrk()
    P0 = init_params(tr, seed=seed)
    t0 = time.perf_counter()
    P, hist = net.train(P0, tr, epochs=epochs, batch_records=batch_records, lr0=lr0, lr1=lr1, seed=seed)
    t_train = time.perf_counter() - t0
    rl = relative_residuals(val, network_r(net, P, val))
    rd = relative_residuals(val, None)
    report = {
        "records_train": int(len(tr)), "records_val": int(len(val)), "epochs": epochs, "batch_records": batch_records,
        "lr0": lr0, "lr1": lr1, "seed": seed, "train_seconds": t_train, "loss_history": [float(x) for x in hist],
        "val_loss": float(net.loss_grad(P, val, grad=False)), "val_loss_debye": float(net.loss_grad(P0, val, grad=False)),
        "val_rel_residual_learned": [float(x) for x in np.nanmean(rl, axis=0)],
        "val_rel_residual_debye": [float(x) for x in np.nanmean(rd, axis=0)],
    }
    print(f"[train_lrk] trained {epochs} epochs in {t_train:.1f} s: val loss {report['val_loss']:.4g} "
          f"(Debye {report['val_loss_debye']:.4g}); rel residual learned {report['val_rel_residual_learned']} "
          f"Debye {report['val_rel_residual_debye']}", file=log, flush=True)
    return P, report


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--config", type=int, choices=[4, 5], default=4,
                    help="chemistry of the training systems: 4 (N 100, charge -0.5 per bead, salt 0.05, B 0.5, lB 1) or 5 "
                         "(N 200, charge -1 on every second bead, salt 0.02, B 1, lB 0.7); --N and --salt override")
    ap.add_argument("--grid", type=int, default=48)
    ap.add_argument("--N", type=int, default=None)
    ap.add_argument("--salt", type=float, default=None)
    ap.add_argument("--dt", type=float, default=0.01)
    ap.add_argument("--trajectories", type=int, default=200)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--burn", type=int, default=100)
    ap.add_argument("--group", type=int, default=50)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch-records", type=int, default=64)
    ap.add_argument("--lr0", type=float, default=3e-3)
    ap.add_argument("--lr1", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=1)
    args = ap.parse_args(argv)
    t0 = time.perf_counter()
    chem = {k: CONFIGS[args.config][k] for k in ("N", "rho", "salt", "B", "lB", "charge", "z")}
    if args.N is not None:
        chem["N"] = args.N
    if args.salt is not None:
        chem["salt"] = args.salt
    rec = generate_records(n=args.grid, dt=args.dt, trajectories=args.trajectories, steps=args.steps, burn=args.burn,
                           group=args.group, **chem)
    t_gen = time.perf_counter() - t0
    P, report = train(rec, epochs=args.epochs, batch_records=args.batch_records, lr0=args.lr0, lr1=args.lr1,
                      seed=args.seed)
    report.update({"config": args.config, "grid": args.grid, **chem, "dt": args.dt, "trajectories": args.trajectories,
                   "steps": args.steps, "burn": args.burn, "records": int(len(rec)), "generate_seconds": t_gen,
                   "data": "plain semi-implicit (Debye kernel) CL, FP32 propagators, synthetic initial fields"})
    save_params(args.out, P, report)
    print(json.dumps(report))


if __name__ == "__main__":
    main()

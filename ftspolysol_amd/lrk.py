"""Host-side helpers of the learned linear-response kernel (FTS_SCHEME_SI_LEARNED).

The network, its training kernels and the per-step judge live in the HIP
library (include/fts.h, FTS_LRK_*); this module only prepares what the host
owns: the initial parameter block (input normalisation from the training
records, Glorot-uniform weights, a zero output layer so that an untrained
network is exactly the Debye kernel), the feature matrix of a set of records,
the relative force residuals of the learned and Debye kernels on recorded steps,
and reading / writing parameter files.
"""
from __future__ import annotations

import os

import numpy as np

from .api import (LRK_BINS, LRK_FEATURES, LRK_HIDDEN, LRK_IN, LRK_MOMENTS, LRK_NPARAMS, LRK_OFF, LRK_OUT,
                  LRK_REC_E, LRK_REC_MOMENTS, LRK_REC_S1, LRK_REC_S2, LRK_RECORD)

BIN_CENTRES = (np.arange(LRK_BINS) + 0.5) / LRK_BINS


def moments(records):
    return np.asarray(records)[:, LRK_REC_MOMENTS:LRK_REC_MOMENTS + LRK_MOMENTS]


def features(records):
    """[n * BINS, FEATURES]: bin centre then the record's field moments."""
    m = moments(records)
    n = m.shape[0]
    f = np.empty((n, LRK_BINS, LRK_FEATURES))
    f[:, :, 0] = BIN_CENTRES[None, :]
    f[:, :, 1:] = m[:, None, :]
    return f.reshape(-1, LRK_FEATURES)


def stats(records):
    """E [n, 2], S1 [n, 2, BINS], S2 [n, 2, BINS] of the recorded steps."""
    r = np.asarray(records)
    E = r[:, LRK_REC_E:LRK_REC_E + 2]
    S1 = r[:, LRK_REC_S1:LRK_REC_S1 + 2 * LRK_BINS].reshape(-1, 2, LRK_BINS)
    S2 = r[:, LRK_REC_S2:LRK_REC_S2 + 2 * LRK_BINS].reshape(-1, 2, LRK_BINS)
    return E, S1, S2


def relative_residuals(records, r):
    """Relative force residuals R_c / E_c of a kernel A_D e^{r} on recorded steps.
    r: [n, BINS, 2] (network output per bin) or None for the Debye kernel (r = 0).
    Returns [n, 2] (NaN where E_c = 0)."""
    E, S1, S2 = stats(records)
    t = np.ones_like(S1) if r is None else np.exp(np.transpose(np.asarray(r, dtype=np.float64), (0, 2, 1)))
    R = E - 2 * (t * S1).sum(-1) + (t * t * S2).sum(-1)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(E > 0, R / E, np.nan)


def init_params(records, seed=0):
    """Initial parameter block: feature normalisation from `records`, Glorot-uniform
    hidden weights (tanh gain 5/3), zero biases and a zero output layer (r = 0)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(LRK_NPARAMS, dtype=np.float64)
    f = features(records)
    mean, std = f.mean(0), f.std(0)
    std = np.where(std > 1e-12 * np.maximum(1.0, np.abs(mean)), std, 1.0)
    p[LRK_OFF["shift"]:LRK_OFF["shift"] + LRK_FEATURES] = mean
    p[LRK_OFF["scale"]:LRK_OFF["scale"] + LRK_FEATURES] = 1.0 / std

    def glorot(key, fan_out, fan_in, fan_in_used):
        a = (5.0 / 3.0) * np.sqrt(6.0 / (fan_in_used + fan_out))
        w = rng.uniform(-a, a, size=(fan_out, fan_in))
        w[:, fan_in_used:] = 0.0
        p[LRK_OFF[key]:LRK_OFF[key] + fan_out * fan_in] = w.reshape(-1)

    glorot("W1", LRK_HIDDEN, LRK_IN, LRK_FEATURES)
    glorot("W2", LRK_HIDDEN, LRK_HIDDEN, LRK_HIDDEN)
    glorot("W3", LRK_HIDDEN, LRK_HIDDEN, LRK_HIDDEN)
    return p.astype(np.float32)


def network_r(net, params, records):
    """Network output per record and bin: [n, BINS, OUT]."""
    return net.forward(params, features(records)).reshape(-1, LRK_BINS, LRK_OUT)


def save_params(path, params, meta=None):
    p = np.asarray(params, dtype=np.float32)
    if p.size != LRK_NPARAMS:
        raise ValueError("parameter block has the wrong size")
    np.save(path, p)
    if meta is not None:
        import json
        with open(os.path.splitext(path)[0] + ".json", "w") as fh:
            json.dump(meta, fh, indent=1, sort_keys=True)


def load_params(path):
    p = np.load(path, allow_pickle=False).astype(np.float32)
    if p.size != LRK_NPARAMS:
        raise ValueError(f"{path}: expected {LRK_NPARAMS} parameters")
    return p


def check_record_block(records):
    r = np.asarray(records)
    if r.ndim != 2 or r.shape[1] != LRK_RECORD:
        raise ValueError("records must be [n, FTS_LRK_RECORD]")
    return r

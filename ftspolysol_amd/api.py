"""ctypes mirror of include/fts.h and a thin Python handle over it.

The engine itself is C++/HIP (libfts_hip.so); this module only declares the
C-ABI structures and wraps the calls, so Python callers (tests, bench.py) drive
exactly what a C or C++ caller would.  `Engine` takes the loaded library and the
symbol prefix, so the same wrapper serves any library that exports this ABI.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

FTS_OK, FTS_NONFINITE, FTS_INVALID, FTS_DEVICE_ERROR = 0, 1, 2, 3
SCHEME_PC_REFERENCE, SCHEME_EM, SCHEME_SEMI_IMPLICIT, SCHEME_SI_LEARNED, SCHEME_SIPC, SCHEME_SIPC_LEARNED = 0, 1, 2, 3, 4, 5
PREC_FP64, PREC_FP32 = 0, 1
SCHED_AUTO, SCHED_PLANE, SCHED_LINE, SCHED_CONV, SCHED_SEPARABLE, SCHED_CONV_MFMA = 0, 1, 2, 3, 4, 5
SYMMETRY_AUTO, SYMMETRY_OFF = 0, 1
MAX_SPECIES = MAX_IONS = MAX_TYPES = 8
STRESS_ALL = 0x3F
STRESS_PISO = 0x40

FIELD_OMEGA, FIELD_PSI, FIELD_RHO_TYPE, FIELD_RHO_P, FIELD_RHO_C = 0, 1, 2, 3, 4
FIELD_RHO_C_POLY, FIELD_RHO_C_ION, FIELD_ION_RHO, FIELD_QF, FIELD_QB = 5, 6, 7, 8, 9

# learned linear-response kernel (FTS_LRK_* in include/fts.h)
LRK_FEATURES, LRK_IN, LRK_HIDDEN, LRK_OUT, LRK_BINS, LRK_MOMENTS = 7, 8, 64, 2, 128, 6
LRK_OFF = {"shift": 0, "scale": 8, "W1": 16, "b1": 528, "W2": 592, "b2": 4688, "W3": 4752, "b3": 8848,
           "W4": 8912, "b4": 9040}
LRK_NPARAMS = 9042
LRK_REC_MOMENTS, LRK_REC_E, LRK_REC_S1, LRK_REC_S2 = 0, 6, 8, 8 + 2 * LRK_BINS
LRK_RECORD = 8 + 4 * LRK_BINS


class SpeciesDesc(C.Structure):
    _fields_ = [("num_chains", C.c_longlong), ("N", C.c_int),
                ("types", C.POINTER(C.c_int)), ("charges", C.POINTER(C.c_double)),
                ("bonds", C.POINTER(C.c_double))]


class CounterionDesc(C.Structure):
    _fields_ = [("enabled", C.c_int), ("auto_number", C.c_int), ("number", C.c_longlong),
                ("valence", C.c_double), ("smearing_length", C.c_double)]


class SaltDesc(C.Structure):
    _fields_ = [("enabled", C.c_int), ("num_salt", C.c_longlong), ("cation_valence", C.c_double),
                ("anion_valence", C.c_double), ("cation_stoich", C.c_int), ("anion_stoich", C.c_int),
                ("smearing_length", C.c_double), ("combine_identical_species", C.c_int)]


class SystemDesc(C.Structure):
    _fields_ = [("m1", C.c_int), ("m2", C.c_int), ("m3", C.c_int), ("dL", C.c_double), ("a", C.c_double),
                ("monomer_types", C.c_int), ("xi", C.c_double), ("chi", C.POINTER(C.c_double)),
                ("eigen_tol_rel", C.c_double), ("lB", C.c_double), ("n_species", C.c_int),
                ("species", C.POINTER(SpeciesDesc)), ("ions_enabled", C.c_int),
                ("default_smearing_length", C.c_double), ("counterions", CounterionDesc), ("salt", SaltDesc),
                ("dt", C.c_double), ("n_lambda_omega", C.c_int), ("lambda_omega", C.POINTER(C.c_double)),
                ("lambda_psi", C.c_double), ("noise", C.c_int), ("scheme", C.c_int), ("precision", C.c_int),
                ("seed", C.c_ulonglong), ("stress_mask", C.c_int), ("fft_schedule", C.c_int),
                ("chain_symmetry", C.c_int)]


class Observables(C.Structure):
    _fields_ = [("step", C.c_longlong), ("n_species", C.c_int), ("n_ions", C.c_int), ("stress_mask", C.c_int),
                ("pad_", C.c_int),
                ("Q_poly", (C.c_double * 2) * MAX_SPECIES), ("Q_ion", (C.c_double * 2) * MAX_IONS),
                ("mu_ex_poly", (C.c_double * 2) * MAX_SPECIES), ("mu_id_poly", C.c_double * MAX_SPECIES),
                ("mu_self_poly", C.c_double * MAX_SPECIES),
                ("mu_ex_ion", (C.c_double * 2) * MAX_IONS), ("mu_id_ion", C.c_double * MAX_IONS),
                ("mu_self_ion", C.c_double * MAX_IONS),
                ("H", C.c_double * 2), ("stress", (C.c_double * 2) * 6), ("p_iso", C.c_double * 2)]

    def as_dict(self):
        ns, ni = self.n_species, self.n_ions
        c = lambda arr, n: np.array([complex(arr[i][0], arr[i][1]) for i in range(n)])
        return {
            "step": self.step,
            "Q_poly": c(self.Q_poly, ns), "Q_ion": c(self.Q_ion, ni),
            "mu_ex_poly": c(self.mu_ex_poly, ns), "mu_id_poly": np.array(self.mu_id_poly[:ns]),
            "mu_self_poly": np.array(self.mu_self_poly[:ns]),
            "mu_ex_ion": c(self.mu_ex_ion, ni), "mu_id_ion": np.array(self.mu_id_ion[:ni]),
            "mu_self_ion": np.array(self.mu_self_ion[:ni]),
            "H": complex(self.H[0], self.H[1]),
            "stress": c(self.stress, 6), "p_iso": complex(self.p_iso[0], self.p_iso[1]),
        }


class Info(C.Structure):
    _fields_ = [("grid_size", C.c_longlong), ("step", C.c_longlong), ("n_omega", C.c_int),
                ("monomer_types", C.c_int), ("n_species", C.c_int), ("n_ions", C.c_int),
                ("n_segment_classes", C.c_int), ("precision", C.c_int), ("scheme", C.c_int), ("fft_schedule", C.c_int),
                ("volume", C.c_double), ("rho0", C.c_double), ("short_range_self_pressure", C.c_double),
                ("mode_eigenvalue", C.c_double * MAX_TYPES), ("omega_coefficient", C.c_double * MAX_TYPES),
                ("gamma_is_imag", C.c_double * MAX_TYPES), ("mode_O", (C.c_double * MAX_TYPES) * MAX_TYPES),
                ("ion_valence", C.c_double * MAX_IONS), ("ion_count", C.c_longlong * MAX_IONS),
                ("ion_smearing", C.c_double * MAX_IONS), ("tune_ms", C.c_double * 2),
                ("mirrored_species", C.c_int), ("restarts", C.c_int), ("noise_epoch", C.c_int),
                ("checkpoint_step", C.c_longlong), ("failed_step", C.c_longlong)]


class Timing(C.Structure):
    _fields_ = [("sweep_ms", C.c_double), ("bead_steps", C.c_longlong), ("sweeps", C.c_longlong),
                ("algorithmic_bytes", C.c_double), ("step_launches", C.c_longlong)]


class LrkStats(C.Structure):
    _fields_ = [("judged", C.c_longlong), ("learned_steps", C.c_longlong), ("fallback_steps", C.c_longlong),
                ("records", C.c_longlong), ("last_excess", C.c_double),
                ("sum_rel_residual_learned", C.c_double * 2), ("sum_rel_residual_debye", C.c_double * 2)]

    def as_dict(self):
        return {"judged": self.judged, "learned_steps": self.learned_steps, "fallback_steps": self.fallback_steps,
                "records": self.records, "last_excess": self.last_excess,
                "sum_rel_residual_learned": np.array(self.sum_rel_residual_learned[:]),
                "sum_rel_residual_debye": np.array(self.sum_rel_residual_debye[:])}


class RunSettings(C.Structure):
    _fields_ = [("Nt", C.c_int), ("Nss", C.c_int), ("Ns", C.c_int), ("Ntmp", C.c_int), ("Nso", C.c_int),
                ("N_stat", C.c_int), ("max_nan_restarts", C.c_int), ("legacy_ab", C.c_int), ("n_ion_names", C.c_int)]


MAX_BATCH_GROUPS = 16
BATCH_DEFAULT_STREAMS = 8
BATCH_MAX_STREAMS = 16


class BatchInfo(C.Structure):
    _fields_ = [("n_replicas", C.c_int), ("n_groups", C.c_int), ("group_size", C.c_int * MAX_BATCH_GROUPS),
                ("group_mesh", (C.c_int * 3) * MAX_BATCH_GROUPS), ("group_precision", C.c_int * MAX_BATCH_GROUPS),
                ("group_schedule", C.c_int * MAX_BATCH_GROUPS), ("group_tune_ms", (C.c_double * 2) * MAX_BATCH_GROUPS)]


class FtsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fts error {code}: {msg}")
        self.code = code


def _bind(lib, prefix):
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "create": (C.c_int, [P(SystemDesc), P(vp)]),
        "destroy": (None, [vp]),
        "last_error": (C.c_char_p, []),
        "get_info": (C.c_int, [vp, P(Info)]),
        "set_fields": (C.c_int, [vp, P(C.c_double), P(C.c_double)]),
        "get_field": (C.c_int, [vp, C.c_int, P(C.c_double), C.c_size_t]),
        "compute_densities": (C.c_int, [vp, P(C.c_double)]),
        "measure": (C.c_int, [vp, P(Observables)]),
        "step": (C.c_int, [vp, C.c_int, C.c_int, P(Observables), C.c_int, P(C.c_int)]),
        "set_step": (C.c_int, [vp, C.c_longlong]),
        "set_restart": (C.c_int, [vp, C.c_int, C.c_int]),
        "inject_nonfinite": (C.c_int, [vp, C.c_longlong]),
        "set_drift_stats": (C.c_int, [vp, C.c_int]),
        "save_state": (C.c_int, [vp, C.c_char_p]),
        "load_state": (C.c_int, [vp, C.c_char_p]),
        "get_drift_rows": (C.c_int, [vp, P(C.c_double), C.c_int, P(C.c_int)]),
        "fft3d": (C.c_int, [vp, P(C.c_double), P(C.c_double), C.c_int]),
        "bead_step": (C.c_int, [vp, P(C.c_double), P(C.c_double), C.c_double, P(C.c_double)]),
        "bead_step_noalloc": (C.c_int, [vp, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]),
        "set_timing": (C.c_int, [vp, C.c_int]),
        "get_timing": (C.c_int, [vp, P(Timing)]),
        "batch_create": (C.c_int, [P(SystemDesc), C.c_int, P(vp)]),
        "batch_destroy": (None, [vp]),
        "batch_get_info": (C.c_int, [vp, P(BatchInfo)]),
        "batch_replica": (C.c_int, [vp, C.c_int, P(vp), P(C.c_int)]),
        "batch_step": (C.c_int, [vp, C.c_int, C.c_int, P(Observables), C.c_int, P(C.c_int)]),
        "batch_set_timing": (C.c_int, [vp, C.c_int]),
        "batch_set_graph": (C.c_int, [vp, C.c_int]),
        "batch_set_streams": (C.c_int, [vp, C.c_int]),
        "batch_get_timing": (C.c_int, [vp, C.c_int, P(Timing)]),
        "deck_read": (C.c_int, [C.c_char_p, C.c_char_p, P(vp)]),
        "deck_destroy": (None, [vp]),
        "deck_system": (C.c_int, [vp, P(SystemDesc)]),
        "deck_settings": (C.c_int, [vp, P(RunSettings)]),
        "deck_string": (C.c_char_p, [vp, C.c_char_p]),
        "rf_read": (C.c_int, [C.c_char_p, P(C.c_int), P(C.c_int), P(C.c_double), C.c_size_t]),
        "rf_write": (C.c_int, [C.c_char_p, P(C.c_int), P(C.c_double), C.c_int, P(C.c_double)]),
        "field_resize": (C.c_int, [P(C.c_double), P(C.c_int), C.c_int, P(C.c_double), P(C.c_int)]),
        "lrk_set_network": (C.c_int, [vp, P(C.c_float), C.c_double]),
        "lrk_set_recording": (C.c_int, [vp, C.c_int]),
        "lrk_get_records": (C.c_int, [vp, P(C.c_double), C.c_int, P(C.c_int)]),
        "lrk_get_stats": (C.c_int, [vp, P(LrkStats)]),
        "lrk_forward": (C.c_int, [P(C.c_float), P(C.c_float), C.c_int, P(C.c_float)]),
        "lrk_forward64": (C.c_int, [P(C.c_double), P(C.c_double), C.c_int, P(C.c_double)]),
        "lrk_loss_grad": (C.c_int, [P(C.c_float), P(C.c_double), C.c_int, P(C.c_double), P(C.c_float)]),
        "lrk_loss_grad64": (C.c_int, [P(C.c_double), P(C.c_double), C.c_int, P(C.c_double), P(C.c_double)]),
        "lrk_trainer_create": (C.c_int, [P(C.c_double), C.c_int, P(vp)]),
        "lrk_trainer_destroy": (None, [vp]),
        "lrk_trainer_run": (C.c_int, [vp, P(C.c_float), C.c_int, C.c_int, C.c_double, C.c_double, C.c_ulonglong,
                                      P(C.c_double)]),
    }
    fns = {}
    for name, (res, args) in sig.items():
        sym = prefix + name
        if not hasattr(lib, sym):
            continue
        f = getattr(lib, sym)
        f.restype = res
        f.argtypes = args
        fns[name] = f
    return fns


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Engine:
    """One engine instance over a library exporting the fts C ABI."""

    def __init__(self, lib, prefix, system, handle=None, owner=None):
        self._fns = _bind(lib, prefix)
        self._keep = system  # keeps the ctypes arrays of the description alive
        self._owner = owner  # a batch that owns the handle (borrowed replica)
        if handle is None:
            h = C.c_void_p()
            rc = self._fns["create"](C.byref(system.desc), C.byref(h))
            self._check(rc)
        else:
            h = handle
        self._h = h
        self.info = self.get_info()
        self.M = int(self.info.grid_size)
        self.shape = (system.desc.m1, system.desc.m2, system.desc.m3)

    def _check(self, rc):
        if rc != FTS_OK:
            msg = self._fns["last_error"]()
            raise FtsError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "_h", None):
            if self._owner is None:
                self._fns["destroy"](self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_info(self):
        info = Info()
        self._check(self._fns["get_info"](self._h, C.byref(info)))
        return info

    def set_fields(self, omega, psi=None):
        om = np.ascontiguousarray(np.asarray(omega, dtype=np.complex128)).view(np.float64)
        ps = None if psi is None else np.ascontiguousarray(np.asarray(psi, dtype=np.complex128)).view(np.float64)
        self._check(self._fns["set_fields"](self._h, _dptr(om), None if ps is None else _dptr(ps)))

    def field_count(self, fid):
        i = self.info
        return {FIELD_OMEGA: i.n_omega, FIELD_RHO_TYPE: i.monomer_types, FIELD_ION_RHO: i.n_ions}.get(fid, 1)

    def get_field(self, fid, count=None):
        if count is None:
            count = self.field_count(fid)
        out = np.zeros(count * self.M, dtype=np.complex128)
        v = out.view(np.float64)
        self._check(self._fns["get_field"](self._h, fid, _dptr(v), v.size))
        return out.reshape(count, *self.shape)

    def compute_densities(self):
        q = np.zeros(2 * MAX_SPECIES)
        self._check(self._fns["compute_densities"](self._h, _dptr(q)))
        n = self.info.n_species
        return q[:2 * n:2] + 1j * q[1:2 * n:2]

    def measure(self):
        rec = Observables()
        self._check(self._fns["measure"](self._h, C.byref(rec)))
        return rec.as_dict()

    def step(self, n, record_every=0, capacity=None):
        if capacity is None:
            capacity = n // record_every if record_every > 0 else 0
        recs = (Observables * max(1, capacity))()
        nrec = C.c_int(0)
        rc = self._fns["step"](self._h, n, record_every, recs, capacity, C.byref(nrec))
        out = [recs[i].as_dict() for i in range(nrec.value)]
        self._check(rc)
        return out

    def set_step(self, step):
        self._check(self._fns["set_step"](self._h, step))

    def set_restart(self, checkpoint_every, max_restarts):
        self._check(self._fns["set_restart"](self._h, checkpoint_every, max_restarts))

    def inject_nonfinite(self, step):
        self._check(self._fns["inject_nonfinite"](self._h, step))

    def save_state(self, path):
        """Fields, step counter and noise key to a binary file (exact resume)."""
        self._check(self._fns["save_state"](self._h, os.fsencode(path)))

    def load_state(self, path):
        self._check(self._fns["load_state"](self._h, os.fsencode(path)))

    def set_drift_stats(self, enabled):
        """Collect the drift statistics (drift_max.dat rows) of every step from now on."""
        self._check(self._fns["set_drift_stats"](self._h, int(enabled)))

    def get_drift_rows(self, capacity=100000):
        """Rows [step, max per mode..., max psi] collected since the last call, oldest first."""
        w = self.info.n_omega + 2
        out = np.zeros((capacity, w))
        n = C.c_int(0)
        self._check(self._fns["get_drift_rows"](self._h, _dptr(out), capacity, C.byref(n)))
        return out[:n.value].copy()

    def fft3d(self, x, sign):
        xin = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).reshape(-1)
        out = np.zeros_like(xin)
        self._check(self._fns["fft3d"](self._h, _dptr(xin.view(np.float64)), _dptr(out.view(np.float64)), sign))
        return out.reshape(self.shape)

    def bead_step(self, q, exp_minus_w, bond=1.0):
        qi = np.ascontiguousarray(np.asarray(q, dtype=np.complex128)).reshape(-1)
        ei = np.ascontiguousarray(np.asarray(exp_minus_w, dtype=np.complex128)).reshape(-1)
        out = np.zeros_like(qi)
        self._check(self._fns["bead_step"](self._h, _dptr(qi.view(np.float64)), _dptr(ei.view(np.float64)),
                                           float(bond), _dptr(out.view(np.float64))))
        return out.reshape(self.shape)

    def bead_step_noalloc(self, q_dev, exp_minus_w_dev, bond, q_out_dev, stream=None):
        """fts_bead_step_noalloc on device addresses (ints); enqueued on `stream` (None: the
        engine's stream) without waiting."""
        self._check(self._fns["bead_step_noalloc"](self._h, q_dev, exp_minus_w_dev, float(bond), q_out_dev, stream))

    def set_timing(self, enabled):
        self._check(self._fns["set_timing"](self._h, int(enabled)))

    def get_timing(self):
        t = Timing()
        self._check(self._fns["get_timing"](self._h, C.byref(t)))
        return _timing_dict(t)

    # ---- learned linear-response kernel ----
    def lrk_set_network(self, params, threshold=1e-3):
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float32))
        if p.size != LRK_NPARAMS:
            raise ValueError(f"expected {LRK_NPARAMS} parameters, got {p.size}")
        self._check(self._fns["lrk_set_network"](self._h, _fptr(p), float(threshold)))

    def lrk_set_recording(self, capacity):
        self._check(self._fns["lrk_set_recording"](self._h, int(capacity)))

    def lrk_get_records(self, max_records):
        out = np.zeros((max(1, max_records), LRK_RECORD))
        n = C.c_int(0)
        self._check(self._fns["lrk_get_records"](self._h, _dptr(out), int(max_records), C.byref(n)))
        return out[:n.value].copy()

    def lrk_get_stats(self):
        st = LrkStats()
        self._check(self._fns["lrk_get_stats"](self._h, C.byref(st)))
        return st.as_dict()


class LrkNet:
    """The learned kernel's network functions of a library exporting the fts ABI
    (no engine needed): forward pass, loss and gradient, and the Adam trainer."""

    def __init__(self, lib, prefix):
        self._fns = _bind(lib, prefix)

    def _check(self, rc):
        if rc != FTS_OK:
            msg = self._fns["last_error"]()
            raise FtsError(rc, msg.decode() if msg else "")

    def forward(self, params, features):
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float32))
        f = np.ascontiguousarray(np.asarray(features, dtype=np.float32).reshape(-1, LRK_FEATURES))
        out = np.zeros((f.shape[0], LRK_OUT), dtype=np.float32)
        self._check(self._fns["lrk_forward"](_fptr(p), _fptr(f), f.shape[0], _fptr(out)))
        return out

    def forward64(self, params, features):
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
        f = np.ascontiguousarray(np.asarray(features, dtype=np.float64).reshape(-1, LRK_FEATURES))
        out = np.zeros((f.shape[0], LRK_OUT))
        self._check(self._fns["lrk_forward64"](_dptr(p), _dptr(f), f.shape[0], _dptr(out)))
        return out

    def loss_grad(self, params, records, grad=True):
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float32))
        r = np.ascontiguousarray(np.asarray(records, dtype=np.float64).reshape(-1, LRK_RECORD))
        loss = C.c_double(0.0)
        g = np.zeros(LRK_NPARAMS, dtype=np.float32) if grad else None
        self._check(self._fns["lrk_loss_grad"](_fptr(p), _dptr(r), r.shape[0], C.byref(loss),
                                               _fptr(g) if grad else None))
        return (loss.value, g) if grad else loss.value

    def loss_grad64(self, params, records):
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
        r = np.ascontiguousarray(np.asarray(records, dtype=np.float64).reshape(-1, LRK_RECORD))
        loss = C.c_double(0.0)
        g = np.zeros(LRK_NPARAMS)
        self._check(self._fns["lrk_loss_grad64"](_dptr(p), _dptr(r), r.shape[0], C.byref(loss), _dptr(g)))
        return loss.value, g

    def train(self, params, records, epochs, batch_records=64, lr0=3e-3, lr1=3e-4, seed=1):
        """Adam on the device; returns (trained params, per-epoch mean loss)."""
        r = np.ascontiguousarray(np.asarray(records, dtype=np.float64).reshape(-1, LRK_RECORD))
        h = C.c_void_p()
        self._check(self._fns["lrk_trainer_create"](_dptr(r), r.shape[0], C.byref(h)))
        try:
            p = np.ascontiguousarray(np.array(params, dtype=np.float32))
            hist = np.zeros(max(1, epochs))
            self._check(self._fns["lrk_trainer_run"](h, _fptr(p), int(epochs), int(batch_records), float(lr0),
                                                     float(lr1), int(seed), _dptr(hist)))
        finally:
            self._fns["lrk_trainer_destroy"](h)
        return p, hist[:epochs]


class Deck:
    """A parameter file plus its sequence file (fts_deck_*): `system` is an object
    with the fts_system description (`desc`), usable wherever a System is."""

    def __init__(self, lib, prefix, param_file, sequence_file=None):
        self._fns = _bind(lib, prefix)
        h = C.c_void_p()
        rc = self._fns["deck_read"](os.fsencode(param_file), None if sequence_file is None else os.fsencode(sequence_file),
                                    C.byref(h))
        if rc != FTS_OK:
            raise FtsError(rc, self._fns["last_error"]().decode())
        self._h = h
        desc = SystemDesc()
        self._fns["deck_system"](self._h, C.byref(desc))
        self.settings = RunSettings()
        self._fns["deck_settings"](self._h, C.byref(self.settings))
        deck = self

        class _System:
            pass

        self.system = _System()
        self.system.desc = desc
        self.system.m = (desc.m1, desc.m2, desc.m3)
        self.system.dL = desc.dL
        self.system._deck = deck

    def string(self, key):
        v = self._fns["deck_string"](self._h, key.encode())
        return None if v is None else v.decode()

    def close(self):
        if getattr(self, "_h", None):
            self._fns["deck_destroy"](self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rf_read(lib, path, mesh=(0, 0, 0), prefix="fts_"):
    """Fields of an .rf table: complex array [n_fields, m1, m2, m3]."""
    fns = _bind(lib, prefix)
    m = (C.c_int * 3)(*mesh)
    nf = C.c_int(0)
    rc = fns["rf_read"](os.fsencode(path), m, C.byref(nf), None, 0)
    if rc != FTS_OK:
        raise FtsError(rc, fns["last_error"]().decode())
    shape = (nf.value, m[0], m[1], m[2])
    out = np.zeros(shape, dtype=np.complex128)
    v = out.reshape(-1).view(np.float64)
    rc = fns["rf_read"](os.fsencode(path), m, C.byref(nf), _dptr(v), v.size)
    if rc != FTS_OK:
        raise FtsError(rc, fns["last_error"]().decode())
    return out


def rf_write(lib, path, fields, L, prefix="fts_"):
    fns = _bind(lib, prefix)
    f = np.ascontiguousarray(np.asarray(fields, dtype=np.complex128))
    if f.ndim == 3:
        f = f[None]
    m = (C.c_int * 3)(*f.shape[1:])
    Ls = (C.c_double * 3)(*L)
    rc = fns["rf_write"](os.fsencode(path), m, Ls, f.shape[0], _dptr(f.reshape(-1).view(np.float64)))
    if rc != FTS_OK:
        raise FtsError(rc, fns["last_error"]().decode())


def field_resize(lib, fields, mesh, prefix="fts_"):
    fns = _bind(lib, prefix)
    f = np.ascontiguousarray(np.asarray(fields, dtype=np.complex128))
    if f.ndim == 3:
        f = f[None]
    out = np.zeros((f.shape[0],) + tuple(mesh), dtype=np.complex128)
    n = (C.c_int * 3)(*f.shape[1:])
    m = (C.c_int * 3)(*mesh)
    rc = fns["field_resize"](_dptr(f.reshape(-1).view(np.float64)), n, f.shape[0], _dptr(out.reshape(-1).view(np.float64)), m)
    if rc != FTS_OK:
        raise FtsError(rc, fns["last_error"]().decode())
    return out


def _timing_dict(t):
    return {"sweep_ms": t.sweep_ms, "bead_steps": t.bead_steps, "sweeps": t.sweeps,
            "algorithmic_bytes": t.algorithmic_bytes, "step_launches": t.step_launches}


class Batch:
    """A replica batch (fts_batch_*): replicas[r] are borrowed Engines on the batch's replicas."""

    def __init__(self, lib, prefix, systems):
        self._fns = _bind(lib, prefix)
        self._keep = list(systems)
        arr = (SystemDesc * len(systems))(*[s.desc for s in systems])
        h = C.c_void_p()
        rc = self._fns["batch_create"](arr, len(systems), C.byref(h))
        self._check(rc)
        self._h = h
        self.replicas = []
        self.group_of = []
        for r, s in enumerate(systems):
            eh, grp = C.c_void_p(), C.c_int()
            self._check(self._fns["batch_replica"](self._h, r, C.byref(eh), C.byref(grp)))
            self.replicas.append(Engine(lib, prefix, s, handle=eh, owner=self))
            self.group_of.append(grp.value)
        self.info = BatchInfo()
        self._check(self._fns["batch_get_info"](self._h, C.byref(self.info)))

    def _check(self, rc):
        if rc != FTS_OK:
            msg = self._fns["last_error"]()
            raise FtsError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "_h", None):
            for e in self.replicas:
                e._h = None
            self._fns["batch_destroy"](self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n, record_every=0, capacity=None):
        """n steps of every replica; returns one list of records per replica."""
        if capacity is None:
            capacity = n // record_every if record_every > 0 else 0
        R = len(self.replicas)
        recs = (Observables * max(1, R * capacity))()
        nrec = (C.c_int * R)()
        rc = self._fns["batch_step"](self._h, n, record_every, recs, capacity, nrec)
        out = [[recs[r * capacity + i].as_dict() for i in range(nrec[r])] for r in range(R)]
        self._check(rc)
        return out

    def set_timing(self, enabled):
        self._check(self._fns["batch_set_timing"](self._h, int(enabled)))

    def set_graph(self, enabled):
        self._check(self._fns["batch_set_graph"](self._h, int(enabled)))

    def set_streams(self, n):
        """Side streams per group for the replicas' own work (0: all on the group stream)."""
        self._check(self._fns["batch_set_streams"](self._h, int(n)))

    def get_timing(self, group=0):
        t = Timing()
        self._check(self._fns["batch_get_timing"](self._h, group, C.byref(t)))
        return _timing_dict(t)


_HIP_LIB = None


def hip_library_path():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfts_hip.so")


def load_hip_library():
    """Load the in-tree HIP engine; raise if it has not been built."""
    global _HIP_LIB
    if _HIP_LIB is None:
        path = hip_library_path()
        if not os.path.exists(path):
            raise RuntimeError(f"HIP engine library missing: {path} (run __graft_entry__.build())")
        _HIP_LIB = C.CDLL(path)
    return _HIP_LIB


def hip_engine(system):
    """An engine on the GPU (libfts_hip.so, symbols fts_*)."""
    return Engine(load_hip_library(), "fts_", system)


def hip_batch(systems):
    """A replica batch on the GPU."""
    return Batch(load_hip_library(), "fts_", systems)


def hip_lrk():
    """The learned kernel's network kernels on the GPU."""
    return LrkNet(load_hip_library(), "fts_")

"""fts_bead_step_noalloc (the allocation-free bead step on device arrays) against
fts_bead_step, through the C ABI: two calls in a row on one GPU, the second on a
smaller grid (the shared scratch array still holds the first call's larger
contents), in both precisions; the results must be equal bit for bit (the same
kernels on the same values).  Device buffers are the caller's, allocated here with
the HIP runtime the engine library loaded."""
import ctypes

import numpy as np
import pytest

from ftspolysol_amd.api import FTS_INVALID, FtsError, PREC_FP32, PREC_FP64, hip_engine, load_hip_library
from ftspolysol_amd.system import System

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipDeviceSynchronize of the runtime already in the process."""

    def __init__(self):
        load_hip_library()
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc})")

    def upload(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        p = ctypes.c_void_p()
        self.ok(self.lib.hipMalloc(ctypes.byref(p), x.nbytes), "hipMalloc")
        self.bufs.append(p)
        self.ok(self.lib.hipMemcpy(p, x.ctypes.data, x.nbytes, 1), "hipMemcpy H2D")
        return p.value

    def empty(self, n):
        p = ctypes.c_void_p()
        self.ok(self.lib.hipMalloc(ctypes.byref(p), 16 * n), "hipMalloc")
        self.bufs.append(p)
        return p.value

    def download(self, ptr, shape):
        out = np.zeros(shape, dtype=np.complex128)
        self.ok(self.lib.hipDeviceSynchronize(), "hipDeviceSynchronize")
        self.ok(self.lib.hipMemcpy(out.ctypes.data, ctypes.c_void_p(ptr), out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def bare(m, precision):
    return System(m=m, dL=0.5, species=[dict(num_chains=1, types=[0, 0], charges=[0.0, 0.0])], xi=0.1,
                  precision=precision)


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_FP32])
def test_noalloc_bead_step_matches_bead_step(precision):
    hip = Hip()
    try:
        # 96^3 and 32^3 on the conv schedule, then 16 x 24 x 48 on the separable one
        for m, bond, seed in (((96, 96, 96), 1.1, 5), ((32, 32, 32), 0.8, 6), ((16, 24, 48), 0.9, 7)):
            print(f"no-alloc bead step {m} precision={precision}", flush=True)
            e = hip_engine(bare(m, precision))
            rng = np.random.default_rng(seed)
            q = 1 + 0.1 * (rng.standard_normal(m) + 1j * rng.standard_normal(m))
            emw = np.exp(-0.2 * (rng.standard_normal(m) + 1j * rng.standard_normal(m)))
            ref = e.bead_step(q, emw, bond)
            dq, de, do = hip.upload(q), hip.upload(emw), hip.empty(q.size)
            e.bead_step_noalloc(dq, de, bond, do)
            got = hip.download(do, m)
            assert np.array_equal(got.view(np.float64), ref.view(np.float64)), np.abs(got - ref).max()
            # the caller's inputs are not modified
            assert np.array_equal(hip.download(dq, m), q)
            e.close()
    finally:
        hip.free()


def test_noalloc_rejects_grids_beyond_its_scratch():
    print("no-alloc bead step rejects 128^3", flush=True)
    m = (128, 128, 128)
    e = hip_engine(bare(m, PREC_FP32))
    hip = Hip()
    try:
        n = m[0] * m[1] * m[2]
        bufs = [hip.empty(n) for _ in range(3)]   # full-size buffers: the call is refused before it reads them
        with pytest.raises(FtsError) as err:
            e.bead_step_noalloc(bufs[0], bufs[1], 1.0, bufs[2])
        assert err.value.code == FTS_INVALID and "exceeds the no-alloc scratch" in str(err.value)
    finally:
        hip.free()

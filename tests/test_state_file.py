"""Binary state files (fts_save_state / fts_load_state): exact resume.

A run of 10 steps must equal 6 steps, a save, a load into a fresh engine (built
from a system with another seed: the file's noise key wins) and 4 more steps, bit
for bit, in fields and records; that includes a state saved after a non-finite
restart (noise epoch 1).  A file of another mesh, a truncated file and a file
that is not a state file are refused.  On the GPU, a state written by the oracle
loads into the engine and the two continue on the same trajectory.
"""
import os

import numpy as np
import pytest

from ftspolysol_amd.api import FIELD_OMEGA, FIELD_PSI, FTS_INVALID, SCHEME_EM, SCHEME_PC_REFERENCE, SCHEME_SIPC, FtsError, hip_engine
from ftspolysol_amd.system import initial_fields, make_system
from oracle_lib import oracle_engine


def small(scheme, seed=5, n=12):
    return make_system(n, 6, rho=0.2, salt=0.05, B=0.5, lB=1.0, dt=0.01, scheme=scheme, seed=seed)


def snapshot(e):
    return e.get_field(FIELD_OMEGA), e.get_field(FIELD_PSI)


@pytest.mark.parametrize("scheme", [SCHEME_PC_REFERENCE, SCHEME_EM, SCHEME_SIPC])
def test_oracle_resume_is_exact(scheme, tmp_path):
    s = small(scheme)
    a = oracle_engine(s)
    a.set_fields(*initial_fields(s, sigma=0.05, seed=3))
    a.set_restart(3, 2)
    a.inject_nonfinite(4)
    a.step(6)
    assert a.get_info().noise_epoch == 1
    path = str(tmp_path / "run.state")
    a.save_state(path)
    ra = a.step(4, record_every=1)
    b = oracle_engine(small(scheme, seed=999))
    b.load_state(path)
    info = b.get_info()
    assert (info.step, info.noise_epoch, info.restarts) == (6, 1, 1)
    rb = b.step(4, record_every=1)
    for x, y in zip(snapshot(a), snapshot(b)):
        assert np.array_equal(x, y)
    assert [r["H"] for r in ra] == [r["H"] for r in rb] and [r["step"] for r in rb] == [7, 8, 9, 10]


def test_oracle_refuses_bad_state_files(tmp_path):
    s = small(SCHEME_EM)
    a = oracle_engine(s)
    a.set_fields(*initial_fields(s, sigma=0.05, seed=3))
    path = str(tmp_path / "a.state")
    a.save_state(path)
    other = oracle_engine(small(SCHEME_EM, n=16))
    with pytest.raises(FtsError) as ei:
        other.load_state(path)
    assert ei.value.code == FTS_INVALID and "mesh" in str(ei.value)
    data = open(path, "rb").read()
    open(path + ".cut", "wb").write(data[:-8])
    with pytest.raises(FtsError) as ei:
        a.load_state(path + ".cut")
    assert "truncated" in str(ei.value)
    open(path + ".txt", "wb").write(b"12 12 12\n" + data[9:])
    with pytest.raises(FtsError):
        a.load_state(path + ".txt")
    assert not os.path.exists(path + ".new")


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", [SCHEME_PC_REFERENCE, SCHEME_EM, SCHEME_SIPC])
def test_engine_resume_is_exact_and_reads_oracle_states(scheme, tmp_path):
    print(f"state file scheme={scheme}", flush=True)
    s = small(scheme)
    g = hip_engine(s)
    g.set_fields(*initial_fields(s, sigma=0.05, seed=3))
    g.step(6)
    path = str(tmp_path / "g.state")
    g.save_state(path)
    g.step(4)
    h = hip_engine(small(scheme, seed=999))
    h.load_state(path)
    h.step(4)
    for x, y in zip(snapshot(g), snapshot(h)):
        assert np.array_equal(x, y)
    # the oracle's state file drives the engine: the same trajectory to the parity bar
    o = oracle_engine(s)
    o.set_fields(*initial_fields(s, sigma=0.05, seed=3))
    o.step(3)
    op = str(tmp_path / "o.state")
    o.save_state(op)
    e = hip_engine(small(scheme, seed=999))
    e.load_state(op)
    assert e.get_info().step == 3
    for x in (o, e):
        x.step(2)
    for x, y in zip(snapshot(e), snapshot(o)):
        assert np.abs(x - y).max() <= 1e-10 * np.abs(y).max()

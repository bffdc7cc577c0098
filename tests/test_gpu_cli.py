"""GPU end-to-end test of the command-line driver (ftspolysol_amd/fts_run, the
counterpart of the reference's main program): a deck in the reference's format
runs Nt predictor-corrector steps; observables.dat has the reference's header,
a row every Nss steps (complexLangevin.cu:687) and the ensemble average over
the records after step Nss (complexLangevin.cu:426-470); the Hamiltonian rows
match the oracle driven by the same deck to the file's 11 significant digits;
checkpoint and final .rf files are written with their mesh headers.
"""
import os
import subprocess

import numpy as np
import pytest

from ftspolysol_amd.api import Deck, load_hip_library, rf_read
from oracle_lib import oracle_engine
from test_frontend import c1_like

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "ftspolysol_amd", "fts_run")


def test_cli_run_matches_oracle(tmp_path):
    ref, deck_path = c1_like(tmp_path)
    # drift statistics after step 4 (complexLangevin.cu:466-497)
    text = open(deck_path).read().replace("Ntmp = 10\n", "Ntmp = 10\nN_stat = 4\n")
    text = text.replace('[output]\n', '[output]\ndrift_max_file = "drift.dat"\n')
    open(deck_path, "w").write(text)
    print("[cli] running fts_run on", deck_path, flush=True)
    r = subprocess.run([CLI, "-p", deck_path, "-i", "run1", "-s", str(tmp_path / "runs")], cwd=str(tmp_path),
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "total time =" in r.stdout
    run = tmp_path / "runs" / "run1"
    assert (run / "run1.param").exists() and (run / "run1.sequence").exists()
    lines = (run / "obs # not a comment.dat").read_text().splitlines()
    assert lines[0] == "# observables Nss=2 Ns=20"
    labels = lines[1].lstrip("#").split()
    assert labels[0] == "step" and labels[1:4] == ["mu_poly_ex", "mu_poly_id", "mu_poly_self"]
    assert labels[-4:] == ["Hmf_density_imag", "stress_xx", "stress_xy", "pressure_iso"]
    rows = [ln.split() for ln in lines[2:12]]
    assert [int(x[0]) for x in rows] == list(range(2, 21, 2))
    vals = np.array([[float(v) for v in x[1:]] for x in rows])
    assert vals.shape[1] == len(labels) - 1
    avg_line = [ln for ln in lines if ln.split() and ln.split()[0] == "average"][0]
    avg = np.array([float(v) for v in avg_line.split()[1:]])
    assert "# ensemble_average Nss=2 Ns=20 count=9" in lines
    np.testing.assert_allclose(avg, vals[1:].mean(axis=0), rtol=1e-9, atol=1e-12)
    # the oracle on the same deck, fields and noise
    deck = Deck(load_hip_library(), "fts_", deck_path)
    o = oracle_engine(deck.system)
    recs = o.step(20, record_every=2)
    iH = labels.index("Hmf_density_real") - 1
    H = np.array([x["H"].real for x in recs])
    np.testing.assert_allclose(vals[:, iH], H, rtol=2e-10)
    np.testing.assert_allclose(vals[:, labels.index("pressure_iso") - 1], [x["p_iso"].real for x in recs], rtol=2e-10)
    # drift_max.dat: header, then one row per step after N_stat, against the oracle's rows
    dl = (run / "drift.dat").read_text().splitlines()
    assert dl[0].lstrip("#").split() == ["step", "u_omega0_max", "u_psi_max"]
    drows = np.array([[float(v) for v in ln.split()] for ln in dl[1:]])
    assert list(drows[:, 0]) == list(range(5, 21))
    o2 = oracle_engine(deck.system)
    o2.set_drift_stats(True)
    o2.step(20)
    orows = o2.get_drift_rows()
    np.testing.assert_allclose(drows[:, 1:], orows[4:, 1:], rtol=2e-9)
    # final fields and checkpoint files (mesh headers, %.10f rows)
    lib = load_hip_library()
    w = rf_read(lib, str(run / "omega.rf"))
    assert w.shape == (1, 8, 8, 8)
    assert np.abs((w[0] - o.get_field(0)[0]).view(np.float64)).max() <= 6e-11
    assert (run / "tmp_omega.rf").exists() and (run / "tmp_psi.rf").exists()
    for name in ("psi.rf", "rho_type.rf", "rhop.rf", "rhoc.rf", "rhoc_poly.rf", "rhoc_ion.rf", "rho_ion_0_counterion.rf",
                 "rho_ion_1_salt_anion.rf"):
        assert (run / name).exists(), name


def test_cli_resume_repeats_the_uninterrupted_run(tmp_path):
    """Nt = 20 in one run, against Nt = 10 and a second run resumed from its
    checkpoint.state: the observables rows of steps 12-20 are the same text."""
    ref, deck_path = c1_like(tmp_path)
    full = subprocess.run([CLI, "-p", deck_path, "-i", "full", "-s", str(tmp_path / "runs")], cwd=str(tmp_path),
                          stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert full.returncode == 0, full.stdout
    half_deck = tmp_path / "half.param"
    half_deck.write_text(open(deck_path).read().replace("Nt = 20\n", "Nt = 10\n"))
    half = subprocess.run([CLI, "-p", str(half_deck), "-i", "half", "-s", str(tmp_path / "runs")], cwd=str(tmp_path),
                          stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert half.returncode == 0, half.stdout
    state = tmp_path / "runs" / "half" / "checkpoint.state"
    assert state.exists()
    print("[cli] resuming from", state, flush=True)
    rest = subprocess.run([CLI, "-p", deck_path, "-i", "rest", "-s", str(tmp_path / "runs"), "--resume", str(state)],
                          cwd=str(tmp_path), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert rest.returncode == 0, rest.stdout

    def rows(run):
        lines = (tmp_path / "runs" / run / "obs # not a comment.dat").read_text().splitlines()
        return {ln.split()[0]: ln for ln in lines if ln.split() and ln.split()[0].isdigit()}

    a, b = rows("full"), rows("rest")
    assert sorted(b, key=int) == [str(k) for k in range(12, 21, 2)]
    for k in b:
        assert a[k] == b[k], k
    assert (tmp_path / "runs" / "full" / "omega.rf").read_text() == (tmp_path / "runs" / "rest" / "omega.rf").read_text()

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
    # final fields and checkpoint files (mesh headers, %.10f rows)
    lib = load_hip_library()
    w = rf_read(lib, str(run / "omega.rf"))
    assert w.shape == (1, 8, 8, 8)
    assert np.abs((w[0] - o.get_field(0)[0]).view(np.float64)).max() <= 6e-11
    assert (run / "tmp_omega.rf").exists() and (run / "tmp_psi.rf").exists()
    for name in ("psi.rf", "rho_type.rf", "rhop.rf", "rhoc.rf", "rhoc_poly.rf", "rhoc_ion.rf", "rho_ion_0_counterion.rf",
                 "rho_ion_1_salt_anion.rf"):
        assert (run / name).exists(), name

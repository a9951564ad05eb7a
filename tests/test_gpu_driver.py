"""The standalone driver on the GPU, every mode end to end through its
command line, with the oracle's solve of the same cells as the caller's CPU
reference: no failed cell, and with the correctly rounded Err**(1/3) on both
sides no differing value (maxrel 0); the sorted mode's two launch orders are
bit-identical."""
import subprocess
import sys

import numpy as np
import pytest

from conftest import PKG
from gcchem_amd import driver, inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", driver.MODES)
def test_driver_modes(tmp_path, mech, oracle_cr, mode):
    d = inputs.random_cells(mech, n=1100, seed=9)
    driver.save_batch(tmp_path / "batch.npz", d)
    ref = oracle_cr.integrate_batch(d["Tstart"], d["Tend"], d["C"], d["RCONST"], d["ATOL"],
                                    d["RTOL"], d["ICNTRL"], d["RCNTRL"], d["Perm"])
    np.savez(tmp_path / "ref.npz", Y=ref[0][:, :mech.nvar])
    print(f"case: driver mode {mode}, 1100 cells", flush=True)
    cmd = [sys.executable, str(PKG / "driver.py"), str(tmp_path / "batch.npz"), mode,
           "--ref", str(tmp_path / "ref.npz"), "--chunk", "1024", "--out", str(tmp_path / "o.npz")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]
    assert "nFail=        0" in r.stdout and "maxrel=   0.000e+00" in r.stdout
    assert "differing values:          0" in r.stdout
    if mode == "sorted":
        assert "sorted vs identity: bit-identical" in r.stdout
    with np.load(tmp_path / "o.npz") as z:
        assert np.array_equal(z["C"][:, :mech.nvar], ref[0][:, :mech.nvar])
        assert np.array_equal(z["ISTAT"], ref[1]) and np.array_equal(z["IERR"], ref[3])

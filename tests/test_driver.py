"""The standalone driver's host logic (no GPU): the reference driver's report
arithmetic (standalone/kpp_batch.F90:262-272, 883-926), its cost-sorted
permutation (Build_Perms, :847-880) and the batch file round trip."""
import numpy as np
import pytest

from gcchem_amd import driver, inputs


def test_report_and_gate_by_hand():
    Yref = np.array([[1.0, 0.0, -2.0], [1e-25, 5.0, 4.0]])
    Y = np.array([[1.0, 1e-22, -2.0 * (1 + 1e-9)], [2e-25, 5.0, 4.0]])
    IE = np.array([1, -6])
    r = driver.report("t", 2.0, Y, Yref, IE)
    # 1e-22 over the 1e-20 floor, 1e-9, 1e-25 over the floor: the largest is 1e-2
    assert r["nFail"] == 1 and r["cells_per_s"] == 1.0
    assert r["maxrel"] == pytest.approx(1e-2, rel=1e-12)
    g = driver.gate_report("t", Y, Yref, IE)
    # three values differ; with the 1e-30 floor the 1e-25 -> 2e-25 change is 1.0
    assert (g["differing"], g["of"], g["nFail"]) == (3, 6, 1)
    assert g["maxrel"] == pytest.approx(1e8, rel=1e-12)   # 1e-22 / 1e-30
    assert driver.gate_report("t", Yref, Yref, IE)["differing"] == 0
    assert driver.report("t", 1.0, Y, None, IE)["maxrel"] is None


def test_build_perms_is_the_stable_ascending_counting_sort():
    rng = np.random.default_rng(3)
    steps = rng.integers(0, 40, 500)
    ident, perm = driver.build_perms(steps)
    assert np.array_equal(ident, np.arange(1, 501))
    assert np.array_equal(perm, np.argsort(steps, kind="stable") + 1)
    assert np.all(np.diff(steps[perm - 1]) >= 0)
    with pytest.raises(ValueError):
        driver.build_perms([3, -1])


def test_batch_file_round_trip(tmp_path, mech):
    d = inputs.random_cells(mech, n=40, seed=2)
    d["Perm"] = d["Perm"][::-1].copy()
    driver.save_batch(tmp_path / "b.npz", d)
    e = driver.load_batch(tmp_path / "b.npz")
    for k in driver.BATCH_KEYS + ("Perm",):
        assert np.array_equal(e[k], d[k]), k
    assert (e["Tstart"], e["Tend"]) == (d["Tstart"], d["Tend"])
    f = driver.load_batch(tmp_path / "b.npz", ncap=7)
    assert f["C"].shape[0] == 7 and np.array_equal(f["Perm"], np.arange(1, 8))
    np.savez(tmp_path / "bad.npz", C=d["C"])
    with pytest.raises(ValueError, match="missing"):
        driver.load_batch(tmp_path / "bad.npz")
    with pytest.raises(ValueError):
        driver.run(d, "nonsense")

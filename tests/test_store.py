"""On-disk formats on the CPU: snapshot .npy and probe CSV round trips (exact
for FP32 values, including subnormals, signed zeros and extremes)."""
import numpy as np

from simulation_amd import store


def test_snapshot_round_trip(tmp_path):
    a = np.random.default_rng(0).standard_normal((7, 9)).astype(np.float32)
    p = store.write_snapshot(str(tmp_path), "ez", 500, a)
    assert p.endswith("ez_t000500.npy")
    b = store.read_snapshot(str(tmp_path), "ez", 500)
    assert b.dtype == np.float32 and np.array_equal(a, b)


def test_probe_csv_round_trip_is_exact(tmp_path):
    rng = np.random.default_rng(1)
    tr = rng.standard_normal((3, 50)).astype(np.float32)
    tr[0, :4] = [np.float32(1e-45), -0.0, np.finfo(np.float32).max, np.finfo(np.float32).tiny]
    path = str(tmp_path / "p" / "probes.csv")
    store.write_probes_csv(path, tr, ["ez[1;2]", "ez[3;4]", "ez[5;6]"], t0=100)
    t, names, back = store.read_probes_csv(path)
    assert names == ["ez[1;2]", "ez[3;4]", "ez[5;6]"]
    assert np.array_equal(t, np.arange(101, 151))
    assert np.array_equal(back.view(np.uint32), tr.view(np.uint32))

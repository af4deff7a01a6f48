"""Host-side monitor API (no GPU): recorded-data statistics and .npz files, following
beamz/devices/monitors.py:549-589 and 787-833."""
import numpy as np
import pytest

from beamz_amd.devices.monitors import Monitor


def filled(n=5, npts=7, seed=0):
    rng = np.random.default_rng(seed)
    m = Monitor(start=(0.0, 1e-6), end=(2e-6, 1e-6), name="line")
    t = [k * 1e-16 for k in range(n)]
    m.fields = {"Ez": [rng.standard_normal(npts) for _ in t], "Hx": [rng.standard_normal(npts) for _ in t],
                "Hy": [rng.standard_normal(npts) for _ in t], "t": t}
    m.power_history = list(rng.random(n))
    m.power_timestamps = t
    return m


def test_statistics_and_field_lookup():
    m = filled()
    s = m.get_field_statistics()
    assert s["total_records"] == 5 and s["time_span"] == pytest.approx(4e-16)
    assert s["avg_power"] == pytest.approx(np.mean(m.power_history)) and s["max_power"] == max(m.power_history)
    assert s["monitor_type"] == "line" and not s["is_3d"] and s["line_start"] == (0.0, 1e-6)
    p = m.get_power_statistics()
    a = np.array(m.power_history)
    assert p["total_energy"] == pytest.approx(a.sum()) and p["std_power"] == pytest.approx(a.std())
    assert p["peak_to_average_ratio"] == pytest.approx(a.max() / a.mean())
    assert m.get_field_at_time("Ez", time_index=2) is m.fields["Ez"][2]
    assert m.get_field_at_time("Ez", time_index=9) is None
    assert m.get_field_at_time("Hx", time_value=2.6e-16) is m.fields["Hx"][3]      # nearest record
    assert m.get_field_at_time("Hy") is m.fields["Hy"][-1]
    assert m.get_field_at_time("Ex") is None
    assert str(m) == "Monitor: line (2D), 5 records"
    empty = Monitor(start=(0.0, 0.0), end=(1e-6, 0.0))
    assert empty.get_field_statistics() == {} and empty.get_power_statistics() == {}


def test_npz_round_trip_without_pickle(tmp_path):
    m = filled()
    m.save_data(str(tmp_path / "mon"))
    with np.load(tmp_path / "mon.npz", allow_pickle=False) as d:        # every entry is a plain array
        assert "fields/Ez" in d.files and d["fields/Ez"].shape == (5, 7)
    r = Monitor(start=(0.0, 1e-6), end=(2e-6, 1e-6))
    r.load_data(str(tmp_path / "mon.npz"))
    for k in ("Ez", "Hx", "Hy", "t"):
        assert np.array_equal(np.asarray(r.fields[k]), np.asarray(m.fields[k])), k
    assert r.power_history == m.power_history and r.power_timestamps == m.power_timestamps
    with pytest.raises(ValueError, match="Unsupported format"):
        m.save_data(str(tmp_path / "x"), format="hdf5")


def test_reference_layout_needs_explicit_pickle(tmp_path):
    """A file in the reference's layout (fields as a pickled dict, no timestamps) is refused unless the caller
    allows unpickling; then it loads, with timestamps 0..n-1 as the reference fills them."""
    m = filled()
    np.savez(tmp_path / "ref.npz", fields=m.fields, power_history=m.power_history,
             monitor_info={"type": "line", "is_3d": False})
    r = Monitor(start=(0.0, 1e-6), end=(2e-6, 1e-6))
    with pytest.raises(ValueError, match="allow_pickle=True"):
        r.load_data(str(tmp_path / "ref.npz"))
    r.load_data(str(tmp_path / "ref.npz"), allow_pickle=True)
    assert np.array_equal(np.asarray(r.fields["Ez"]), np.asarray(m.fields["Ez"]))
    assert r.power_timestamps == list(range(5))

"""Mass tracking summary (lucky_droplets_amd/tracking.py), host side."""
import numpy as np

from lucky_droplets_amd.tracking import summary


def test_summary_reports_gain_and_drift():
    hist = np.array([[0, 1.0, 9.0, 10.0], [10, 3.0, 7.0 + 1e-6, 10.0 + 1e-6], [20, 6.0, 4.0 - 2e-6, 10.0 - 2e-6]])
    s = summary(hist)
    assert s["steps"] == 20 and s["droplet_mass_gained"] == 5.0
    assert abs(s["max_total_drift"] - 2e-6) < 1e-15
    assert abs(s["max_total_drift_per_gained"] - 4e-7) < 1e-15
    assert abs(s["max_total_drift_rel"] - 2e-7) < 1e-15

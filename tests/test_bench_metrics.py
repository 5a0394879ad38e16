"""The benchmark's accuracy metrics on hand-made traces (no GPU): the onsets
of events (max V reaching 1e-3 m/s after having been below 1e-5 m/s since
the previous onset; Functions_Plot.jl:469-490 uses a threshold on recorded
max V the same way) and the first event's peak V (largest max V from its
onset until max V is back below 1e-5 m/s)."""
import importlib.util
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_first_event_peak():
    b = _bench()
    v = np.array([1e-9, 1e-6, 2e-3, 0.5, 1.2, 0.3, 1e-4, 5e-6, 1e-6, 2e-3, 3.0])
    # the second event (3.0) is past the end of the first one and does not count
    assert b.first_event_peak(v) == (1.2, True)
    # a run that stops inside its first event: the peak so far, not ended
    assert b.first_event_peak(v[:6]) == (1.2, False)
    assert b.first_event_peak(v[:2]) == (None, False)
    # the onset step itself can hold the peak
    assert b.first_event_peak(np.array([1e-9, 5.0, 1e-6])) == (5.0, True)


def test_onsets_rearm_and_interpolation():
    b = _bench()
    told = np.arange(8, dtype=float)
    dt = np.ones(8)
    # oscillation around the threshold inside one event counts once; the
    # second onset needs V below 1e-5 first
    v = np.array([1e-9, 1e-2, 5e-4, 2e-3, 1e-6, 1e-9, 1e-1, 1e-9])
    on = b.onsets(told, dt, v)
    assert len(on) == 2
    # log-linear interpolation between step ends: 1e-9 -> 1e-2 crosses 1e-3 at 6/7 of the step
    np.testing.assert_allclose(on[0], 1.0 + 6.0 / 7.0, rtol=1e-12)
    np.testing.assert_allclose(on[1], 6.0 + 6.0 / 8.0, rtol=1e-12)

"""Pins of the action-potential marker oracle (legacy/calculate_local_activation_time.py)
by hand-worked traces and an independent run-by-run formulation."""
import numpy as np

from oracle import aplat_oracle as ao

THR = 0.13


def test_upstroke_marker_of_a_clean_beat():
    s = np.zeros(40)
    s[11:25] = [0.5, 0.9, 1.0, 0.97, 0.9, 0.8, 0.7, 0.6, 0.5, 0.4, 0.3, 0.2, 0.15, 0.14]
    s[10] = 0.1
    # one run [11, 24]; diffs inside: 0.4 at 11 is the largest
    np.testing.assert_array_equal(ao.markers(s, THR), [11])


def test_one_sample_run_shifts_the_pairing():
    s = np.zeros(25)
    s[3:7] = [0.2, 0.6, 0.9, 0.8]        # run [3, 6]
    s[10] = 0.3                          # one-sample run [10, 10]
    s[14:19] = [0.4, 0.7, 1.0, 0.9, 0.5]  # run [14, 18]
    s[11:14] = [0.05, 0.0, 0.02]
    # boundary points 3, 6, 10, 14, 18 -> pairs (3, 6), (10, 14), 18 dropped
    d1 = np.diff(s[3:7])
    d2 = np.diff(s[10:15])
    want = [3 + int(np.argmax(d1)), 10 + int(np.argmax(d2))]
    assert want == [3, 13]               # 0.2 -> 0.6 opens run 1; 0.02 -> 0.4 (below -> run 3)
    np.testing.assert_array_equal(ao.markers(s, THR), want)


def test_decaying_start_is_dropped_and_no_activation_is_nan():
    s = np.concatenate([[0.9, 0.7, 0.5, 0.3], np.zeros(6), [0.2, 0.8, 1.0, 0.6], np.zeros(3)])
    # runs [0, 3] (only falling: max diff < 0, dropped) and [10, 13] (marker 10)
    np.testing.assert_array_equal(ao.markers(s, THR), [10])
    assert np.isnan(ao.markers(np.zeros(10), THR)).all()
    # a single sample above: one boundary point, dropped as odd; empty, not nan
    one = np.zeros(10)
    one[4] = 0.5
    m = ao.markers(one, THR)
    assert m.size == 0


def test_window_and_shift_of_execute():
    T = 200
    a = np.zeros((T, 3))
    beat = np.concatenate([[0.3, 0.9, 1.0], np.linspace(0.95, 0.2, 17)])   # steepest rise at +0
    for col, starts in ((0, (10, 60, 110)), (1, (20, 70))):
        for t0 in starts:
            a[t0:t0 + 20, col] = beat
    lat, cl = ao.lat_from_markers(5.0, a, THR, 0.8)
    # markers: col 0 at 10, 60, 110; col 1 at 20, 70; col 2 none -> cycle lengths 50, 50, 50,
    # CL 50, window [5, 45]: col 0 -> 10, col 1 -> 20; shifted so that the earliest is 1
    np.testing.assert_array_equal(ao.markers(a[:, 0], THR), [10, 60, 110])
    np.testing.assert_array_equal(cl, [50.0, 50.0, 50.0])
    assert cl.dtype == np.float64         # col 2 contributed a NaN placeholder
    np.testing.assert_array_equal(lat[:2], [1.0, 11.0])
    assert np.isnan(lat[2])
    lat2, cl2 = ao.lat_from_markers(5.0, a[:, :2], THR, 0.8)
    assert cl2.dtype.kind == "i"          # every signal has a marker: integer differences


def _runs(s, thr):
    out, t, T = [], 0, len(s)
    while t < T:
        if s[t] > thr:
            a = t
            while t + 1 < T and s[t + 1] > thr:
                t += 1
            out.append((a, t))
        t += 1
    return out


def test_matches_run_by_run_formulation_without_one_sample_runs():
    rng = np.random.default_rng(3)
    for _ in range(50):
        s = np.convolve(rng.random(300), np.ones(7) / 7, mode="same")
        runs = _runs(s, 0.5)
        if any(a == b for a, b in runs):
            continue
        want = []
        for a, b in runs:
            d = np.diff(s[a:b + 1])
            if d.max() > 0:
                want.append(a + int(np.argmax(d)))
        got = ao.markers(s, 0.5)
        if not runs:
            assert np.isnan(got).all()
        else:
            np.testing.assert_array_equal(got, want)

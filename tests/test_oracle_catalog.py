"""Oracle of the event catalog (get_event_fragments / merge_fragments,
scripts/Functions_Plot.jl:430-579), pinned by hand-worked histories and
fragment lists whose fragments, event numbering and moments are computed
here by hand, and by invariants computed independently with numpy
(rising-edge counts, total moment, merge limits)."""
import numpy as np
import pytest


def test_fragments_hand_case(orc):
    # one element: rates 0, 1, 2, 0, 1 over unit steps, threshold 0.5
    t = np.array([0.0, 1, 2, 3, 4, 5])
    d = np.array([[0.0], [0], [1], [3], [3], [4]])
    ft, fl, fm, fe = orc.event_fragments(t, d, np.array([[7.0, 8.0, 9.0]]), np.array([2.0]), 3.0, 0.5)
    np.testing.assert_array_equal(ft, [2.0, 5.0])        # the later record of the first fast interval
    np.testing.assert_array_equal(fm, [1 * 2 * 3 + 2 * 2 * 3, 1 * 2 * 3])
    np.testing.assert_array_equal(fl, [[7, 8, 9], [7, 8, 9]])
    np.testing.assert_array_equal(fe, [0, 0])


def test_fragments_rate_uses_record_spacing_and_strict_threshold(orc):
    # equal slip increments over unequal intervals; rate == threshold is not an event
    t = np.array([0.0, 1.0, 3.0, 3.5])
    d = np.array([[0.0, 0.0], [1.0, 0.5], [2.0, 1.5], [3.0, 1.5]])
    ft, fl, fm, fe = orc.event_fragments(t, d, np.zeros((2, 3)), np.ones(2), 1.0, 1.0)
    # element 0: rates 1 (not > 1), 0.5, 2 -> one fragment at t=3.5 of moment 1
    # element 1: rates 0.5, 0.5, 0 -> none
    np.testing.assert_array_equal(ft, [3.5])
    np.testing.assert_array_equal(fm, [1.0])
    np.testing.assert_array_equal(fe, [0])


def test_fragments_sorted_stably_by_time(orc):
    # element 0 opens at t=3; element 1 at t=2 and (after a pause) t=4; element 2 at t=3
    t = np.arange(6.0)
    d = np.zeros((6, 3))
    d[3:, 0] = 10
    d[2:, 1] = 10
    d[4:, 1] = 20
    d[3:, 2] = 10
    ctr = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.0, 0, 0]])
    ft, fl, fm, fe = orc.event_fragments(t, d, ctr, np.ones(3), 1.0, 0.5)
    np.testing.assert_array_equal(ft, [2, 3, 3, 4])
    np.testing.assert_array_equal(fe, [1, 0, 2, 1])      # ties keep element order
    np.testing.assert_array_equal(fl[:, 0], [1, 0, 2, 1])
    np.testing.assert_array_equal(fm, [10, 10, 10, 10])


def _merge_case():
    ft = np.array([0.0, 1.0, 2.0, 10.0, 10.5, 11.0, 11.2, 12.0])
    x = np.array([0.0, 100.0, 2000.0, 0.0, 3000.0, 2990.0, 2980.0, 10.0])
    fl = np.stack([x, np.zeros(8), np.zeros(8)], axis=1)
    fm = 10.0 ** (9.05 + 1.5 * np.array([2.0, 1.0, 3.0, 2.5, 1.5, 1.2, 1.1, 0.5]))
    return ft, fl, fm


def test_merge_hand_case(orc):
    ft, fl, fm = _merge_case()
    mt, mw, ml, mm, fid = orc.merge_fragments(ft, fl, fm, 5.0, 500.0)
    # f1 joins f0 (1 s, 100 m); f2 is 1900 / 2000 m away -> new; f3 is 8 s after f2 -> new;
    # f4 is 3000 m from f3 and the scan stops at f2 -> new; f5, f6 chain onto f4;
    # f7 passes the far f6, f5, f4 and joins f3 (2 s, 10 m)
    np.testing.assert_array_equal(fid, [0, 0, 1, 2, 3, 3, 3, 2])
    np.testing.assert_array_equal(mt, [0.0, 2.0, 10.0, 10.5])
    np.testing.assert_array_equal(ml[:, 0], [0.0, 2000.0, 0.0, 3000.0])
    np.testing.assert_array_equal(mm, [0.0 + fm[0] + fm[1], 0.0 + fm[2], 0.0 + fm[3] + fm[7],
                                       0.0 + fm[4] + fm[5] + fm[6]])
    np.testing.assert_allclose(mw[1], 3.0, rtol=1e-14)
    np.testing.assert_allclose(mw, (np.log10(mm) - 9.05) / 1.5, rtol=1e-15)


def test_merge_limits(orc):
    ft, fl, fm = _merge_case()
    mt, mw, ml, mm, fid = orc.merge_fragments(ft, fl, fm, 0.0, 1e30)   # no time window
    np.testing.assert_array_equal(fid, np.arange(8))
    mt, mw, ml, mm, fid = orc.merge_fragments(ft, fl, fm, np.inf, np.inf)   # everything merges
    np.testing.assert_array_equal(fid, np.zeros(8))
    assert mm.size == 1 and mm[0] == pytest.approx(fm.sum(), rel=1e-15)
    out = orc.merge_fragments(np.zeros(0), np.zeros((0, 3)), np.zeros(0), 5.0, 500.0)
    assert out[0].size == 0 and out[4].size == 0


def _bursty_history(seed, n_rec, n):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.uniform(0.1, 2.0, n_rec))
    fast = rng.random((n_rec, n)) < 0.15
    fast |= np.roll(fast, 1, axis=0) & (rng.random((n_rec, n)) < 0.6)
    inc = np.where(fast, rng.uniform(0.05, 1.0, (n_rec, n)), rng.uniform(0, 1e-4, (n_rec, n)))
    inc[0] = 0
    dt = np.diff(t, prepend=t[0] - 1.0)
    d = np.cumsum(inc * dt[:, None], axis=0)
    return t, d


def test_fragments_invariants_random(orc):
    t, d = _bursty_history(7, 300, 64)
    rng = np.random.default_rng(8)
    ctr = rng.uniform(-5e3, 5e3, (64, 3))
    area = rng.uniform(1e3, 1e4, 64)
    mu, crit = 3.2e10, 1e-2
    ft, fl, fm, fe = orc.event_fragments(t, d, ctr, area, mu, crit)
    rate = np.diff(d, axis=0) / np.diff(t)[:, None]
    fast = rate > crit
    rises = fast & ~np.vstack([np.zeros((1, 64), bool), fast[:-1]])
    assert ft.size == rises.sum() > 50
    np.testing.assert_array_equal(np.bincount(fe, minlength=64), rises.sum(0))
    assert np.all(np.diff(ft) >= 0)
    np.testing.assert_array_equal(fl, ctr[fe])
    total = (np.diff(d, axis=0) * fast * area[None, :] * mu).sum()
    assert fm.sum() == pytest.approx(total, rel=1e-12)
    mt, mw, ml, mm, fid = orc.merge_fragments(ft, fl, fm, 5.0, 3000.0)
    assert mm.sum() == pytest.approx(fm.sum(), rel=1e-12)
    # events are numbered in order of their first fragment, whose time and place they take
    prev_max = np.maximum.accumulate(np.r_[-1, fid[:-1]])
    assert np.all(fid <= prev_max + 1)
    first = np.flatnonzero(fid > prev_max)
    np.testing.assert_array_equal(fid[first], np.arange(mt.size))
    np.testing.assert_array_equal(mt, ft[first])
    np.testing.assert_array_equal(ml, fl[first])
    assert fid.max() + 1 == mt.size and mt.size < ft.size

"""GPU parity of the event catalog (get_event_fragments / merge_fragments,
scripts/Functions_Plot.jl:430-579) against the oracle: fragments and merged
events are index work plus sums taken in the reference's order, so times,
locations, element / event indices and moments must be bitwise equal; Mw
differs only by the log10 implementation (1e-14 absolute)."""
import numpy as np
import pytest

from tests.test_oracle_catalog import _bursty_history, _merge_case

pytestmark = pytest.mark.gpu

MW_TOL = 1e-14


def _check_same(qd, orc, t, d, ctr, area, mu, crit, mt, md):
    ref = orc.event_fragments(t, d, ctr, area, mu, crit)
    with qd.EventCatalog.from_history(t, d, ctr, area, mu, crit) as c:
        got = c.fragments()
        for g, r in zip(got, ref):
            np.testing.assert_array_equal(g, r)
        E = c.merge(mt, md)
        ev = c.events()
    rm = orc.merge_fragments(ref[0], ref[1], ref[2], mt, md)
    assert E == rm[0].size
    np.testing.assert_array_equal(ev[0], rm[0])            # times
    np.testing.assert_allclose(ev[1], rm[1], rtol=0, atol=MW_TOL)
    np.testing.assert_array_equal(ev[2], rm[2])            # locations
    np.testing.assert_array_equal(ev[3], rm[3])            # moments
    np.testing.assert_array_equal(ev[4], rm[4])            # fragment -> event
    return ref, rm


def test_catalog_random_histories(qd, orc):
    t, d = _bursty_history(11, 400, 700)
    rng = np.random.default_rng(12)
    ctr = rng.uniform(-5e3, 5e3, (700, 3))
    area = rng.uniform(1e3, 1e4, 700)
    ref, rm = _check_same(qd, orc, t, d, ctr, area, 3.2e10, 1e-2, 5.0, 3000.0)
    assert ref[0].size > 10000 and 1 < rm[0].size < ref[0].size


def test_catalog_hand_merge_case(qd, orc):
    ft, fl, fm = _merge_case()
    with qd.EventCatalog.from_fragments(ft, fl, fm) as c:
        assert c.merge(5.0, 500.0) == 4
        t, mw, loc, m, fid = c.events()
    np.testing.assert_array_equal(fid, [0, 0, 1, 2, 3, 3, 3, 2])
    np.testing.assert_array_equal(t, [0.0, 2.0, 10.0, 10.5])
    tm, mwm, lm = qd.merge_fragments(ft, fl, fm, 8, 5.0, 500.0)
    np.testing.assert_array_equal(tm, t)
    np.testing.assert_array_equal(mwm, mw)


def test_catalog_edges(qd, orc):
    # fewer than two records, no element over the threshold, and no fragments to merge
    t, loc, m, cnt = qd.get_event_fragments(np.array([0.0]), np.zeros((1, 4)), np.zeros((4, 3)), np.ones(4), 1.0)
    assert cnt == 0
    t, loc, m, cnt = qd.get_event_fragments(np.arange(5.0), np.zeros((5, 4)), np.zeros((4, 3)), np.ones(4), 1.0)
    assert cnt == 0
    tm, mw, lm = qd.merge_fragments(t, loc, m, cnt, 5.0, 500.0)
    np.testing.assert_array_equal(tm, [0.0])              # the reference's placeholders
    assert mw.size == 0 and lm.shape == (1, 3)
    with qd.EventCatalog.from_fragments(np.zeros(0), np.zeros((0, 3)), np.zeros(0)) as c:
        assert c.merge(5.0, 500.0) == 0


def test_catalog_large_properties(qd):
    # 20,000 elements x 600 records: counts and total moment against numpy
    t, d = _bursty_history(13, 600, 20000)
    rng = np.random.default_rng(14)
    ctr = rng.uniform(-5e3, 5e3, (20000, 3))
    area = rng.uniform(1e3, 1e4, 20000)
    with qd.EventCatalog.from_history(t, d, ctr, area, 3.2e10, 1e-2) as c:
        ft, fl, fm, fe = c.fragments()
        E = c.merge(5.0, 500.0)
        tm, mw, lm, mm, fid = c.events()
    fast = np.diff(d, axis=0) / np.diff(t)[:, None] > 1e-2
    rises = fast & ~np.vstack([np.zeros((1, 20000), bool), fast[:-1]])
    assert ft.size == rises.sum()
    np.testing.assert_array_equal(np.bincount(fe, minlength=20000), rises.sum(0))
    assert np.all(np.diff(ft) >= 0)
    total = (np.diff(d, axis=0) * fast * area[None, :] * 3.2e10).sum()
    assert fm.sum() == pytest.approx(total, rel=1e-12)
    assert mm.sum() == pytest.approx(total, rel=1e-12)
    assert E == tm.size == fid.max() + 1


def test_catalog_from_run_matches_history(qd, orc):
    # a single-fault run through its first event, recorded every step; the
    # catalog built on the device-resident records equals the oracle's on the
    # copied-back histories
    from qdfn import configs
    c = configs.single_fault()
    seg = c["seg"]
    op = qd.DenseOperator.from_rect_segments(seg, c["mu"], c["nu"], qd.PREC_FP64)
    ks, _ = op.kself()
    ksum, knsum = op.matvec(np.ones(seg.shape[0]), loading_form=False)
    prob = configs.problem_from_segments(c, ks, ksum, knsum)
    steps = 4000
    sim = qd.Simulation(op, prob, record_step=1, record_capacity=steps)
    sim.run(steps)
    h = sim.history()
    ctr = seg[:, 0:3]
    area = qd.fault_area("R", seg[:, 3], seg[:, 4])
    crit = 1e-3  # events on this fault are small: about 60 fragments at 1e-3 m/s in 4000 steps
    with sim.event_catalog(ctr, area, c["mu"], crit) as cat:
        got = cat.fragments()
        cat.merge(5.0, 500.0)
        ev = cat.events()
    sim.close()
    op.close()
    ref = orc.event_fragments(h["History_Time"], h["History_Disp"], ctr, area, c["mu"], crit)
    assert ref[0].size >= 20
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(g, r)
    rm = orc.merge_fragments(ref[0], ref[1], ref[2], 5.0, 500.0)
    np.testing.assert_array_equal(ev[4], rm[4])
    np.testing.assert_array_equal(ev[3], rm[3])
    np.testing.assert_allclose(ev[1], rm[1], rtol=0, atol=MW_TOL)

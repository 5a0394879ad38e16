"""GPU checks of the sampled H-matrix construction (qdfn_hmat_build_*_sampled:
partially pivoted cross approximation that evaluates only the rows and columns
it pivots on, in place of HmatBuild_T / HmatBuild_R's full-block assembly +
pqrfact, scripts/Functions_BuildInputFile.jl:674-716, :754-797).

* Inadmissible blocks come from the same per-pair code as the full build:
  bitwise equal to the full build's blocks.  Blocks whose shear part reaches
  the rank cap are evaluated in full and take the full build's path with the
  same (exact) scale: both parts bitwise equal to the full build's.
* Sampled parts against the oracle's entries: the stop test is an estimate, so
  the bound is checked in distribution: every part within 2 rtol ||K_S,b||_F
  except isolated entries of the triangular-dislocation normal traction that
  no row or column sample meets (at most 2 % of the parts, each within
  100 rtol ||K_S,b||_F).
* Products within 2 rtol ||K_S||_F ||x||_2 of the dense products (the full
  build's bound, doubled for the estimate).
* Fewer kernel evaluations than the n^2 of the full build on a mesh with
  leaves of 64 (rough fault, 2,400 triangles), where the full build's
  products are also checked against the dense ones.
* main_H on the sampled operator: first-event time within 1e-6 of the run on
  the full build's operator.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-7


def _parts_equal(Bf, Bs, q, w):
    if Bf[w]["kind"][q] != Bs[w]["kind"][q] or Bf[w]["rank"][q] != Bs[w]["rank"][q]:
        return False
    if not np.array_equal(Bf[w]["a"][q], Bs[w]["a"][q]):
        return False
    if Bf[w]["kind"][q] == 1 and Bf[w]["rank"][q] > 0:
        return np.array_equal(Bf[w]["b"][q], Bs[w]["b"][q])
    return True


def _sampled_errors(Bs, KS, KN, rs, ad, rtol):
    """Per low-rank part: ||U V - K_b||_F / (rtol ||K_S,b||_F)."""
    out = []
    for w, K in enumerate((KS, KN)):
        for q, r in enumerate(rs):
            if Bs[w]["kind"][q] != 1:
                continue
            assert ad[q] == 1
            Kb = K[r[2] - 1:r[3], r[0] - 1:r[1]]
            sb = np.linalg.norm(KS[r[2] - 1:r[3], r[0] - 1:r[1]])
            approx = Bs[w]["a"][q] @ Bs[w]["b"][q] if Bs[w]["rank"][q] > 0 else np.zeros_like(Kb)
            out.append(np.linalg.norm(approx - Kb) / (rtol * sb) if sb > 0 else 0.0)
    return np.array(out)


def _check_against_full(Hf, Hs, rs, ad):
    """Dense blocks bitwise; blocks with a full-evaluated shear part bitwise."""
    Bf, Bs = Hf.blocks(rs), Hs.blocks(rs)
    n_same = 0
    for q in range(len(ad)):
        # inadmissible blocks, and admissible blocks stored dense (only a
        # shear part that reached the rank cap twice is)
        if ad[q] != 1 or Bs[0]["kind"][q] == 0:
            assert _parts_equal(Bf, Bs, q, 0) and _parts_equal(Bf, Bs, q, 1), q
            n_same += 1
    return Bf, Bs, n_same


def _check_case(qd, Hf, Hs, KS, KN, rs, ad, rtol, seed):
    Bf, Bs, _ = _check_against_full(Hf, Hs, rs, ad)
    err = _sampled_errors(Bs, KS, KN, rs, ad, rtol)
    if err.size:
        assert np.mean(err > 2.0) <= 0.02, np.sort(err)[-10:]
        assert err.max() <= 100.0, np.sort(err)[-10:]
    n = KS.shape[0]
    x = np.random.default_rng(seed).standard_normal(n)
    yS, yN = Hs.matvec(x, loading_form=False)
    bound = 2 * rtol * np.linalg.norm(KS) * np.linalg.norm(x) + 1e-9
    assert np.linalg.norm(yS - KS @ x) <= bound
    assert np.linalg.norm(yN - KN @ x) <= bound
    # the loading form uses the same diagonal (dense diagonal blocks are shared)
    np.testing.assert_array_equal(Hs.kself()[0], Hf.kself()[0])
    return err


def test_rect_network_sampled(orc, qd):
    from qdfn import configs
    cfg = configs.fault_network(n_faults=3, nx=16, nz=12)
    seg = cfg["seg"]
    st = qd.hmat_structure(seg[:, :3], seg[:, :3], min_elements_to_cut=16, max_levels=8, dist_diam_ratio=1.0,
                           arrange_point=(0, 0, 0))
    segs = np.ascontiguousarray(seg[st["perm"]])
    KS, KN = orc.stiffness_rect(segs, segs, cfg["mu"], cfg["nu"])
    rs, ad = st["range_SR"], st["admissible"]
    Hf = qd.HMatrix.build_rect(segs, cfg["mu"], cfg["nu"], rs, ad, aca_rtol=RTOL)
    Hs = qd.HMatrix.build_rect(segs, cfg["mu"], cfg["nu"], rs, ad, aca_rtol=RTOL, sampled=True)
    _check_case(qd, Hf, Hs, KS, KN, rs, ad, RTOL, 11)
    bi = Hs.build_info()
    assert bi["parts_sampled"] > 0 and bi["steps"] > 0
    assert Hf.build_info()["entries_evaluated"] == segs.shape[0] ** 2
    Hf.close()
    Hs.close()


def _tri_case(orc, qd, nx, nz, leaf, h=50.0, ratio=1e-2, oracle_entries=True):
    from qdfn import configs
    m = configs.rough_fault(nx=nx, nz=nz, h=h, ratio=ratio)
    P1, P2, P3, nrm, _, _, slip = orc.rotverts_unitvectors(m["P1"], m["P2"], m["P3"], m["rake"])
    ctr = (P1 + P2 + P3) / 3
    st = qd.hmat_structure(ctr, P1, min_elements_to_cut=leaf, max_levels=20, dist_diam_ratio=1.0,
                           arrange_point=(0, 0, 0))
    p = st["perm"]
    P1, P2, P3, nrm, slip, ctr, rake = P1[p], P2[p], P3[p], nrm[p], slip[p], ctr[p], m["rake"][p]
    lam = 2 * m["mu"] * m["nu"] / (1 - 2 * m["nu"])
    geo = (P1, P2, P3, rake, ctr, nrm, slip)
    if oracle_entries:
        KS, KN = orc.stiffness_tri(*geo, m["mu"], lam)
    else:  # device assembly (equal to the oracle's to 1e-10 in the assembly tests)
        KS, KN = qd.StiffnessMatrix_ByParts_Calculation_Tri(P1, P2, P3, rake, ctr, m["mu"], lam, nrm, slip)
    return geo, m["mu"], lam, st, KS, KN


def test_tri_rough_sampled(orc, qd):
    geo, mu, lam, st, KS, KN = _tri_case(orc, qd, 16, 12, 16, h=60.0, ratio=2e-2)
    rs, ad = st["range_SR"], st["admissible"]
    Hf = qd.HMatrix.build_tri(*geo, mu, lam, rs, ad, aca_rtol=RTOL)
    Hs = qd.HMatrix.build_tri(*geo, mu, lam, rs, ad, aca_rtol=RTOL, sampled=True)
    _check_case(qd, Hf, Hs, KS, KN, rs, ad, RTOL, 12)
    Hf.close()
    Hs.close()


def test_tri_leaf64_fewer_evaluations(orc, qd):
    rtol = 1e-6
    geo, mu, lam, st, KS, KN = _tri_case(orc, qd, 40, 30, 64, oracle_entries=False)
    rs, ad = st["range_SR"], st["admissible"]
    Hf = qd.HMatrix.build_tri(*geo, mu, lam, rs, ad, aca_rtol=rtol)
    Hs = qd.HMatrix.build_tri(*geo, mu, lam, rs, ad, aca_rtol=rtol, sampled=True)
    err = _check_case(qd, Hf, Hs, KS, KN, rs, ad, rtol, 13)
    n = KS.shape[0]
    # the full build (rank ceiling 1024: large admissible blocks are low rank)
    # within its bound rtol ||K_S||_F ||x||_2 of the dense products
    x = np.random.default_rng(14).standard_normal(n)
    fS, fN = Hf.matvec(x, loading_form=False)
    bound = rtol * np.linalg.norm(KS) * np.linalg.norm(x) * 1.000001
    assert np.linalg.norm(fS - KS @ x) <= bound
    assert np.linalg.norm(fN - KN @ x) <= bound
    rk = Hf.ranks()
    assert int((rk["kind_S"] == 1).sum()) > 0
    bi = Hs.build_info()
    print(f"sampled build n={n}: {bi}, parts {err.size}, err/rtol median {np.median(err):.3g} "
          f"p98 {np.percentile(err, 98):.3g} max {err.max():.3g}", flush=True)
    assert bi["entries_evaluated"] < n * n
    Hf.close()
    Hs.close()


def test_main_h_on_sampled_operator(orc, qd):
    from qdfn import configs
    cfg = configs.single_fault()
    seg = cfg["seg"]
    st = qd.hmat_structure(seg[:, :3], seg[:, :3], min_elements_to_cut=16, max_levels=7, dist_diam_ratio=1.0,
                           arrange_point=(0, 0, 0))
    segs = np.ascontiguousarray(seg[st["perm"]])
    cfg = dict(cfg, seg=segs)
    rs, ad = st["range_SR"], st["admissible"]
    Hf = qd.HMatrix.build_rect(segs, cfg["mu"], cfg["nu"], rs, ad, aca_rtol=RTOL)
    Hs = qd.HMatrix.build_rect(segs, cfg["mu"], cfg["nu"], rs, ad, aca_rtol=RTOL, sampled=True)
    # problem terms from the full build's diagonal and unit-slip responses
    ks, _ = Hf.kself()
    one = np.ones(segs.shape[0])
    sS, sN = Hf.matvec(one, loading_form=False)
    prob = configs.problem_from_segments(cfg, ks, sS, sN)
    prob["ext_at_told"] = 1
    steps = 1200
    first = []
    for H in (Hf, Hs):
        sim = qd.Simulation(H, prob, record_step=10, record_capacity=steps // 10, trace_capacity=steps)
        sim.run(steps)
        stats, tr = sim.stats(), sim.trace()
        sim.close()
        assert stats["status"] == 0 and stats["steps"] == steps
        first.append(configs.first_event_time(tr["told"], tr["dt"], tr["vmax"]))
    assert first[0] is not None and first[1] is not None
    assert abs(first[1] - first[0]) <= 1e-6 * first[0], first
    Hf.close()
    Hs.close()


def test_tri_mixed_near_field_as_full_build(orc, qd):
    """Tail storage (QDFN_PREC_MIXED): both builds compress the inadmissible
    blocks off the diagonal from the same fully evaluated entries, so every
    inadmissible block (dense or near-field low-rank) is bitwise the full
    build's; diagonal blocks stay dense; the sampled operator stays within
    its bound of the assembled stiffness."""
    rtol = 1e-6
    geo, mu, lam, st, KS, KN = _tri_case(orc, qd, 40, 30, 64, oracle_entries=False)
    rs, ad = st["range_SR"], st["admissible"]
    Hf = qd.HMatrix.build_tri(*geo, mu, lam, rs, ad, aca_rtol=rtol, prec=qd.PREC_MIXED)
    Hs = qd.HMatrix.build_tri(*geo, mu, lam, rs, ad, aca_rtol=rtol, prec=qd.PREC_MIXED, sampled=True)
    try:
        offdiag = (rs[:, 0] != rs[:, 2]) | (rs[:, 1] != rs[:, 3])
        rkf, rks = Hf.ranks(), Hs.ranks()
        assert ((ad == 0) & offdiag & (rks["kind_S"] == 1)).any()
        assert not ((ad == 0) & ~offdiag & (rks["kind_S"] == 1)).any()
        BF, BS = Hf.blocks(rs), Hs.blocks(rs)
        for w in range(2):
            for q in np.nonzero(ad == 0)[0]:
                assert BF[w]["kind"][q] == BS[w]["kind"][q]
                np.testing.assert_array_equal(BF[w]["a"][q], BS[w]["a"][q])
                if BF[w]["kind"][q] == 1 and BF[w]["rank"][q] > 0:
                    np.testing.assert_array_equal(BF[w]["b"][q], BS[w]["b"][q])
        SS, SN = Hs.blocks(rs, stored=True)
        worst = 0.0
        for q in np.nonzero(ad == 0)[0]:
            r = rs[q]
            sl = (slice(r[2] - 1, r[3]), slice(r[0] - 1, r[1]))
            tol = rtol * np.linalg.norm(KS[sl])
            for B, K in ((SS, KS), (SN, KN)):
                approx = B["a"][q] if B["kind"][q] == 0 else (
                    B["a"][q] @ B["b"][q] if B["rank"][q] > 0 else np.zeros(K[sl].shape))
                err = np.linalg.norm(approx - K[sl])
                # dense blocks: FP32 rounding (2^-24 per entry) on top
                allow = tol * (1 + 1e-9) if B["kind"][q] == 1 else tol + 2.0 ** -24 * np.linalg.norm(K[sl])
                assert err <= allow + 1e-9 * np.abs(K).max(), (q, err, tol)
                worst = max(worst, err / tol if tol > 0 else 0.0)
        print(f"[sampled mixed] worst near-field block error / tolerance {worst:.3f}")
    finally:
        Hf.close()
        Hs.close()

"""GPU parity of the H-matrix path (HmatBuild_R/T, HmatSolver_Pararllel, main_H).

* Build (device assembly + full-pivot ACA): dense blocks equal the oracle's
  stiffness to 1e-10 of max |K|; every low-rank block satisfies the
  approximation bound ||U V - K_b||_F <= rtol ||K_S,b||_F (+ the 1e-10 entry
  tolerance) against the oracle's entries, for the shear and normal parts of
  the block.
* Product: the device H product equals the oracle's HmatSolver on the same
  stored blocks to 1e-12 (FP64 storage), in both the loading form of main_H
  and the plain form, and is within the approximation bound of the dense
  product K x; an operator imported from the exported blocks
  (qdfn_hmat_create) reproduces the built one bit for bit; FP32 storage stays
  within 2e-6 of FP64 storage.
* main_H: the device loop with the H operator tracks the oracle's main_H on
  the same blocks step by step (told 1e-9, first-event time 1e-9), and its
  first-event time is within 1e-5 of the dense FP64 run's.
* Triangles: the same checks of the build on a rough triangle mesh.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-7
STEPS = 1200


def _rect_case(orc, qd, cfg, **kw):
    seg = cfg["seg"]
    st = qd.hmat_structure(seg[:, :3], seg[:, :3], **kw)
    perm = st["perm"]
    segs = np.ascontiguousarray(seg[perm])
    KS, KN = orc.stiffness_rect(segs, segs, cfg["mu"], cfg["nu"])
    return dict(seg=segs, st=st, KS=KS, KN=KN, cfg=dict(cfg, seg=segs))


@pytest.fixture(scope="module")
def net(orc, qd):
    from qdfn import configs
    cfg = configs.fault_network(n_faults=3, nx=16, nz=12)
    c = _rect_case(orc, qd, cfg, min_elements_to_cut=16, max_levels=8, dist_diam_ratio=1.0,
                   arrange_point=(0, 0, 0))
    c["H"] = qd.HMatrix.build_rect(c["seg"], cfg["mu"], cfg["nu"], c["st"]["range_SR"], c["st"]["admissible"],
                                   aca_rtol=RTOL)
    c["blocks"] = c["H"].blocks(c["st"]["range_SR"])
    yield c
    c["H"].close()


def _check_blocks(blocks, KS, KN, rs, ad):
    scale = max(np.abs(KS).max(), np.abs(KN).max())
    nlr = 0
    for w, K in enumerate((KS, KN)):
        B = blocks[w]
        for q, r in enumerate(rs):
            Kb = K[r[2] - 1:r[3], r[0] - 1:r[1]]
            sb = np.linalg.norm(KS[r[2] - 1:r[3], r[0] - 1:r[1]])
            if B["kind"][q] == 0:
                assert np.abs(B["a"][q] - Kb).max() <= 1e-10 * scale
            else:
                assert ad[q] == 1
                nlr += 1
                approx = B["a"][q] @ B["b"][q] if B["rank"][q] > 0 else np.zeros_like(Kb)
                err = np.linalg.norm(approx - Kb)
                assert err <= RTOL * sb * (1 + 1e-6) + 1e-10 * scale * np.sqrt(Kb.size), (q, w, err)
                assert B["rank"][q] < min(Kb.shape)
    return nlr


def test_rect_build_blocks(net):
    rs, ad = net["st"]["range_SR"], net["st"]["admissible"]
    nlr = _check_blocks(net["blocks"], net["KS"], net["KN"], rs, ad)
    assert nlr > 0
    info = net["H"].info()
    assert info["stored_entries"] < 2 * net["seg"].shape[0] ** 2


def test_rect_product_matches_oracle(net, orc):
    H, (BS, BN) = net["H"], net["blocks"]
    n = net["seg"].shape[0]
    rng = np.random.default_rng(7)
    x = rng.standard_normal(n)
    yS, yN = H.matvec(x, loading_form=False)
    oS, oN = orc.hmat_matvec(BS, x), orc.hmat_matvec(BN, x)
    assert np.abs(yS - oS).max() <= 1e-12 * np.abs(oS).max()
    assert np.abs(yN - oN).max() <= 1e-12 * np.abs(oN).max()
    # loading form of main_H: (H_load (0 - x)) ./ -K_Self
    HL, ks = orc.hmat_transition_to_loading(BS)
    tl, _ = H.matvec(x, loading_form=True)
    otl = orc.hmat_matvec(HL, 0.0 - x) / -ks
    assert np.abs(tl - otl).max() <= 1e-12 * np.abs(otl).max()
    hks, hdg = H.kself()
    np.testing.assert_array_equal(hks, ks)
    # approximation bound against the dense product
    # |(H - K) x|_i <= ||H - K||_F ||x||_2 with ||H - K||_F^2 <= sum_b (rtol ||K_S,b||_F)^2 <= rtol^2 ||K_S||_F^2
    bound = RTOL * np.linalg.norm(net["KS"]) * np.linalg.norm(x) * 1.000001 + 1e-9
    assert np.abs(yS - net["KS"] @ x).max() <= bound
    assert np.abs(yN - net["KN"] @ x).max() <= bound


def test_imported_blocks_reproduce_built_operator(net, qd):
    rs, ad = net["st"]["range_SR"], net["st"]["admissible"]
    parts = []
    for B in net["blocks"]:
        parts.append([B["a"][q] if B["kind"][q] == 0 else (B["a"][q], B["b"][q]) for q in range(len(B["a"]))])
    H2 = qd.HMatrix.from_blocks(net["seg"].shape[0], rs, ad, parts[0], parts[1])
    x = np.random.default_rng(8).standard_normal(net["seg"].shape[0])
    for form in (True, False):
        a = net["H"].matvec(x, form)
        b = H2.matvec(x, form)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
    H2.close()


def test_fp32_storage(net, qd):
    rs, ad = net["st"]["range_SR"], net["st"]["admissible"]
    H32 = qd.HMatrix.build_rect(net["seg"], net["cfg"]["mu"], net["cfg"]["nu"], rs, ad, aca_rtol=RTOL,
                                prec=qd.PREC_FP32)
    x = np.random.default_rng(9).standard_normal(net["seg"].shape[0])
    for form in (True, False):
        a = net["H"].matvec(x, form)
        b = H32.matvec(x, form)
        for u, v in zip(a, b):
            assert np.abs(u - v).max() <= 2e-6 * np.abs(u).max()
    assert H32.info()["bytes_per_product"] < net["H"].info()["bytes_per_product"]
    assert H32.time_matvec(5) > 0
    H32.close()


@pytest.fixture(scope="module")
def single(orc, qd):
    from qdfn import configs
    cfg = configs.single_fault()
    c = _rect_case(orc, qd, cfg, min_elements_to_cut=16, max_levels=7, dist_diam_ratio=1.0,
                   arrange_point=(0, 0, 0))
    E, ks = orc.dense_prepare(c["KS"])
    prob = configs.problem_from_segments(c["cfg"], ks, c["KS"].sum(1), c["KN"].sum(1))
    prob["ext_at_told"] = 1
    c["prob"] = prob
    return c


def test_main_h_tracks_oracle(single, orc, qd):
    from qdfn import configs
    rs, ad = single["st"]["range_SR"], single["st"]["admissible"]
    H = qd.HMatrix.build_rect(single["seg"], single["cfg"]["mu"], single["cfg"]["nu"], rs, ad, aca_rtol=RTOL)
    BS, BN = H.blocks(rs)
    st, ref = orc.sim_hmat(single["prob"], BS, BN, total_step=STEPS, record_step=10)
    assert st == 0
    sim = qd.Simulation(H, single["prob"], record_step=10, record_capacity=STEPS // 10, trace_capacity=STEPS)
    sim.run(STEPS)
    stats, tr = sim.stats(), sim.trace()
    sim.close()
    assert stats["status"] == 0 and stats["steps"] == STEPS
    k = configs.first_event_time(ref["trace_told"], ref["trace_dt"], ref["trace_vmax"])
    kg = configs.first_event_time(tr["told"], tr["dt"], tr["vmax"])
    assert k is not None and kg is not None
    assert abs(kg - k) <= 1e-9 * k
    first = int(np.nonzero(ref["trace_vmax"] >= 1e-3)[0][0])
    np.testing.assert_allclose(tr["told"][:first], ref["trace_told"][:first], rtol=1e-9)
    np.testing.assert_allclose(tr["dt"][:first], ref["trace_dt"][:first], rtol=1e-7)
    # against the dense FP64 oracle run (same external-stress timing)
    st, dref = orc.sim_dense(single["prob"], single["KS"], single["KN"], total_step=STEPS, record_step=0)
    kd = configs.first_event_time(dref["trace_told"], dref["trace_dt"], dref["trace_vmax"])
    assert abs(kg - kd) <= 1e-5 * kd, (kg, kd)
    H.close()


def test_tri_build_and_product(orc, qd):
    from qdfn import configs
    m = configs.rough_fault(nx=16, nz=12, h=60.0, ratio=2e-2)
    P1, P2, P3, nrm, _, _, slip = orc.rotverts_unitvectors(m["P1"], m["P2"], m["P3"], m["rake"])
    ctr = (P1 + P2 + P3) / 3
    st = qd.hmat_structure(ctr, P1, min_elements_to_cut=16, max_levels=8, dist_diam_ratio=1.0,
                           arrange_point=(0, 0, 0))
    p = st["perm"]
    P1, P2, P3, nrm, slip, ctr, rake = P1[p], P2[p], P3[p], nrm[p], slip[p], ctr[p], m["rake"][p]
    mu, nu = m["mu"], m["nu"]
    lam = 2 * mu * nu / (1 - 2 * nu)
    KS, KN = orc.stiffness_tri(P1, P2, P3, rake, ctr, nrm, slip, mu, lam)
    H = qd.HMatrix.build_tri(P1, P2, P3, rake, ctr, nrm, slip, mu, lam, st["range_SR"], st["admissible"],
                             aca_rtol=RTOL)
    blocks = H.blocks(st["range_SR"])
    assert _check_blocks(blocks, KS, KN, st["range_SR"], st["admissible"]) > 0
    x = np.random.default_rng(10).standard_normal(P1.shape[0])
    yS, yN = H.matvec(x, loading_form=False)
    assert np.abs(yS - orc.hmat_matvec(blocks[0], x)).max() <= 1e-12 * np.abs(yS).max()
    assert np.abs(yN - orc.hmat_matvec(blocks[1], x)).max() <= 1e-12 * np.abs(yN).max()
    H.close()


@pytest.mark.parametrize("prec_name", ["PREC_FP64", "PREC_BF16_FAR"])
def test_phase1_lanes_per_row(net, qd, monkeypatch, prec_name):
    """The phase-1 layout with 1 to 64 lanes per V-panel row (QDFN_HMAT_VLPR;
    row sums by DPP moves within 16 lanes, shuffles across rows) gives the
    default layout's products to FP64 summation order."""
    rs, ad = net["st"]["range_SR"], net["st"]["admissible"]
    parts = []
    for B in net["blocks"]:
        parts.append([B["a"][q] if B["kind"][q] == 0 else (B["a"][q], B["b"][q]) for q in range(len(B["a"]))])
    n = net["seg"].shape[0]
    prec = getattr(qd, prec_name)
    x = np.random.default_rng(10).standard_normal(n)
    H0 = qd.HMatrix.from_blocks(n, rs, ad, parts[0], parts[1], prec=prec)
    ref = H0.matvec(x, False)
    H0.close()
    for lanes in ("1", "2", "4", "8", "32", "64"):
        monkeypatch.setenv("QDFN_HMAT_VLPR", lanes)
        H = qd.HMatrix.from_blocks(n, rs, ad, parts[0], parts[1], prec=prec)
        y = H.matvec(x, False)
        H.close()
        for u, v in zip(ref, y):
            assert np.abs(u - v).max() <= 1e-12 * np.abs(u).max(), lanes

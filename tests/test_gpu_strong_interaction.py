"""GPU: ReduceTooStrongInteraction_Hmat (scripts/Functions_Solvers.jl:546-579)
on an H operator built on the device (qdfn_hmat_reduce_strong_interaction):
the changed inadmissible blocks equal the oracle's loop-for-loop restatement
on the same blocks bit for bit (FP64 storage), admissible blocks are
untouched, the count matches, the rebuilt product equals the oracle's
HmatSolver on the changed blocks, and the call is refused once a run exists
on the operator.  The oracle's rule is pinned in tests/test_strong_interaction.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_hmat_reduction_matches_oracle(qd, orc):
    from qdfn import configs
    c = configs.single_fault()
    seg = c["seg"]
    hs = qd.hmat_structure(seg[:, :3], seg[:, :3], min_elements_to_cut=16, max_levels=7, arrange_point=(0, 0, 0))
    segs = seg[hs["perm"]]
    rs, adm = hs["range_SR"], hs["admissible"]
    H = qd.HMatrix.build_rect(segs, c["mu"], c["nu"], rs, adm, aca_rtol=1e-7, prec=qd.PREC_FP64)
    try:
        BS, BN = H.blocks(rs)
        mult = 0.15
        S2, N2, pairs = orc.reduce_too_strong_interaction_hmat(mult, adm, BS, BN)
        assert 0 < len(pairs) < seg.shape[0] ** 2 // 10
        cnt = H.reduce_too_strong_interaction(mult)
        assert cnt == len(pairs)
        CS, CN = H.blocks(rs)
        for q in range(len(adm)):
            for got, ref, before in ((CS, S2, BS), (CN, N2, BN)):
                if adm[q] == 0:
                    np.testing.assert_array_equal(got["a"][q], ref["a"][q])
                else:
                    np.testing.assert_array_equal(got["a"][q], before["a"][q])
                    np.testing.assert_array_equal(got["b"][q], before["b"][q])
        x = np.random.default_rng(4).standard_normal(seg.shape[0])
        yS, yN = H.matvec(x, loading_form=False)
        rS, rN = orc.hmat_matvec(S2, x), orc.hmat_matvec(N2, x)
        scale = max(np.abs(rS).max(), np.abs(rN).max())
        assert np.abs(yS - rS).max() <= 1e-12 * scale
        assert np.abs(yN - rN).max() <= 1e-12 * scale
        # runs read the operator's layout: no further change once one exists
        ks, _ = H.kself()
        hsum, hnsum = H.matvec(np.ones(seg.shape[0]), loading_form=False)
        prob = configs.problem_from_segments(dict(c, seg=segs), ks, hsum, hnsum)
        sim = qd.Simulation(H, prob, record_step=0)
        with pytest.raises(qd.QdfnError):
            H.reduce_too_strong_interaction(mult)
        sim.close()
        assert H.reduce_too_strong_interaction(mult) == 0  # nothing left above the bound
    finally:
        H.close()


def test_hmat_reduction_expands_low_rank_near_field(qd, orc):
    """Tail storage compresses inadmissible blocks off the diagonal too; the
    reduction first expands such a block to its FP64 product U V (components
    added in order) and then scales it, bit for bit as the oracle's rule on
    the expanded blocks."""
    from qdfn import configs
    m = configs.rough_fault(nx=40, nz=30, h=50.0)
    n = m["P1"].shape[0]
    P1, P2, P3, nrm, _, _, slip = qd.RotVerts_UnitVectors(np.hstack([m["P1"], m["P2"], m["P3"]]), n, m["rake"])
    ctr = (P1 + P2 + P3) / 3
    st = qd.hmat_structure(ctr, P1, min_elements_to_cut=64, max_levels=40, dist_diam_ratio=1.0,
                           arrange_point=(0.0, 0.0, 0.0))
    p = st["perm"]
    geo = [a[p] for a in (P1, P2, P3, m["rake"], ctr, nrm, slip)]
    lam = 2 * m["mu"] * m["nu"] / (1 - 2 * m["nu"])
    rs, adm = st["range_SR"], st["admissible"]
    H = qd.HMatrix.build_tri(*geo, m["mu"], lam, rs, adm, aca_rtol=1e-6, prec=qd.PREC_MIXED)
    try:
        rk = H.ranks()
        offdiag = (rs[:, 0] != rs[:, 2]) | (rs[:, 1] != rs[:, 3])
        near_lr = (adm == 0) & offdiag & (rk["kind_S"] == 1)
        assert near_lr.any(), "no near-field block was stored low-rank"
        assert not ((adm == 0) & ~offdiag & (rk["kind_S"] == 1)).any()  # diagonal blocks stay dense
        BS, BN = H.blocks(rs)
        for B in (BS, BN):
            for q in np.nonzero(adm == 0)[0]:
                if B["kind"][q] == 1:
                    U, V = B["a"][q], B["b"][q]
                    d = np.zeros((U.shape[0], V.shape[1]))
                    for a in range(U.shape[1]):
                        d += np.outer(U[:, a], V[a, :])
                    B["a"][q], B["b"][q], B["kind"][q], B["rank"][q] = d, None, 0, 0
        mult = 0.15
        S2, N2, pairs = orc.reduce_too_strong_interaction_hmat(mult, adm, BS, BN)
        cnt = H.reduce_too_strong_interaction(mult)
        assert cnt == len(pairs) > 0
        CS, CN = H.blocks(rs)
        for q in np.nonzero(adm == 0)[0]:
            for got, ref in ((CS, S2), (CN, N2)):
                assert got["kind"][q] == 0
                np.testing.assert_array_equal(got["a"][q], ref["a"][q])
        SS, SN = H.blocks(rs, stored=True)
        x = np.random.default_rng(5).standard_normal(n)
        yS, yN = H.matvec(x, loading_form=False)
        rS, rN = orc.hmat_matvec(SS, x), orc.hmat_matvec(SN, x)
        scale = max(np.abs(rS).max(), np.abs(rN).max())
        assert np.abs(yS - rS).max() <= 1e-12 * scale
        assert np.abs(yN - rN).max() <= 1e-12 * scale
    finally:
        H.close()

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

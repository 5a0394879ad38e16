"""ReduceTooStrongInteraction (scripts/Functions_Solvers.jl:586-607) on dense
matrices: the package's elementwise version equals the oracle's loop-for-loop
restatement bit for bit, and the oracle is pinned by what the rule fixes:
every reduced off-diagonal shear entry ends at magnitude multiple |K_ii| with
its sign (to one rounding), its normal entry scaled by the same ratio, every
other entry (the diagonal included) unchanged, and no entry above the bound
afterwards."""
import numpy as np

from qdfn import assembly


def _mats(n, seed):
    rng = np.random.default_rng(seed)
    KS = rng.standard_normal((n, n))
    KS[np.diag_indices(n)] = -(np.abs(KS[np.diag_indices(n)]) + 1.0)
    return np.asfortranarray(KS), np.asfortranarray(rng.standard_normal((n, n)))


def test_dense_reduction_matches_oracle_and_its_rule(orc, capsys):
    n, c = 60, 0.7
    KS, KN = _mats(n, 1)
    S0, N0, cnt = orc.reduce_too_strong_interaction(c, n, KS, KN)
    S1, N1 = assembly.ReduceTooStrongInteraction(c, n, KS.copy(order="F"), KN.copy(order="F"))
    np.testing.assert_array_equal(S1, S0)
    np.testing.assert_array_equal(N1, N0)
    assert f"{cnt}/{n * n} interaction is smoothened" in capsys.readouterr().out
    assert cnt > 0
    lim = np.abs(np.diag(KS)).reshape(-1, 1) * c
    off = ~np.eye(n, dtype=bool)
    hit = (np.abs(KS) > lim) & off
    assert hit.sum() == cnt
    np.testing.assert_allclose(np.abs(S0[hit]), np.broadcast_to(lim, KS.shape)[hit], rtol=4e-16)
    assert np.all(np.sign(S0[hit]) == np.sign(KS[hit]))
    np.testing.assert_allclose(N0[hit] / KN[hit], S0[hit] / KS[hit], rtol=1e-15)
    np.testing.assert_array_equal(S0[~hit], KS[~hit])
    np.testing.assert_array_equal(N0[~hit], KN[~hit])
    assert np.all(np.abs(S0[off]) <= np.broadcast_to(lim, KS.shape)[off] * (1 + 4e-16))

"""H-matrix structure and the oracle's H-matrix product (CPU only).

* qdfn_hmat_structure (host code of the library, GroupAndSort_AllLevel +
  BuildHierarchy, scripts/Functions_Hmatrix.jl:1-212) against the oracle's
  line-by-line BuildHierarchy on the same cluster list, and against the
  invariants the partition must satisfy: blocks tile the non-loading square
  exactly once, admissible blocks couple disjoint ranges, each level
  partitions its parent, children are ordered by distance to ArrangePoint,
  Block_Ctr_Diam is the mean / bounding-box diagonal of the members.
* orc_hmat_matvec (HmatSolver_Pararllel, :631-646) against a plain dense
  product on a matrix cut into the same blocks, with exact integer-valued
  entries and exact low-rank products, and StiffnessTransitionToLoading
  (:459-477) against its definition.
"""
import numpy as np
import pytest


def _rough(nx=24, nz=16):
    from qdfn import configs
    m = configs.rough_fault(nx=nx, nz=nz, h=60.0, ratio=2e-2)
    ctr = (m["P1"] + m["P2"] + m["P3"]) / 3
    return ctr, m["P1"]


def _network():
    from qdfn import configs
    c = configs.fault_network(n_faults=3, nx=16, nz=12)
    seg = c["seg"]
    return seg[:, 0:3].copy(), seg[:, 0:3].copy()


CASES = [("rough", dict(min_elements_to_cut=32, max_levels=8, dist_diam_ratio=1.0, arrange_point=(0, 0, 0))),
         ("rough", dict(min_elements_to_cut=16, max_levels=10, dist_diam_ratio=0.5, arrange_point=(5e3, 0, 0))),
         ("network", dict(min_elements_to_cut=10, max_levels=8, dist_diam_ratio=1.0,
                          arrange_point=(10000, 6000, 10000))),
         ("network", dict(min_elements_to_cut=20, max_levels=6, dist_diam_ratio=2.0, arrange_point=(0, 0, -1000),
                          n_loading=40))]


def _geom(kind):
    return _rough() if kind == "rough" else _network()


@pytest.mark.parametrize("kind,kw", CASES)
def test_structure_matches_oracle_and_invariants(qd, orc, kind, kw):
    ctr, dpts = _geom(kind)
    n = ctr.shape[0]
    L = kw.get("n_loading", 0)
    st = qd.hmat_structure(ctr, dpts, **kw)
    perm, ni, ng, rs, ad = st["perm"], st["node_info"], st["node_geo"], st["range_SR"], st["admissible"]
    N0 = n - L
    # permutation, loading rows kept last and in place
    assert sorted(perm.tolist()) == list(range(n))
    assert np.array_equal(perm[N0:], np.arange(N0, n))
    # the oracle's BuildHierarchy on the same clusters gives the same blocks in the same order
    lr = (N0 + 1, n) if L else None
    ors, oad = orc.build_hierarchy(ni, ng, kw["max_levels"], kw["dist_diam_ratio"], lr)
    assert np.array_equal(ors, rs)
    assert np.array_equal(oad, ad)
    # tiling: every (receiver, source) pair of the non-loading square once; loading block
    cov = np.zeros((n, n), dtype=np.int32)
    for r in rs:
        cov[r[2] - 1:r[3], r[0] - 1:r[1]] += 1
    assert (cov[:N0, :N0] == 1).all()
    if L:
        assert (cov[:N0, N0:] == 1).all() and (cov[N0:, :] == 0).all()
        assert list(rs[0]) == [N0 + 1, n, 1, N0] and ad[0] == 0
    # admissible blocks couple disjoint clusters
    for r, a in zip(rs, ad):
        if a:
            assert r[1] < r[2] or r[3] < r[0]
    # cluster geometry and ordering
    c_sorted = ctr[perm]
    d_sorted = dpts[perm]
    ap = np.array(kw["arrange_point"], dtype=float)
    for q in range(ni.shape[0]):
        b, e, lev, par = ni[q]
        mem = c_sorted[b - 1:e]
        np.testing.assert_allclose(ng[q, :3], mem.mean(axis=0), rtol=1e-12, atol=1e-9)
        box = d_sorted[b - 1:e]
        np.testing.assert_allclose(ng[q, 3], np.linalg.norm(box.max(0) - box.min(0)), rtol=1e-12, atol=1e-9)
        if par >= 0:
            pb, pe = ni[par, :2]
            assert pb <= b and e <= pe and ni[par, 2] == lev - 1
    for lev in range(1, kw["max_levels"] + 1):
        rows = ni[ni[:, 2] == lev]
        # a level's clusters partition the non-loading rows in order
        assert rows[0, 0] == 1 and rows[-1, 1] == N0
        assert (rows[1:, 0] == rows[:-1, 1] + 1).all()
    # siblings: the first is the one closer to ArrangePoint; unsplit clusters carry over
    for par in np.unique(ni[:, 3]):
        kids = np.where(ni[:, 3] == par)[0]
        if len(kids) == 2:
            d = [np.linalg.norm(ng[k, :3] - ap) for k in kids]
            assert d[0] <= d[1]
        if par >= 0 and len(kids) == 1:
            assert (ni[kids[0], :2] == ni[par, :2]).all()
        if par >= 0 and len(kids) == 2:
            assert ni[par, 1] - ni[par, 0] >= kw["min_elements_to_cut"]


def test_structure_rejects_bad_arguments(qd):
    ctr, dpts = _rough(4, 2)
    with pytest.raises(qd.QdfnError):
        qd.hmat_structure(ctr, dpts, max_levels=0)
    with pytest.raises(qd.QdfnError):
        qd.hmat_structure(ctr, dpts, n_loading=ctr.shape[0] - 1)


def _blocks_of(K, rs, ad, lowrank=None):
    """Cut K into the blocks of rs; admissible blocks given as exact (U, V) pairs when supplied."""
    H = {"n": K.shape[0], "range": rs, "kind": [], "rank": [], "a": [], "b": []}
    for q, r in enumerate(rs):
        blk = K[r[2] - 1:r[3], r[0] - 1:r[1]]
        if ad[q] and lowrank is not None:
            U, V = lowrank[q]
            H["kind"].append(1)
            H["rank"].append(U.shape[1])
            H["a"].append(U)
            H["b"].append(V)
        else:
            H["kind"].append(0)
            H["rank"].append(0)
            H["a"].append(blk.copy())
            H["b"].append(None)
    return H


def test_oracle_hmat_matvec_exact(qd, orc):
    ctr, dpts = _rough(12, 8)
    st = qd.hmat_structure(ctr, dpts, min_elements_to_cut=12, max_levels=6, dist_diam_ratio=1.0,
                           arrange_point=(0, 0, 0))
    rs, ad = st["range_SR"], st["admissible"]
    n = ctr.shape[0]
    rng = np.random.default_rng(3)
    # integer-valued matrix; admissible blocks replaced by exact integer rank-2 products
    K = rng.integers(-50, 50, (n, n)).astype(float)
    lowrank = {}
    for q, r in enumerate(rs):
        if ad[q]:
            m, k = r[3] - r[2] + 1, r[1] - r[0] + 1
            U = rng.integers(-5, 5, (m, 2)).astype(float)
            V = rng.integers(-5, 5, (2, k)).astype(float)
            K[r[2] - 1:r[3], r[0] - 1:r[1]] = U @ V
            lowrank[q] = (U, V)
    x = rng.integers(-20, 20, n).astype(float)
    y_dense = K @ x
    for lr in (None, lowrank):
        H = _blocks_of(K, rs, ad, lr)
        assert np.array_equal(orc.hmat_matvec(H, x), y_dense)
    # StiffnessTransitionToLoading: zero the diagonal of blocks with equal ranges, K_Self = -diag
    H = _blocks_of(K, rs, ad, lowrank)
    HL, ks = orc.hmat_transition_to_loading(H)
    np.testing.assert_array_equal(ks, -np.diag(K))
    K0 = K - np.diag(np.diag(K))
    assert np.array_equal(orc.hmat_matvec(HL, x), K0 @ x)

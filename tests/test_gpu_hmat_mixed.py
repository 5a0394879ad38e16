"""GPU: H operators with QDFN_PREC_MIXED storage — FP32 dense blocks and
leading low-rank components, the trailing components of each recompressed
low-rank part as FP16 unit vectors with an FP64 scale, as many as fit in the
part's unused error budget — and QDFN_PREC_MIXED8, whose smallest trailing
components are FP8 (E4M3) unit vectors scaled to [128, 256) (no reference
counterpart: the reference stores FP64; HmatSolver_Pararllel,
scripts/Functions_Hmatrix.jl:631-646, is the product, main_H,
Functions_RSFDFN3DMain_H.jl:241-248, the loop).

On the 2,400-triangle rough fault (ACA tolerance 1e-6):
* every block as stored (rounded) is within the ACA bound of the stiffness
  assembled on the device: ||H_b - K_b||_F <= rtol ||K_S,b||_F, for both
  matrices — the FP16 tail costs no accuracy class;
* a large share of the factor entries is FP16 (and FP8), and the operator
  moves fewer bytes than FP32 storage (MIXED8: fewer than MIXED);
* the product equals the oracle's HmatSolver on the stored blocks to 1e-12
  (both forms);
* main_H on the stored operator tracks the oracle's main_H on the same stored
  blocks (told 1e-9, onset 1e-9) on the patch-friction mesh, and its onset is
  within 1e-6 of the FP64-storage run's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
RTOL = 1e-6


def _mesh(qd, m):
    n = m["P1"].shape[0]
    P1, P2, P3, nrm, _, _, slip = qd.RotVerts_UnitVectors(np.hstack([m["P1"], m["P2"], m["P3"]]), n, m["rake"])
    ctr = (P1 + P2 + P3) / 3
    st = qd.hmat_structure(ctr, P1, min_elements_to_cut=64, max_levels=40, dist_diam_ratio=1.0,
                           arrange_point=(0.0, 0.0, 0.0))
    p = st["perm"]
    geo = [a[p] for a in (P1, P2, P3, m["rake"], ctr, nrm, slip)]
    lam = 2 * m["mu"] * m["nu"] / (1 - 2 * m["nu"])
    return dict(m=dict(m, P1=P1, P2=P2, P3=P3), st=st, p=p, geo=geo, lam=lam, n=n)


@pytest.fixture(scope="module")
def rough_mesh(qd):
    from qdfn import configs
    return _mesh(qd, configs.rough_fault(nx=40, nz=30, h=50.0))


@pytest.fixture(scope="module", params=["mixed", "mixed8"])
def rough(qd, rough_mesh, request):
    c = dict(rough_mesh)
    st = c["st"]
    c["prec"] = {"mixed": qd.PREC_MIXED, "mixed8": qd.PREC_MIXED8}[request.param]
    c["H"] = qd.HMatrix.build_tri(*c["geo"], c["m"]["mu"], c["lam"], st["range_SR"], st["admissible"],
                                  aca_rtol=RTOL, prec=c["prec"])
    yield c
    c["H"].close()


def test_mixed_blocks_within_tolerance(rough, qd):
    H, st = rough["H"], rough["st"]
    geo = rough["geo"]
    KS, KN = qd.StiffnessMatrix_ByParts_Calculation_Tri(geo[0], geo[1], geo[2], geo[3], geo[4], rough["m"]["mu"],
                                                         rough["lam"], geo[5], geo[6])
    BS, BN = H.blocks(st["range_SR"], stored=True)
    sS, sN = H.split()
    eS, eN = H.split8()
    lo_entries = lr_entries = lo8_entries = 0
    worst = 0.0
    for q, r in enumerate(st["range_SR"]):
        sl = (slice(r[2] - 1, r[3]), slice(r[0] - 1, r[1]))
        tol = RTOL * np.linalg.norm(KS[sl])
        mb, kb = r[3] - r[2] + 1, r[1] - r[0] + 1
        for B, K, rhi, r8 in ((BS, KS, sS, eS), (BN, KN, sN, eN)):
            if B["kind"][q] == 0:
                approx = B["a"][q]
            else:
                rk = B["rank"][q]
                approx = B["a"][q] @ B["b"][q] if rk > 0 else np.zeros((mb, kb))
                lr_entries += rk * (mb + kb)
                lo_entries += (rk - rhi[q]) * (mb + kb)
                lo8_entries += (rk - r8[q]) * (mb + kb)
                assert rhi[q] <= r8[q] <= rk
            err = np.linalg.norm(approx - K[sl])
            worst = max(worst, err / tol if tol > 0 else 0.0)
            # FP32 rounding of dense blocks is outside the ACA budget: 2^-24 per entry
            allow = tol * (1 + 1e-9) if B["kind"][q] == 1 else tol + 2.0 ** -24 * np.linalg.norm(K[sl])
            assert err <= allow + 1e-9 * np.abs(K).max(), (q, err, tol)
    # near-field (inadmissible, off-diagonal) blocks are compressed too where that saves bytes;
    # diagonal blocks stay dense
    rs, adm, rk = st["range_SR"], st["admissible"], H.ranks()
    offdiag = (rs[:, 0] != rs[:, 2]) | (rs[:, 1] != rs[:, 3])
    assert ((adm == 0) & offdiag & (rk["kind_S"] == 1)).any()
    assert not ((adm == 0) & ~offdiag & (rk["kind_S"] == 1)).any()
    print(f"[mixed] worst block error / tolerance {worst:.3f}; 16- and 8-bit share of factor entries "
          f"{lo_entries / lr_entries:.3f}, 8-bit {lo8_entries / lr_entries:.3f}")
    assert lo_entries > 0.3 * lr_entries
    if rough["prec"] == qd.PREC_MIXED8:
        assert lo8_entries > 0.05 * lr_entries
    else:
        assert lo8_entries == 0


def test_mixed_fewer_bytes(rough, qd):
    st = rough["st"]
    H32 = qd.HMatrix.build_tri(*rough["geo"], rough["m"]["mu"], rough["lam"], st["range_SR"], st["admissible"],
                               aca_rtol=RTOL, prec=qd.PREC_FP32)
    b32, bmx = H32.info()["bytes_per_product"], rough["H"].info()["bytes_per_product"]
    H32.close()
    print(f"[mixed] bytes per product: FP32 {b32 / 1e6:.1f} MB, precision {rough['prec']} {bmx / 1e6:.1f} MB")
    assert bmx < 0.9 * b32
    if rough["prec"] == qd.PREC_MIXED8:
        H16 = qd.HMatrix.build_tri(*rough["geo"], rough["m"]["mu"], rough["lam"], st["range_SR"],
                                   st["admissible"], aca_rtol=RTOL, prec=qd.PREC_MIXED)
        b16 = H16.info()["bytes_per_product"]
        H16.close()
        print(f"[mixed] bytes per product: MIXED {b16 / 1e6:.1f} MB")
        assert bmx < b16


def test_mixed_product_matches_oracle(rough, orc):
    H, st = rough["H"], rough["st"]
    BS, BN = H.blocks(st["range_SR"], stored=True)
    x = np.random.default_rng(12).standard_normal(rough["n"])
    yS, yN = H.matvec(x, loading_form=False)
    oS, oN = orc.hmat_matvec(BS, x), orc.hmat_matvec(BN, x)
    scale = np.abs(BS["a"][0]).max() * np.abs(x).sum()
    assert np.abs(yS - oS).max() <= 1e-12 * np.abs(oS).max()
    assert np.abs(yN - oN).max() <= 1e-12 * max(np.abs(oN).max(), 1e-300)
    HL, ks = orc.hmat_transition_to_loading(BS)
    tl, _ = H.matvec(x, loading_form=True)
    otl = orc.hmat_matvec(HL, 0.0 - x) / -ks
    assert np.abs(tl - otl).max() <= 1e-12 * np.abs(otl).max()
    assert scale > 0


@pytest.mark.parametrize("tiered", ["mixed", "mixed8"])
def test_mixed_main_h_tracks_oracle(qd, orc, tiered):
    from qdfn import configs
    c = _mesh(qd, configs.rough_fault_patch(nx=40, nz=30))
    st = c["st"]
    steps = 1200
    runs = {}
    mp = {"mixed": qd.PREC_MIXED, "mixed8": qd.PREC_MIXED8}[tiered]
    for prec in (mp, qd.PREC_FP64):
        H = qd.HMatrix.build_tri(*c["geo"], c["m"]["mu"], c["lam"], st["range_SR"], st["admissible"],
                                 aca_rtol=1e-7, prec=prec)
        ks, _ = H.kself()
        s, nn = H.matvec(np.ones(c["n"]), loading_form=False)
        prob = configs.problem_from_triangles(c["m"], ks, s, nn, order=c["p"])
        prob["ext_at_told"] = 1
        sim = qd.Simulation(H, prob, record_step=0, record_capacity=0, trace_capacity=steps)
        sim.run(steps)
        tr = sim.trace()
        sim.close()
        runs[prec] = tr
        if prec == mp:
            BS, BN = H.blocks(st["range_SR"], stored=True)
            status, ref = orc.sim_hmat(prob, BS, BN, total_step=steps, record_step=0)
            assert status == 0
            k = int(np.nonzero(ref["trace_vmax"] >= 1e-3)[0][0])
            np.testing.assert_allclose(tr["told"][:k], ref["trace_told"][:k], rtol=1e-9)
            e_ref = configs.first_event_time(ref["trace_told"], ref["trace_dt"], ref["trace_vmax"])
            e_gpu = configs.first_event_time(tr["told"], tr["dt"], tr["vmax"])
            assert abs(e_gpu - e_ref) <= 1e-9 * e_ref
        H.close()
    on = {p: configs.first_event_time(t["told"], t["dt"], t["vmax"]) for p, t in runs.items()}
    rel = abs(on[mp] - on[qd.PREC_FP64]) / on[qd.PREC_FP64]
    print(f"[mixed] onset {tiered} {on[mp]:.9e} FP64 {on[qd.PREC_FP64]:.9e} rel {rel:.2e}")
    assert rel <= 1e-6

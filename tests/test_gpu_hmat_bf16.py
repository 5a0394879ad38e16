"""GPU checks of H operators with BF16 low-rank factors (QDFN_PREC_BF16_FAR:
U and V in BF16, dense blocks in FP32, FP64 accumulation), the product of
HmatSolver_Pararllel (scripts/Functions_Hmatrix.jl:631-646) and main_H
(scripts/Functions_RSFDFN3DMain_H.jl:241-248) at that storage.

The reference has no BF16 path, so the expected values are the reference's
FP64 algorithms applied to the rounded blocks: every U and V entry rounded once
to BF16 (8 significant bits, ties to even), every dense entry to FP32, the
diagonals of the shear matrix's diagonal blocks kept in FP64 (the product reads
K_S,ii from the FP64 diagonal).  The rounding is restated here with numpy's
frexp / rint / ldexp, independently of the library's.

* A synthetic operator that exercises every layout case (row groups with dense
  columns only, low-rank columns only, both, rank-0 factors, and panels wider
  than one tile on either side): both products within 1e-12 of the emulation,
  relative to the row's sum of |entry| |x|.
* The 2,400-triangle rough fault (leaves of 64, ACA 1e-6): the same in both
  forms; the difference from the FP32-storage operator within 2^-8 of the
  low-rank terms' magnitude; fewer bytes per product; the error against the
  dense FP64 product is printed next to FP32 storage's.
* main_H on the single fault: the device loop with BF16 factors tracks the
  oracle's main_H on the rounded blocks step by step (told 1e-9, first-event
  time 1e-9).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bf16(a):
    a = np.asarray(a, dtype=np.float64)
    m, e = np.frexp(a)
    r = np.ldexp(np.rint(np.ldexp(m, 8)), e - 8)
    return np.where(np.abs(a) < 2.0 ** -126, 0.0 * a, r)


def _fp32(a):
    return np.asarray(a, dtype=np.float64).astype(np.float32).astype(np.float64)


def test_bf16_rounding_restatement():
    # ties to even at 8 significant bits, and a value that one rounding to
    # FP32 followed by one to BF16 would move (double rounding)
    one = 1.0
    assert _bf16(one + 2.0 ** -8) == one
    assert _bf16(one + 3 * 2.0 ** -8) == one + 2.0 ** -6
    x = one + 2.0 ** -8 + 2.0 ** -30
    assert np.float32(x) == np.float32(one + 2.0 ** -8)
    assert _bf16(x) == one + 2.0 ** -7
    assert _bf16(-x) == -(one + 2.0 ** -7)
    assert _bf16(3.0e-40) == 0.0


def _round_parts(parts, rs, shear):
    """Rounded copies of from_blocks-style parts (array or (U, V))."""
    out = []
    for q, blk in enumerate(parts):
        if isinstance(blk, tuple):
            out.append((_bf16(blk[0]), _bf16(blk[1])))
            continue
        a = _fp32(blk)
        r = rs[q]
        if shear and r[0] == r[2] and r[1] == r[3]:
            d = np.arange(a.shape[0])
            a[d, d] = np.asarray(blk)[d, d]
        out.append(a)
    return out


def _assemble(parts, rs, n, absval=False):
    M = np.zeros((n, n))
    for q, blk in enumerate(parts):
        s0, s1, r0, r1 = (int(v) for v in rs[q])
        if isinstance(blk, tuple):
            U, V = (np.abs(blk[0]), np.abs(blk[1])) if absval else blk
            if U.shape[1] > 0:
                M[r0 - 1:r1, s0 - 1:s1] += U @ V
        else:
            M[r0 - 1:r1, s0 - 1:s1] += np.abs(blk) if absval else blk
    return M


def _check_against_emulation(H, S, N, rs, x, forms=(False, True)):
    n = x.size
    Sr, Nr = _round_parts(S, rs, True), _round_parts(N, rs, False)
    MS, MN = _assemble(Sr, rs, n), _assemble(Nr, rs, n)
    scS = _assemble(Sr, rs, n, True) @ np.abs(x)
    scN = _assemble(Nr, rs, n, True) @ np.abs(x)
    kd = np.diag(MS).copy()
    for form in forms:
        yS, yN = H.matvec(x, loading_form=form)
        if form:
            ref = -(MS @ x - kd * x) / kd
            tol = 1e-12 * scS / np.abs(kd)
        else:
            ref, tol = MS @ x, 1e-12 * scS
        assert np.all(np.abs(yS - ref) <= tol + 1e-300), (form, np.max(np.abs(yS - ref) / (tol + 1e-300)))
        assert np.all(np.abs(yN - MN @ x) <= 1e-12 * scN + 1e-300), form
    return MS, MN


def _parts(blocks):
    out = []
    for B in blocks:
        out.append([B["a"][q] if B["kind"][q] == 0 else (B["a"][q], B["b"][q]) for q in range(len(B["a"]))])
    return out


def test_synthetic_layout_cases(qd, monkeypatch):
    # phase-2 units of 64 KiB so that moderate panels split into several units
    # on both sides (several slots per row group)
    monkeypatch.setenv("QDFN_HMAT_UNIT", "65536")
    rng = np.random.default_rng(31)
    nr, w = 5, 600
    n = nr * w
    rs, ad = [], []
    for bi in range(nr):          # receiver block row
        for bj in range(nr):      # source block column
            rs.append([bj * w + 1, (bj + 1) * w, bi * w + 1, (bi + 1) * w])
            ad.append(0 if bi == bj else 1)
    rs, ad = np.array(rs, dtype=np.int64), np.array(ad, dtype=np.int64)
    # block row 0: dense only (3,000 FP32 columns: several units, no BF16 panel)
    # block row 1: dense diagonal + four factors of rank 300 (1,200 BF16 columns: two units)
    # block row 2: dense diagonal and (2, 3), factors of ranks 0, 7 and 64
    # block row 3: dense diagonal + factors of ranks 1..4 with scales 1e-6..1e6
    # block row 4: factors only (normal matrix; the shear diagonal stays dense)
    ranks = {1: [300, 300, 300, 300], 2: [0, 7, None, 64], 3: [1, 2, 3, 4], 4: [5, 9, 3, 2, 11]}
    S, N = [], []
    for w_, parts in ((0, S), (1, N)):
        for q in range(nr * nr):
            bi, bj = divmod(q, nr)
            scale = 10.0 ** rng.integers(-3, 4)
            dense = bi == 0 or (bi == bj and not (bi == 4 and w_ == 1)) or (bi == 2 and bj == 3)
            if dense:
                parts.append(scale * rng.standard_normal((w, w)))
                continue
            lst = ranks[bi]
            r = lst[(bj if bi == 4 else (bj if bj < bi else bj - 1)) % len(lst)]
            if bi == 3:
                scale = 10.0 ** (6 * (2 * r - 5) / 3)
            U = scale * rng.standard_normal((w, r))
            V = rng.standard_normal((r, w))
            parts.append((U, V))
    H = qd.HMatrix.from_blocks(n, rs, ad, S, N, prec=qd.PREC_BF16_FAR)
    x = rng.standard_normal(n)
    _check_against_emulation(H, S, N, rs, x, forms=(False, True))
    H32 = qd.HMatrix.from_blocks(n, rs, ad, S, N, prec=qd.PREC_FP32)
    assert H.info()["bytes_per_product"] < H32.info()["bytes_per_product"]
    H32.close()
    H.close()


def test_rough_fault_2400(qd):
    from qdfn import configs
    rtol = 1e-6
    m = configs.rough_fault(nx=40, nz=30)
    P1, P2, P3, nrm, _, _, slip = qd.RotVerts_UnitVectors(np.hstack([m["P1"], m["P2"], m["P3"]]),
                                                          m["P1"].shape[0], m["rake"])
    ctr = (P1 + P2 + P3) / 3
    st = qd.hmat_structure(ctr, P1, min_elements_to_cut=64, max_levels=20, dist_diam_ratio=1.0,
                           arrange_point=(0, 0, 0))
    p = st["perm"]
    geo = [v[p] for v in (P1, P2, P3, m["rake"], ctr, nrm, slip)]
    lam = 2 * m["mu"] * m["nu"] / (1 - 2 * m["nu"])
    rs, ad = st["range_SR"], st["admissible"]
    H32 = qd.HMatrix.build_tri(*geo, m["mu"], lam, rs, ad, aca_rtol=rtol, prec=qd.PREC_FP32)
    S, N = _parts(H32.blocks(rs))
    n = geo[0].shape[0]
    assert any(isinstance(b, tuple) and b[0].shape[1] > 0 for b in S)
    Hb = qd.HMatrix.from_blocks(n, rs, ad, S, N, prec=qd.PREC_BF16_FAR)
    x = np.random.default_rng(32).standard_normal(n)
    _check_against_emulation(Hb, S, N, rs, x)
    # difference from FP32 storage: bounded by the factors' rounding
    lowS = [b if isinstance(b, tuple) else np.zeros_like(b) for b in S]
    lowN = [b if isinstance(b, tuple) else np.zeros_like(b) for b in N]
    magS = _assemble(lowS, rs, n, True) @ np.abs(x)
    magN = _assemble(lowN, rs, n, True) @ np.abs(x)
    bS, bN = Hb.matvec(x, loading_form=False)
    fS, fN = H32.matvec(x, loading_form=False)
    # |U~ V~ - U V| <= (2^-8 + 2^-18) |U| |V| entry by entry; FP64 sums in another order
    assert np.all(np.abs(bS - fS) <= 1.001 * 2.0 ** -8 * magS + 1e-12 * np.abs(fS).max())
    assert np.all(np.abs(bN - fN) <= 1.001 * 2.0 ** -8 * magN + 1e-12 * np.abs(fN).max())
    assert Hb.info()["bytes_per_product"] < H32.info()["bytes_per_product"]
    KS, KN = qd.StiffnessMatrix_ByParts_Calculation_Tri(*geo[:5], m["mu"], lam, geo[5], geo[6])
    dS, dN = KS @ x, KN @ x
    rel = lambda y, d: float(np.linalg.norm(y - d) / np.linalg.norm(d))
    print(f"n={n}: product error vs dense FP64, shear / normal: BF16 factors {rel(bS, dS):.3e} / "
          f"{rel(bN, dN):.3e}, FP32 storage {rel(fS, dS):.3e} / {rel(fN, dN):.3e}; bytes per product "
          f"{Hb.info()['bytes_per_product']} vs {H32.info()['bytes_per_product']}", flush=True)
    Hb.close()
    H32.close()


def test_main_h_bf16_tracks_oracle(orc, qd):
    from qdfn import configs
    steps = 1200
    cfg = configs.single_fault()
    seg = cfg["seg"]
    st = qd.hmat_structure(seg[:, :3], seg[:, :3], min_elements_to_cut=16, max_levels=7, dist_diam_ratio=1.0,
                           arrange_point=(0, 0, 0))
    segs = np.ascontiguousarray(seg[st["perm"]])
    cfg = dict(cfg, seg=segs)
    KS, KN = orc.stiffness_rect(segs, segs, cfg["mu"], cfg["nu"])
    _, ks = orc.dense_prepare(KS)
    prob = configs.problem_from_segments(cfg, ks, KS.sum(1), KN.sum(1))
    prob["ext_at_told"] = 1
    rs, ad = st["range_SR"], st["admissible"]
    H64 = qd.HMatrix.build_rect(segs, cfg["mu"], cfg["nu"], rs, ad, aca_rtol=1e-7)
    BS, BN = H64.blocks(rs)
    S, N = _parts((BS, BN))
    H64.close()
    assert any(isinstance(b, tuple) and b[0].shape[1] > 0 for b in S)
    # the oracle's main_H on the rounded blocks
    rounded = []
    for B, parts in ((BS, _round_parts(S, rs, True)), (BN, _round_parts(N, rs, False))):
        R = dict(B, a=[b[0] if isinstance(b, tuple) else b for b in parts],
                 b=[b[1] if isinstance(b, tuple) else None for b in parts])
        rounded.append(R)
    code, ref = orc.sim_hmat(prob, rounded[0], rounded[1], total_step=steps, record_step=10)
    assert code == 0
    H = qd.HMatrix.from_blocks(segs.shape[0], rs, ad, S, N, prec=qd.PREC_BF16_FAR)
    sim = qd.Simulation(H, prob, record_step=10, record_capacity=steps // 10, trace_capacity=steps)
    sim.run(steps)
    stats, tr = sim.stats(), sim.trace()
    sim.close()
    H.close()
    assert stats["status"] == 0 and stats["steps"] == steps
    k = configs.first_event_time(ref["trace_told"], ref["trace_dt"], ref["trace_vmax"])
    kg = configs.first_event_time(tr["told"], tr["dt"], tr["vmax"])
    assert k is not None and kg is not None
    assert abs(kg - k) <= 1e-9 * k
    first = int(np.nonzero(ref["trace_vmax"] >= 1e-3)[0][0])
    np.testing.assert_allclose(tr["told"][:first], ref["trace_told"][:first], rtol=1e-9)
    np.testing.assert_allclose(tr["dt"][:first], ref["trace_dt"][:first], rtol=1e-7)

"""GPU parity of main_H on a triangle mesh — the loop the rough-fault
benchmark times (scripts/Functions_RSFDFN3DMain_H.jl:180-383 with the
HmatSolver coupling at :241-248).

Mesh: a rough strike-slip surface of 40 x 30 cells of 50 m (2,400 TDstressHS
triangles), with a central velocity-weakening patch (a = 0.010, b = 0.015
inside, 0.006 outside, sigma_n = 50 MPa) so that the first event is a
well-conditioned function of the operator.  What this loop adds over the
rectangle loops: the tension-positive triangle K_N feeding the effective
normal stress (Functions_BuildInputFile.jl:575-580, main :218-224), the
triangle FaultLengthStrike mass (:320-321, QUAKEDFN_RunARG.jl:159-166) and
back-slip tables from the H product itself.

* FP64 storage: the device main_H tracks the oracle's main_H on the same
  stored blocks (told 1e-9 and Dt 1e-7 up to the first event, first-event
  onset 1e-9, the run's peak V 1e-6).
* The dense FP64 device run (main) reaches the onset within 1e-7 of the H run
  (ACA tolerance 1e-7; measured 1.5e-9).
* FP32 storage (the benchmark's operator) reaches it within 1e-6 of FP64
  storage (measured 3.7e-8).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 1200


@pytest.fixture(scope="module")
def tri(qd):
    from qdfn import configs
    m = configs.rough_fault_patch(nx=40, nz=30)
    n = m["P1"].shape[0]
    P1, P2, P3, nrm, _, _, slip = qd.RotVerts_UnitVectors(np.hstack([m["P1"], m["P2"], m["P3"]]), n, m["rake"])
    ctr = (P1 + P2 + P3) / 3
    st = qd.hmat_structure(ctr, P1, min_elements_to_cut=64, max_levels=40, dist_diam_ratio=1.0,
                           arrange_point=(0.0, 0.0, 0.0))
    p = st["perm"]
    m = dict(m, P1=P1, P2=P2, P3=P3)
    geo = [a[p] for a in (P1, P2, P3, m["rake"], ctr, nrm, slip)]
    lam = 2 * m["mu"] * m["nu"] / (1 - 2 * m["nu"])
    return dict(m=m, st=st, p=p, geo=geo, lam=lam, n=n)


def _prob(qd, tri, op):
    from qdfn import configs
    ks, _ = op.kself()
    s, nn = op.matvec(np.ones(tri["n"]), loading_form=False)
    prob = configs.problem_from_triangles(tri["m"], ks, s, nn, order=tri["p"])
    prob["ext_at_told"] = 1
    return prob


def _run(qd, op, prob, steps=STEPS):
    sim = qd.Simulation(op, prob, record_step=0, record_capacity=0, trace_capacity=steps)
    sim.run(steps)
    st, tr = sim.stats(), sim.trace()
    sim.close()
    assert st["status"] == 0 and st["steps"] == steps
    return tr


def _onset(tr):
    from qdfn import configs
    return configs.first_event_time(tr["told"], tr["dt"], tr["vmax"])


@pytest.fixture(scope="module")
def h64(qd, tri):
    st = tri["st"]
    H = qd.HMatrix.build_tri(*tri["geo"], tri["m"]["mu"], tri["lam"], st["range_SR"], st["admissible"],
                             aca_rtol=1e-7, prec=qd.PREC_FP64)
    prob = _prob(qd, tri, H)
    tr = _run(qd, H, prob)
    yield dict(H=H, prob=prob, tr=tr)
    H.close()


def test_tri_main_h_tracks_oracle(h64, tri, orc):
    H, prob, tr = h64["H"], h64["prob"], h64["tr"]
    assert not H.normal_zero()
    BS, BN = H.blocks(tri["st"]["range_SR"])
    status, ref = orc.sim_hmat(prob, BS, BN, total_step=STEPS, record_step=0)
    assert status == 0
    k = _onset({"told": ref["trace_told"], "dt": ref["trace_dt"], "vmax": ref["trace_vmax"]})
    kg = _onset(tr)
    assert k is not None and kg is not None
    assert abs(kg - k) <= 1e-9 * k, (kg, k)
    first = int(np.nonzero(ref["trace_vmax"] >= 1e-3)[0][0])
    assert first > 20
    np.testing.assert_allclose(tr["told"][:first], ref["trace_told"][:first], rtol=1e-9)
    np.testing.assert_allclose(tr["dt"][:first], ref["trace_dt"][:first], rtol=1e-7)
    assert abs(tr["vmax"].max() - ref["trace_vmax"].max()) <= 1e-6 * ref["trace_vmax"].max()
    # the effective normal stress really varies (tension-positive K_N of a rough surface)
    _, sN = H.matvec(np.ones(tri["n"]), loading_form=False)
    assert np.abs(sN).max() > 1e-3 * np.abs(H.matvec(np.ones(tri["n"]), loading_form=False)[0]).max()


def test_tri_dense_onset_matches_h(h64, tri, qd):
    d = qd.DenseOperator.from_triangles(*tri["geo"], tri["m"]["mu"], tri["lam"], qd.PREC_FP64)
    tr = _run(qd, d, _prob(qd, tri, d))
    d.close()
    kh, kd = _onset(h64["tr"]), _onset(tr)
    print(f"[tri] onset H {kh:.9e} dense {kd:.9e} rel {abs(kh - kd) / kd:.2e}")
    assert abs(kh - kd) <= 1e-7 * kd, (kh, kd)


def test_tri_fp32_storage_onset(h64, tri, qd):
    st = tri["st"]
    H = qd.HMatrix.build_tri(*tri["geo"], tri["m"]["mu"], tri["lam"], st["range_SR"], st["admissible"],
                             aca_rtol=1e-7, prec=qd.PREC_FP32)
    tr = _run(qd, H, _prob(qd, tri, H))
    H.close()
    k64, k32 = _onset(h64["tr"]), _onset(tr)
    print(f"[tri] onset FP64 {k64:.9e} FP32 {k32:.9e} rel {abs(k64 - k32) / k64:.2e}")
    assert abs(k64 - k32) <= 1e-6 * k64, (k64, k32)

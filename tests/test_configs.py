"""CPU: run set-up of the triangle workloads against the reference's
definitions, on hand-worked cases.

* FaultLengthStrike of a triangle = the largest |component| of the edge
  vectors P1-P2, P2-P3, P1-P3 (Functions_BuildInputFile.jl:320-321).
* FaultMass = min(FaultLengthStrike, FaultLengthDip) * RockDensity / 2 on
  every element, RecTimeStep = min sqrt(M / |K_Self|) / 10, RuptureDt =
  RuptureTimeStepMultiple * RecTimeStep (QUAKEDFN_RunARG.jl:159-201).
* The central velocity-weakening patch of rough_fault_patch.
"""
import numpy as np


def test_triangle_length_scale_hand_worked():
    from qdfn import configs
    P1 = np.array([[0.0, 0.0, 0.0], [10.0, 1.0, -5.0]])
    P2 = np.array([[3.0, 0.0, 0.0], [10.0, 1.0, -7.5]])
    P3 = np.array([[0.0, 0.0, -4.0], [12.0, -1.0, -5.0]])
    # edges of the first: (-3,0,0), (3,0,4), (0,0,4) -> 4; second: (0,0,2.5), (-2,2,-2.5), (-2,2,0) -> 2.5
    np.testing.assert_array_equal(configs.triangle_lengths(P1, P2, P3), [4.0, 2.5])


def test_triangle_problem_mass_and_time_steps():
    from qdfn import configs
    m = configs.rough_fault_patch(nx=4, nz=3, h=50.0)
    n = m["P1"].shape[0]
    kself = np.linspace(1e8, 3e8, n)
    prob = configs.problem_from_triangles(m, kself, np.zeros(n), np.zeros(n))
    L = configs.triangle_lengths(m["P1"], m["P2"], m["P3"]).min()
    np.testing.assert_allclose(prob["mass"], L * m["rho"] / 2, rtol=0, atol=0)
    rec = np.sqrt(L * m["rho"] / 2 / kself).min() / 10
    assert prob["rec_time_step"] == rec
    assert prob["rupture_dt"] == 3.0 * rec
    # steady state at V_i = 1e-9 expressing f0 = 0.6 at 1e-6 under the hard-coded V0 = 1e-9
    np.testing.assert_allclose(prob["friction_i"], 0.6 + (prob["a"] - prob["b"]) * np.log(1e-3), rtol=1e-15)
    np.testing.assert_allclose(prob["theta_i"], 1e-3 / 1e-9, rtol=1e-15)


def test_rough_fault_patch_friction():
    from qdfn import configs
    m = configs.rough_fault_patch(nx=40, nz=30)
    ctr = (m["P1"] + m["P2"] + m["P3"]) / 3
    vw = m["b"] > m["a"]
    assert set(np.unique(m["b"])) == {0.006, 0.015} and np.all(m["a"] == 0.010)
    assert np.abs(ctr[vw, 0]).max() <= 0.3 * 2000.0
    assert np.abs(-ctr[vw, 2] - (100.0 + 750.0)).max() <= 0.3 * 1500.0
    assert 0.25 < vw.mean() < 0.45

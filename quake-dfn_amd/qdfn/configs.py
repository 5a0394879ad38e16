"""Run set-up helpers and the seeded synthetic workloads.

The bulk-geometry subdivision (BulkToSegment, LRtoRake) lives in qdfn.inputs.
Set-up (reference file:line):
  run_setup          scripts/QUAKEDFN_RunARG.jl:159-201 (mass, RecTimeStep, RuptureDt)
  backslip_tables    Tools/ExternalStress/UniformBackslip.jl:75-94 (uniform back-slip loading)

Workloads (all arrays seeded with numpy PCG64):
  single_fault   one vertical strike-slip fault 3.2 x 1.6 km, 32 x 16 elements of 100 m
  fault_network  8 rectangular faults of 32 x 32 elements (8,192 elements)
  rough_fault    a non-planar strike-slip surface with sinusoidal roughness meshed
                 into triangles (100 x 100 cells of 50 m -> 20,000 triangles)
  rough_fault_patch  the same mesh with a central velocity-weakening patch
  ensemble_members  256 planar faults of 500-4,000 elements of 50 m, seeded
                 per-element friction (the batched-ensemble workload)
"""
import numpy as np

SEC_PER_YEAR = 3.15576e7

# RUN3_QUAKEDFN.jl:13-17 defaults
RUN3_DEFAULTS = dict(dt_cut=5.0, switch_v=1e-2, rupture_multiple=3.0, maximum_dt=1e7, vertical_length_scale=0.0)


from .inputs import BulkToSegment, LRtoRake  # noqa: E402,F401
from .inputs import sind as _sind  # noqa: E402


def run_setup(seg, rho, kself, rupture_multiple=3.0, vertical_length_scale=0.0):
    """FaultMass, RecTimeStep and RuptureDt as RunRSFDFN3D computes them."""
    n = seg.shape[0]
    vs = vertical_length_scale if vertical_length_scale != 0 else min(seg[:, 3].min(), seg[:, 4].min())
    mass = np.ones(n) * vs * rho / 2
    period = np.sqrt(mass / np.abs(kself))
    rec = period.min() / 10
    return mass, rec, rec * rupture_multiple


def backslip_tables(ks_row_sums, kn_row_sums, rate=1e-9, t_end=1e11):
    """Back-slip loading tables: stress after t_end seconds of uniform slip
    deficit at `rate` (UniformBackslip.jl:81-86), linearly interpolated."""
    n = ks_row_sums.size
    ext_s = np.zeros((2, n))
    ext_n = np.zeros((2, n))
    ext_s[1] = -ks_row_sums * t_end * rate
    ext_n[1] = -kn_row_sums * t_end * rate
    return {"t": np.array([0.0, t_end]), "n": ext_n, "s": ext_s}


def _friction_init(a, b, v_i=1e-9, f_ref=0.6, v_ref=1e-6):
    """Friction_i at steady state for V_i, expressing f0 = f_ref at V_ref under
    the reference's hard-coded V0 = 1e-9."""
    return f_ref + (a - b) * np.log(v_i / v_ref)


def single_fault(nx=32, nz=16, elem=100.0, mu=32e9, nu=0.25, vs=3464.0, sigma=50e6):
    """Config: one vertical strike-slip fault 3.2 km x 1.6 km, top at the surface."""
    L, W = nx * elem, nz * elem
    bulk = np.zeros((1, 18))
    bulk[0, :8] = [0.0, 0.0, W / 2, L, W, 90.0, 90.0, 1.0]
    bulk[0, 8:18] = [0.010, 0.006, 1e-3, 1e6, 1e-9, 0.6, sigma, 0.0, 0.0, elem]
    seg = BulkToSegment(LRtoRake(1, bulk))
    n = seg.shape[0]
    # velocity-weakening 2 km x 1 km central patch
    x_along = seg[:, 1]  # strike 90 deg: along-strike is the y axis
    depth = seg[:, 2]
    b = np.full(n, 0.006)
    vw = (np.abs(x_along) <= 1000.0) & (np.abs(depth - W / 2) <= 500.0)
    b[vw] = 0.015
    a = np.full(n, 0.010)
    seg[:, 8], seg[:, 9] = a, b
    dc = np.full(n, 1e-3)
    v_i = np.full(n, 1e-9)
    seg[:, 10], seg[:, 12] = dc, v_i
    seg[:, 11] = dc / v_i
    seg[:, 13] = _friction_init(a, b)
    seg[:, 14] = sigma
    rho = mu / vs ** 2
    return dict(name="single_fault_512", seg=seg, mu=mu, nu=nu, rho=rho, min_normal_stress=2e6)


def fault_network(seed=20240611, n_faults=8, nx=32, nz=32, mu=32e9, nu=0.25, vs=3464.0, min_gap=150.0):
    """Config: 8 rectangular faults (length U(2,5) km, depth extent U(1,3) km,
    strike U(0,180), dip U(60,90), centres U in a 10 x 10 km box, tops buried
    >= 100 m), each 32 x 32 elements -> 8,192 elements.  Faults closer than
    min_gap to an already placed fault are redrawn (seeded rejection), since
    intersecting rectangles have no finite interaction."""
    rng = np.random.Generator(np.random.PCG64(seed))
    placed = []
    segs = []
    while len(segs) < n_faults:
        L = rng.uniform(2000, 5000)
        W = rng.uniform(1000, 3000)
        st = rng.uniform(0, 180)
        dp = rng.uniform(60, 90)
        cx, cy = rng.uniform(-5000, 5000, size=2)
        top = rng.uniform(100, 1000)
        cz = top + W / 2 * _sind(dp)
        bulk = np.zeros((1, 18))
        bulk[0, :8] = [cx, cy, cz, L, W, st, dp, rng.choice([-1.0, 1.0])]
        bulk[0, 8:18] = [0.01, 0.015, 1e-3, 1e6, 1e-9, 0.6, 50e6, 0.0, 0.0, max(L / nx, W / nz)]
        seg = BulkToSegment(LRtoRake(1, bulk), counts=[(nx, nz)])
        ok = True
        for other in placed:
            d = np.sqrt(((seg[:, None, :3] - other[None, :, :3]) ** 2).sum(-1)).min()
            if d < min_gap:
                ok = False
                break
        if not ok:
            continue
        placed.append(seg)
        segs.append(seg)
    seg = np.vstack(segs)
    n = seg.shape[0]
    a = rng.uniform(0.008, 0.012, n)
    b = a + rng.uniform(-0.004, 0.006, n)
    sigma = rng.uniform(30e6, 80e6, n)
    seg[:, 8], seg[:, 9] = a, b
    seg[:, 10] = 1e-3
    seg[:, 12] = 1e-9
    seg[:, 11] = seg[:, 10] / seg[:, 12]
    seg[:, 13] = _friction_init(a, b)
    seg[:, 14] = sigma
    rho = mu / vs ** 2
    return dict(name="fault_network_8192", seg=seg, mu=mu, nu=nu, rho=rho, min_normal_stress=2e6)


def ensemble_members(m=256, seed=20240613, n_min=500, n_max=4000, elem=50.0, mu=32e9, nu=0.25, vs=3464.0):
    """Config: m independent planar faults, N_k ~ U{n_min..n_max} (seeded), each
    a vertical strike-slip rectangle of elem-sized elements on a grid
    round(sqrt(2 N)) x round(sqrt(N / 2)) with its top at the surface;
    per-element a ~ U(0.008, 0.012), b - a ~ U(-0.004, 0.006),
    Dc ~ U(5e-4, 5e-3) m, sigma_n ~ U(30, 80) MPa.  Members with the same grid
    have the same geometry (key "grid") and can share one operator."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    rho = mu / vs ** 2
    for k in range(m):
        N = int(rng.integers(n_min, n_max + 1))
        nx, nz = max(2, int(round(np.sqrt(2 * N)))), max(1, int(round(np.sqrt(N / 2))))
        L, W = nx * elem, nz * elem
        bulk = np.zeros((1, 18))
        bulk[0, :8] = [0.0, 0.0, W / 2, L, W, 90.0, 90.0, 1.0]
        bulk[0, 8:18] = [0.01, 0.015, 1e-3, 1e6, 1e-9, 0.6, 50e6, 0.0, 0.0, elem]
        seg = BulkToSegment(LRtoRake(1, bulk), counts=[(nx, nz)])
        n = seg.shape[0]
        a = rng.uniform(0.008, 0.012, n)
        b = a + rng.uniform(-0.004, 0.006, n)
        seg[:, 8], seg[:, 9] = a, b
        seg[:, 10] = rng.uniform(5e-4, 5e-3, n)
        seg[:, 12] = 1e-9
        seg[:, 11] = seg[:, 10] / seg[:, 12]
        seg[:, 13] = _friction_init(a, b)
        seg[:, 14] = rng.uniform(30e6, 80e6, n)
        out.append(dict(name=f"ensemble_member_{k}", seg=seg, mu=mu, nu=nu, rho=rho, min_normal_stress=2e6,
                        grid=(nx, nz)))
    return out


def adjoint_fault(nx=64, nz=64, elem=50.0, seed=20240614, mu=32e9, nu=0.25, vs=3464.0):
    """Config: one vertical strike-slip fault of nx x nz elements of elem m (64 x 64
    of 50 m = 4,096 elements), top at the surface, seeded per-element
    a ~ U(0.008, 0.012), b - a ~ U(-0.004, 0.006), Dc ~ U(5e-4, 5e-3) m,
    sigma_n ~ U(30, 80) MPa (the parameters of the adjoint workload)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L, W = nx * elem, nz * elem
    bulk = np.zeros((1, 18))
    bulk[0, :8] = [0.0, 0.0, W / 2, L, W, 90.0, 90.0, 1.0]
    bulk[0, 8:18] = [0.01, 0.015, 1e-3, 1e6, 1e-9, 0.6, 50e6, 0.0, 0.0, elem]
    seg = BulkToSegment(LRtoRake(1, bulk), counts=[(nx, nz)])
    n = seg.shape[0]
    a = rng.uniform(0.008, 0.012, n)
    b = a + rng.uniform(-0.004, 0.006, n)
    seg[:, 8], seg[:, 9] = a, b
    seg[:, 10] = rng.uniform(5e-4, 5e-3, n)
    seg[:, 12] = 1e-9
    seg[:, 11] = seg[:, 10] / seg[:, 12]
    seg[:, 13] = _friction_init(a, b)
    seg[:, 14] = rng.uniform(30e6, 80e6, n)
    rho = mu / vs ** 2
    return dict(name=f"adjoint_fault_{n}", seg=seg, mu=mu, nu=nu, rho=rho, min_normal_stress=2e6)


def problem_from_segments(cfg, kself, ks_row_sums, kn_row_sums, rate=1e-9, **overrides):
    """Element arrays and run parameters for a rectangle workload: the
    Input_Segment columns plus RUN3 defaults and back-slip loading."""
    seg = cfg["seg"]
    n = seg.shape[0]
    d = dict(RUN3_DEFAULTS)
    d.update(overrides)
    mass, rec, rup = run_setup(seg, cfg["rho"], kself, d["rupture_multiple"], d["vertical_length_scale"])
    prob = dict(n=n, n_loading=0, mass=mass, a=seg[:, 8].copy(), b=seg[:, 9].copy(), dc=seg[:, 10].copy(),
                theta_i=seg[:, 11].copy(), v_i=seg[:, 12].copy(), friction_i=seg[:, 13].copy(),
                sigma0=seg[:, 14].copy(), loading_rate=seg[:, 15].copy(), mu=cfg["mu"], rho=cfg["rho"],
                switch_v=d["switch_v"], dt_cut=d["dt_cut"], rupture_dt=rup, maximum_dt=d["maximum_dt"],
                min_normal_stress=cfg["min_normal_stress"], alpha_evo=0.0, evolution_dr=1, planar=0,
                ext_at_told=0, rec_time_step=rec)
    prob["ext"] = backslip_tables(ks_row_sums, kn_row_sums, rate)
    return prob


def problem_from_bulk(seg, n_loading, mu, rho, min_normal_stress, kself, **overrides):
    """Element arrays and run parameters of RunRSFDFN3D for an Input_Segment
    table as ReadBulkInput returns it (QUAKEDFN_RunARG.jl:36-262): friction and
    stress columns as written, loading faults last (n_loading rows moving at
    their V_Const), FaultMass from the smallest element length, RUN3 defaults,
    no external stress (the loading faults drive the run)."""
    seg = np.asarray(seg, dtype=np.float64)
    n = seg.shape[0]
    d = dict(RUN3_DEFAULTS)
    d.update(overrides)
    mass, rec, rup = run_setup(seg, rho, kself, d["rupture_multiple"], d["vertical_length_scale"])
    return dict(n=n, n_loading=int(n_loading), mass=mass, a=seg[:, 8].copy(), b=seg[:, 9].copy(),
                dc=seg[:, 10].copy(), theta_i=seg[:, 11].copy(), v_i=seg[:, 12].copy(),
                friction_i=seg[:, 13].copy(), sigma0=seg[:, 14].copy(), loading_rate=seg[:, 15].copy(), mu=mu,
                rho=rho, switch_v=d["switch_v"], dt_cut=d["dt_cut"], rupture_dt=rup, maximum_dt=d["maximum_dt"],
                min_normal_stress=min_normal_stress, alpha_evo=0.0, evolution_dr=1, planar=0, ext_at_told=0,
                rec_time_step=rec, ext=None)


def first_event_time(told, dt, vmax, threshold=1e-3):
    """Onset of the first event: the first accepted step whose max V reaches
    `threshold`, placed by linear interpolation in log V between that step and
    the one before (time axis = end of each step)."""
    t_end = np.asarray(told) + np.asarray(dt)
    v = np.asarray(vmax)
    idx = np.nonzero(v >= threshold)[0]
    if idx.size == 0:
        return None
    k = idx[0]
    if k == 0:
        return float(t_end[0])
    l0, l1 = np.log10(v[k - 1]), np.log10(v[k])
    lt = np.log10(threshold)
    f = (lt - l0) / (l1 - l0) if l1 != l0 else 1.0
    return float(t_end[k - 1] + f * (t_end[k] - t_end[k - 1]))


def rough_fault(nx=100, nz=100, h=50.0, top=100.0, seed=20240612, ratio=1e-2, n_modes=3, mu=32e9, nu=0.25,
                vs=3464.0):
    """Config: a vertical strike-slip surface y = sum_m A_m sin(2 pi k_m . (x, z) + phi_m)
    (A_m / lambda_m = ratio, n_modes seeded modes), nx x nz cells of h metres,
    each split into two triangles; top edge at depth `top` (triangle vertices
    need z <= 0).  Returns vertex arrays P1, P2, P3 (n x 3, EFCS: z up) and
    per-element friction drawn as in the fault network."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Lx, Lz = nx * h, nz * h
    lam = rng.uniform(0.2, 1.0, n_modes) * max(Lx, Lz)
    ang = rng.uniform(0, np.pi, n_modes)
    phi = rng.uniform(0, 2 * np.pi, n_modes)
    amp = ratio * lam

    def yoff(x, z):
        out = np.zeros_like(x)
        for m in range(n_modes):
            kx, kz = np.cos(ang[m]) / lam[m], np.sin(ang[m]) / lam[m]
            out += amp[m] * np.sin(2 * np.pi * (kx * x + kz * z) + phi[m])
        return out

    xs = np.arange(nx + 1) * h - Lx / 2
    zs = -top - np.arange(nz + 1) * h
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    Y = yoff(X, Z)
    V = np.stack([X, Y, Z], -1)
    tris = []
    for i in range(nx):
        for j in range(nz):
            a, b, c, d = V[i, j], V[i + 1, j], V[i + 1, j + 1], V[i, j + 1]
            tris.append((a, b, c))
            tris.append((a, c, d))
    P = np.array(tris)
    P1, P2, P3 = P[:, 0].copy(), P[:, 1].copy(), P[:, 2].copy()
    n = P1.shape[0]
    a = rng.uniform(0.008, 0.012, n)
    b = a + rng.uniform(-0.004, 0.006, n)
    sigma = rng.uniform(30e6, 80e6, n)
    rake = np.zeros(n)
    return dict(name=f"rough_fault_{n}", P1=P1, P2=P2, P3=P3, rake=rake, a=a, b=b, sigma=sigma,
                dc=np.full(n, 1e-3), mu=mu, nu=nu, rho=mu / vs ** 2, min_normal_stress=2e6, h=h)


def rough_fault_patch(nx=40, nz=30, h=50.0, top=100.0, seed=20240612, ratio=1e-2, patch=0.3, sigma=50e6):
    """The rough-fault mesh of rough_fault() with the friction of the single
    fault: a = 0.010 everywhere, b = 0.015 on a central velocity-weakening
    patch (triangle centres within patch * length of the mid-point along
    strike and along depth) and 0.006 elsewhere, uniform sigma_n, Dc = 1e-3 m.
    Its first event is a well-conditioned function of the operator (the
    loop parity tests on triangles use it)."""
    m = rough_fault(nx=nx, nz=nz, h=h, top=top, seed=seed, ratio=ratio)
    ctr = (m["P1"] + m["P2"] + m["P3"]) / 3
    n = ctr.shape[0]
    Lx, Lz = nx * h, nz * h
    vw = (np.abs(ctr[:, 0]) <= patch * Lx) & (np.abs(-ctr[:, 2] - (top + Lz / 2)) <= patch * Lz)
    b = np.full(n, 0.006)
    b[vw] = 0.015
    return dict(m, name=f"rough_fault_patch_{n}", a=np.full(n, 0.010), b=b, sigma=np.full(n, float(sigma)),
                dc=np.full(n, 1e-3))


def problem_from_triangles(m, kself, ks_row_sums, kn_row_sums, rate=1e-9, order=None, **overrides):
    """Element arrays and run parameters for a triangle workload (rows in
    `order` when given): FaultMass from the smallest triangle length scale
    (FaultLengthStrike = FaultLengthDip of ReadSegmentInput,
    Functions_BuildInputFile.jl:320-321), friction initialised at steady state
    with V_i = 1e-9, back-slip loading."""
    idx = np.arange(m["a"].size) if order is None else np.asarray(order)
    n = idx.size
    d = dict(RUN3_DEFAULTS)
    d.update(overrides)
    lengths = triangle_lengths(m["P1"][idx], m["P2"][idx], m["P3"][idx])
    vs = d["vertical_length_scale"] if d["vertical_length_scale"] != 0 else lengths.min()
    mass = np.ones(n) * vs * m["rho"] / 2
    rec = np.sqrt(mass / np.abs(kself)).min() / 10
    a, b = m["a"][idx].copy(), m["b"][idx].copy()
    dc = m["dc"][idx].copy()
    v_i = np.full(n, 1e-9)
    prob = dict(n=n, n_loading=0, mass=mass, a=a, b=b, dc=dc, theta_i=dc / v_i, v_i=v_i,
                friction_i=_friction_init(a, b), sigma0=m["sigma"][idx].copy(), loading_rate=np.zeros(n),
                mu=m["mu"], rho=m["rho"], switch_v=d["switch_v"], dt_cut=d["dt_cut"],
                rupture_dt=rec * d["rupture_multiple"], maximum_dt=d["maximum_dt"],
                min_normal_stress=m["min_normal_stress"], alpha_evo=0.0, evolution_dr=1, planar=0,
                ext_at_told=0, rec_time_step=rec)
    prob["ext"] = backslip_tables(ks_row_sums, kn_row_sums, rate)
    return prob


def triangle_lengths(P1, P2, P3):
    """FaultLengthStrike of a triangle as ReadSegmentInput defines it
    (Functions_BuildInputFile.jl:320): the largest |component| of the three
    edge vectors."""
    return np.max(np.abs(np.hstack([P1 - P2, P2 - P3, P1 - P3])), axis=1)

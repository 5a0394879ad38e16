# QuakeDFNHIP.jl — ccall bindings of libqdfn_hip.so with the signatures of the
# Quake-DFN Julia functions they replace.  `include` this file instead of
# scripts/Functions_OKADA3D.jl / the assembly part of
# scripts/Functions_BuildInputFile.jl / scripts/Functions_Solvers.jl, or call
# RunDense(...) in place of main(...).  Julia is not available in this build's
# container, so this file is not exercised by the test suite; the C ABI it
# binds is (through ctypes).
module QuakeDFNHIP

const LIB = get(ENV, "QDFN_LIB", joinpath(@__DIR__, "..", "qdfn", "libqdfn_hip.so"))

struct QdfnError <: Exception
    code::Cint
    msg::String
end

function check(rc)
    if rc < 0
        msg = unsafe_string(ccall((:qdfn_last_error, LIB), Cstring, ()))
        throw(QdfnError(Cint(rc), msg))
    end
    return rc
end

# Functions_OKADA3D.jl:464
function Okada_DC3D_Vector(ALPHA, X, Y, Z, DEPTH, DIP_Angle, AL1, AL2, AW1, AW2, DISL1, DISL2, DISL3)
    X = Vector{Float64}(X); Y = Vector{Float64}(Y); Z = Vector{Float64}(Z)
    n = length(X)
    U = zeros(n, 12)
    IRET = zeros(n)
    GC.@preserve X Y Z U IRET begin
        check(ccall((:qdfn_okada_dc3d, LIB), Cint,
                    (Float64, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Float64, Float64,
                     Float64, Float64, Float64, Float64, Float64, Ptr{Float64}, Ptr{Float64}),
                    ALPHA, n, X, Y, Z, DEPTH, DIP_Angle, AL1, AL2, AW1, AW2, DISL1, DISL2, DISL3, U, IRET))
    end
    return (U[:, 1], U[:, 2], U[:, 3], U[:, 4], U[:, 5], U[:, 6], U[:, 7], U[:, 8], U[:, 9], U[:, 10], U[:, 11],
            U[:, 12], IRET)
end

# Functions_BuildInputFile.jl:386
function StiffnessMatrix_ByParts_Calculation_Rec(Input_SegmentSource, Input_SegmentReceiver, ShearModulus,
                                                 PoissonRatio, CurrentPart=1, TotalParts=1)
    src = Matrix{Float64}(Input_SegmentSource); rcv = Matrix{Float64}(Input_SegmentReceiver)
    ns, nr = size(src, 1), size(rcv, 1)
    KS = zeros(nr, ns); KN = zeros(nr, ns)
    GC.@preserve src rcv KS KN begin
        check(ccall((:qdfn_okada_stiffness_rect, LIB), Cint,
                    (Ptr{Float64}, Int64, Int64, Ptr{Float64}, Int64, Int64, Float64, Float64, Ptr{Float64},
                     Ptr{Float64}, Int64),
                    src, ns, ns, rcv, nr, nr, ShearModulus, PoissonRatio, KS, KN, nr))
    end
    return KS, KN
end

# Functions_BuildInputFile.jl:516 (one device launch covers the whole matrix)
BuildMatrixByParts_Even_Rec(FaultCount, Input_Segment, ShearModulus, PoissonRatio) =
    StiffnessMatrix_ByParts_Calculation_Rec(Input_Segment, Input_Segment, ShearModulus, PoissonRatio)

# Functions_TDstressHS.jl:2 (P1..P3 are not mutated)
function TDstressHS(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, mu, lambda)
    X = Vector{Float64}(X); Y = Vector{Float64}(Y); Z = Vector{Float64}(Z)
    p1 = Vector{Float64}(P1); p2 = Vector{Float64}(P2); p3 = Vector{Float64}(P3)
    n = length(X)
    Stress = zeros(n, 6); Strain = zeros(n, 6)
    GC.@preserve X Y Z p1 p2 p3 Stress Strain begin
        check(ccall((:qdfn_tde_stress_hs, LIB), Cint,
                    (Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                     Float64, Float64, Float64, Float64, Float64, Ptr{Float64}, Ptr{Float64}),
                    n, X, Y, Z, p1, p2, p3, Ss, Ds, Ts, mu, lambda, Stress, Strain))
    end
    return Stress, Strain
end

# Functions_BuildInputFile.jl:551
function StiffnessMatrix_ByParts_Calculation_Tri(P1, P2, P3, Rake, FaultCenter, ShearModulus, lambda,
                                                 UnitVector_Normal, UnitVector_Slip)
    P1 = Matrix{Float64}(P1); P2 = Matrix{Float64}(P2); P3 = Matrix{Float64}(P3)
    Rake = Vector{Float64}(vec(Rake)); C = Matrix{Float64}(FaultCenter)
    Nrm = Matrix{Float64}(UnitVector_Normal); Slp = Matrix{Float64}(UnitVector_Slip)
    ns, nr = size(P1, 1), size(C, 1)
    KS = zeros(nr, ns); KN = zeros(nr, ns)
    GC.@preserve P1 P2 P3 Rake C Nrm Slp KS KN begin
        check(ccall((:qdfn_tde_stiffness_tri, LIB), Cint,
                    (Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64},
                     Ptr{Float64}, Ptr{Float64}, Float64, Float64, Ptr{Float64}, Ptr{Float64}, Int64),
                    ns, P1, P2, P3, Rake, nr, C, Nrm, Slp, ShearModulus, lambda, KS, KN, nr))
    end
    return KS, KN
end

struct StepParams
    conv_crit::Float64
    v0::Float64
    mu::Float64
    rho::Float64
    switch_v::Float64
    alpha_evo::Float64
    evolution_dr::Int32
    pad::Int32
end

# Functions_Solvers.jl:334 — same argument list; fills the output arrays in place
function SolveOneTimeStep(ConvergenceCrit, DispOld, FrictionOld, ThetaOld, VOld, Friction0, a, b, Dc, V0, K_Self,
                          Dt, Omega, Mass, ShearModulus, EffNormalStress_i, Total_Loading_Disp, SolverSwitch, V,
                          Friction, Disp, Theta, EffNormalStress, Dt_All, InstabilityThistime, LoadingRate,
                          LoadingFaultCount, FaultCount, RockDensity, D_EffStress_Shear, SwitchV, Alpha_Evo,
                          EffNormalStress_Old, EvolutionDR, InitialShearStress)
    f(x) = Vector{Float64}(vec(x))
    p = Ref(StepParams(ConvergenceCrit, V0, ShearModulus, RockDensity, SwitchV, Alpha_Evo, Int32(EvolutionDR), 0))
    ins = map(f, (DispOld, FrictionOld, ThetaOld, VOld, Friction0, a, b, Dc, K_Self))
    om, ms, si, tl = map(f, (Omega, Mass, EffNormalStress_i, Total_Loading_Disp))
    sw = Vector{Int32}(vec(SolverSwitch))
    lr, dt, so, t0 = map(f, (LoadingRate, D_EffStress_Shear, EffNormalStress_Old, InitialShearStress))
    n = FaultCount
    outs = [zeros(n) for _ in 1:6]
    inst = zeros(Int32, n); flag = zeros(Int32, n); its = zeros(Int32, n)
    GC.@preserve p ins om ms si tl sw lr dt so t0 outs inst flag its begin
        check(ccall((:qdfn_solve_one_timestep, LIB), Int64,
                    (Ptr{StepParams}, Int64, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                     Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Ptr{Float64},
                     Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Int32}, Ptr{Float64}, Ptr{Float64},
                     Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                     Ptr{Float64}, Ptr{Float64}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}),
                    p, n, LoadingFaultCount, ins..., Dt, om, ms, si, tl, sw, lr, dt, so, t0, outs..., inst, flag,
                    its))
    end
    V .= outs[1]; Friction .= outs[2]; Disp .= outs[3]; Theta .= outs[4]; EffNormalStress .= outs[5]
    Dt_All .= outs[6]; InstabilityThistime .= inst
    return V, Friction, Disp, Theta, EffNormalStress, Dt_All, InstabilityThistime, Float64.(flag)
end

# ---- H-matrix (Hmatrix = true) --------------------------------------------

# H operator from the reference's ShearStiffness_H / NormalStiffness_H (RUN2_BUILD_INPUT.jl:362):
# dense blocks pass as they are; a LowRankApprox PartialQR F (A[:, F[:p]] ~ F[:Q] F[:R]) passes as
# U = F[:Q] and V with V[:, F[:p]] = F[:R].  prec: 0 FP64, 1 FP32 storage, 2 BF16 low-rank
# factors with FP32 dense blocks (QDFN_PREC_BF16_FAR), 3 FP32 with the trailing components of each
# recompressed low-rank block as FP16 unit vectors within the block's ACA budget (QDFN_PREC_MIXED;
# imported blocks are not recompressed and stay FP32), 4 the same with the smallest of those
# components in FP8 (QDFN_PREC_MIXED8); FP64 accumulation in every case.
function HmatCreate(FaultCount, ElementRange_SR, Admissible, ShearStiffness_H, NormalStiffness_H; prec=0)
    nb = size(ElementRange_SR, 1)
    rng = Matrix{Int64}(permutedims(ElementRange_SR))          # 4 x nb = nb x 4 row-major
    adm = Vector{Int64}(vec(Admissible))
    keep = Any[]
    parts = map((ShearStiffness_H, NormalStiffness_H)) do H
        kind = zeros(Int32, nb); rank = zeros(Int64, nb)
        A = Vector{Ptr{Float64}}(undef, nb); B = fill(Ptr{Float64}(0), nb)
        for q in 1:nb
            blk = H[q]
            if blk isa AbstractMatrix
                M = Matrix{Float64}(blk); push!(keep, M); A[q] = pointer(M)
            else
                k = size(blk[:R], 2); r = size(blk[:Q], 2)
                U = Matrix{Float64}(blk[:Q]); V = zeros(r, k); V[:, blk[:p]] = blk[:R]
                push!(keep, U); push!(keep, V)
                kind[q] = 1; rank[q] = r; A[q] = pointer(U); B[q] = pointer(V)
            end
        end
        push!(keep, kind, rank, A, B)
        (kind, rank, A, B)
    end
    h = GC.@preserve rng adm keep ccall((:qdfn_hmat_create, LIB), Ptr{Cvoid},
        (Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int32}, Ptr{Int64}, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}},
         Ptr{Int32}, Ptr{Int64}, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}, Int32),
        FaultCount, nb, rng, adm, parts[1]..., parts[2]..., Int32(prec))
    h == C_NULL && check(-1)
    return h
end

# HmatBuild_R (Functions_BuildInputFile.jl:725) on the device: returns an operator handle
# (ACA with full pivoting + SVD recompression, Frobenius tolerance relative to each block's
# shear norm; Tolerance in the reference's absolute sense can be passed as atol with rtol = 0).
# sampled = true: admissible blocks from sampled rows and columns, never assembled in full
# (qdfn_hmat_build_rect_sampled; n is then not bounded by n^2 device memory).  prec as for HmatCreate.
function HmatBuild_R_Device(ShearModulus, PoissonRatio, ElementRange_SR, Input_Segment, Admissible;
                            rtol=1e-6, atol=0.0, prec=1, sampled=false)
    seg = Matrix{Float64}(Input_Segment); n = size(seg, 1)
    rng = Matrix{Int64}(permutedims(ElementRange_SR)); adm = Vector{Int64}(vec(Admissible))
    fn = sampled ? :qdfn_hmat_build_rect_sampled : :qdfn_hmat_build_rect
    h = GC.@preserve seg rng adm ccall((fn, LIB), Ptr{Cvoid},
        (Ptr{Float64}, Int64, Int64, Float64, Float64, Int64, Ptr{Int64}, Ptr{Int64}, Float64, Float64, Int32),
        seg, n, n, ShearModulus, PoissonRatio, length(adm), rng, adm, rtol, atol, Int32(prec))
    h == C_NULL && check(-1)
    return h
end

HmatDestroy(h) = ccall((:qdfn_hmat_destroy, LIB), Cvoid, (Ptr{Cvoid},), h)

# HmatSolver_Pararllel (Functions_Hmatrix.jl:631) for both matrices: (shear, normal) products.
# ReduceTooStrongInteraction_Hmat (Functions_Solvers.jl:546) on a device H operator h, in place of the
# block arrays (QUAKEDFN_RunARG.jl:126-131); before runs are created on h.
function ReduceTooStrongInteraction_Hmat(StrongInteractionCriteriaMultiple, h, FaultCount)
    cnt = Ref{Int64}(0)
    check(ccall((:qdfn_hmat_reduce_strong_interaction, LIB), Cint, (Ptr{Cvoid}, Float64, Ref{Int64}), h,
                StrongInteractionCriteriaMultiple, cnt))
    println(cnt[], "/", FaultCount^2, " interaction is smoothened")
    return h
end

function HmatSolver(h, x, n; loading_form=false)
    xv = Vector{Float64}(vec(x)); yS = zeros(n); yN = zeros(n)
    GC.@preserve xv yS yN check(ccall((:qdfn_hmat_matvec, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int32), h, xv, yS, yN, Int32(loading_form)))
    return yS, yN
end

# ---- event catalog (Functions_Plot.jl:430-579) ------------------------------

function _catalog_fragments(c)
    cnt = Ref{Int64}(0)
    check(ccall((:qdfn_catalog_fragments, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Float64}, Ptr{Float64},
          Ptr{Float64}, Ptr{Int32}), c, cnt, C_NULL, C_NULL, C_NULL, C_NULL))
    F = cnt[]; t = zeros(F); loc = zeros(F, 3); m = zeros(F)
    F > 0 && check(ccall((:qdfn_catalog_fragments, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Float64},
                         Ptr{Float64}, Ptr{Float64}, Ptr{Int32}), c, cnt, t, loc, m, C_NULL))
    return t, loc, m, F
end

# get_event_fragments with the arrays the reference loads from Result.jld2 / Result_Input.jld2
# (History_Time, History_Disp, FaultCenter, FaultArea as :453-465 computes it, ShearModulus).
function get_event_fragments(ResultTime, ResultDisp, FaultCenter, FaultArea, ShearModulus, DispRateCrits)
    t = Vector{Float64}(vec(ResultTime)); D = Matrix{Float64}(ResultDisp)
    ctr = Matrix{Float64}(FaultCenter); A = Vector{Float64}(vec(FaultArea))
    c = GC.@preserve t D ctr A ccall((:qdfn_catalog_from_history, LIB), Ptr{Cvoid},
        (Int64, Int64, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}, Float64, Float64),
        length(t), size(D, 2), t, D, max(size(D, 1), 1), ctr, A, ShearModulus, DispRateCrits)
    c == C_NULL && check(-1)
    out = _catalog_fragments(c)
    ccall((:qdfn_catalog_destroy, LIB), Cvoid, (Ptr{Cvoid},), c)
    return out
end

# merge_fragments: same arguments and returns; with no fragments, the reference's placeholders.
function merge_fragments(EventTime_Fragment, EventLocation_Fragment, EventMoment_Fragment, EventCount_Fragment,
                         MergeTimeCriteria, MergeDistanceCriteria)
    F = EventCount_Fragment
    F == 0 && return [0;], zeros(0), [0 0 0]
    t = Vector{Float64}(EventTime_Fragment[1:F]); loc = Matrix{Float64}(EventLocation_Fragment[1:F, :])
    m = Vector{Float64}(EventMoment_Fragment[1:F])
    c = GC.@preserve t loc m ccall((:qdfn_catalog_from_fragments, LIB), Ptr{Cvoid},
        (Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), F, t, loc, m)
    c == C_NULL && check(-1)
    E = Ref{Int64}(0)
    check(ccall((:qdfn_catalog_merge, LIB), Cint, (Ptr{Cvoid}, Float64, Float64, Ref{Int64}), c,
                MergeTimeCriteria, MergeDistanceCriteria, E))
    tm = zeros(E[]); mw = zeros(E[]); lm = zeros(E[], 3)
    check(ccall((:qdfn_catalog_events, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Float64}, Ptr{Float64},
          Ptr{Float64}, Ptr{Float64}, Ptr{Int32}), c, E, tm, mw, lm, C_NULL, C_NULL))
    ccall((:qdfn_catalog_destroy, LIB), Cvoid, (Ptr{Cvoid},), c)
    return tm, mw, lm
end

# ---- device-resident runs (main, Functions_RSFDFN3DMain_D.jl:2) and ensembles ----

# qdfn_sim_params, field for field (include/qdfn.h)
struct SimParams
    n::Int64; n_loading::Int64
    mu::Float64; rho::Float64; switch_v::Float64; dt_cut::Float64; rupture_dt::Float64; maximum_dt::Float64
    min_normal_stress::Float64; alpha_evo::Float64
    evolution_dr::Int32; planar::Int32; ext_at_told::Int32; pad::Int32
    record_step::Int64; record_capacity::Int64; trace_capacity::Int64
end

# Dense operator from an Input_Segment table (assembled on the device); prec 0 FP64, 1 FP32.
function DenseCreateRect(Input_Segment, ShearModulus, PoissonRatio; prec=1)
    seg = Matrix{Float64}(Input_Segment); n = size(seg, 1)
    h = GC.@preserve seg ccall((:qdfn_dense_create_rect, LIB), Ptr{Cvoid},
        (Ptr{Float64}, Int64, Int64, Float64, Float64, Int32), seg, n, n, ShearModulus, PoissonRatio, Int32(prec))
    h == C_NULL && check(-1)
    return h
end

# Column splits of an operator's coupling product (0: the single-run default); operators of
# runs advanced together by EnsembleRun take about 1.5 MB of matrix per split (qdfn.ensemble).
DenseSetSplits(op, splits) = check(ccall((:qdfn_dense_set_splits, LIB), Cint, (Ptr{Cvoid}, Int32), op, Int32(splits)))

# main()'s set-up on the device: element arrays as QUAKEDFN_RunARG.jl passes them.
function SimCreateDense(op, p::SimParams, FaultMass, Fault_a, Fault_b, Fault_Dc, Fault_Theta_i, Fault_V_i,
                        Fault_Friction_i, Fault_NormalStress, Fault_V_Const)
    arrs = map(a -> Vector{Float64}(vec(a)), (FaultMass, Fault_a, Fault_b, Fault_Dc, Fault_Theta_i, Fault_V_i,
                                              Fault_Friction_i, Fault_NormalStress, Fault_V_Const))
    pr = Ref(p)
    h = GC.@preserve arrs pr ccall((:qdfn_sim_create_dense, LIB), Ptr{Cvoid},
        (Ptr{Cvoid}, Ref{SimParams}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), op, pr, arrs...)
    h == C_NULL && check(-1)
    return h
end

# Input_ExternalStressChange tables: times (nt), ext_n / ext_s nt x n.
function SimSetExternal(sim, times, ext_n, ext_s)
    t = Vector{Float64}(vec(times)); en = Matrix{Float64}(ext_n); es = Matrix{Float64}(ext_s)
    GC.@preserve t en es check(ccall((:qdfn_sim_set_external, LIB), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), sim, length(t), t, en, es))
end

SimRun(sim, n_steps; mode=0, graph_attempts=32) =
    check(ccall((:qdfn_sim_run, LIB), Cint, (Ptr{Cvoid}, Int64, Int32, Int32), sim, n_steps, Int32(mode),
                Int32(graph_attempts)))

# main()'s loop to a simulated time with its periodic save (Functions_RSFDFN3DMain_D.jl:265-297):
# on_records(first, History_Time, History_Dt, History_V, History_Disp, History_Theta,
# History_NormalStress) receives every batch of records (copies); return true to stop.
function _record_trampoline(user::Ptr{Cvoid}, first::Int64, count::Int64, t::Ptr{Float64}, dt::Ptr{Float64},
                            V::Ptr{Float64}, D::Ptr{Float64}, Th::Ptr{Float64}, S::Ptr{Float64})::Cint
    f = unsafe_pointer_to_objref(user)::Base.RefValue{Any}
    n = f[][2]
    w(p, k) = copy(unsafe_wrap(Array, p, k))
    m(p) = reshape(w(p, count * n), count, n)
    stop = f[][1](first, w(t, count), w(dt, count), m(V), m(D), m(Th), m(S))
    return stop === true ? Cint(1) : Cint(0)
end

function SimRunUntil(sim, t_end, max_steps, n; on_records=nothing, graph_attempts=32)
    if on_records === nothing
        return check(ccall((:qdfn_sim_run_until, LIB), Cint, (Ptr{Cvoid}, Float64, Int64, Int32, Ptr{Cvoid}, Ptr{Cvoid}),
                           sim, t_end, max_steps, Int32(graph_attempts), C_NULL, C_NULL))
    end
    box = Ref{Any}((on_records, n))
    cb = @cfunction(_record_trampoline, Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                                               Ptr{Float64}, Ptr{Float64}, Ptr{Float64}))
    GC.@preserve box check(ccall((:qdfn_sim_run_until, LIB), Cint,
                                 (Ptr{Cvoid}, Float64, Int64, Int32, Ptr{Cvoid}, Ptr{Cvoid}),
                                 sim, t_end, max_steps, Int32(graph_attempts), cb, pointer_from_objref(box)))
end

# History_Time, History_Dt, History_V, History_Disp, History_Theta, History_NormalStress (count x n).
function SimHistory(sim, count, n)
    t = zeros(count); dt = zeros(count); V = zeros(count, n); D = zeros(count, n); Th = zeros(count, n)
    S = zeros(count, n)
    GC.@preserve t dt V D Th S check(ccall((:qdfn_sim_get_history, LIB), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
        sim, count, t, dt, V, D, Th, S))
    return t, dt, V, D, Th, S
end

SimDestroy(sim) = ccall((:qdfn_sim_destroy, LIB), Cvoid, (Ptr{Cvoid},), sim)

# Many runs (one per input file) advanced together; runs on the same operator share its reads.
function EnsembleCreate(sims::Vector{Ptr{Cvoid}}; max_rhs=0)
    e = GC.@preserve sims ccall((:qdfn_ensemble_create, LIB), Ptr{Cvoid}, (Int64, Ptr{Ptr{Cvoid}}, Int32),
                                length(sims), sims, Int32(max_rhs))
    e == C_NULL && check(-1)
    return e
end

EnsembleRun(e, n_steps; t_end=0.0, graph_attempts=16) =
    check(ccall((:qdfn_ensemble_run, LIB), Cint, (Ptr{Cvoid}, Int64, Float64, Int32), e, n_steps, t_end,
                Int32(graph_attempts)))

EnsembleDestroy(e) = ccall((:qdfn_ensemble_destroy, LIB), Cvoid, (Ptr{Cvoid},), e)

# Gradient of sum_o (Disp[obs_o] - target_o)^2 + w_time (t_first - t_target)^2 in a, b, Dc
# for every step the run (recorded every step) has taken; obs are 1-based element indices.
struct AdjointInfo
    steps::Int64; output_mismatches::Int64; select_mismatches::Int64; onset_step::Int64; retried_steps::Int64
    t_first::Float64; ms_backward::Float64
    lambda_growth::Float64; growth_step::Int64
end

function SimAdjoint(sim, V_i, Theta_i, obs, target; w_time=0.0, t_target=0.0, v_onset=1e-3)
    n = length(V_i)
    vi = Vector{Float64}(vec(V_i)); th = Vector{Float64}(vec(Theta_i))
    o = Vector{Int64}(vec(obs) .- 1); tg = Vector{Float64}(vec(target))
    loss = Ref{Float64}(0.0); ga = zeros(n); gb = zeros(n); gd = zeros(n)
    info = Ref{AdjointInfo}()
    GC.@preserve vi th o tg ga gb gd check(ccall((:qdfn_sim_adjoint, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Int64}, Ptr{Float64}, Float64, Float64, Float64,
         Ref{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{AdjointInfo}),
        sim, vi, th, length(o), o, tg, w_time, t_target, v_onset, loss, ga, gb, gd, info))
    return loss[], ga, gb, gd, info[]
end

# ConvergenceCrit of the local solves (1e-9 in main()); before the run starts.
SimSetConvergence(sim, conv_crit) =
    check(ccall((:qdfn_sim_set_convergence, LIB), Cint, (Ptr{Cvoid}, Float64), sim, conv_crit))

# Forward of n_steps from the run's start keeping a checkpoint every `every` steps
# (p.trace_capacity >= n_steps); SimAdjoint then recomputes each segment from its checkpoint.
SimRunCheckpointed(sim, n_steps, every; graph_attempts=32) =
    check(ccall((:qdfn_sim_run_checkpointed, LIB), Cint, (Ptr{Cvoid}, Int64, Int64, Int32), sim, n_steps, every,
                Int32(graph_attempts)))

# Slip / event-time misfit of a dense run of K steps as a function of the per-element
# a, b, Dc (the other inputs as main() takes them):
#   m = CycleMisfit(op, p, FaultMass, Theta_i, V_i, Friction_i, NormalStress, V_Const, K, obs, target)
#   L = m(a, b, Dc)
# With ChainRulesCore available, rrule(m, a, b, Dc) runs the forward with a checkpoint every
# m.every steps and its pullback returns the adjoint gradient (qdfn_sim_adjoint), so Zygote or
# any ChainRules-aware AD takes gradients through the GPU loop.
struct CycleMisfit
    op::Ptr{Cvoid}; p::SimParams
    FaultMass::Vector{Float64}; Theta_i::Vector{Float64}; V_i::Vector{Float64}
    Friction_i::Vector{Float64}; NormalStress::Vector{Float64}; V_Const::Vector{Float64}
    K::Int64; obs::Vector{Int64}; target::Vector{Float64}
    w_time::Float64; t_target::Float64; v_onset::Float64; every::Int64
end
CycleMisfit(op, p, FaultMass, Theta_i, V_i, Friction_i, NormalStress, V_Const, K, obs, target;
            w_time=0.0, t_target=0.0, v_onset=1e-3, every=50) =
    CycleMisfit(op, p, map(x -> Vector{Float64}(vec(x)), (FaultMass, Theta_i, V_i, Friction_i, NormalStress,
                                                            V_Const))..., Int64(K), Vector{Int64}(vec(obs)),
                Vector{Float64}(vec(target)), Float64(w_time), Float64(t_target), Float64(v_onset), Int64(every))

function _misfit_run(m::CycleMisfit, a, b, Dc)
    p = SimParams(m.p.n, m.p.n_loading, m.p.mu, m.p.rho, m.p.switch_v, m.p.dt_cut, m.p.rupture_dt, m.p.maximum_dt,
                  m.p.min_normal_stress, m.p.alpha_evo, m.p.evolution_dr, m.p.planar, m.p.ext_at_told, Int32(0),
                  0, 0, m.K)
    sim = SimCreateDense(m.op, p, m.FaultMass, a, b, Dc, m.Theta_i, m.V_i, m.Friction_i, m.NormalStress, m.V_Const)
    try
        SimRunCheckpointed(sim, m.K, m.every)
    catch
        SimDestroy(sim); rethrow()
    end
    return sim
end

# the loss from the run's final Disp and its per-step trace (first onset of max V >= v_onset,
# interpolated in log10 V between step end times), as qdfn_sim_adjoint defines it
function _misfit_value(m::CycleMisfit, sim)
    n = m.p.n; K = m.K
    D = zeros(n)
    GC.@preserve D check(ccall((:qdfn_sim_get_state, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
        sim, C_NULL, D, C_NULL, C_NULL, C_NULL, C_NULL))
    L = sum(abs2, D[m.obs] .- m.target)
    if m.w_time > 0
        told = zeros(K); dt = zeros(K); vmax = zeros(K)
        GC.@preserve told dt vmax check(ccall((:qdfn_sim_get_trace, LIB), Cint,
            (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), sim, K, told, dt, vmax))
        k = findfirst(>=(m.v_onset), vmax)
        if k !== nothing
            te = told .+ dt
            tf = if k == 1
                te[1]
            else
                l0, l1, lt = log10(vmax[k-1]), log10(vmax[k]), log10(m.v_onset)
                f = l1 != l0 ? (lt - l0) / (l1 - l0) : 1.0
                te[k-1] + f * (te[k] - te[k-1])
            end
            L += m.w_time * (tf - m.t_target)^2
        end
    end
    return L
end

function (m::CycleMisfit)(a, b, Dc)
    sim = _misfit_run(m, a, b, Dc)
    try
        return _misfit_value(m, sim)
    finally
        SimDestroy(sim)
    end
end

if Base.find_package("ChainRulesCore") !== nothing
    @eval import ChainRulesCore
    @eval function ChainRulesCore.rrule(m::CycleMisfit, a, b, Dc)
        sim = _misfit_run(m, a, b, Dc)
        L = try
            _misfit_value(m, sim)
        catch
            SimDestroy(sim); rethrow()
        end
        function cycle_misfit_pullback(dL)
            try
                _, ga, gb, gd, _ = SimAdjoint(sim, m.V_i, m.Theta_i, m.obs, m.target; w_time=m.w_time,
                                             t_target=m.t_target, v_onset=m.v_onset)
                d = ChainRulesCore.unthunk(dL)
                return (ChainRulesCore.NoTangent(), d .* ga, d .* gb, d .* gd)
            finally
                SimDestroy(sim)
            end
        end
        return L, cycle_misfit_pullback
    end
end

end # module

/* qdfn — MI355X (gfx950) earthquake-cycle hot path of Quake-DFN behind a C ABI.
 *
 * libqdfn_hip.so exports the functions below.  Every pointer argument is a
 * HOST pointer read or written during the call only, unless stated otherwise;
 * matrices are column-major (Julia layout) with an explicit leading
 * dimension.  No function throws or aborts across the ABI: each returns
 * QDFN_OK (0) or a negative QDFN_ERR_* code and leaves a thread-local
 * message for qdfn_last_error().  Handles are not thread-safe; each owns one
 * HIP stream and calls on it are synchronous.
 *
 * Reference = limkjae/Quake-DFN (Julia).  Each declaration cites the
 * reference interface it replaces.
 */
#ifndef QDFN_H
#define QDFN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QDFN_OK 0
#define QDFN_ERR_ARG (-1)
#define QDFN_ERR_HIP (-2)
#define QDFN_ERR_NOMEM (-3)
#define QDFN_ERR_STATE (-4)
/* more than 100 retries of one step: the reference blocks in readline()
 * (scripts/Functions_RSFDFN3DMain_D.jl:255-258, _H.jl same loop) */
#define QDFN_ERR_RETRIES (-10)
/* empty candidate set for min |V/Accel|: Julia's minimum() of an empty
 * collection throws (Functions_RSFDFN3DMain_D.jl:165-168) */
#define QDFN_ERR_EMPTY_DTMIN (-11)

const char* qdfn_last_error(void);

/* Device name, compute-unit count and HBM bytes of the current HIP device. */
int qdfn_device_info(char* name, int64_t name_len, int32_t* cu_count, int64_t* hbm_bytes);

/* ---- Okada rectangles -------------------------------------------------- */

/* Replaces Okada_DC3D_Vector(ALPHA,X,Y,Z,DEPTH,DIP,AL1,AL2,AW1,AW2,DISL1,DISL2,DISL3)
 * scripts/Functions_OKADA3D.jl:464.  U is n x 12 column-major (columns
 * UX,UY,UZ,UXX,UYX,UZX,UXY,UYY,UZY,UXZ,UYZ,UZZ); iret[n] as IRET. */
int qdfn_okada_dc3d(double alpha, int64_t n, const double* X, const double* Y, const double* Z, double depth,
                    double dip, double al1, double al2, double aw1, double aw2, double disl1, double disl2,
                    double disl3, double* U, double* iret);

/* Replaces StiffnessMatrix_ByParts_Calculation_Rec(Input_SegmentSource,
 * Input_SegmentReceiver, ShearModulus, PoissonRatio, ...) and
 * BuildMatrixByParts_Even_Rec, scripts/Functions_BuildInputFile.jl:386,516.
 * Segment tables use the reference's Input_Segment columns (0-based here):
 * 0-2 centre (z = +depth), 3 strike length, 4 dip length, 5 strike deg,
 * 6 dip deg, 7 rake deg (further columns ignored).  KS/KN are nr x ns. */
int qdfn_okada_stiffness_rect(const double* src_seg, int64_t ns, int64_t ld_src, const double* rcv_seg,
                              int64_t nr, int64_t ld_rcv, double mu, double nu, double* KS, double* KN,
                              int64_t ldk);

/* ---- triangular dislocations (Nikkhoo & Walter 2015, half-space) -------- */

/* Replaces TDstressHS(X,Y,Z,P1,P2,P3,Ss,Ds,Ts,mu,lambda),
 * scripts/Functions_TDstressHS.jl:2.  Stress and Strain are n x 6
 * column-major (xx,yy,zz,xy,xz,yz).  P1..P3 are 3-vectors (z <= 0). */
int qdfn_tde_stress_hs(int64_t n, const double* X, const double* Y, const double* Z, const double* P1,
                       const double* P2, const double* P3, double Ss, double Ds, double Ts, double mu,
                       double lambda, double* Stress, double* Strain);

/* Replaces StiffnessMatrix_ByParts_Calculation_Tri(P1,P2,P3,Rake,FaultCenter,
 * ShearModulus,lambda,UnitVector_Normal,UnitVector_Slip) and
 * BuildMatrixByParts_Even_Tri, scripts/Functions_BuildInputFile.jl:551,599.
 * P1..P3, ctr, nrm, slip are n x 3 column-major, already processed by
 * RotVerts_UnitVectors (qdfn_rotverts_unitvectors). */
int qdfn_tde_stiffness_tri(int64_t ns, const double* P1, const double* P2, const double* P3, const double* rake,
                           int64_t nr, const double* ctr, const double* nrm, const double* slip, double mu,
                           double lambda, double* KS, double* KN, int64_t ldk);

/* Replaces RotVerts_UnitVectors(Input_Segment, FaultCount, Rake),
 * scripts/Functions_BuildInputFile.jl:338.  P1..P3 (n x 3) are reordered in
 * place; outputs n x 3.  Returns the number of reversed triangles (>= 0). */
int qdfn_rotverts_unitvectors(int64_t n, double* P1, double* P2, double* P3, const double* rake, double* nrm,
                              double* strike, double* dip, double* slip);

/* ---- dense stiffness operator (device resident) ------------------------- */

#define QDFN_PREC_FP64 0 /* FP64 storage, FP64 accumulation */
#define QDFN_PREC_FP32 1 /* FP32 storage, FP64 accumulation */
/* H operators only: low-rank factors U, V in BF16, dense blocks in FP32, FP64
 * accumulation (about 2^-9 relative per factor entry; the error against the
 * FP64 product is reported per result, not bounded by the ACA tolerance) */
#define QDFN_PREC_BF16_FAR 2
/* H operators only: dense blocks and the leading components of each low-rank
 * block in FP32; its trailing components (the smallest singular values) as
 * FP16 unit vectors with an FP64 scale each, as many as fit in the part of
 * the block's error budget (aca_rtol ||K_S,b||_F) that the ACA residual and
 * the truncation leave unused, by the measured FP32 and FP16 rounding of each
 * component.  FP64 accumulation.  Blocks that were not recompressed (rank
 * below 2, aca_rtol = 0, or imported through qdfn_hmat_create) stay FP32.
 * The device builds (qdfn_hmat_build_*, the sampled ones too, which evaluate
 * such blocks in full) also compress the inadmissible blocks off the diagonal
 * (near-field neighbours, which the reference keeps dense) where their tiered
 * factors take fewer bytes than the dense block, within the same bound;
 * diagonal blocks stay dense. */
#define QDFN_PREC_MIXED 3
/* H operators only: QDFN_PREC_MIXED with a third tier: the smallest trailing
 * components as FP8 (E4M3) vectors scaled to a largest entry in [128, 256)
 * with an FP64 scale each, then FP16, then FP32, as many in each as the
 * block's unused error budget admits by the measured rounding of each
 * component.  FP64 accumulation. */
#define QDFN_PREC_MIXED8 4

typedef struct qdfn_dense qdfn_dense_t;

/* Upload K_S, K_N (n x n, leading dimension ld) and build the operator of
 * scripts/Functions_RSFDFN3DMain_D.jl:54-63: E_ij = K_S,ij/K_S,ii (i != j),
 * E_ii = 0, K_Self = |K_S,ii|, stored at precision prec.  FP64 operators of
 * 6,144 elements or more are stored split-tiled for the coupling kernel (each
 * workgroup's rows x columns contiguous), the others column-major; the
 * environment variable QDFN_DENSE_TILED=0/1 overrides.  Products are the
 * same bit for bit either way. */
qdfn_dense_t* qdfn_dense_create(const double* KS, const double* KN, int64_t n, int64_t ld, int32_t prec);
/* Assemble the stiffness of n rectangular segments on the device (no host
 * round trip) and build the operator as above. */
qdfn_dense_t* qdfn_dense_create_rect(const double* seg, int64_t n, int64_t ld_seg, double mu, double nu,
                                     int32_t prec);
/* Assemble the stiffness of n triangles (inputs as qdfn_tde_stiffness_tri). */
qdfn_dense_t* qdfn_dense_create_tri(int64_t n, const double* P1, const double* P2, const double* P3,
                                    const double* rake, const double* ctr, const double* nrm, const double* slip,
                                    double mu, double lambda, int32_t prec);
void qdfn_dense_destroy(qdfn_dense_t* op);
/* K_Self (n) as the reference computes it (|K_S,ii|); diag = signed K_S,ii. */
int qdfn_dense_kself(qdfn_dense_t* op, double* kself, double* diag);
/* 1 if every K_N entry is zero (the reference's NormalStiffnessZero). */
int qdfn_dense_normal_zero(qdfn_dense_t* op);
/* Column splits of the coupling product (no reference counterpart: the
 * reference's product is one BLAS call).  By default an operator takes the
 * split count that best fills the GPU when it runs alone; runs advanced
 * together (qdfn_ensemble_*) fill it with their number instead, and fewer
 * splits mean fewer partial sums written and read per product.  splits = 0
 * restores the default.  Sums keep a fixed order for any count, but a run's
 * arithmetic depends on it: change it before runs on the operator start.
 * Returns the split count taken (>= 1), or an error for an FP64 operator of
 * 6,144 or more elements, whose split-tiled layout is fixed at its build. */
int qdfn_dense_set_splits(qdfn_dense_t* op, int32_t splits);

/* Coupling product (Functions_RSFDFN3DMain_D.jl:199-202).  loading_form=1:
 * yS = E*(0-x) (Elastic_Load_Disp), yN = K_N*x.  loading_form=0: yS = K_S*x,
 * yN = K_N*x. */
int qdfn_matvec(qdfn_dense_t* op, const double* x, double* yS, double* yN, int32_t loading_form);

/* The coupling product of qdfn_matvec on DEVICE vectors (d_x in, d_yS and
 * d_yN out, n doubles each), enqueued on `stream` (a hipStream_t; NULL is the
 * null stream) without waiting for it, and allocating nothing: no hipMalloc,
 * hipFree or host copy, so it can be called under HIP graph capture or in a
 * loop that must not allocate.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory.  The split partial sums live
 * in one static device array per GPU of 8 MiB (2 x splits x ld doubles must
 * fit: QDFN_ERR_ARG otherwise), shared by every call on that GPU — calls that
 * overlap on one GPU (several streams or threads) race on it.  The operator
 * must live on the current device.  Results equal qdfn_matvec's bit for bit. */
int qdfn_matvec_device_noalloc(qdfn_dense_t* op, const double* d_x, double* d_yS, double* d_yN,
                               int32_t loading_form, void* stream);

/* Average duration (ms, HIP events on the operator's stream) of `reps`
 * launches of the fused coupling kernel. */
int qdfn_dense_time_coupling(qdfn_dense_t* op, int32_t reps, double* avg_ms);

/* ---- hierarchical (H-matrix) stiffness operator ----------------------- */

/* Cluster-tree and admissibility parameters (RUN1_HMatrix_Structure.jl:28-33). */
typedef struct {
    int64_t min_elements_to_cut; /* MinimumElementsToCut: a cluster is split only when
                                    (end - begin) >= this (leaf size) */
    int64_t max_levels;          /* TotalHierarchyLevel */
    double dist_diam_ratio;      /* DistDiamRatioCrit: admissible iff |c_i - c_j| / max(d_i, d_j) > this */
    double arrange_point[3];     /* ArrangePoint: children ordered by distance of their centre to it */
    int64_t n_loading;           /* loading elements: the last n_loading rows, kept out of the tree */
    int64_t kmeans_max_iter;     /* Lloyd iterations of the 2-means split (200 in the reference) */
} qdfn_hmat_params;

/* Replaces GroupAndSort_AllLevel + BuildHierarchy (scripts/Functions_Hmatrix.jl:
 * 1-212).  Host-only (no device work).  ctr: element centres (n x 3,
 * column-major); dpts: the points whose bounding box gives a cluster's
 * diameter (Input_Segment columns 1:3: the centre for rectangles, the first
 * vertex for triangles).  Outputs:
 *   perm[n]         element at sorted position i is input element perm[i] (0-based)
 *   node_info       cap_nodes x 4 row-major: begin, end (1-based, inclusive, sorted order),
 *                   level, parent node (-1 for level 1)
 *   node_geo        cap_nodes x 4 row-major: centre xyz, diameter (Block_Ctr_Diam)
 *   range_SR        cap_blocks x 4 row-major: src_i src_f rcv_i rcv_f (1-based,
 *                   ElementRange_SR), admissible[cap_blocks] (Admissible)
 * The counts are written to *n_nodes / *n_blocks; when a capacity is too small
 * nothing else is written and QDFN_ERR_ARG is returned with the counts set. */
int qdfn_hmat_structure(const qdfn_hmat_params* p, int64_t n, const double* ctr, const double* dpts, int64_t* perm,
                        int64_t cap_nodes, int64_t* node_info, double* node_geo, int64_t* n_nodes,
                        int64_t cap_blocks, int64_t* range_SR, int64_t* admissible, int64_t* n_blocks);

typedef struct qdfn_hmat qdfn_hmat_t;

#define QDFN_HBLK_DENSE 0
#define QDFN_HBLK_LOWRANK 1

/* Replaces HmatBuild_R (scripts/Functions_BuildInputFile.jl:725): per block the
 * shear/normal stiffness of segment rows (already in sorted order, n x >=8,
 * column-major, leading dimension ld); admissible blocks are compressed to
 * U V (m x r, r x k) by adaptive cross approximation with full pivoting,
 * stopping when ||residual||_F <= aca_rtol ||K_S,b||_F or max |residual| <=
 * aca_atol, K_S,b the block's shear entries, for the shear and the normal part
 * alike (aca_rtol = 0, aca_atol = Tolerance mirrors the reference's
 * pqrfact(block, atol = Tolerance)).  The whole matrix is
 * assembled in device memory first, so n is bounded by HBM (about 130k). */
qdfn_hmat_t* qdfn_hmat_build_rect(const double* seg, int64_t n, int64_t ld, double mu, double nu, int64_t nblk,
                                  const int64_t* range_SR, const int64_t* admissible, double aca_rtol,
                                  double aca_atol, int32_t prec);
/* Replaces HmatBuild_T (:643); triangle inputs as qdfn_tde_stiffness_tri, sorted order. */
qdfn_hmat_t* qdfn_hmat_build_tri(int64_t n, const double* P1, const double* P2, const double* P3,
                                 const double* rake, const double* ctr, const double* nrm, const double* slip,
                                 double mu, double lambda, int64_t nblk, const int64_t* range_SR,
                                 const int64_t* admissible, double aca_rtol, double aca_atol, int32_t prec);
/* Sampled construction: the same operator as qdfn_hmat_build_rect / _tri
 * (same arguments and block structure) without assembling any admissible
 * block.  Replaces the full-block assembly + pqrfact of HmatBuild_T / HmatBuild_R
 * (scripts/Functions_BuildInputFile.jl:674-716, :754-797): each admissible
 * part is approximated by partially pivoted cross approximation, which
 * evaluates only the rows and columns it pivots on, and stops when the
 * estimate |u_r| |v_r| <= (aca_rtol / 2) ||S||_F holds on two consecutive
 * steps (||S||_F the block's shear-norm estimate; with aca_atol > 0 also when
 * max|u_r| max|v_r| <= aca_atol), then is recompressed to the other half of
 * aca_rtol.  Parts that reach the rank cap are evaluated in full and take the
 * full build's path; inadmissible blocks are evaluated in full.  Memory and
 * entry evaluations grow like the H-matrix, so n is not bounded by n^2
 * device memory. */
qdfn_hmat_t* qdfn_hmat_build_rect_sampled(const double* seg, int64_t n, int64_t ld, double mu, double nu,
                                          int64_t nblk, const int64_t* range_SR, const int64_t* admissible,
                                          double aca_rtol, double aca_atol, int32_t prec);
qdfn_hmat_t* qdfn_hmat_build_tri_sampled(int64_t n, const double* P1, const double* P2, const double* P3,
                                         const double* rake, const double* ctr, const double* nrm,
                                         const double* slip, double mu, double lambda, int64_t nblk,
                                         const int64_t* range_SR, const int64_t* admissible, double aca_rtol,
                                         double aca_atol, int32_t prec);
typedef struct {
    int64_t entries_evaluated; /* stiffness entries computed by the build (n*n for the full builds) */
    int64_t steps;             /* line-evaluation / approximation launch pairs (sampled builds) */
    int64_t parts_sampled;     /* admissible parts finished by the sampled approximation */
    int64_t blocks_fallback;   /* parts evaluated in full after reaching the rank cap */
    double ms_build;           /* wall time of the build (ms) */
    double ms_phase[5];        /* of which: [0] entry evaluation of dense and fully evaluated blocks (the
                                  full builds: the assembly), [1] sampled approximation, [2] full-pivot
                                  approximation, [3] host recompression, [4] product layout */
} qdfn_hmat_build_stats;
/* Statistics of the build that made h (zeros for qdfn_hmat_create). */
int qdfn_hmat_build_info(qdfn_hmat_t* h, qdfn_hmat_build_stats* out);
/* Operator from given blocks (ShearStiffness_H / NormalStiffness_H as written
 * by RUN2_BUILD_INPUT.jl:362): per block kind_S[b], rank_S[b], S_a[b] (dense
 * m x k, or U m x r), S_b[b] (NULL, or V r x k), all column-major; the same for
 * N.  A LowRankApprox PartialQR (A[:, p] ~ Q R) is passed as U = Q and
 * V = R with its columns put back in place (V[:, p] = R). */
qdfn_hmat_t* qdfn_hmat_create(int64_t n, int64_t nblk, const int64_t* range_SR, const int64_t* admissible,
                              const int32_t* kind_S, const int64_t* rank_S, const double* const* S_a,
                              const double* const* S_b, const int32_t* kind_N, const int64_t* rank_N,
                              const double* const* N_a, const double* const* N_b, int32_t prec);
void qdfn_hmat_destroy(qdfn_hmat_t* h);
/* n, block count, HBM bytes one product moves, stored entries (both matrices),
 * and phase_bytes[2]: algorithmic bytes of the phase-1 (t = V x) and phase-2
 * (y = panels [x; t]) kernels. */
int qdfn_hmat_info(qdfn_hmat_t* h, int64_t* n, int64_t* nblk, int64_t* device_bytes, int64_t* stored_entries,
                   int64_t* phase_bytes);
/* Algorithmic bytes of one product by part: out[0] phase-1 stored entries,
 * out[1] phase-1 x slices and t rows, out[2] phase-2 stored entries, out[3]
 * phase-2 index list, out[4] phase-2 gathered right-hand side, out[5] outputs. */
int qdfn_hmat_bytes(qdfn_hmat_t* h, int64_t* out);
/* Per block: kind and rank of the shear and normal parts (Ranks_Shear/Ranks_Normal). */
int qdfn_hmat_get_ranks(qdfn_hmat_t* h, int32_t* kind_S, int64_t* rank_S, int32_t* kind_N, int64_t* rank_N);
/* One block's stored factors (which = 0 shear, 1 normal): dense m x k into a,
 * or U (m x r) into a and V (r x k) into b; FP64 values as stored. */
int qdfn_hmat_get_block(qdfn_hmat_t* h, int64_t b, int32_t which, double* a, double* bv);
/* Block b as the product applies it: the stored entries after rounding to the
 * operator's storage (FP32 / BF16 / FP16 unit vectors times their FP64 scale
 * folded into U; the shear diagonal of diagonal blocks stays FP64, as the
 * product applies K_S,ii), in FP64, same shapes as qdfn_hmat_get_block.  No
 * reference counterpart (the reference stores FP64 only). */
int qdfn_hmat_get_block_stored(qdfn_hmat_t* h, int64_t b, int32_t which, double* a, double* bv);
/* Per block, the first low-rank component held in the 16-bit pool (r_hi;
 * rank: none, 0: all). */
int qdfn_hmat_get_split(qdfn_hmat_t* h, int64_t* rhi_S, int64_t* rhi_N);
/* Per block, the first low-rank component held in the 8-bit pool
 * (QDFN_PREC_MIXED8; rank: none).  No reference counterpart. */
int qdfn_hmat_get_split8(qdfn_hmat_t* h, int64_t* r8_S, int64_t* r8_N);
/* y[i] = x[i] rounded to FP8 E4M3 (OCP: bias 7, no infinities, largest
 * finite 448; round to nearest even, saturating), the storage of the 8-bit
 * pool of QDFN_PREC_MIXED8.  Host only.  No reference counterpart. */
int qdfn_fp8_round(const double* x, double* y, int64_t count);
/* K_Self = -K_S,ii from the diagonal blocks (StiffnessTransitionToLoading,
 * Functions_Hmatrix.jl:459); diag = K_S,ii (0 where no diagonal block). */
int qdfn_hmat_kself(qdfn_hmat_t* h, double* kself, double* diag);
int qdfn_hmat_normal_zero(qdfn_hmat_t* h);
/* ReduceTooStrongInteraction_Hmat (scripts/Functions_Solvers.jl:546-579,
 * applied at QUAKEDFN_RunARG.jl:126-131 when StrongInteractionCriteriaMultiple
 * > 0): in every inadmissible block, each entry of the loading-form shear
 * matrix (diagonal of diagonal blocks zeroed) whose magnitude exceeds
 * multiple * |K_Self| of its receiver has its shear and normal entries scaled
 * by multiple * |K_Self| / |entry|; the operator's product layout is rebuilt.
 * An inadmissible block stored low-rank (QDFN_PREC_MIXED / _MIXED8) is first
 * expanded to its FP64 product U V.  n_pairs (optional): entries changed.
 * Before runs are created on h. */
int qdfn_hmat_reduce_strong_interaction(qdfn_hmat_t* h, double multiple, int64_t* n_pairs);
/* HmatSolver_Pararllel (Functions_Hmatrix.jl:631) for both matrices.
 * loading_form=1: yS = (H_load * (0 - x)) ./ (-K_Self) (main_H :241-243),
 * yN = H_N * x; loading_form=0: yS = H_S * x, yN = H_N * x. */
int qdfn_hmat_matvec(qdfn_hmat_t* h, const double* x, double* yS, double* yN, int32_t loading_form);
/* Average duration (ms, HIP events) of `reps` H products (both phases). */
int qdfn_hmat_time_matvec(qdfn_hmat_t* h, int32_t reps, double* avg_ms);
/* Average durations (ms, HIP events on the operator's stream) of phase 1
 * (with its finish), the phase-2 kernel, and the whole product. */
int qdfn_hmat_time_phases(qdfn_hmat_t* h, int32_t reps, double* ms_phase1, double* ms_phase2, double* ms_total);

/* ---- rate-and-state local solve ---------------------------------------- */

typedef struct {
    double conv_crit; /* ConvergenceCrit, 1e-9 in the reference */
    double v0;        /* V0, 1e-9 */
    double mu, rho, switch_v, alpha_evo;
    int32_t evolution_dr; /* 1: aging (Dieterich) law, else Ruina */
    int32_t pad_;
} qdfn_step_params;

/* Replaces SolveOneTimeStep(...), scripts/Functions_Solvers.jl:334-380, with
 * Solver_HighV_D (:26) / Solver_LowV_D (:185) / EvolutionLaw (:2).  Inputs
 * (length n): DispOld, FrictionOld, ThetaOld, VOld, Friction0, a, b, Dc,
 * K_Self, Omega, Mass, EffNormalStress_i, Total_Loading_Disp, SolverSwitch,
 * LoadingRate, D_EffStress_Shear, EffNormalStress_Old, InitialShearStress and
 * the scalar Dt.  Outputs (length n): V, Friction, Disp, Theta,
 * EffNormalStress, Dt_All, InstabilityThistime, FLAG_GoodToGo; iters
 * (optional) = secant iterations used.  The last n_loading rows are loading
 * faults.  Returns the number of readline() stalls of Solver_HighV_D (>= 0). */
int64_t qdfn_solve_one_timestep(const qdfn_step_params* p, int64_t n, int64_t n_loading, const double* DispOld,
                                const double* FrictionOld, const double* ThetaOld, const double* VOld,
                                const double* Friction0, const double* a, const double* b, const double* Dc,
                                const double* K_Self, double Dt, const double* Omega, const double* Mass,
                                const double* EffNormalStress_i, const double* Total_Loading_Disp,
                                const int32_t* SolverSwitch, const double* LoadingRate,
                                const double* D_EffStress_Shear, const double* EffNormalStress_Old,
                                const double* InitialShearStress, double* V, double* Friction, double* Disp,
                                double* Theta, double* EffNormalStress, double* Dt_All,
                                int32_t* InstabilityThistime, int32_t* FLAG_GoodToGo, int32_t* iters);

/* ---- device-resident earthquake-cycle run (main / main_H) --------------- */

typedef struct {
    int64_t n, n_loading;
    double mu, rho;
    double switch_v, dt_cut, rupture_dt, maximum_dt, min_normal_stress, alpha_evo;
    int32_t evolution_dr;
    int32_t planar;       /* NormalStiffnessZero: skip the K_N product */
    int32_t ext_at_told;  /* 0: external stress at TOld+Dt (main), 1: at TOld (main_H) */
    int32_t pad_;
    int64_t record_step;  /* RecordStep; histories every record_step accepted steps */
    int64_t record_capacity; /* rows of history kept on the device (0: none) */
    int64_t trace_capacity;  /* per-step (TOld, Dt, max V) entries kept (0: none) */
} qdfn_sim_params;

typedef struct qdfn_sim qdfn_sim_t;

typedef struct {
    int64_t steps, rhs_evals, retry_passes, stalls, n_recorded, n_traced;
    int32_t status; /* QDFN_OK or QDFN_ERR_RETRIES / QDFN_ERR_EMPTY_DTMIN */
    int32_t pad_;
    double t, dt;   /* simulated time after the last accepted step, last Dt */
    double coupling_ms_avg; /* mean coupling-kernel duration (timed runs) */
    int64_t coupling_timed; /* launches averaged */
} qdfn_sim_stats;

/* Create a run over a dense operator (which must outlive the run).  Element
 * arrays (length n): Mass, a, b, Dc, Theta_i, V_i, Friction_i,
 * InitialNormalStress, LoadingRate (V_Const).  Mirrors the arguments of
 * main(), scripts/Functions_RSFDFN3DMain_D.jl:2. */
qdfn_sim_t* qdfn_sim_create_dense(qdfn_dense_t* op, const qdfn_sim_params* p, const double* mass, const double* a,
                                  const double* b, const double* dc, const double* theta_i, const double* v_i,
                                  const double* friction_i, const double* sigma0, const double* loading_rate);
/* main_H (scripts/Functions_RSFDFN3DMain_H.jl:2): the same loop with the
 * H-matrix coupling; K_Self = -K_S,ii from the diagonal blocks and the
 * external tables read at TOld (set p->ext_at_told = 1 for the reference's
 * behaviour). */
qdfn_sim_t* qdfn_sim_create_hmat(qdfn_hmat_t* h, const qdfn_sim_params* p, const double* mass, const double* a,
                                 const double* b, const double* dc, const double* theta_i, const double* v_i,
                                 const double* friction_i, const double* sigma0, const double* loading_rate);
/* External stress tables (Input_ExternalStressChange.jld2; InterpolateFromStressChange,
 * Functions_Solvers.jl:465): times[nt], non-decreasing (QDFN_ERR_ARG otherwise: the
 * device finds the reference's sample index by binary search); ext_n/ext_s are
 * nt x n column-major. */
int qdfn_sim_set_external(qdfn_sim_t* sim, int64_t nt, const double* times, const double* ext_n,
                          const double* ext_s);
/* Advance n_steps accepted steps (the reference's TotalStep loop body).
 * mode 0: attempts replayed from a captured HIP graph (graph_attempts per
 * replay); mode 1: direct launches with HIP events around every coupling
 * kernel (fills coupling_ms_avg; for an H operator the phase-2 kernel).
 * Returns QDFN_OK or the run's status. */
int qdfn_sim_run(qdfn_sim_t* sim, int64_t n_steps, int32_t mode, int32_t graph_attempts);
/* Record callback of qdfn_sim_run_until: records first .. first + count - 1
 * of the run (counted from the first record handed over), as host arrays
 * laid out like the reference's History_* (time[count], dt[count]; V, Disp,
 * Theta, Sigma count x n column-major), valid during the call.  Return 0 to
 * continue, nonzero to stop the run after this batch. */
typedef int (*qdfn_record_cb)(void* user, int64_t first, int64_t count, const double* time, const double* dt,
                              const double* V, const double* Disp, const double* Theta, const double* Sigma);
/* main()'s TotalStep loop with its periodic save (Functions_RSFDFN3DMain_D.jl:
 * 146-297, main_H the same): advance until TOld reaches t_end (> 0; the step
 * that crosses it is completed) or max_steps accepted steps, replaying
 * captured attempts (graph_attempts per replay).  A record is taken every
 * record_step steps as in the reference; with cb, the run's history buffer
 * (record_capacity records, at least graph_attempts / record_step + 1) is a
 * staging area handed to cb whenever the next replay could overflow it and at
 * the end, so runs of any length keep every record (the reference's
 * save(SaveResultFileName, ...) every SaveStep steps).  cb = NULL keeps the
 * records on the device up to record_capacity.  Returns QDFN_OK or the run's
 * status. */
int qdfn_sim_run_until(qdfn_sim_t* sim, double t_end, int64_t max_steps, int32_t graph_attempts, qdfn_record_cb cb,
                       void* user);
/* ConvergenceCrit of the local solves (Solver_LowV_D / Solver_HighV_D stop
 * when |V_trial / V - 1| falls below it; 1e-9 in the reference,
 * Functions_RSFDFN3DMain_D.jl, the default here).  Before the run starts.
 * A tighter value makes the run a smoother function of its parameters (the
 * finite-difference checks of the adjoint use it). */
int qdfn_sim_set_convergence(qdfn_sim_t* sim, double conv_crit);
/* Advance a dense run from its start by n_steps accepted steps, keeping the
 * run state every `every` steps (checkpoints; 11 n doubles each) and the
 * per-step trace (trace_capacity >= n_steps), for qdfn_sim_adjoint's
 * checkpointed recomputation: each segment of `every` steps is replayed from
 * its checkpoint (bitwise the forward's steps) just before its backward
 * sweep, so the tape takes `every` records instead of n_steps.  No reference
 * counterpart (the reference has no gradients). */
int qdfn_sim_run_checkpointed(qdfn_sim_t* sim, int64_t n_steps, int64_t every, int32_t graph_attempts);
int qdfn_sim_get_stats(qdfn_sim_t* sim, qdfn_sim_stats* st);
/* Current state (after the last accepted step): V, Disp, Theta, Friction,
 * EffNormalStress, Accel (each length n, any may be NULL). */
int qdfn_sim_get_state(qdfn_sim_t* sim, double* V, double* Disp, double* Theta, double* Friction, double* Sigma,
                       double* Accel);
/* Histories recorded so far: time/dt [count], V/Disp/Theta/Sigma count x n
 * column-major (History_* of the reference, rows = records). */
int qdfn_sim_get_history(qdfn_sim_t* sim, int64_t count, double* time, double* dt, double* V, double* Disp,
                         double* Theta, double* Sigma);
/* Solver used by each element in each recorded step's accepted attempt
 * (count x n column-major, 1 = Solver_HighV_D, 0 = Solver_LowV_D; the
 * per-element SolverSwitch of main(), Functions_RSFDFN3DMain_D.jl:152-158, 242-249). */
int qdfn_sim_get_solver_history(qdfn_sim_t* sim, int64_t count, int8_t* solver);
/* Per-step trace: TOld, Dt and max V over non-loading elements. */
int qdfn_sim_get_trace(qdfn_sim_t* sim, int64_t count, double* told, double* dt, double* vmax);
void qdfn_sim_destroy(qdfn_sim_t* sim);

/* ---- adjoint of the step loop ---------------------------------------------
 * Gradient with respect to the per-element a, b and Dc of
 *   L = sum_o (Disp_K[obs_o] - d_target_o)^2 + w_time (t_first - t_target)^2
 * for the K accepted steps a dense run has taken since its creation (it must
 * record every step: record_step = 1, record_capacity >= K, or have been
 * advanced by qdfn_sim_run_checkpointed, whose segments are recomputed from
 * their checkpoints; aging law with alpha = 0).  t_first: the first step whose max V over the
 * non-loading elements reaches v_onset, interpolated in log10 V between the
 * end times of that step and the one before (the term is dropped if
 * w_time = 0 or no onset occurs).  v_i / theta_i: the run's initial V and
 * Theta (as given to qdfn_sim_create_dense).  A discrete adjoint of the
 * forward's own arithmetic (Functions_RSFDFN3DMain_D.jl:146-319,
 * Functions_Solvers.jl:26-327): the implicit-function derivative of each
 * converged local solve, transposed coupling products, the Dt controller's
 * argmin and growth rules, with the forward's discrete choices held fixed.
 * It has no counterpart in the reference (which has no gradients).
 * Validity: through a rupture the discrete step map amplifies perturbations
 * (RuptureDt oscillator steps), and so does its adjoint: lambda_growth is the
 * largest max|adjoint of Disp| over the backward sweep divided by its seed
 * value, growth_step the forward step at which it first exceeded 1e8 (-1:
 * never).  Interseismic windows and windows that end shortly after an onset
 * stay below that; beyond it the gradient is that of a chaotic discrete path,
 * which no difference quotient can check (verified: see the tests). */
typedef struct {
    int64_t steps, output_mismatches, select_mismatches, onset_step, retried_steps;
    double t_first, ms_backward;
    double lambda_growth;
    int64_t growth_step;
} qdfn_adjoint_info;
int qdfn_sim_adjoint(qdfn_sim_t* sim, const double* v_i, const double* theta_i, int64_t n_obs, const int64_t* obs,
                     const double* d_target, double w_time, double t_target, double v_onset, double* loss,
                     double* grad_a, double* grad_b, double* grad_dc, qdfn_adjoint_info* info);

/* ---- batched ensemble -----------------------------------------------------
 * Many independent runs of main() (scripts/Functions_RSFDFN3DMain_D.jl:2-324,
 * the loop QUAKEDFN_RunARG.jl:247-262 starts once per input file), advanced
 * together: one attempt of every run per batched launch sequence, attempts
 * replayed from one captured HIP graph.  Runs must be dense runs with FP32
 * operators; runs created on the same operator handle (same geometry) are
 * read as one multi-right-hand-side product.  Each run's trajectory is
 * bitwise the one it has when run alone.  The runs are borrowed (they must
 * outlive the ensemble) and keep their getters (stats, state, history, trace,
 * catalog); do not advance them with qdfn_sim_run while in an ensemble.
 * max_rhs: runs per operator pass (1..4), 0 = pick by timing (autotune). */
typedef struct qdfn_ensemble qdfn_ensemble_t;
qdfn_ensemble_t* qdfn_ensemble_create(int64_t m, qdfn_sim_t* const* sims, int32_t max_rhs);
/* Advance every run by up to n_steps accepted steps, stopping a run early
 * once its simulated time reaches t_end (t_end = 0: no time limit). */
int qdfn_ensemble_run(qdfn_ensemble_t* e, int64_t n_steps, double t_end, int32_t graph_attempts);
/* Operator groups, runs per pass, matrix bytes of a sweep in which every run needs its coupling. */
int qdfn_ensemble_info(qdfn_ensemble_t* e, int64_t* groups, int32_t* max_rhs, int64_t* bytes_per_sweep);
void qdfn_ensemble_destroy(qdfn_ensemble_t* e);

/* ---- event catalog (scripts/Functions_Plot.jl:430-579) -------------------
 * get_event_fragments (:430-510): per element, a fragment opens at every
 * record where the slip rate (D[t]-D[t-1])/(T[t]-T[t-1]) rises above
 * DispRateCrits and sums dD * FaultArea * ShearModulus while it stays above;
 * fragments are returned sorted by time (stable).  merge_fragments (:513-579)
 * joins fragment i to the event of the nearest earlier fragment within
 * MergeTimeCriteria and MergeDistanceCriteria and gives the events' time and
 * location (those of their first fragment) and Mw = (log10 M0 - 9.05) / 1.5.
 * Locations are count x 3 column-major (FaultCenter rows); fault_center is
 * n x 3 column-major; element / fragment_event indices are 0-based. */
typedef struct qdfn_catalog qdfn_catalog_t;
/* Fragments from a run's device-resident History_Time / History_Disp. */
qdfn_catalog_t* qdfn_catalog_from_sim(qdfn_sim_t* sim, const double* fault_center, const double* fault_area,
                                      double shear_modulus, double disp_rate_crit);
/* Fragments from host histories: time[n_rec], disp n_rec x n column-major (leading dimension ld). */
qdfn_catalog_t* qdfn_catalog_from_history(int64_t n_rec, int64_t n, const double* time, const double* disp,
                                          int64_t ld, const double* fault_center, const double* fault_area,
                                          double shear_modulus, double disp_rate_crit);
/* A catalog holding given (time-sorted) fragments, for qdfn_catalog_merge alone. */
qdfn_catalog_t* qdfn_catalog_from_fragments(int64_t count, const double* time, const double* location,
                                            const double* moment);
/* Fragment count, then (any pointer may be NULL) the arrays, sized by count. */
int qdfn_catalog_fragments(qdfn_catalog_t* c, int64_t* count, double* time, double* location, double* moment,
                           int32_t* element);
/* merge_fragments; sets the event count. */
int qdfn_catalog_merge(qdfn_catalog_t* c, double merge_time, double merge_distance, int64_t* count);
/* Merged events (count), and each fragment's event (fragment count entries). */
int qdfn_catalog_events(qdfn_catalog_t* c, int64_t* count, double* time, double* magnitude, double* location,
                        double* moment, int32_t* fragment_event);
void qdfn_catalog_destroy(qdfn_catalog_t* c);

#ifdef __cplusplus
}
#endif
#endif

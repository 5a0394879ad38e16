// Stand-alone SolveOneTimeStep (scripts/Functions_Solvers.jl:334-380) over
// host arrays: one lane per element, the same device solver the run loop uses.
#include <vector>

#include "friction_device.h"
#include "qdfn_internal.h"

namespace qdfn {

struct SolveArgs {
    int64_t n, n_loading;
    double dt, conv, v0, mu, rho, switch_v, alpha_evo;
    int32_t dr;
    // inputs
    const double *d_old, *f_old, *th_old, *v_old, *f0, *a, *b, *dc, *k, *omega, *mass, *sig_i, *tl, *lrate,
        *dtau, *sig_old, *tau0;
    const int32_t* sw;
    // outputs
    double *v, *f, *d, *th, *sig, *dt_all;
    int32_t *instab, *ok, *iters;
    unsigned long long* stalls;
};

__global__ __launch_bounds__(256) void k_solve_standalone(SolveArgs A) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    if (i < A.n - A.n_loading) {
        ElemIn e;
        e.conv = A.conv; e.d_old = A.d_old[i]; e.f_old = A.f_old[i]; e.th_old = A.th_old[i]; e.v_old = A.v_old[i];
        e.f0 = A.f0[i]; e.a = A.a[i]; e.b = A.b[i]; e.dc = A.dc[i]; e.v0 = A.v0; e.k = A.k[i]; e.omega = A.omega[i];
        e.mass = A.mass[i]; e.mu = A.mu; e.sig = A.sig_i[i]; e.tl = A.tl[i]; e.rho = A.rho; e.dtau = A.dtau[i];
        e.switch_v = A.switch_v; e.alpha_evo = A.alpha_evo; e.sig_old = A.sig_old[i]; e.tau0 = A.tau0[i];
        e.dr = A.dr;
        const ElemOut r = (A.sw[i] == 1) ? solve_high_v(e, A.dt) : solve_low_v(e, A.dt);
        A.v[i] = r.v; A.f[i] = r.f; A.d[i] = r.d; A.th[i] = r.th; A.sig[i] = r.sig; A.dt_all[i] = r.dt;
        A.instab[i] = r.instab; A.ok[i] = r.ok; A.iters[i] = r.iters;
        if (r.stall) atomicAdd(A.stalls, (unsigned long long)r.stall);
    } else {
        const double v = A.lrate[i];
        A.v[i] = v; A.f[i] = A.f_old[i]; A.d[i] = A.d_old[i] + v * A.dt; A.th[i] = A.th_old[i]; A.sig[i] = 1e10;
        A.dt_all[i] = A.dt; A.instab[i] = 0; A.ok[i] = 1; A.iters[i] = 0;
    }
}

}  // namespace qdfn

using namespace qdfn;

extern "C" int64_t qdfn_solve_one_timestep(const qdfn_step_params* p, int64_t n, int64_t n_loading,
                                           const double* DispOld, const double* FrictionOld, const double* ThetaOld,
                                           const double* VOld, const double* Friction0, const double* a,
                                           const double* b, const double* Dc, const double* K_Self, double Dt,
                                           const double* Omega, const double* Mass, const double* EffNormalStress_i,
                                           const double* Total_Loading_Disp, const int32_t* SolverSwitch,
                                           const double* LoadingRate, const double* D_EffStress_Shear,
                                           const double* EffNormalStress_Old, const double* InitialShearStress,
                                           double* V, double* Friction, double* Disp, double* Theta,
                                           double* EffNormalStress, double* Dt_All, int32_t* InstabilityThistime,
                                           int32_t* FLAG_GoodToGo, int32_t* iters) {
    if (!p || n < 0 || n_loading < 0 || n_loading > n) return fail(QDFN_ERR_ARG, "qdfn_solve_one_timestep: bad sizes");
    const double* in[17] = {DispOld, FrictionOld, ThetaOld, VOld, Friction0, a, b, Dc, K_Self, Omega, Mass,
                            EffNormalStress_i, Total_Loading_Disp, LoadingRate, D_EffStress_Shear,
                            EffNormalStress_Old, InitialShearStress};
    double* out[6] = {V, Friction, Disp, Theta, EffNormalStress, Dt_All};
    for (auto q : in) if (!q) return fail(QDFN_ERR_ARG, "qdfn_solve_one_timestep: null input");
    for (auto q : out) if (!q) return fail(QDFN_ERR_ARG, "qdfn_solve_one_timestep: null output");
    if (!SolverSwitch || !InstabilityThistime || !FLAG_GoodToGo)
        return fail(QDFN_ERR_ARG, "qdfn_solve_one_timestep: null int array");
    if (n == 0) return 0;
    const size_t b8 = sizeof(double) * n, b4 = sizeof(int32_t) * n;
    DevBuf din, dout, dsw, dints, dst;
    QDFN_HIP(din.alloc(b8 * 17));
    QDFN_HIP(dout.alloc(b8 * 6));
    QDFN_HIP(dsw.alloc(b4));
    QDFN_HIP(dints.alloc(b4 * 3));
    QDFN_HIP(dst.alloc(sizeof(unsigned long long)));
    for (int k = 0; k < 17; ++k) QDFN_HIP(hipMemcpy(din.as<double>() + k * n, in[k], b8, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemcpy(dsw.p, SolverSwitch, b4, hipMemcpyHostToDevice));
    QDFN_HIP(hipMemset(dst.p, 0, sizeof(unsigned long long)));
    SolveArgs A;
    A.n = n; A.n_loading = n_loading; A.dt = Dt; A.conv = p->conv_crit; A.v0 = p->v0; A.mu = p->mu; A.rho = p->rho;
    A.switch_v = p->switch_v; A.alpha_evo = p->alpha_evo; A.dr = p->evolution_dr;
    const double* di = din.as<double>();
    A.d_old = di; A.f_old = di + n; A.th_old = di + 2 * n; A.v_old = di + 3 * n; A.f0 = di + 4 * n; A.a = di + 5 * n;
    A.b = di + 6 * n; A.dc = di + 7 * n; A.k = di + 8 * n; A.omega = di + 9 * n; A.mass = di + 10 * n;
    A.sig_i = di + 11 * n; A.tl = di + 12 * n; A.lrate = di + 13 * n; A.dtau = di + 14 * n; A.sig_old = di + 15 * n;
    A.tau0 = di + 16 * n;
    A.sw = dsw.as<int32_t>();
    double* dd = dout.as<double>();
    A.v = dd; A.f = dd + n; A.d = dd + 2 * n; A.th = dd + 3 * n; A.sig = dd + 4 * n; A.dt_all = dd + 5 * n;
    int32_t* ii = dints.as<int32_t>();
    A.instab = ii; A.ok = ii + n; A.iters = ii + 2 * n;
    A.stalls = dst.as<unsigned long long>();
    hipLaunchKernelGGL(k_solve_standalone, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, A);
    QDFN_HIP(hipGetLastError());
    for (int k = 0; k < 6; ++k) QDFN_HIP(hipMemcpy(out[k], dd + k * n, b8, hipMemcpyDeviceToHost));
    QDFN_HIP(hipMemcpy(InstabilityThistime, ii, b4, hipMemcpyDeviceToHost));
    QDFN_HIP(hipMemcpy(FLAG_GoodToGo, ii + n, b4, hipMemcpyDeviceToHost));
    if (iters) QDFN_HIP(hipMemcpy(iters, ii + 2 * n, b4, hipMemcpyDeviceToHost));
    unsigned long long stalls = 0;
    QDFN_HIP(hipMemcpy(&stalls, dst.p, sizeof(stalls), hipMemcpyDeviceToHost));
    return (int64_t)stalls;
}

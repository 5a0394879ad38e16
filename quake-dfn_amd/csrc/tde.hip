// Triangular dislocation (TDstressHS) kernels: see tde_device.h.
#include "assembly.h"
#include "qdfn_internal.h"

namespace qdfn {
int assemble_tri_device(int64_t, const double*, const double*, const double*, const double*, int64_t, const double*,
                        const double*, const double*, double, double, double*, double*, int64_t, hipStream_t) {
    return fail(QDFN_ERR_STATE, "triangular assembly is not built yet");
}
}  // namespace qdfn

extern "C" int qdfn_tde_stress_hs(int64_t, const double*, const double*, const double*, const double*, const double*,
                                  const double*, double, double, double, double, double, double*, double*) {
    return qdfn::fail(QDFN_ERR_STATE, "qdfn_tde_stress_hs is not built yet");
}
extern "C" int qdfn_tde_stiffness_tri(int64_t, const double*, const double*, const double*, const double*, int64_t,
                                      const double*, const double*, const double*, double, double, double*, double*,
                                      int64_t) {
    return qdfn::fail(QDFN_ERR_STATE, "qdfn_tde_stiffness_tri is not built yet");
}
extern "C" int qdfn_rotverts_unitvectors(int64_t, double*, double*, double*, const double*, double*, double*, double*,
                                         double*) {
    return qdfn::fail(QDFN_ERR_STATE, "qdfn_rotverts_unitvectors is not built yet");
}

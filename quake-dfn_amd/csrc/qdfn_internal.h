// Internal helpers shared by the HIP translation units of libqdfn_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/qdfn.h"

namespace qdfn {

// Thread-local last-error text returned by qdfn_last_error().
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define QDFN_HIP(call)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return ::qdfn::fail(QDFN_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// RAII device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t nbytes) {
        release();
        bytes = nbytes;
        if (nbytes == 0) return hipSuccess;
        return hipMalloc(&p, nbytes);
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Degree trigonometry with exact zeros at multiples of 90 degrees; the
// argument is reduced in degrees before conversion, the convention of Julia's
// sind/cosd used throughout the reference assembly
// (Functions_BuildInputFile.jl:430-431, :450-451, :480-497).
__host__ __device__ inline double deg_to_rad(double x) { return x * (3.141592653589793 / 180.0); }

__host__ __device__ inline double sin_deg(double x) {
    if (__builtin_isinf(x) || __builtin_isnan(x)) return NAN;
    const double rx = copysign(fmod(x, 360.0), x);
    const double ax = fabs(rx);
    if (rx == 0.0) return rx;
    if (ax < 45.0) return sin(deg_to_rad(rx));
    if (ax <= 135.0) return copysign(cos(deg_to_rad(90.0 - ax)), rx);
    if (ax == 180.0) return copysign(0.0, rx);
    if (ax < 225.0) return sin(deg_to_rad((180.0 - ax) * (rx > 0.0 ? 1.0 : -1.0)));
    if (ax <= 315.0) return -copysign(cos(deg_to_rad(270.0 - ax)), rx);
    return sin(deg_to_rad(rx - copysign(360.0, rx)));
}

__host__ __device__ inline double cos_deg(double x) {
    if (__builtin_isinf(x) || __builtin_isnan(x)) return NAN;
    const double ax = fabs(fmod(x, 360.0));
    if (ax <= 45.0) return cos(deg_to_rad(ax));
    if (ax < 135.0) return sin(deg_to_rad(90.0 - ax));
    if (ax <= 225.0) return -cos(deg_to_rad(ax - 180.0));
    if (ax < 315.0) return sin(deg_to_rad(ax - 270.0));
    return cos(deg_to_rad(360.0 - ax));
}

inline int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// Control words — coupling gates, the run state's buffer parity, step count,
// Dt and decision flags — are written by one kernel (plain vector stores) and
// read by the next kernels on the stream.  Their addresses are uniform, so a
// plain load may be compiled to a scalar load served by the scalar data cache,
// which vector stores do not update; whether a kernel boundary invalidates that
// cache depends on the acquire fence of the dispatch, which the runtime may
// narrow between kernels of one queue or graph.  Every such read is therefore
// a relaxed atomic load at agent scope: a vector load that the memory model
// makes coherent with stores of earlier kernels on the device.
template <typename T>
__device__ __forceinline__ T ctrl_load(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Device-resident histories of a run (History_Time / History_Disp): time[n_rec],
// disp column-major n_rec x n with leading dimension ld.  Valid until the run
// is destroyed; work on them is ordered after the run's stream.
struct SimHistoryView {
    const double* time = nullptr;
    const double* disp = nullptr;
    int64_t n_rec = 0, ld = 0, n = 0;
    hipStream_t stream = nullptr;
};
int sim_history_view(qdfn_sim_t* s, SimHistoryView* v);

}  // namespace qdfn

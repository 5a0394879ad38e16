// Hierarchical (H-matrix) stiffness operator: build, device layout, product.
//
// Reference: HmatBuild_R / HmatBuild_T (scripts/Functions_BuildInputFile.jl:
// 643-803) build every block densely and compress admissible ones with
// pqrfact; HmatSolver_Pararllel (scripts/Functions_Hmatrix.jl:631-646) adds
// y[rcv] += B * x[src] over blocks; main_H (Functions_RSFDFN3DMain_H.jl:
// 241-248) uses the shear blocks with zeroed diagonals (StiffnessTransition-
// ToLoading, Functions_Hmatrix.jl:459-477) divided by -K_Self.
//
// Build on the device: the full K_S and K_N are assembled in HBM (FP64), dense
// blocks are read out of them and admissible blocks (with tail storage also
// the near-field blocks off the diagonal) are compressed by an adaptive cross
// approximation with full pivoting, one workgroup per block and matrix.  The residual is held explicitly, so the stopping test is exact:
// ||residual||_F <= rtol ||K_S,b||_F or max |residual| <= atol, K_S,b the
// block's shear entries (for the normal part too: both are stresses per unit
// slip, and the normal entries of a near-planar fault are far smaller, so a
// tolerance relative to their own size would spend rank on noise).  atol alone
// gives the reference's pqrfact(block, atol = Tolerance).
//
// Product layout (HBM): the rows are cut into row groups of at most 64 rows
// whose boundaries include every block's receiver boundaries.  Each row group
// owns column-major panels per matrix holding, side by side, the columns of
// every block that covers it: dense blocks contribute their columns, low-rank
// blocks the matching rows of U.  A product is then
//   phase 1: t = V_b x[src_b] for every low-rank factor; the V rows of all
//            factors that share a source range form one panel over that
//            (contiguous) range of x, stored as column-major sub-panels;
//   phase 2: y = panel * [x; t] per row group, the right-hand side gathered
//            through an index list.
// Phase 1 runs one wavefront per task (k_vpanel*), phase 2 one workgroup per
// unit, a column range of one panel of one row group and matrix (k_hunit),
// whose row sums go to the unit's slot; the slots are added in slot order
// (k_hsum, or the step kernel).  Both phases stream contiguous column runs
// with wide loads and FP64 accumulation; sums are combined over fixed shuffle
// trees and (phase 2) through LDS in wavefront order, with no atomics, so
// results are the same run to run.
// Storage: FP64, FP32, FP32 with BF16 low-rank factors (QDFN_PREC_BF16_FAR),
// FP32 with FP16 trailing components (QDFN_PREC_MIXED), and with FP8 ones as
// well (QDFN_PREC_MIXED8); each tier's columns form their own panels in their
// own pool.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <map>
#include <tuple>
#include <type_traits>
#include <thread>
#include <vector>

#include "assembly.h"
#include "dense.h"
#include "hmat.h"

namespace qdfn {

namespace {

constexpr int kRowsPerGroup = 64;
constexpr int kP2Waves = 4;    // phase 2: wavefronts per workgroup (one unit of a row group)
constexpr int64_t kP2Unit = 192 << 10; // phase 2: unit size (panel bytes) a row group is split into
constexpr int kP2MaxSlots = 8;          // phase 2: most units of one row group (larger units beyond)
constexpr int kP2Threads = 64 * kP2Waves;
constexpr int kP2Depth = 4;    // phase 2: steps in flight per wavefront
constexpr int kVThreads = 256; // phase 1: four independent wavefronts per workgroup
constexpr int kVDepth = 4;     // phase 1: steps in flight per wavefront
constexpr int kVTaskEntries = 12288;  // phase 1: target entries per task (any pool)

int64_t round4(int64_t v) { return (v + 3) / 4 * 4; }

// d rounded once to BF16 (8 significant bits, ties to even); magnitudes below
// the smallest normal FP32 become signed zeros
uint16_t to_bf16(double d) {
    if (!std::isfinite(d) || std::fabs(d) < 0x1p-126) {
        const float f = std::isfinite(d) ? (d < 0.0 ? -0.0f : 0.0f) : (float)d;
        uint32_t b;
        std::memcpy(&b, &f, 4);
        return (uint16_t)(b >> 16);
    }
    int e;
    std::frexp(d, &e);  // |d| = f 2^e, 0.5 <= f < 1
    const float r = (float)std::ldexp(std::nearbyint(std::ldexp(d, 8 - e)), e - 8);  // exact in FP32
    uint32_t b;
    std::memcpy(&b, &r, 4);
    return (uint16_t)(b >> 16);
}

}  // namespace

// ---------------------------------------------------------------- kernels ----

// BF16 entry (the upper half of an FP32 word)
struct bf16_t {
    uint16_t u;
};
// FP16 entry (IEEE binary16)
struct fp16_t {
    uint16_t u;
};
// FP8 entry (OCP E4M3: 3 mantissa bits, exponent bias 7, largest 448, quantum
// 2^-9 below 2^-6), decoded by the gfx950 packed conversion
struct fp8_t {
    uint8_t u;
};

// the two 16-bit entries of a 32-bit word (entry 0 in the low half) as FP32
template <class T>
struct Dec16;
template <>
struct Dec16<bf16_t> {
    __device__ static float lo(uint32_t w) { return __uint_as_float(w << 16); }
    __device__ static float hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
};
template <>
struct Dec16<fp16_t> {
    __device__ static float lo(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu)); }
    __device__ static float hi(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)); }
};

template <class T>
__device__ inline void load4raw(const T* p, T v[4]) {
    if constexpr (sizeof(T) == 4) {
        const float4 q = *reinterpret_cast<const float4*>(p);
        v[0] = q.x;
        v[1] = q.y;
        v[2] = q.z;
        v[3] = q.w;
    } else {
        const double2 a = reinterpret_cast<const double2*>(p)[0];
        const double2 b = reinterpret_cast<const double2*>(p)[1];
        v[0] = a.x;
        v[1] = a.y;
        v[2] = b.x;
        v[3] = b.y;
    }
}

// W consecutive entries of one row as loaded (one 16-byte load for FP32 and
// BF16, two for FP64) and their FP64 values
template <class T>
struct RowLoad;
// Panel entries are read once per product: FP32 and 16-bit loads are
// nontemporal (streamed past the caches that keep x, t and the indices;
// measured 5-13% faster than plain loads here), FP64 loads plain (its
// nontemporal pair of 16-byte loads measured slower).
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef unsigned int nt_u4 __attribute__((ext_vector_type(4)));
template <>
struct RowLoad<float> {
    static constexpr int W = 4;
    float4 v;
    __device__ void load(const float* p) {
        const nt_f4 q = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
        v = make_float4(q.x, q.y, q.z, q.w);
    }
    __device__ double get(int j) const { return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w; }
};
template <>
struct RowLoad<double> {
    static constexpr int W = 4;
    double2 a, b;
    __device__ void load(const double* p) {
        a = reinterpret_cast<const double2*>(p)[0];
        b = reinterpret_cast<const double2*>(p)[1];
    }
    __device__ double get(int j) const { return j == 0 ? a.x : j == 1 ? a.y : j == 2 ? b.x : b.y; }
};
template <class T16>
struct RowLoad16 {
    static constexpr int W = 8;
    uint4 v;
    __device__ void load(const T16* p) {
        const nt_u4 q = __builtin_nontemporal_load(reinterpret_cast<const nt_u4*>(p));
        v = make_uint4(q.x, q.y, q.z, q.w);
    }
    // entry 2k is the low half of word k, entry 2k + 1 the high half
    __device__ double get(int j) const {
        const uint32_t w = (j >> 1) == 0 ? v.x : (j >> 1) == 1 ? v.y : (j >> 1) == 2 ? v.z : v.w;
        return (double)((j & 1) ? Dec16<T16>::hi(w) : Dec16<T16>::lo(w));
    }
};
template <>
struct RowLoad<bf16_t> : RowLoad16<bf16_t> {};
template <>
struct RowLoad<fp16_t> : RowLoad16<fp16_t> {};
// eight FP8 entries per 8-byte load (entry j in byte j)
typedef unsigned int nt_u2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
template <>
struct RowLoad<fp8_t> {
    static constexpr int W = 8;
    uint2 v;
    __device__ void load(const fp8_t* p) {
        const nt_u2 q = __builtin_nontemporal_load(reinterpret_cast<const nt_u2*>(p));
        v = make_uint2(q.x, q.y);
    }
    __device__ double get(int j) const {
        const uint32_t w = j < 4 ? v.x : v.y;
        const f32x2 pr = (j & 2) ? __builtin_amdgcn_cvt_pk_f32_fp8((int)w, true)
                                 : __builtin_amdgcn_cvt_pk_f32_fp8((int)w, false);
        return (double)((j & 1) ? pr.y : pr.x);
    }
};

// x moved between lanes by the DPP control CTRL (both 32-bit halves)
template <int CTRL>
__device__ inline double dpp_f64(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over aligned groups of lpr lanes (a power of two), valid in every lane of
// the group: within 16-lane rows by DPP moves (quad permutations [1,0,3,2] and
// [2,3,0,1], then the half-row and row mirrors, lane i with 7 - i and 15 - i:
// after each round all lanes of the group so far hold its sum), across rows
// by shuffles
__device__ inline double group_sum(double v, int lpr) {
    if (lpr >= 2) v += dpp_f64<0xB1>(v);
    if (lpr >= 4) v += dpp_f64<0x4E>(v);
    if (lpr >= 8) v += dpp_f64<0x141>(v);
    if (lpr >= 16) v += dpp_f64<0x140>(v);
    if (lpr >= 32) v += __shfl_xor(v, 16, 64);
    if (lpr >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Output of one row (main_H :241-243 for the shear loading form:
// (H_load * (0 - x)) ./ (-K_Self) with -K_Self = K_ii).
__device__ inline void hmat_out(int kind, int64_t i, double s, const double* __restrict__ x,
                                const double* __restrict__ kdiag, int32_t form, double* __restrict__ yS,
                                double* __restrict__ yN) {
    if (kind == 1)
        yS[i] = form ? (-s) / kdiag[i] : s + kdiag[i] * x[i];
    else
        yN[i] = s;
}

// Phase 2 (y = panels [x; t], one workgroup per unit: a column range of one
// panel of one row group and matrix, at most kP2Unit panel bytes, so that no
// workgroup holds up the end of the launch).  A row group's panels are
// column-major: panel 0 (FP64 / FP32 pool) holds the
// columns of its dense blocks and the leading components of its low-rank
// parts with column stride ld0 = m rounded up to four, panel 1 (16-bit pool)
// the trailing components with ld1 = m rounded up to eight (padding rows
// zero).  Column c of a panel multiplies w[idx[z + c]], w = [x; t].  Each of
// the kP2Waves wavefronts streams one contiguous column slice of one panel,
// as phase 1 does: L = ld / W lanes cover a column (W = 4 entries per lane
// load in FP32 and FP64, 8 in 16 bits; lane j rows [jW, jW + W)), G = 64 / L
// lane groups take G consecutive columns per step.  The slice's right-hand
// side is gathered into the wavefront's LDS chunk by chunk (double-buffered,
// the next chunk's gather in flight while the current one is streamed), so a
// step issues one panel load per lane; steps are software-pipelined
// kP2Depth deep (half that in FP64), loads branch-free.  The lane groups'
// sums are added over a fixed shuffle tree, the wavefronts' sums through LDS
// in wavefront order, the units' sums in unit order, so every row is formed
// in a fixed order (no finish kernel, results the same run to run).
template <class T>
struct P2Step {
    RowLoad<T> w;
};

constexpr int kP2Chunk = 256;  // right-hand-side values staged per wavefront and buffer

// Columns [0, k) of a panel slice at `panel` (gather indices zi); zs: this
// wavefront's two LDS buffers of kP2Chunk values.  acc[v]: this lane's sum for
// row Wj + v over its lane group's columns, then (lanes of group 0) over the
// whole slice.
template <class T>
__device__ __forceinline__ void p2_slice(const T* __restrict__ panel, const int32_t* __restrict__ zi, int32_t k,
                                         int32_t ld, const double* __restrict__ x, const double* __restrict__ t,
                                         int64_t n, double (*zs)[kP2Chunk], double acc[RowLoad<T>::W]) {
    constexpr int W = RowLoad<T>::W;
    constexpr int D = sizeof(T) == 8 ? kP2Depth / 2 : kP2Depth;  // FP64: 32 bytes per lane and step
    constexpr int kPer = kP2Chunk / 64;                          // gathered values per lane and chunk
    const int lane = threadIdx.x & 63;
    const int L = ld / W, G = 64 / L;
    const int g = lane / L, j = lane - g * L;  // g == G: idle lane (j < L still)
    // a chunk is SC steps (a multiple of D) of G columns
    const int32_t SC = D * max(1, kP2Chunk / (G * D)), C = SC * G;
    // steps rounded up to the pipeline depth only: the last chunk stops there
    const int32_t ns = (k + G - 1) / G, nsd = (ns + D - 1) / D * D, nchunk = (ns + SC - 1) / SC;
    const T* base = panel + j * W;
    // gather of chunk q: indices, then values (zero past the slice)
    auto load_idx = [&](int32_t q, int32_t iv[kPer]) {
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int32_t c = q * C + lane + 64 * i;
            iv[i] = zi[c < k ? c : k - 1];
        }
    };
    auto load_vals = [&](int32_t q, const int32_t iv[kPer], double zv[kPer]) {
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int32_t c = q * C + lane + 64 * i;
            const double* zp = iv[i] < n ? x + iv[i] : t + (iv[i] - n);
            const long long bits = __double_as_longlong(*zp);
            zv[i] = __longlong_as_double((c < k && lane + 64 * i < C) ? bits : 0LL);
        }
    };
    auto store_vals = [&](int32_t q, const double zv[kPer]) {
#pragma unroll
        for (int i = 0; i < kPer; ++i) zs[q & 1][lane + 64 * i] = zv[i];
    };
    auto load = [&](int32_t s, P2Step<T>& S) {
        const int32_t c = s * G + g;
        S.w.load(base + (uint32_t)((c < k && g < G) ? c : k - 1) * (uint32_t)ld);
    };
    // step s of chunk q reads its column's value at zq[s G] (zq set per chunk);
    // columns past the slice read a staged zero, idle lanes a zero weight
    const double gm = g < G ? 1.0 : 0.0;
    auto consume = [&](const double* zq, int32_t s, const P2Step<T>& S) {
        const double z = gm * zq[s * G];
#pragma unroll
        for (int v = 0; v < W; ++v) acc[v] = fma(S.w.get(v), z, acc[v]);
    };
    int32_t iv[kPer];
    double zv[kPer];
    load_idx(0, iv);
    P2Step<T> S[D];
#pragma unroll
    for (int d = 0; d < D; ++d) load(d, S[d]);
    load_vals(0, iv, zv);
    load_idx(1, iv);
    store_vals(0, zv);
    for (int32_t q = 0; q < nchunk; ++q) {
        // gather of chunk q + 1 in flight while chunk q is streamed
        load_vals(q + 1, iv, zv);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int32_t s_beg = q * SC, s_end = min(s_beg + SC, nsd);
        const double* zq = zs[q & 1] + min(g, G - 1) - s_beg * G;
        for (int32_t s0 = s_beg; s0 < s_end; s0 += D) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                consume(zq, s0 + d, S[d]);
                load(min(s0 + D + d, nsd - 1), S[d]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        store_vals(q + 1, zv);
        load_idx(q + 2, iv);
    }
    // lane (g, j) adds lane (g + o, j) where g is a multiple of 2o
    for (int o = 1; o < G; o <<= 1) {
        const int src = min(lane + o * L, 63);
        const bool take = (g & (2 * o - 1)) == 0 && g + o < G;
#pragma unroll
        for (int v = 0; v < W; ++v) {
            const double u = __shfl(acc[v], src, 64);
            if (take) acc[v] += u;
        }
    }
}

// One workgroup per unit: the unit's columns in kP2Waves equal slices, the
// slices' row sums added through LDS in wavefront order and written to the
// unit's slot: part[u 2n + r] for row r of [y_S; y_N] (r = out + row, plus n
// for the normal matrix).  Slots of units a group does not have stay zero, so
// a row is the sum over every slot in slot order (k_hsum, or the step kernel;
// no atomics, results the same run to run).  T0: panel-0 entries (float,
// double), T1: panel-1 entries (16-bit; unused when no unit is in panel 1),
// F8: units of panel 2 (FP8) exist.
template <class T>
__device__ __forceinline__ void hunit_slice(const HUnit& R, const T* __restrict__ pool,
                                            const int32_t* __restrict__ idx, int32_t c0, int32_t c1,
                                            const double* __restrict__ x, const double* __restrict__ t, int64_t n,
                                            double (*zs)[kP2Chunk], double* __restrict__ red) {
    constexpr int W = RowLoad<T>::W;
    const int lane = threadIdx.x & 63;
    double acc[W];
#pragma unroll
    for (int v = 0; v < W; ++v) acc[v] = 0.0;
    if (c1 > c0) p2_slice<T>(pool + R.off + (int64_t)c0 * R.ld, idx + R.z + c0, c1 - c0, R.ld, x, t, n, zs, acc);
    if (lane < R.ld / W) {
#pragma unroll
        for (int v = 0; v < W; ++v) red[W * lane + v] = acc[v];
    }
}

template <class T0, class T1, bool F8>
__global__ __launch_bounds__(kP2Threads, 6) void k_hunit(const HUnit* __restrict__ units,
                                                        const T0* __restrict__ pool, const T1* __restrict__ pool_lo,
                                                        const fp8_t* __restrict__ pool_8,
                                                        const int32_t* __restrict__ idx,
                                                        const double* __restrict__ x0, const double* __restrict__ t,
                                                        int64_t n, double* __restrict__ part, CouplingGate g) {
    if (g.gate && ctrl_load(g.gate) == 0) return;
    __shared__ double red[kP2Waves][kRowsPerGroup];
    __shared__ double zs[kP2Waves][2][kP2Chunk];
    const double* x = g.parity ? x0 + (int64_t)ctrl_load(g.parity) * g.parity_stride : x0;
    const HUnit R = units[blockIdx.x];
    const int wave = threadIdx.x >> 6;
    const int32_t c0 = (int32_t)((int64_t)R.k * wave / kP2Waves), c1 = (int32_t)((int64_t)R.k * (wave + 1) / kP2Waves);
    // lanes of lane group 0 hold the slice's row sums (a slice with no columns: zeros)
    if (F8 && R.lo == 2)
        hunit_slice<fp8_t>(R, pool_8, idx, c0, c1, x, t, n, zs[wave], red[wave]);
    else if (R.lo)
        hunit_slice<T1>(R, pool_lo, idx, c0, c1, x, t, n, zs[wave], red[wave]);
    else
        hunit_slice<T0>(R, pool, idx, c0, c1, x, t, n, zs[wave], red[wave]);
    __syncthreads();
    const int row = (int)threadIdx.x;
    if (row < R.m) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kP2Waves; ++w) s += red[w][row];
        part[(int64_t)R.u * 2 * n + (R.kind == 2 ? n : 0) + R.out + row] = s;
    }
    if (g.ran_log && blockIdx.x == 0 && threadIdx.x == 0) g.ran_log[g.slot] = 1;
}

// Rows of the product from the unit slots (slot order), one lane per row of
// [y_S; y_N].
__global__ __launch_bounds__(256) void k_hsum(const double* __restrict__ part, int32_t slots, int64_t n,
                                              const double* __restrict__ x0, const double* __restrict__ kdiag,
                                              int32_t form, double* __restrict__ yS, double* __restrict__ yN,
                                              CouplingGate g) {
    if (g.gate && ctrl_load(g.gate) == 0) return;
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= 2 * n) return;
    const double* x = g.parity ? x0 + (int64_t)ctrl_load(g.parity) * g.parity_stride : x0;
    double s = 0.0;
    for (int32_t u = 0; u < slots; ++u) s += part[(int64_t)u * 2 * n + r];
    hmat_out(r < n ? 1 : 2, r < n ? r : r - n, s, x, kdiag, form, yS, yN);
}

// Phase 1 (t = V x over the V panels).  A V panel (the components, in one
// pool, of the low-rank parts that share a source cluster) is stored as
// column-major sub-panels of at most 64 W rows, W = RowLoad<T>::W entries per
// lane load (4 in FP32 and FP64, 8 in 16 bits), rows padded to a multiple of
// W with zeros: column c of a sub-panel is its ld rows at c ld, so any column
// range is one contiguous run.  One wavefront per task (a sub-panel's columns,
// or a segment of them): L = ld / W lanes cover a column (lane j its rows
// [jW, jW + W)), and the wavefront's G = 64 / L lane groups take G consecutive
// columns per step, so a step reads G ld contiguous entries and G values of x
// (each lane the one of its column).  Steps are software-pipelined kVDepth
// deep: each register slot is reloaded as soon as it is consumed.  Loads are
// branch-free (column clamped, x masked to zero past the task's columns and in
// the lanes past G L).  At the end the G group sums of each row are added over
// a fixed tree of lane shuffles (the same order every run), and the lanes of
// group 0 write W consecutive t rows (times the row's scale for the FP16
// panels of QDFN_PREC_MIXED), or partial sums for column-split panels.
template <class T>
struct VColStep {
    RowLoad<T> w;
    double z;
};

// t row `out` of a finished phase-1 row: the sum, times the row's scale when
// the panel holds FP16 unit vectors (QDFN_PREC_MIXED)
__device__ inline void vpanel_out(const VTask& tk, int32_t row, double v, double* __restrict__ t,
                                  double* __restrict__ vpart, const double* __restrict__ tscale) {
    if (tk.part < 0)
        t[tk.out + row] = tscale ? v * tscale[tk.out + row] : v;
    else
        vpart[tk.part + row] = v;
}

template <class T>
__device__ __forceinline__ void vcol_body(const VTask& tk, const T* __restrict__ pool, const double* __restrict__ x,
                                          double* __restrict__ t, double* __restrict__ vpart,
                                          const double* __restrict__ tscale) {
    constexpr int W = RowLoad<T>::W;
    const int lane = threadIdx.x & 63;
    const int L = tk.lpr, G = 64 / L;
    const int g = lane / L, j = lane - g * L;  // g == G: idle lane (j < L still)
    const int32_t K = tk.K, ns = (K + G - 1) / G;
    const int32_t nsd = (ns + kVDepth - 1) / kVDepth * kVDepth;  // steps, masked past ns
    const T* base = pool + tk.data_off + j * W;
    const double* xs = x + tk.xoff;
    auto load = [&](int32_t s, VColStep<T>& S) {
        const int32_t c = s * G + g;
        const bool ok = c < K && g < G;
        const int32_t cc = ok ? c : K - 1;
        S.w.load(base + (uint32_t)cc * (uint32_t)tk.ld);
        const long long bits = __double_as_longlong(xs[cc]);
        S.z = __longlong_as_double(ok ? bits : 0LL);
    };
    double acc[W];
#pragma unroll
    for (int v = 0; v < W; ++v) acc[v] = 0.0;
    auto consume = [&](const VColStep<T>& S) {
#pragma unroll
        for (int v = 0; v < W; ++v) acc[v] = fma(S.w.get(v), S.z, acc[v]);
    };
    VColStep<T> S[kVDepth];
#pragma unroll
    for (int d = 0; d < kVDepth; ++d) load(d, S[d]);
    for (int32_t s0 = 0; s0 + kVDepth < nsd; s0 += kVDepth) {
#pragma unroll
        for (int d = 0; d < kVDepth; ++d) {
            consume(S[d]);
            load(s0 + kVDepth + d, S[d]);
        }
    }
#pragma unroll
    for (int d = 0; d < kVDepth; ++d) consume(S[d]);
    // lane (g, j) adds lane (g + o, j) where g is a multiple of 2o
    for (int o = 1; o < G; o <<= 1) {
        const int src = min(lane + o * L, 63);
        const bool take = (g & (2 * o - 1)) == 0 && g + o < G;
#pragma unroll
        for (int v = 0; v < W; ++v) {
            const double u = __shfl(acc[v], src, 64);
            if (take) acc[v] += u;
        }
    }
    if (g == 0) {
#pragma unroll
        for (int v = 0; v < W; ++v) {
            const int32_t row = j * W + v;
            if (row < tk.m) vpanel_out(tk, row, acc[v], t, vpart, tscale);
        }
    }
}

template <class T>
__global__ __launch_bounds__(kVThreads) void k_vpanel(const VTask* __restrict__ tasks, int32_t ntasks,
                                                      const T* __restrict__ pool, const double* __restrict__ x0,
                                                      double* __restrict__ t, double* __restrict__ vpart,
                                                      CouplingGate g) {
    if (g.gate && ctrl_load(g.gate) == 0) return;
    const int32_t ti = (int32_t)blockIdx.x * (kVThreads / 64) + (int32_t)(threadIdx.x >> 6);
    if (ti >= ntasks) return;
    const double* x = g.parity ? x0 + (int64_t)ctrl_load(g.parity) * g.parity_stride : x0;
    vcol_body<T>(tasks[ti], pool, x, t, vpart, nullptr);
}

// Phase 1 of QDFN_PREC_MIXED(8): FP32 sub-panels (leading components), FP16
// and (F8) FP8 sub-panels (scaled unit vectors of the trailing components,
// their t rows scaled back) in one launch; the task's flag picks the body
// (uniform per wavefront).
template <bool F8>
__global__ __launch_bounds__(kVThreads, 6) void k_vpanel_mixed(const VTask* __restrict__ tasks, int32_t ntasks,
                                                            const float* __restrict__ pool,
                                                            const fp16_t* __restrict__ pool_lo,
                                                            const fp8_t* __restrict__ pool_8,
                                                            const double* __restrict__ tscale,
                                                            const double* __restrict__ x0, double* __restrict__ t,
                                                            double* __restrict__ vpart, CouplingGate g) {
    if (g.gate && ctrl_load(g.gate) == 0) return;
    const int32_t ti = (int32_t)blockIdx.x * (kVThreads / 64) + (int32_t)(threadIdx.x >> 6);
    if (ti >= ntasks) return;
    const double* x = g.parity ? x0 + (int64_t)ctrl_load(g.parity) * g.parity_stride : x0;
    const VTask tk = tasks[ti];
    if (F8 && tk.lo == 2)
        vcol_body<fp8_t>(tk, pool_8, x, t, vpart, tscale);
    else if (tk.lo)
        vcol_body<fp16_t>(tk, pool_lo, x, t, vpart, tscale);
    else
        vcol_body<float>(tk, pool, x, t, vpart, nullptr);
}

// Phase-1 finish: t rows of column-split panels, segments added in order.
__global__ __launch_bounds__(256) void k_vfinish(const VFin* __restrict__ fins, int32_t nfin,
                                                 const double* __restrict__ vpart, double* __restrict__ t,
                                                 const double* __restrict__ tscale, CouplingGate g) {
    if (g.gate && ctrl_load(g.gate) == 0) return;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    // fins are per row: i indexes a row of some split row chunk
    if (i >= nfin) return;
    const VFin f = fins[i];
    double s = 0.0;
    for (int32_t q = 0; q < f.nseg; ++q) s += vpart[f.part0 + (int64_t)q * f.stride];
    t[f.out] = tscale ? s * tscale[f.out] : s;
}

__global__ void k_mark_skipped(CouplingGate g) {
    if (g.ran_log && g.gate && ctrl_load(g.gate) == 0) g.ran_log[g.slot] = 0;
}

// Adaptive cross approximation with full pivoting of one block (one workgroup).
struct AcaJob {
    const double* src;  // block in the full matrix (column-major, leading dimension ld)
    const double* scale;  // ||K_S,b||_F of the block
    int64_t ld;
    double* R;          // m x k residual workspace
    double* U;          // m x rmax
    double* V;          // rmax x k, row-major
    int32_t m, k, rmax, pad;
    double rtol, atol;
};

constexpr int kAcaThreads = 1024;

__device__ inline void better(double& v, int32_t& id, double& sv, double v2, int32_t id2, double sv2) {
    if (v2 > v || (v2 == v && id2 < id)) {
        v = v2;
        id = id2;
        sv = sv2;
    }
}

// workgroup reduction: argmax of |R| (smallest index on ties, with the signed
// value) and the sum of squares, in a fixed order
__device__ void aca_reduce(double& v, int32_t& id, double& sv, double& ss, double* sh_v, int32_t* sh_i,
                           double* sh_s, double* sh_q) {
    for (int o = 32; o > 0; o >>= 1) {
        const double v2 = __shfl_xor(v, o, 64);
        const int32_t i2 = __shfl_xor(id, o, 64);
        const double s2 = __shfl_xor(sv, o, 64);
        better(v, id, sv, v2, i2, s2);
        ss += __shfl_xor(ss, o, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        sh_v[wave] = v;
        sh_i[wave] = id;
        sh_s[wave] = sv;
        sh_q[wave] = ss;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kAcaThreads / 64; ++w) {
            better(sh_v[0], sh_i[0], sh_s[0], sh_v[w], sh_i[w], sh_s[w]);
            sh_q[0] += sh_q[w];
        }
    }
    __syncthreads();
    v = sh_v[0];
    id = sh_i[0];
    sv = sh_s[0];
    ss = sh_q[0];
}

// Frobenius norm of blocks (one workgroup per block, fixed-order sums): the
// scale of the ACA threshold
struct MaxJob {
    const double* src;
    int64_t ld;
    int32_t m, k;
};

__global__ __launch_bounds__(kAcaThreads) void k_blocknorm(const MaxJob* __restrict__ jobs, double* __restrict__ out) {
    __shared__ double sh[kAcaThreads / 64];
    const MaxJob J = jobs[blockIdx.x];
    double v = 0.0;
    const int64_t mk = (int64_t)J.m * J.k;
    for (int64_t e = threadIdx.x; e < mk; e += kAcaThreads) {
        const int64_t i = e % J.m, j = e / J.m;
        const double a = J.src[i + j * J.ld];
        v += a * a;
    }
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kAcaThreads / 64; ++w) v += sh[w];
        out[blockIdx.x] = sqrt(v);
    }
}

__global__ __launch_bounds__(kAcaThreads) void k_aca(const AcaJob* __restrict__ jobs, int32_t* __restrict__ rank,
                                                     int32_t* __restrict__ converged, double* __restrict__ resid) {
    __shared__ double sh_v[kAcaThreads / 64], sh_s[kAcaThreads / 64], sh_q[kAcaThreads / 64];
    __shared__ int32_t sh_i[kAcaThreads / 64];
    const AcaJob J = jobs[blockIdx.x];
    const int32_t m = J.m, k = J.k, mk = m * k;
    double amax = -1.0, piv = 0.0, ss = 0.0;
    int32_t aid = 0x7fffffff;
    for (int32_t e = threadIdx.x; e < mk; e += kAcaThreads) {
        const int32_t i = e % m, j = e / m;
        const double v = J.src[i + (int64_t)j * J.ld];
        J.R[e] = v;
        better(amax, aid, piv, fabs(v), e, v);
        ss += v * v;
    }
    aca_reduce(amax, aid, piv, ss, sh_v, sh_i, sh_s, sh_q);
    const double fthr = J.rtol * *J.scale;  // Frobenius threshold
    auto done = [&]() { return amax <= J.atol || sqrt(ss) <= fthr || amax == 0.0; };
    int32_t r = 0;
    while (!done() && r < J.rmax) {
        const int32_t is = aid % m, js = aid / m;
        double* u = J.U + (int64_t)r * m;
        double* v = J.V + (int64_t)r * k;
        for (int32_t i = threadIdx.x; i < m; i += kAcaThreads) u[i] = J.R[i + (int64_t)js * m];
        for (int32_t j = threadIdx.x; j < k; j += kAcaThreads) v[j] = J.R[is + (int64_t)j * m] / piv;
        __syncthreads();
        amax = -1.0;
        aid = 0x7fffffff;
        piv = 0.0;
        ss = 0.0;
        for (int32_t e = threadIdx.x; e < mk; e += kAcaThreads) {
            const int32_t i = e % m, j = e / m;
            const double nv = J.R[e] - u[i] * v[j];
            J.R[e] = nv;
            better(amax, aid, piv, fabs(nv), e, nv);
            ss += nv * nv;
        }
        aca_reduce(amax, aid, piv, ss, sh_v, sh_i, sh_s, sh_q);
        ++r;
    }
    if (threadIdx.x == 0) {
        rank[blockIdx.x] = r;
        converged[blockIdx.x] = done() ? 1 : 0;
        resid[blockIdx.x] = sqrt(ss);
    }
}

// ------------------------------------------------------ host-side layout ----

// d rounded once to FP16 (11 significant bits, ties to even; quantum 2^-24
// below 2^-14), for |d| <= 1
uint16_t to_fp16(double d) {
    double r;
    if (std::fabs(d) < 0x1p-14) {
        r = std::ldexp(std::nearbyint(std::ldexp(d, 24)), -24);
    } else {
        int e;
        std::frexp(d, &e);
        r = std::ldexp(std::nearbyint(std::ldexp(d, 11 - e)), e - 11);
    }
    const _Float16 f = (_Float16)r;  // exact: r is an FP16 value
    uint16_t b;
    std::memcpy(&b, &f, 2);
    return b;
}

// d rounded once to FP8 E4M3 (OCP: 4 significant bits, ties to even; quantum
// 2^-9 below 2^-6), for |d| < 448; the byte the gfx950 conversion decodes
uint8_t to_fp8(double d) {
    const uint8_t sgn = std::signbit(d) ? 0x80 : 0x00;
    const double a = std::fabs(d);
    if (a < 0x1p-6) {
        const double q = std::nearbyint(std::ldexp(a, 9));  // 0..8 (8 = the smallest normal)
        if (q >= 8.0) return sgn | 0x08;
        return sgn | (uint8_t)q;
    }
    int e;
    std::frexp(a, &e);  // a = f 2^e, 0.5 <= f < 1: the leading bit weighs 2^(e-1)
    double m = std::nearbyint(std::ldexp(a, 4 - e));  // 8..16: four significant bits
    if (m >= 16.0) {
        m = 8.0;
        ++e;
    }
    const int ef = (e - 1) + 7;  // biased exponent of 2^(e-1)
    if (ef > 15 || (ef == 15 && m - 8.0 > 6.0)) return sgn | 0x7E;  // saturate at 448
    return sgn | (uint8_t)(ef << 3) | (uint8_t)(m - 8.0);
}

extern "C" int qdfn_fp8_round(const double* x, double* y, int64_t count) {
    if ((!x || !y) && count > 0) return fail(QDFN_ERR_ARG, "qdfn_fp8_round: null array");
    for (int64_t i = 0; i < count; ++i) y[i] = from_fp8(to_fp8(x[i]));
    return QDFN_OK;
}

double from_fp8(uint8_t b) {
    const int e = (b >> 3) & 15, mt = b & 7;
    double v;
    if (e == 15 && mt == 7) return NAN;
    if (e == 0) v = std::ldexp((double)mt, -9);
    else v = std::ldexp(1.0 + mt / 8.0, e - 7);
    return (b & 0x80) ? -v : v;
}

// QDFN_PREC_MIXED / _MIXED8: the storage tiers of low-rank part w of b, which
// is in SVD form (column a of U has norm sigma_a, decreasing; V has
// orthonormal rows), within b.slack[w] (tail_tiers): b.r_hi[w] and b.r_8[w].
// Per tail component: fu, fv scale its U column and V row to unit vectors
// whose largest entry lies in [0.5, 1) (FP16) or [128, 256) (FP8), fc =
// 1 / (fu fv) scales its t row.
void choose_tails(HBlock& b, int w, bool fp8) {
    const int64_t m = b.r1 - b.r0 + 1, k = b.s1 - b.s0 + 1, r = b.rank[w];
    const std::vector<double>& U = b.a[w];
    const std::vector<double>& V = b.b[w];
    std::vector<double> sig(r);
    for (int64_t a = 0; a < r; ++a) {
        double q = 0.0;
        for (int64_t i = 0; i < m; ++i) q += U[i + a * m] * U[i + a * m];
        sig[a] = std::sqrt(q);
    }
    int64_t r_hi, r_8;
    tail_tiers(component_rounding(U.data(), V.data(), m, k, r, r), r, b.slack[w], fp8, r_hi, r_8);
    b.r_hi[w] = r_hi;
    b.r_8[w] = r_8;
    std::vector<double>& fu = b.fu[w];
    std::vector<double>& fv = b.fv[w];
    std::vector<double>& fc = b.fc[w];
    fu.assign(r - r_hi, 0.0);
    fv.assign(r - r_hi, 0.0);
    fc.assign(r - r_hi, 0.0);
    for (int64_t a = r_hi; a < r; ++a) {
        if (!(sig[a] > 0.0)) continue;  // a zero component stays zero
        double mu = 0.0, mv = 0.0;
        for (int64_t i = 0; i < m; ++i) mu = std::max(mu, std::fabs(U[i + a * m]));
        for (int64_t c = 0; c < k; ++c) mv = std::max(mv, std::fabs(V[a + c * r]));
        int eu, ev;
        std::frexp(mu / sig[a], &eu);  // largest entry of the unit column = f 2^eu, f in [0.5, 1)
        std::frexp(mv, &ev);
        const int up = a >= r_8 ? 8 : 0;  // FP8: largest entry in [128, 256)
        fu[a - r_hi] = std::ldexp(1.0 / sig[a], up - eu);
        fv[a - r_hi] = std::ldexp(1.0, up - ev);
        fc[a - r_hi] = std::ldexp(sig[a], eu + ev - 2 * up);
    }
}

int hmat_layout(qdfn_hmat* h) {
    ScopeMs timer(&h->build_phase_ms[4]);
    const int64_t n = h->n;
    auto& B = h->blocks;
    // K_S,ii from the diagonal blocks (src range == rcv range)
    h->kdiag_h.assign(n, 0.0);
    for (const HBlock& b : B) {
        if (b.s0 == b.r0 && b.s1 == b.r1 && b.kind[0] == QDFN_HBLK_DENSE) {
            const int64_t m = b.r1 - b.r0 + 1;
            for (int64_t q = 0; q < m; ++q) h->kdiag_h[b.r0 + q] = b.a[0][q + q * m];
        }
    }
    int nz = 1;
    for (const HBlock& b : B)
        for (const auto* v : {&b.a[1], &b.b[1]})
            for (double e : *v)
                if (e != 0.0) nz = 0;
    h->normal_zero = nz;
    // entries of the FP64 / FP32 panels (pool) and of the 16-bit panels (lo):
    // BF16 low-rank factors (QDFN_PREC_BF16_FAR: every component of every
    // low-rank part) or FP16 unit vectors (QDFN_PREC_MIXED: the trailing
    // components of each recompressed part, choose_tails) and FP8 unit vectors
    // (QDFN_PREC_MIXED8: the smallest of those components)
    const int tmode = tail_mode(h->prec);
    const bool lo_far = h->prec == QDFN_PREC_BF16_FAR, mixed = tmode > 0, fp8 = tmode == 2;
    const bool has_lo = lo_far || mixed;
    const int ntier = fp8 ? 3 : 2;  // pools: FP64 / FP32, 16-bit, 8-bit
    std::vector<double> pool, lo, lo8;
    std::vector<int32_t> idx;
    for (HBlock& b : B)
        for (int w = 0; w < 2; ++w) {
            b.r_hi[w] = lo_far ? 0 : b.rank[w];
            b.r_8[w] = b.rank[w];
            b.fu[w].clear();
            b.fv[w].clear();
            b.fc[w].clear();
            if (mixed && b.kind[w] == QDFN_HBLK_LOWRANK && b.rank[w] >= 2 && b.slack[w] > 0.0)
                choose_tails(b, w, fp8);
        }
    // components of tier glo of a low-rank part: [0, r_hi), [r_hi, r_8), [r_8, rank)
    auto tier_span = [&](const HBlock& b, int w, int glo, int64_t& a0, int64_t& a1) {
        a0 = glo == 0 ? 0 : glo == 1 ? b.r_hi[w] : b.r_8[w];
        a1 = glo == 0 ? b.r_hi[w] : glo == 1 ? b.r_8[w] : b.rank[w];
    };
    std::vector<HUnit> units;
    int64_t slots = 1, n_gid = 0;  // unit slots (most units of a row group), row groups
    std::vector<VTask> vtasks;
    std::vector<VFin> vfins;
    std::vector<double> tsc;  // scale of each t row (QDFN_PREC_MIXED / _MIXED8)
    int64_t v_entries = 0, v_entries_lo = 0, v_entries_8 = 0, v_entries_lo_rows = 0, v_cols = 0, vpart_total = 0;
    // phase 1: V panels, one per source range and pool (components [0, r_hi)
    // in FP64 / FP32, [r_hi, r_8) in 16 bits, [r_8, rank) in 8 bits), in order
    // of first appearance; toffs[tier]: first t row of each part's components
    // in that tier's pool
    std::vector<std::vector<int64_t>> toffs(ntier, std::vector<int64_t>(2 * B.size(), -1));
    {
        std::vector<std::tuple<int64_t, int64_t, int>> keys;
        std::vector<std::vector<std::pair<size_t, int>>> members;
        std::map<std::tuple<int64_t, int64_t, int>, size_t> key_index;
        for (size_t bi = 0; bi < B.size(); ++bi)
            for (int w = 0; w < 2; ++w) {
                const HBlock& b = B[bi];
                if (b.kind[w] != QDFN_HBLK_LOWRANK || b.rank[w] == 0) continue;
                for (int glo = 0; glo < ntier; ++glo) {
                    int64_t a0, a1;
                    tier_span(b, w, glo, a0, a1);
                    if (a1 == a0) continue;
                    const auto key = std::make_tuple(b.s0, b.s1, glo);
                    auto it = key_index.find(key);
                    size_t gk;
                    if (it == key_index.end()) {
                        gk = keys.size();
                        key_index[key] = gk;
                        keys.push_back(key);
                        members.emplace_back();
                    } else {
                        gk = it->second;
                    }
                    members[gk].push_back({bi, w});
                }
            }
        int64_t task_env = 0;
        if (const char* e = getenv("QDFN_HMAT_VTASK")) {
            const int64_t v = atoll(e);
            if (v >= 256 && v <= (1 << 20)) task_env = v;
        }
        // lanes per column: the largest count <= the target whose lane groups
        // fill at least 60 of a wavefront's 64 lanes
        auto good_lanes = [](int64_t L) {
            L = std::max<int64_t>(1, std::min<int64_t>(64, L));
            while (L > 1 && (64 / L) * L < 60) --L;
            return L;
        };
        int64_t tbase = 0;
        std::vector<std::pair<size_t, int64_t>> rowsrc;  // (member, component) of each panel row
        for (size_t gk = 0; gk < keys.size(); ++gk) {
            const int64_t s0 = std::get<0>(keys[gk]), k = std::get<1>(keys[gk]) - s0 + 1;
            const int glo = std::get<2>(keys[gk]);
            // 16- and 8-bit panels are read eight entries per lane load, the others four
            const int64_t al = glo ? 8 : 4;
            auto round_al = [&](int64_t v) { return (v + al - 1) / al * al; };
            const int64_t task_entries = task_env ? task_env : kVTaskEntries;
            std::vector<double>& vdst = glo == 2 ? lo8 : glo ? lo : pool;
            rowsrc.clear();
            for (size_t f = 0; f < members[gk].size(); ++f) {
                const HBlock& b = B[members[gk][f].first];
                const int w = members[gk][f].second;
                toffs[glo][2 * members[gk][f].first + w] = tbase + (int64_t)rowsrc.size();
                int64_t a0, a1;
                tier_span(b, w, glo, a0, a1);
                for (int64_t a = a0; a < a1; ++a) rowsrc.push_back({f, a});
            }
            const int64_t rows = (int64_t)rowsrc.size();
            tsc.resize(tbase + rows, 1.0);
            // sub-panels of L al rows, L chosen for about task_entries entries per task
            const int64_t L = good_lanes(task_entries / (k * al));
            const int64_t sub_rows = L * al;
            for (int64_t ra = 0; ra < rows; ra += sub_rows) {
                const int64_t rs = std::min<int64_t>(sub_rows, rows - ra), ld = round_al(rs);
                const int64_t off = (int64_t)vdst.size();
                vdst.resize(off + ld * k, 0.0);
                for (int64_t q = 0; q < rs; ++q) {
                    const auto& m = members[gk][rowsrc[ra + q].first];
                    const HBlock& b = B[m.first];
                    const int w = m.second;
                    const int64_t r = b.rank[w], a = rowsrc[ra + q].second, a0 = b.r_hi[w];
                    const double mult = (glo && mixed) ? b.fv[w][a - a0] : 1.0;
                    const std::vector<double>& V = b.b[w];
                    for (int64_t c = 0; c < k; ++c) vdst[off + c * ld + q] = V[a + c * r] * mult;
                    if (glo && mixed) tsc[tbase + ra + q] = b.fc[w][a - a0];
                }
                // wavefront tasks: all columns, or (sub-panels of W rows past four
                // times the target) column segments whose partial sums k_vfinish adds
                const int64_t nseg = std::max<int64_t>(1, (ld * k + 4 * task_entries - 1) / (4 * task_entries));
                const int64_t segw = (k + nseg - 1) / nseg;
                for (int64_t sg = 0; sg < nseg; ++sg) {
                    const int64_t cs = sg * segw, ce = std::min<int64_t>(k, cs + segw);
                    if (cs >= ce) return fail(QDFN_ERR_STATE, "H-matrix: empty phase-1 segment");
                    VTask T;
                    T.data_off = off + cs * ld;
                    T.xoff = (int32_t)(s0 + cs);
                    T.out = (int32_t)(tbase + ra);
                    T.m = (int32_t)rs;
                    T.K = (int32_t)(ce - cs);
                    T.ld = (int32_t)ld;
                    T.lpr = (int32_t)(ld / al);
                    T.part = nseg > 1 ? (int32_t)(vpart_total + sg * rs) : -1;
                    T.lo = glo;
                    vtasks.push_back(T);
                }
                if (nseg > 1) {
                    for (int64_t q = 0; q < rs; ++q) {
                        VFin F;
                        F.part0 = vpart_total + q;
                        F.stride = (int32_t)rs;
                        F.nseg = (int32_t)nseg;
                        F.out = (int32_t)(tbase + ra + q);
                        F.pad = 0;
                        vfins.push_back(F);
                    }
                    vpart_total += nseg * rs;
                }
            }
            (glo == 2 ? v_entries_8 : glo ? v_entries_lo : v_entries) += rows * k;
            if (glo) v_entries_lo_rows += rows;
            v_cols += k;
            tbase += rows;
        }
        h->n_t = tbase;
    }
    if (n + h->n_t > 0x7fffffffLL) return fail(QDFN_ERR_ARG, "H-matrix: index space exceeds int32");
    // phase 2: row groups of at most 64 rows whose boundaries include every receiver boundary
    std::vector<int64_t> cut = {0, n};
    for (const HBlock& b : B) {
        cut.push_back(b.r0);
        cut.push_back(b.r1 + 1);
    }
    std::sort(cut.begin(), cut.end());
    cut.erase(std::unique(cut.begin(), cut.end()), cut.end());
    std::vector<int64_t> g0;
    for (size_t q = 0; q + 1 < cut.size(); ++q)
        for (int64_t r = cut[q]; r < cut[q + 1]; r += kRowsPerGroup) g0.push_back(r);
    const int64_t G = (int64_t)g0.size();
    g0.push_back(n);
    std::vector<std::vector<int32_t>> cover(G);
    for (int32_t bi = 0; bi < (int32_t)B.size(); ++bi) {
        const HBlock& b = B[bi];
        int64_t gi = std::lower_bound(g0.begin(), g0.begin() + G, b.r0) - g0.begin();
        for (; gi < G && g0[gi + 1] - 1 <= b.r1; ++gi) cover[gi].push_back(bi);
    }
    int64_t out_rows2 = 0, b2 = 0, k2 = 0;  // output rows, stored entry bytes, columns
    const int64_t es = h->prec == QDFN_PREC_FP64 ? sizeof(double) : sizeof(float);  // pool entries
    int64_t unit_bytes = kP2Unit;
    if (const char* e = getenv("QDFN_HMAT_UNIT")) {
        const long long v = atoll(e);
        if (v >= 4096 && v <= (1LL << 30)) unit_bytes = v;
    }
    for (int64_t gi = 0; gi < G; ++gi)
        for (int w = 0; w < 2; ++w) {
            const int64_t m = g0[gi + 1] - g0[gi], row0 = g0[gi];
            out_rows2 += m;
            // the row group's panels (column-major, column stride ld): in the
            // FP64 / FP32 pool (tier 0) the columns of its dense blocks and the
            // components [0, r_hi) of its low-rank parts, in the 16-bit pool
            // (tier 1) the components [r_hi, r_8), in the 8-bit pool (tier 2)
            // [r_8, rank); returns (offset, first index, columns)
            auto panel = [&](int tier, std::vector<double>& dst, int64_t& off, int64_t& i0) {
                const int64_t ld = tier ? (m + 7) / 8 * 8 : round4(m);
                auto span = [&](const HBlock& b, int64_t& a0, int64_t& a1) {
                    if (b.kind[w] == QDFN_HBLK_DENSE) {
                        a0 = 0;
                        a1 = tier ? 0 : b.s1 - b.s0 + 1;
                    } else {
                        tier_span(b, w, tier, a0, a1);
                    }
                };
                i0 = (int64_t)idx.size();
                for (int32_t bi : cover[gi]) {
                    const HBlock& b = B[bi];
                    int64_t a0, a1;
                    span(b, a0, a1);
                    for (int64_t c = a0; c < a1; ++c) {
                        if (b.kind[w] == QDFN_HBLK_DENSE)
                            idx.push_back((int32_t)(b.s0 + c));
                        else
                            idx.push_back((int32_t)(n + toffs[tier][2 * bi + w] + (c - a0)));
                    }
                }
                const int64_t K = (int64_t)idx.size() - i0;
                off = (int64_t)dst.size();
                dst.resize(off + K * ld, 0.0);
                int64_t col = 0;
                for (int32_t bi : cover[gi]) {
                    const HBlock& b = B[bi];
                    int64_t a0, a1;
                    span(b, a0, a1);
                    if (a1 == a0) continue;
                    const int64_t mb = b.r1 - b.r0 + 1, roff = row0 - b.r0;
                    const bool diag = (w == 0 && b.s0 == b.r0 && b.s1 == b.r1 && b.kind[0] == QDFN_HBLK_DENSE);
                    const bool scaled = tier && mixed && b.kind[w] == QDFN_HBLK_LOWRANK;
                    const std::vector<double>& su = b.fu[w];
                    const std::vector<double>& A = b.a[w];
                    for (int64_t c = a0; c < a1; ++c)
                        for (int64_t i = 0; i < m; ++i) {
                            double v = A[(roff + i) + c * mb];
                            if (diag && row0 + i == b.s0 + c) v = 0.0;  // loading form: zero diagonal
                            if (scaled) v *= su[c - b.r_hi[w]];
                            dst[off + (col + (c - a0)) * ld + i] = v;
                        }
                    col += a1 - a0;
                }
                return K;
            };
            const int64_t ldt[3] = {round4(m), (m + 7) / 8 * 8, (m + 7) / 8 * 8};
            const int64_t est[3] = {es, (int64_t)sizeof(uint16_t), 1};  // entry bytes per tier
            int64_t offt[3] = {0, 0, 0}, zt[3] = {0, 0, 0}, kt[3] = {0, 0, 0};
            kt[0] = panel(0, pool, offt[0], zt[0]);
            if (has_lo) kt[1] = panel(1, lo, offt[1], zt[1]);
            if (fp8) kt[2] = panel(2, lo8, offt[2], zt[2]);
            // units of at most unit_bytes panel bytes (at least one per group);
            // unit u of the group writes slot u; the tiers' units in tier order
            int64_t ub = unit_bytes, St[3];
            for (;;) {
                int64_t S = 0;
                for (int q = 0; q < 3; ++q) S += St[q] = (ldt[q] * kt[q] * est[q] + ub - 1) / ub;
                if (S <= kP2MaxSlots) break;
                ub += ub / 4;
            }
            if (St[0] + St[1] + St[2] == 0) St[0] = 1;
            const int64_t S = St[0] + St[1] + St[2];
            slots = std::max(slots, S);
            for (int64_t u = 0; u < S; ++u) {
                const int tr = u < St[0] ? 0 : u < St[0] + St[1] ? 1 : 2;
                const int64_t Sp = St[tr], up = u - (tr > 0 ? St[0] : 0) - (tr > 1 ? St[1] : 0), kp = kt[tr];
                const int64_t ca = kp * up / Sp, cb = kp * (up + 1) / Sp, ldp = ldt[tr];
                HUnit U;
                U.off = offt[tr] + ca * ldp;
                U.z = zt[tr] + ca;
                U.k = (int32_t)(cb - ca);
                U.ld = (int32_t)ldp;
                U.lo = tr;
                U.out = (int32_t)row0;
                U.m = (int32_t)m;
                U.kind = 1 + w;
                U.u = (int32_t)u;
                U.pad = 0;
                units.push_back(U);
            }
            ++n_gid;
            b2 += m * (kt[0] * es + kt[1] * (int64_t)sizeof(uint16_t) + kt[2]);
            k2 += kt[0] + kt[1] + kt[2];
        }
    // largest units first (the dispatcher hands out workgroups in order, so
    // the tail of a launch is made of small units)
    auto unit_bytes_of = [&](const HUnit& U) {
        return (int64_t)U.ld * U.k * (U.lo == 2 ? 1 : U.lo ? (int64_t)sizeof(uint16_t) : es);
    };
    // XCD-aware order: the dispatcher hands workgroup i to XCD i mod 8, each
    // XCD with its own L2.  Row groups are split into 8 contiguous row ranges
    // of equal panel bytes, one per XCD, so that an XCD's units gather the x
    // and t entries of their own receiver range (the blocks below the top
    // levels of the tree lie within one range) instead of every XCD fetching
    // all of x and t into its L2.  Position p of the launch takes the next
    // largest unit of XCD p mod 8's list (another list's once it is empty).
    {
        constexpr int kXcds = 8;
        int64_t total = 0;
        for (const HUnit& U : units) total += unit_bytes_of(U);
        std::vector<int64_t> rowbytes(n + 1, 0);  // panel bytes of the units starting at each row
        for (const HUnit& U : units) rowbytes[U.out] += unit_bytes_of(U);
        std::vector<int8_t> xcd_of_row(n, 0);
        int64_t acc = 0;
        for (int64_t r = 0; r < n; ++r) {
            xcd_of_row[r] = (int8_t)std::min<int64_t>(kXcds - 1, total > 0 ? acc * kXcds / total : 0);
            acc += rowbytes[r];
        }
        std::vector<std::vector<HUnit>> lists(kXcds);
        for (const HUnit& U : units) lists[xcd_of_row[U.out]].push_back(U);
        for (auto& L : lists)
            std::stable_sort(L.begin(), L.end(),
                             [&](const HUnit& a, const HUnit& b) { return unit_bytes_of(a) > unit_bytes_of(b); });
        std::vector<size_t> next(kXcds, 0);
        const size_t nu = units.size();
        units.clear();
        for (size_t pos = 0; pos < nu; ++pos) {
            int x = (int)(pos % kXcds);
            if (next[x] == lists[x].size()) {  // this XCD's list is done: the largest unit left anywhere
                int best = -1;
                for (int y = 0; y < kXcds; ++y)
                    if (next[y] < lists[y].size() &&
                        (best < 0 || unit_bytes_of(lists[y][next[y]]) > unit_bytes_of(lists[best][next[best]])))
                        best = y;
                x = best;
            }
            units.push_back(lists[x][next[x]++]);
        }
    }
    h->n_units = (int64_t)units.size();
    h->n_slots = slots;
    int64_t slot_rows = 0;  // rows the units write to their slots
    for (const HUnit& U : units) slot_rows += U.m;
    std::stable_sort(vtasks.begin(), vtasks.end(), [](const VTask& a, const VTask& b) {
        return (int64_t)a.K * a.ld > (int64_t)b.K * b.ld;
    });
    h->n_vtasks = (int64_t)vtasks.size();
    h->n_vfins = (int64_t)vfins.size();
    if (vpart_total > 0x7fffffffLL) return fail(QDFN_ERR_ARG, "H-matrix: phase-1 partial space exceeds int32");
    if (getenv("QDFN_HMAT_VERBOSE") && !vtasks.empty()) {
        std::vector<int64_t> sz;
        for (const VTask& T : vtasks) sz.push_back((int64_t)T.m * T.K);
        std::sort(sz.begin(), sz.end());
        fprintf(stderr, "[hmat] phase 1: %zu wavefront tasks, %lld entries, per task min %lld median %lld p90 %lld "
                "max %lld\n", sz.size(), (long long)v_entries, (long long)sz.front(), (long long)sz[sz.size() / 2],
                (long long)sz[sz.size() * 9 / 10], (long long)sz.back());
    }
    if (getenv("QDFN_HMAT_VERBOSE") && !units.empty()) {
        std::vector<int64_t> sz;
        for (const HUnit& U : units) sz.push_back(unit_bytes_of(U));
        std::sort(sz.begin(), sz.end());
        fprintf(stderr, "[hmat] phase 2: %lld row groups, %zu units, %lld bytes, per unit min %lld median %lld p90 %lld "
                "max %lld\n", (long long)n_gid, sz.size(), (long long)b2, (long long)sz.front(),
                (long long)sz[sz.size() / 2], (long long)sz[sz.size() * 9 / 10], (long long)sz.back());
    }
    h->stored_entries = 0;
    for (const HBlock& b : B)
        for (int w = 0; w < 2; ++w) h->stored_entries += (int64_t)(b.a[w].size() + b.b[w].size());
    // upload
    if (pool.size() < 4) pool.resize(4, 0.0);
    QDFN_HIP(h->pool.alloc(es * pool.size()));
    if (has_lo) {
        std::vector<uint16_t> b(std::max<size_t>(lo.size(), 4), 0);
        for (size_t q = 0; q < lo.size(); ++q) b[q] = lo_far ? to_bf16(lo[q]) : to_fp16(lo[q]);
        QDFN_HIP(h->pool_lo.alloc(sizeof(uint16_t) * b.size()));
        QDFN_HIP(hipMemcpy(h->pool_lo.p, b.data(), sizeof(uint16_t) * b.size(), hipMemcpyHostToDevice));
    }
    if (fp8) {
        std::vector<uint8_t> b(std::max<size_t>(lo8.size(), 8), 0);
        for (size_t q = 0; q < lo8.size(); ++q) b[q] = to_fp8(lo8[q]);
        QDFN_HIP(h->pool_8.alloc(b.size()));
        QDFN_HIP(hipMemcpy(h->pool_8.p, b.data(), b.size(), hipMemcpyHostToDevice));
    }
    if (mixed) {
        tsc.resize(std::max<int64_t>(h->n_t, 1), 1.0);
        QDFN_HIP(h->tscale.alloc(sizeof(double) * tsc.size()));
        QDFN_HIP(hipMemcpy(h->tscale.p, tsc.data(), sizeof(double) * tsc.size(), hipMemcpyHostToDevice));
    }
    if (h->prec != QDFN_PREC_FP64) {
        std::vector<float> f(pool.begin(), pool.end());
        QDFN_HIP(hipMemcpy(h->pool.p, f.data(), es * f.size(), hipMemcpyHostToDevice));
    } else {
        QDFN_HIP(hipMemcpy(h->pool.p, pool.data(), es * pool.size(), hipMemcpyHostToDevice));
    }
    QDFN_HIP(h->units.alloc(sizeof(HUnit) * std::max<size_t>(units.size(), 1)));
    if (!units.empty())
        QDFN_HIP(hipMemcpy(h->units.p, units.data(), sizeof(HUnit) * units.size(), hipMemcpyHostToDevice));
    // unit slots: zero once here; a slot no unit writes stays zero
    QDFN_HIP(h->part.alloc(sizeof(double) * slots * 2 * n));
    QDFN_HIP(hipMemset(h->part.p, 0, sizeof(double) * slots * 2 * n));
    QDFN_HIP(h->vtasks.alloc(sizeof(VTask) * std::max<size_t>(vtasks.size(), 1)));
    if (!vtasks.empty())
        QDFN_HIP(hipMemcpy(h->vtasks.p, vtasks.data(), sizeof(VTask) * vtasks.size(), hipMemcpyHostToDevice));
    QDFN_HIP(h->vfins.alloc(sizeof(VFin) * std::max<size_t>(vfins.size(), 1)));
    if (!vfins.empty())
        QDFN_HIP(hipMemcpy(h->vfins.p, vfins.data(), sizeof(VFin) * vfins.size(), hipMemcpyHostToDevice));
    QDFN_HIP(h->vpart.alloc(sizeof(double) * std::max<int64_t>(vpart_total, 1)));
    QDFN_HIP(h->idx.alloc(sizeof(int32_t) * std::max<size_t>(idx.size(), 1)));
    if (!idx.empty()) QDFN_HIP(hipMemcpy(h->idx.p, idx.data(), sizeof(int32_t) * idx.size(), hipMemcpyHostToDevice));
    QDFN_HIP(h->t.alloc(sizeof(double) * std::max<int64_t>(h->n_t, 1)));
    QDFN_HIP(h->kdiag.alloc(sizeof(double) * n));
    QDFN_HIP(hipMemcpy(h->kdiag.p, h->kdiag_h.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    QDFN_HIP(h->x.alloc(sizeof(double) * n));
    QDFN_HIP(h->yS.alloc(sizeof(double) * n));
    QDFN_HIP(h->yN.alloc(sizeof(double) * n));
    // HBM bytes of one product: panel entries, indices, descriptors, t (write + read), x, yS, yN
    h->phase_bytes = (int64_t)(es * pool.size() + sizeof(uint16_t) * lo.size() + lo8.size() + sizeof(int32_t) * idx.size() +
                               sizeof(HUnit) * units.size() + sizeof(VTask) * vtasks.size() +
                               sizeof(VFin) * vfins.size() + 8 * slot_rows + 8 * slots * 2 * n +
                               16 * vpart_total + 16 * h->n_t +
                               (mixed ? 8 * h->n_t : 0) + 8 * 3 * n);
    h->device_bytes = h->phase_bytes;
    // algorithmic bytes of each phase's main kernel: stored entries (unpadded),
    // index list (phase 2), gathered right-hand side, outputs
    // phase 1: V entries, each panel's x slice once, t
    h->p1_entries = v_entries * es + v_entries_lo * (int64_t)sizeof(uint16_t) + v_entries_8;
    h->p1_other = 8 * v_cols + 8 * h->n_t + (mixed ? 8 * v_entries_lo_rows : 0);
    h->p2_entries = b2;
    h->p2_index = 4 * k2;
    h->p2_gather = 8 * k2;
    h->p2_out = 8 * out_rows2;
    h->alg_bytes1 = h->p1_entries + h->p1_other;
    h->alg_bytes2 = b2 + 12 * k2 + 8 * out_rows2;
    return QDFN_OK;
}

int launch_hmat(const qdfn_hmat* h, const double* d_x, const CouplingGate& g, int32_t form, double* d_yS,
                double* d_yN, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, hipEvent_t ev2, bool finish) {
    const double* kd = h->kdiag.as<double>();
    double* t = h->t.as<double>();
    const bool f32 = h->prec == QDFN_PREC_FP32, bf = h->prec == QDFN_PREC_BF16_FAR;
    const bool f8 = h->prec == QDFN_PREC_MIXED8, mx = h->prec == QDFN_PREC_MIXED || f8;
    const double* tsc = mx ? h->tscale.as<double>() : nullptr;
    const fp8_t* p8 = h->pool_8.as<fp8_t>();
    if (ev0) QDFN_HIP(hipEventRecord(ev0, s));
    if (h->n_vtasks > 0) {
        const unsigned grid = (unsigned)((h->n_vtasks + kVThreads / 64 - 1) / (kVThreads / 64));
        const VTask* vt = h->vtasks.as<VTask>();
        double* vp = h->vpart.as<double>();
        if (f8)
            hipLaunchKernelGGL(k_vpanel_mixed<true>, dim3(grid), dim3(kVThreads), 0, s, vt, (int32_t)h->n_vtasks,
                               h->pool.as<float>(), h->pool_lo.as<fp16_t>(), p8, tsc, d_x, t, vp, g);
        else if (mx)
            hipLaunchKernelGGL(k_vpanel_mixed<false>, dim3(grid), dim3(kVThreads), 0, s, vt, (int32_t)h->n_vtasks,
                               h->pool.as<float>(), h->pool_lo.as<fp16_t>(), p8, tsc, d_x, t, vp, g);
        else if (f32)
            hipLaunchKernelGGL((k_vpanel<float>), dim3(grid), dim3(kVThreads), 0, s, vt, (int32_t)h->n_vtasks,
                               h->pool.as<float>(), d_x, t, vp, g);
        else if (bf)
            hipLaunchKernelGGL((k_vpanel<bf16_t>), dim3(grid), dim3(kVThreads), 0, s, vt, (int32_t)h->n_vtasks,
                               h->pool_lo.as<bf16_t>(), d_x, t, vp, g);
        else
            hipLaunchKernelGGL((k_vpanel<double>), dim3(grid), dim3(kVThreads), 0, s, vt, (int32_t)h->n_vtasks,
                               h->pool.as<double>(), d_x, t, vp, g);
        QDFN_HIP(hipGetLastError());
        if (h->n_vfins > 0) {
            hipLaunchKernelGGL(k_vfinish, dim3((unsigned)((h->n_vfins + 255) / 256)), dim3(256), 0, s,
                               h->vfins.as<VFin>(), (int32_t)h->n_vfins, vp, t, tsc, g);
            QDFN_HIP(hipGetLastError());
        }
    }
    if (g.ran_log) hipLaunchKernelGGL(k_mark_skipped, dim3(1), dim3(1), 0, s, g);
    if (ev1) QDFN_HIP(hipEventRecord(ev1, s));
    const HUnit* un = h->units.as<HUnit>();
    const unsigned nu = (unsigned)h->n_units;
    const int32_t* ix = h->idx.as<int32_t>();
    double* pt = h->part.as<double>();
    if (nu == 0) return fail(QDFN_ERR_STATE, "H-matrix: no phase-2 units");
    if (bf)
        hipLaunchKernelGGL((k_hunit<float, bf16_t, false>), dim3(nu), dim3(kP2Threads), 0, s, un,
                           h->pool.as<float>(), h->pool_lo.as<bf16_t>(), p8, ix, d_x, t, h->n, pt, g);
    else if (f8)
        hipLaunchKernelGGL((k_hunit<float, fp16_t, true>), dim3(nu), dim3(kP2Threads), 0, s, un,
                           h->pool.as<float>(), h->pool_lo.as<fp16_t>(), p8, ix, d_x, t, h->n, pt, g);
    else if (mx)
        hipLaunchKernelGGL((k_hunit<float, fp16_t, false>), dim3(nu), dim3(kP2Threads), 0, s, un,
                           h->pool.as<float>(), h->pool_lo.as<fp16_t>(), p8, ix, d_x, t, h->n, pt, g);
    else if (f32)
        hipLaunchKernelGGL((k_hunit<float, fp16_t, false>), dim3(nu), dim3(kP2Threads), 0, s, un,
                           h->pool.as<float>(), (const fp16_t*)nullptr, p8, ix, d_x, t, h->n, pt, g);
    else
        hipLaunchKernelGGL((k_hunit<double, fp16_t, false>), dim3(nu), dim3(kP2Threads), 0, s, un,
                           h->pool.as<double>(), (const fp16_t*)nullptr, p8, ix, d_x, t, h->n, pt, g);
    QDFN_HIP(hipGetLastError());
    if (ev2) QDFN_HIP(hipEventRecord(ev2, s));
    if (finish) {
        hipLaunchKernelGGL(k_hsum, dim3((unsigned)((2 * h->n + 255) / 256)), dim3(256), 0, s, pt,
                           (int32_t)h->n_slots, h->n, d_x, kd, form, d_yS, d_yN, g);
        QDFN_HIP(hipGetLastError());
    }
    return QDFN_OK;
}

// ---------------------------------------------------------------- build ----

qdfn_hmat* new_hmat(int64_t n, int32_t prec) {
    if (n <= 0 || (prec != QDFN_PREC_FP64 && prec != QDFN_PREC_FP32 && prec != QDFN_PREC_BF16_FAR &&
                   prec != QDFN_PREC_MIXED && prec != QDFN_PREC_MIXED8)) {
        fail(QDFN_ERR_ARG, "H-matrix: bad size or precision");
        return nullptr;
    }
    qdfn_hmat* h = new qdfn_hmat();
    h->n = n;
    h->prec = prec;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        fail(QDFN_ERR_HIP, "H-matrix: stream creation failed");
        delete h;
        return nullptr;
    }
    return h;
}

int check_blocks(int64_t n, int64_t nblk, const int64_t* range, const int64_t* adm) {
    if (nblk <= 0 || !range || !adm) return fail(QDFN_ERR_ARG, "H-matrix: empty block list");
    for (int64_t q = 0; q < nblk; ++q) {
        const int64_t* r = range + 4 * q;
        if (r[0] < 1 || r[1] < r[0] || r[1] > n || r[2] < 1 || r[3] < r[2] || r[3] > n)
            return fail(QDFN_ERR_ARG, "H-matrix: block range out of bounds");
    }
    return QDFN_OK;
}

// ---- recompression of ACA factors (host): U V = Q_u (R_u R_v^T) Q_v^T, SVD of
// the r x r core, truncation to the remaining Frobenius budget.

// thin Householder QR of A (m x r, column-major, m >= r): Q (m x r), R (r x r)
static void qr_thin(int64_t m, int64_t r, const double* A, std::vector<double>& Q, std::vector<double>& R) {
    std::vector<double> W(A, A + m * r), tau(r, 0.0), diag(r, 0.0);
    for (int64_t j = 0; j < r; ++j) {
        double* a = W.data() + j * m;
        double nrm = 0.0;
        for (int64_t i = j; i < m; ++i) nrm += a[i] * a[i];
        nrm = std::sqrt(nrm);
        if (nrm == 0.0) {
            diag[j] = 0.0;
            continue;
        }
        const double alpha = a[j] >= 0 ? -nrm : nrm;
        a[j] -= alpha;  // v = x - alpha e_j, stored in place
        double vn = 0.0;
        for (int64_t i = j; i < m; ++i) vn += a[i] * a[i];
        tau[j] = vn > 0 ? 2.0 / vn : 0.0;
        diag[j] = alpha;
        for (int64_t c = j + 1; c < r; ++c) {
            double* b = W.data() + c * m;
            double sdot = 0.0;
            for (int64_t i = j; i < m; ++i) sdot += a[i] * b[i];
            sdot *= tau[j];
            for (int64_t i = j; i < m; ++i) b[i] -= sdot * a[i];
        }
    }
    R.assign(r * r, 0.0);
    for (int64_t j = 0; j < r; ++j) {
        for (int64_t i = 0; i < j; ++i) R[i + j * r] = W[i + j * m];
        R[j + j * r] = diag[j];
    }
    Q.assign(m * r, 0.0);
    for (int64_t j = 0; j < r; ++j) Q[j + j * m] = 1.0;
    for (int64_t j = r - 1; j >= 0; --j) {
        const double* a = W.data() + j * m;
        if (tau[j] == 0.0) continue;
        for (int64_t c = 0; c < r; ++c) {
            double* q = Q.data() + c * m;
            double sdot = 0.0;
            for (int64_t i = j; i < m; ++i) sdot += a[i] * q[i];
            sdot *= tau[j];
            for (int64_t i = j; i < m; ++i) q[i] -= sdot * a[i];
        }
    }
}

// one-sided Jacobi SVD of C (r x r): C = W diag(sig) Z^T
static void svd_jacobi(int64_t r, std::vector<double> C, std::vector<double>& W, std::vector<double>& sig,
                       std::vector<double>& Z) {
    Z.assign(r * r, 0.0);
    for (int64_t j = 0; j < r; ++j) Z[j + j * r] = 1.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        for (int64_t p = 0; p < r; ++p)
            for (int64_t q = p + 1; q < r; ++q) {
                double al = 0, be = 0, ga = 0;
                for (int64_t i = 0; i < r; ++i) {
                    const double x = C[i + p * r], y = C[i + q * r];
                    al += x * x;
                    be += y * y;
                    ga += x * y;
                }
                if (std::fabs(ga) <= 1e-15 * std::sqrt(al * be) || ga == 0.0) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
                for (int64_t i = 0; i < r; ++i) {
                    const double x = C[i + p * r], y = C[i + q * r];
                    C[i + p * r] = c * x - sn * y;
                    C[i + q * r] = sn * x + c * y;
                    const double zx = Z[i + p * r], zy = Z[i + q * r];
                    Z[i + p * r] = c * zx - sn * zy;
                    Z[i + q * r] = sn * zx + c * zy;
                }
            }
        if (!rotated) break;
    }
    sig.assign(r, 0.0);
    W.assign(r * r, 0.0);
    for (int64_t j = 0; j < r; ++j) {
        double nn = 0.0;
        for (int64_t i = 0; i < r; ++i) nn += C[i + j * r] * C[i + j * r];
        sig[j] = std::sqrt(nn);
        if (sig[j] > 0)
            for (int64_t i = 0; i < r; ++i) W[i + j * r] = C[i + j * r] / sig[j];
    }
}

std::vector<CompRound> component_rounding(const double* U, const double* V, int64_t m, int64_t k, int64_t r,
                                          int64_t ldv) {
    std::vector<CompRound> out(r);
    auto fp16_value = [](double x) {
        const uint16_t b = to_fp16(x);
        _Float16 f;
        std::memcpy(&f, &b, 2);
        return (double)f;
    };
    for (int64_t a = 0; a < r; ++a) {
        const double* u = U + a * m;
        double s2 = 0.0, mu = 0.0, mv = 0.0;
        for (int64_t i = 0; i < m; ++i) {
            s2 += u[i] * u[i];
            mu = std::max(mu, std::fabs(u[i]));
        }
        for (int64_t c = 0; c < k; ++c) mv = std::max(mv, std::fabs(V[a + c * ldv]));
        CompRound& o = out[a];
        o.sig = std::sqrt(s2);
        o.eu32 = o.ev32 = o.eu16 = o.ev16 = o.eu8 = o.ev8 = 0.0;
        if (!(o.sig > 0.0)) continue;  // a zero component stays zero in any pool
        // FP16: the unit vectors scaled so their largest entry lies in [0.5, 1);
        // FP8: in [128, 256) (choose_tails); errors measured in the unit
        // vectors' scale
        int eu, ev;
        std::frexp(mu / o.sig, &eu);
        std::frexp(mv, &ev);
        const double fu = std::ldexp(1.0 / o.sig, -eu), fv = std::ldexp(1.0, -ev);
        const double fu8 = std::ldexp(fu, 8), fv8 = std::ldexp(fv, 8);
        double du32 = 0.0, du16 = 0.0, dv32 = 0.0, dv16 = 0.0, du8 = 0.0, dv8 = 0.0;
        for (int64_t i = 0; i < m; ++i) {
            const double d32 = (double)(float)u[i] - u[i];
            const double x = u[i] * fu, d16 = fp16_value(x) - x;
            const double y = u[i] * fu8, d8 = from_fp8(to_fp8(y)) - y;
            du32 += d32 * d32;
            du16 += d16 * d16;
            du8 += d8 * d8;
        }
        for (int64_t c = 0; c < k; ++c) {
            const double v = V[a + c * ldv];
            const double d32 = (double)(float)v - v;
            const double x = v * fv, d16 = fp16_value(x) - x;
            const double y = v * fv8, d8 = from_fp8(to_fp8(y)) - y;
            dv32 += d32 * d32;
            dv16 += d16 * d16;
            dv8 += d8 * d8;
        }
        o.eu32 = std::sqrt(du32) / o.sig;
        o.ev32 = std::sqrt(dv32);
        o.eu16 = std::sqrt(du16) / (fu * o.sig);
        o.ev16 = std::sqrt(dv16) / fv;
        o.eu8 = std::sqrt(du8) / (fu8 * o.sig);
        o.ev8 = std::sqrt(dv8) / fv8;
    }
    return out;
}

// The stored factors of a part in SVD form are sum_a sigma_a (u_a + du_a)
// (v_a + dv_a)^T with u_a orthonormal columns and v_a orthonormal rows, so
// their error is at most sqrt(sum sigma_a^2 |du_a|^2) + sqrt(sum sigma_a^2
// |dv_a|^2) + sum sigma_a |du_a| |dv_a| in the Frobenius norm (the first two
// terms exactly, by the orthonormality; the last by the triangle
// inequality), with the measured rounding of each component in its pool.
// With fp8 the trailing components first move to FP8, from the smallest up,
// while that bound stays within the slack; then the components before them
// move to FP16 the same way (none when even FP32 rounding exceeds it):
// [0, r_hi) FP32, [r_hi, r_8) FP16, [r_8, r) FP8.
void tail_tiers(const std::vector<CompRound>& c, int64_t r, double slack, bool fp8, int64_t& r_hi, int64_t& r_8) {
    double su = 0.0, sv = 0.0, sx = 0.0;
    for (int64_t a = 0; a < r; ++a) {
        su += c[a].sig * c[a].sig * c[a].eu32 * c[a].eu32;
        sv += c[a].sig * c[a].sig * c[a].ev32 * c[a].ev32;
        sx += c[a].sig * c[a].eu32 * c[a].ev32;
    }
    r_hi = r_8 = r;
    if (!(std::sqrt(su) + std::sqrt(sv) + sx <= slack)) return;
    // component a moves from FP32 to the tier with errors (eu, ev) if the bound holds
    auto try_move = [&](const CompRound& q, double eu, double ev) {
        const double s2 = q.sig * q.sig;
        const double su2 = su + s2 * (eu * eu - q.eu32 * q.eu32);
        const double sv2 = sv + s2 * (ev * ev - q.ev32 * q.ev32);
        const double sx2 = sx + q.sig * (eu * ev - q.eu32 * q.ev32);
        if (!(std::sqrt(std::max(su2, 0.0)) + std::sqrt(std::max(sv2, 0.0)) + sx2 <= slack)) return false;
        su = su2;
        sv = sv2;
        sx = sx2;
        return true;
    };
    // the smallest components to FP8, then the next ones to FP16
    if (fp8)
        while (r_8 > 0 && try_move(c[r_8 - 1], c[r_8 - 1].eu8, c[r_8 - 1].ev8)) --r_8;
    r_hi = r_8;
    while (r_hi > 0 && try_move(c[r_hi - 1], c[r_hi - 1].eu16, c[r_hi - 1].ev16)) --r_hi;
}

// U (m x r), V (r x k) -> the shortest factors whose product differs from U V
// by at most `budget` in the Frobenius norm.
double recompress(int64_t m, int64_t k, int64_t& r, std::vector<double>& U, std::vector<double>& V,
                  double budget, int tail_mode) {
    if (r < 2 || budget <= 0) return 0.0;
    std::vector<double> Vt(k * r);
    for (int64_t i = 0; i < r; ++i)
        for (int64_t c = 0; c < k; ++c) Vt[c + i * k] = V[i + c * r];
    std::vector<double> Qu, Ru, Qv, Rv;
    qr_thin(m, r, U.data(), Qu, Ru);
    qr_thin(k, r, Vt.data(), Qv, Rv);
    std::vector<double> C(r * r, 0.0);  // Ru Rv^T
    for (int64_t j = 0; j < r; ++j)
        for (int64_t l = 0; l < r; ++l) {
            const double rv = Rv[j + l * r];
            if (rv == 0.0) continue;
            for (int64_t i = 0; i < r; ++i) C[i + j * r] += Ru[i + l * r] * rv;
        }
    std::vector<double> W, sig, Z;
    svd_jacobi(r, C, W, sig, Z);
    std::vector<int64_t> ord(r);
    for (int64_t i = 0; i < r; ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return sig[a] > sig[b]; });
    int64_t keep = r;
    double tail = 0.0;
    while (keep > 0) {
        const double s2 = sig[ord[keep - 1]] * sig[ord[keep - 1]];
        if (std::sqrt(tail + s2) > budget) break;
        tail += s2;
        --keep;
    }
    // components in SVD form (U = Q_u W Sigma, V = Z^T Q_v^T); with FP16 (and
    // FP8) trailing components, keeping a few more of them leaves more of the
    // budget for rounding: the count whose storage (4 bytes per FP32 entry, 2
    // per FP16 entry, 1 per FP8 entry) is smallest, by the measured rounding of
    // each component
    const int64_t kmax = tail_mode ? std::min<int64_t>(r, keep + 32) : keep;
    std::vector<double> U2(m * kmax, 0.0), V2(kmax * k, 0.0);
    for (int64_t a = 0; a < kmax; ++a) {
        const int64_t j = ord[a];
        for (int64_t l = 0; l < r; ++l) {
            const double wl = W[l + j * r] * sig[j];
            if (wl != 0.0)
                for (int64_t i = 0; i < m; ++i) U2[i + a * m] += Qu[i + l * m] * wl;
            const double zl = Z[l + j * r];
            if (zl != 0.0)
                for (int64_t c = 0; c < k; ++c) V2[a + c * kmax] += Qv[c + l * k] * zl;
        }
    }
    if (tail_mode) {
        // tail2[kk]: squared Frobenius norm of the components dropped when kk are kept
        std::vector<double> tail2(kmax + 1, 0.0);
        for (int64_t a = r - 1; a >= kmax; --a) tail2[kmax] += sig[ord[a]] * sig[ord[a]];
        for (int64_t a = kmax - 1; a >= 0; --a) tail2[a] = tail2[a + 1] + sig[ord[a]] * sig[ord[a]];
        const std::vector<CompRound> cr = component_rounding(U2.data(), V2.data(), m, k, kmax, kmax);
        auto bytes = [&](int64_t kk) {
            int64_t rh, r8;
            tail_tiers(cr, kk, budget - std::sqrt(tail2[kk]), tail_mode == 2, rh, r8);
            return tier_bytes(m, k, kk, rh, r8);
        };
        int64_t best = keep, best_b = bytes(keep);
        for (int64_t kk = keep + 1; kk <= kmax; ++kk) {
            const int64_t b = bytes(kk);
            if (b < best_b) {
                best = kk;
                best_b = b;
            }
        }
        keep = best;
        tail = tail2[keep];
        if (keep < kmax) {
            std::vector<double> V3(keep * k);
            for (int64_t c = 0; c < k; ++c)
                for (int64_t a = 0; a < keep; ++a) V3[a + c * keep] = V2[a + c * kmax];
            V2.swap(V3);
            U2.resize(m * keep);
        }
    }
    U.swap(U2);
    V.swap(V2);
    r = keep;
    return std::max(0.0, budget - std::sqrt(tail));
}

__global__ __launch_bounds__(256) void k_pack(const PackItem* __restrict__ items, double* __restrict__ dst) {
    const PackItem it = items[blockIdx.x];
    const int64_t n = (int64_t)it.rows * it.cols;
    for (int64_t e = threadIdx.x; e < n; e += 256) {
        const int64_t i = e % it.rows, j = e / it.rows;
        dst[it.dst + e] = it.src[i + j * it.ld];
    }
}

int pack_to_host(std::vector<PackItem>& items, std::vector<double>& host, hipStream_t stream) {
    int64_t total = 0;
    for (PackItem& it : items) {
        it.dst = total;
        total += (int64_t)it.rows * it.cols;
    }
    host.resize(total);
    if (items.empty() || total == 0) return QDFN_OK;
    DevBuf d, di;
    QDFN_HIP(d.alloc(sizeof(double) * total));
    QDFN_HIP(di.alloc(sizeof(PackItem) * items.size()));
    QDFN_HIP(hipMemcpyAsync(di.p, items.data(), sizeof(PackItem) * items.size(), hipMemcpyHostToDevice, stream));
    for (size_t b = 0; b < items.size(); b += 0x7fffffff) {
        const size_t nb = std::min<size_t>(items.size() - b, 0x7fffffff);
        hipLaunchKernelGGL(k_pack, dim3((unsigned)nb), dim3(256), 0, stream, di.as<PackItem>() + b, d.as<double>());
        QDFN_HIP(hipGetLastError());
    }
    QDFN_HIP(hipMemcpyAsync(host.data(), d.p, sizeof(double) * total, hipMemcpyDeviceToHost, stream));
    QDFN_HIP(hipStreamSynchronize(stream));
    return QDFN_OK;
}

// part w of b as dense from m x k packed entries
static void set_dense(HBlock& b, int w, const double* a) {
    const int64_t m = b.r1 - b.r0 + 1, k = b.s1 - b.s0 + 1;
    b.kind[w] = QDFN_HBLK_DENSE;
    b.rank[w] = 0;
    b.a[w].assign(a, a + m * k);
    b.b[w].clear();
}

// Rank cap of the cross approximation: mk/(m+k) (beyond it the dense block is
// smaller) and a ceiling (QDFN_HMAT_RANK_CAP, default 1024).  A lower ceiling
// stores large admissible blocks dense: 256 made the 20,000-triangle operator
// 811 MB instead of 632 MB (ranks up to ~1000 at ACA 1e-6 with
// DistDiamRatioCrit 1.0); 2048 and 4096 give the 1024 operator there.
double aca_share(bool fp16_tail) {
    static const double share = [] {
        const char* e = getenv("QDFN_HMAT_ACA_SHARE");
        const double v = e ? atof(e) : 0.0;
        return v >= 0.05 && v <= 0.95 ? v : 0.0;
    }();
    // FP16 tails: a smaller share leaves more of the budget to round
    // components to FP16 (20,000 triangles, crit 1.5, near-field blocks
    // compressed: 0.4 / 0.3 / 0.2 / 0.15 / 0.12 / 0.1 -> 530.3 / 515.0 / 501.7 /
    // 495.6 / 492.3 / 494.1 MB per product, 8,151 / 8,451 / 8,629 / 8,762 /
    // 8,780 / 8,552 RHS evals/s); FP32 storage gains nothing from it (more
    // parts reach the rank ceiling and are stored dense)
    return share > 0.0 ? share : (fp16_tail ? 0.15 : 0.5);
}

int64_t aca_rank_cap(int64_t m, int64_t k, bool fp16_tail) {
    static const int64_t ceiling = [] {
        const char* e = getenv("QDFN_HMAT_RANK_CAP");
        const long long v = e ? atoll(e) : 0;
        return v >= 16 && v <= 4096 ? (int64_t)v : int64_t(1024);
    }();
    // FP32 factors break even with the dense FP32 block at mk/(m+k); with an
    // FP16 tail they can be cheaper up to twice that (all-FP16 components)
    const int64_t even = m * k / (m + k);
    return std::min<int64_t>(ceiling, fp16_tail ? std::min<int64_t>(std::min(m, k), 2 * even) : even);
}

int block_norms(const std::vector<const double*>& src, const std::vector<int64_t>& ld,
                const std::vector<int32_t>& m, const std::vector<int32_t>& k, std::vector<double>& out,
                hipStream_t stream) {
    const size_t nb = src.size();
    out.assign(nb, 0.0);
    if (nb == 0) return QDFN_OK;
    std::vector<MaxJob> mj(nb);
    for (size_t q = 0; q < nb; ++q) {
        mj[q].src = src[q];
        mj[q].ld = ld[q];
        mj[q].m = m[q];
        mj[q].k = k[q];
    }
    DevBuf dmax, dmj;
    QDFN_HIP(dmax.alloc(sizeof(double) * nb));
    QDFN_HIP(dmj.alloc(sizeof(MaxJob) * nb));
    QDFN_HIP(hipMemcpy(dmj.p, mj.data(), sizeof(MaxJob) * nb, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_blocknorm, dim3((unsigned)nb), dim3(kAcaThreads), 0, stream, dmj.as<MaxJob>(),
                       dmax.as<double>());
    QDFN_HIP(hipGetLastError());
    QDFN_HIP(hipStreamSynchronize(stream));
    QDFN_HIP(hipMemcpy(out.data(), dmax.p, sizeof(double) * nb, hipMemcpyDeviceToHost));
    return QDFN_OK;
}

int compress_device_blocks(qdfn_hmat* h, const std::vector<DeviceBlock>& lr, const std::vector<double>& bnorm,
                           double rtol, double atol) {
    const double rc0 = h->build_phase_ms[3];
    const int tmode = tail_mode(h->prec);
    const bool tail = tmode > 0;
    double total = 0.0;  // full-pivot time: this scope minus the recompression inside it
    {
        ScopeMs whole(&total);
    // threshold scales on the device, indexed by block
    DevBuf dscale;
    QDFN_HIP(dscale.alloc(sizeof(double) * std::max<size_t>(1, bnorm.size())));
    if (!bnorm.empty())
        QDFN_HIP(hipMemcpy(dscale.p, bnorm.data(), sizeof(double) * bnorm.size(), hipMemcpyHostToDevice));
    // admissible blocks in batches bounded by a workspace budget
    size_t free_b = 0, total_b = 0;
    QDFN_HIP(hipMemGetInfo(&free_b, &total_b));
    const int64_t budget = std::max<int64_t>(int64_t(256) << 20, std::min<int64_t>(int64_t(16) << 30, (int64_t)free_b / 3));
    size_t pos = 0;
    while (pos < lr.size()) {
        std::vector<AcaJob> jobs;
        std::vector<int64_t> off;  // workspace offsets (doubles): R, U, V per job
        int64_t used = 0;
        size_t end = pos;
        while (end < lr.size()) {
            const HBlock& b = h->blocks[lr[end].q];
            const int64_t m = b.r1 - b.r0 + 1, k = b.s1 - b.s0 + 1;
            const int64_t rmax = aca_rank_cap(m, k, tail);
            const int64_t need = m * k + m * rmax + rmax * k;
            if (end > pos && (used + need) * 8 > budget) break;
            off.push_back(used);
            used += need;
            ++end;
        }
        DevBuf ws, djobs, drank, dconv, dres;
        QDFN_HIP(ws.alloc(sizeof(double) * used));
        for (size_t q = pos; q < end; ++q) {
            const HBlock& b = h->blocks[lr[q].q];
            const int64_t m = b.r1 - b.r0 + 1, k = b.s1 - b.s0 + 1;
            const int64_t rmax = aca_rank_cap(m, k, tail);
            AcaJob J;
            J.src = lr[q].src;
            J.scale = dscale.as<double>() + lr[q].q;
            J.ld = lr[q].ld;
            J.R = ws.as<double>() + off[q - pos];
            J.U = J.R + m * k;
            J.V = J.U + m * rmax;
            J.m = (int32_t)m;
            J.k = (int32_t)k;
            J.rmax = (int32_t)rmax;
            J.pad = 0;
            // a share of the budget for the cross approximation (aca_share,
            // or QDFN_HMAT_ACA_SHARE in [0.05, 0.95]), the rest for
            // recompression and rounding
            J.rtol = aca_share(tail) * rtol;
            J.atol = atol;
            jobs.push_back(J);
        }
        const size_t nj = jobs.size();
        QDFN_HIP(djobs.alloc(sizeof(AcaJob) * nj));
        QDFN_HIP(drank.alloc(sizeof(int32_t) * nj));
        QDFN_HIP(dconv.alloc(sizeof(int32_t) * nj));
        QDFN_HIP(dres.alloc(sizeof(double) * nj));
        QDFN_HIP(hipMemcpy(djobs.p, jobs.data(), sizeof(AcaJob) * nj, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_aca, dim3((unsigned)nj), dim3(kAcaThreads), 0, h->stream, djobs.as<AcaJob>(),
                           drank.as<int32_t>(), dconv.as<int32_t>(), dres.as<double>());
        QDFN_HIP(hipGetLastError());
        QDFN_HIP(hipStreamSynchronize(h->stream));
        std::vector<int32_t> rk(nj), cv(nj);
        std::vector<double> res(nj);
        QDFN_HIP(hipMemcpy(rk.data(), drank.p, sizeof(int32_t) * nj, hipMemcpyDeviceToHost));
        QDFN_HIP(hipMemcpy(cv.data(), dconv.p, sizeof(int32_t) * nj, hipMemcpyDeviceToHost));
        QDFN_HIP(hipMemcpy(res.data(), dres.p, sizeof(double) * nj, hipMemcpyDeviceToHost));
        // one packed copy: capped parts dense (from the source), the others U and V
        std::vector<PackItem> items;
        for (size_t q = pos; q < end; ++q) {
            const AcaJob& J = jobs[q - pos];
            const int32_t r = rk[q - pos];
            if (!cv[q - pos]) {
                items.push_back({lr[q].src, lr[q].ld, J.m, J.k, 0});
            } else if (r > 0) {
                items.push_back({J.U, J.m, J.m, r, 0});
                items.push_back({J.V, J.k, J.k, r, 0});  // V row-major r x k, read as k x r
            }
        }
        std::vector<double> hp;
        int rc = pack_to_host(items, hp, h->stream);
        if (rc) return rc;
        std::vector<size_t> to_recompress;
        size_t ii = 0;
        for (size_t q = pos; q < end; ++q) {
            HBlock& b = h->blocks[lr[q].q];
            const int w = lr[q].w;
            const AcaJob& J = jobs[q - pos];
            const int64_t m = J.m, k = J.k, r = rk[q - pos];
            if (!cv[q - pos]) {  // rank cap reached: keep the block dense
                set_dense(b, w, hp.data() + items[ii++].dst);
                continue;
            }
            b.kind[w] = QDFN_HBLK_LOWRANK;
            b.rank[w] = r;
            b.a[w].resize(m * r);
            b.b[w].resize(r * k);
            if (r > 0) {
                const double* u = hp.data() + items[ii++].dst;
                const double* vr = hp.data() + items[ii++].dst;
                std::copy(u, u + m * r, b.a[w].begin());
                for (int64_t i = 0; i < r; ++i)
                    for (int64_t c = 0; c < k; ++c) b.b[w][i + c * r] = vr[i * k + c];
                if (rtol > 0 && r > 1) to_recompress.push_back(q);
            }
        }
        // recompression on the host threads (blocks are independent)
        std::vector<double> budgets(to_recompress.size());
        for (size_t i = 0; i < to_recompress.size(); ++i) {
            const size_t q = to_recompress[i];
            budgets[i] = rtol * bnorm[lr[q].q] - res[q - pos];
        }
        recompress_parallel(h, lr, to_recompress, budgets);
        if (tail) {
            // parts whose tiered factors (FP32 and FP16 / FP8 tails) would take at
            // least the dense block's FP32 bytes are stored dense (from the
            // assembled source)
            std::vector<PackItem> dense_items;
            std::vector<size_t> dense_q;
            for (size_t q = pos; q < end; ++q) {
                HBlock& b = h->blocks[lr[q].q];
                const int w = lr[q].w;
                if (b.kind[w] != QDFN_HBLK_LOWRANK) continue;
                const int64_t m = b.r1 - b.r0 + 1, k = b.s1 - b.s0 + 1, r = b.rank[w];
                int64_t r_hi = r, r_8 = r;
                if (r >= 2 && b.slack[w] > 0.0)
                    tail_tiers(component_rounding(b.a[w].data(), b.b[w].data(), m, k, r, r), r, b.slack[w],
                               tmode == 2, r_hi, r_8);
                if (tier_bytes(m, k, r, r_hi, r_8) >= 4 * m * k) {
                    dense_items.push_back({lr[q].src, lr[q].ld, (int32_t)m, (int32_t)k, 0});
                    dense_q.push_back(q);
                }
            }
            if (!dense_items.empty()) {
                std::vector<double> hd;
                int rcd = pack_to_host(dense_items, hd, h->stream);
                if (rcd) return rcd;
                for (size_t i = 0; i < dense_q.size(); ++i) {
                    HBlock& b = h->blocks[lr[dense_q[i]].q];
                    set_dense(b, lr[dense_q[i]].w, hd.data() + dense_items[i].dst);
                    b.slack[lr[dense_q[i]].w] = 0.0;
                }
            }
        }
        pos = end;
    }
    }
    h->build_phase_ms[2] += total - (h->build_phase_ms[3] - rc0);
    return QDFN_OK;
}

void recompress_parallel(qdfn_hmat* h, const std::vector<DeviceBlock>& lr, const std::vector<size_t>& items,
                         const std::vector<double>& budgets) {
    ScopeMs timer(&h->build_phase_ms[3]);
    std::atomic<size_t> next(0), done(0);
    const bool verbose = getenv("QDFN_HMAT_VERBOSE") != nullptr;
    auto t_last = std::chrono::steady_clock::now();
    auto worker = [&](bool main_thread) {
        for (size_t i = next++; i < items.size(); i = next++) {
            HBlock& b = h->blocks[lr[items[i]].q];
            const int w = lr[items[i]].w;
            const int64_t m = b.r1 - b.r0 + 1, k = b.s1 - b.s0 + 1;
            b.slack[w] = recompress(m, k, b.rank[w], b.a[w], b.b[w], budgets[i], tail_mode(h->prec));
            ++done;
            if (main_thread && verbose &&
                std::chrono::steady_clock::now() - t_last > std::chrono::seconds(10)) {
                t_last = std::chrono::steady_clock::now();
                fprintf(stderr, "[hmat build] recompression: %zu of %zu parts\n", done.load(), items.size());
                fflush(stderr);
            }
        }
    };
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned i = 1; i < nt; ++i) pool.emplace_back(worker, false);
    worker(true);
    while (verbose && done.load() < items.size()) {  // the other threads' last parts
        std::this_thread::sleep_for(std::chrono::milliseconds(200));
        if (std::chrono::steady_clock::now() - t_last > std::chrono::seconds(10)) {
            t_last = std::chrono::steady_clock::now();
            fprintf(stderr, "[hmat build] recompression: %zu of %zu parts\n", done.load(), items.size());
            fflush(stderr);
        }
    }
    for (auto& th : pool) th.join();
}

void init_blocks(qdfn_hmat* h, int64_t nblk, const int64_t* range, const int64_t* adm) {
    h->blocks.resize(nblk);
    for (int64_t q = 0; q < nblk; ++q) {
        HBlock& b = h->blocks[q];
        b.s0 = range[4 * q] - 1;
        b.s1 = range[4 * q + 1] - 1;
        b.r0 = range[4 * q + 2] - 1;
        b.r1 = range[4 * q + 3] - 1;
        b.adm = adm[q] > 0 ? 1 : 0;
    }
}

// Compress / read out every block from full device matrices KS, KN (n x n, ld n).
// Admissible blocks are compressed; with tail storage (QDFN_PREC_MIXED / _MIXED8)
// so are the inadmissible blocks off the diagonal (near-field neighbours, kept
// low-rank only where the tiered factors take fewer bytes than the dense block,
// like any other part, and within the same bound rtol ||K_S,b||_F); diagonal
// blocks stay dense (K_S,ii and the loading form's zero diagonal).
static int blocks_from_full(qdfn_hmat* h, const double* dKS, const double* dKN, int64_t nblk, const int64_t* range,
                            const int64_t* adm, double rtol, double atol) {
    const int64_t n = h->n;
    init_blocks(h, nblk, range, adm);
    const bool near_lr = tail_mode(h->prec) > 0;
    const double* full[2] = {dKS, dKN};
    // dense blocks (one packed copy)
    std::vector<DeviceBlock> lr;  // parts to compress
    std::vector<PackItem> items;
    std::vector<std::pair<int64_t, int>> dense_parts;
    for (int64_t q = 0; q < nblk; ++q)
        for (int w = 0; w < 2; ++w) {
            HBlock& b = h->blocks[q];
            const int64_t m = b.r1 - b.r0 + 1, k = b.s1 - b.s0 + 1;
            const double* src = full[w] + b.r0 + b.s0 * n;
            const bool diag = b.s0 == b.r0 && b.s1 == b.r1;
            const bool compress = b.adm || (near_lr && !diag);
            if (!compress || aca_rank_cap(m, k) < 1 || m * k > 0x7fffffffLL) {
                items.push_back({src, n, (int32_t)m, (int32_t)k, 0});
                dense_parts.push_back({q, w});
            } else {
                lr.push_back({q, w, src, n});
            }
        }
    {
        std::vector<double> hp;
        int rc = pack_to_host(items, hp, h->stream);
        if (rc) return rc;
        for (size_t i = 0; i < items.size(); ++i)
            set_dense(h->blocks[dense_parts[i].first], dense_parts[i].second, hp.data() + items[i].dst);
    }
    // threshold scale: ||K_S,b||_F of every block (device reduction)
    std::vector<const double*> ns(nblk);
    std::vector<int64_t> nld(nblk, n);
    std::vector<int32_t> nm(nblk), nk(nblk);
    for (int64_t q = 0; q < nblk; ++q) {
        const HBlock& b = h->blocks[q];
        ns[q] = dKS + b.r0 + b.s0 * n;
        nm[q] = (int32_t)(b.r1 - b.r0 + 1);
        nk[q] = (int32_t)(b.s1 - b.s0 + 1);
    }
    std::vector<double> bnorm;
    int rc = block_norms(ns, nld, nm, nk, bnorm, h->stream);
    if (rc) return rc;
    return compress_device_blocks(h, lr, bnorm, rtol, atol);
}

}  // namespace qdfn

using namespace qdfn;

extern "C" qdfn_hmat_t* qdfn_hmat_build_rect(const double* seg, int64_t n, int64_t ld, double mu, double nu,
                                             int64_t nblk, const int64_t* range_SR, const int64_t* admissible,
                                             double aca_rtol, double aca_atol, int32_t prec) {
    if (!seg || ld < n) {
        fail(QDFN_ERR_ARG, "qdfn_hmat_build_rect: null table or ld < n");
        return nullptr;
    }
    if (check_blocks(n, nblk, range_SR, admissible)) return nullptr;
    const auto t_build = std::chrono::steady_clock::now();
    qdfn_hmat* h = new_hmat(n, prec);
    if (!h) return nullptr;
    DevBuf drcv, dks, dkn;
    const size_t b = sizeof(double) * n * n;
    int rc = QDFN_OK;
    if (drcv.alloc(sizeof(double) * ld * 8) || dks.alloc(b) || dkn.alloc(b))
        rc = fail(QDFN_ERR_NOMEM, "qdfn_hmat_build_rect: out of device memory for the full matrices");
    if (rc == QDFN_OK && hipMemcpy(drcv.p, seg, sizeof(double) * ld * 8, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(QDFN_ERR_HIP, "qdfn_hmat_build_rect: upload failed");
    if (rc == QDFN_OK && (hipMemsetAsync(dks.p, 0, b, h->stream) || hipMemsetAsync(dkn.p, 0, b, h->stream)))
        rc = fail(QDFN_ERR_HIP, "memset");
    if (rc == QDFN_OK) {
        ScopeMs timer(&h->build_phase_ms[0]);
        rc = assemble_rect_device(seg, n, ld, drcv.as<double>(), n, ld, mu, nu, dks.as<double>(), dkn.as<double>(), n,
                                  h->stream);
    }
    if (rc == QDFN_OK && hipStreamSynchronize(h->stream) != hipSuccess) rc = fail(QDFN_ERR_HIP, "assembly");
    if (rc == QDFN_OK)
        rc = blocks_from_full(h, dks.as<double>(), dkn.as<double>(), nblk, range_SR, admissible, aca_rtol, aca_atol);
    dks.release();
    dkn.release();
    h->build_entries = n * n;
    h->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_build).count();
    if (rc == QDFN_OK) rc = hmat_layout(h);
    if (rc != QDFN_OK) {
        qdfn_hmat_destroy(h);
        return nullptr;
    }
    return h;
}

extern "C" qdfn_hmat_t* qdfn_hmat_build_tri(int64_t n, const double* P1, const double* P2, const double* P3,
                                            const double* rake, const double* ctr, const double* nrm,
                                            const double* slip, double mu, double lambda, int64_t nblk,
                                            const int64_t* range_SR, const int64_t* admissible, double aca_rtol,
                                            double aca_atol, int32_t prec) {
    if (!P1 || !P2 || !P3 || !rake || !ctr || !nrm || !slip) {
        fail(QDFN_ERR_ARG, "qdfn_hmat_build_tri: null input");
        return nullptr;
    }
    if (check_blocks(n, nblk, range_SR, admissible)) return nullptr;
    for (int64_t s = 0; s < n; ++s)
        if (P1[s + 2 * n] > 0 || P2[s + 2 * n] > 0 || P3[s + 2 * n] > 0) {
            fail(QDFN_ERR_ARG, "Half-space solution: Z coordinates must be negative!");
            return nullptr;
        }
    const auto t_build = std::chrono::steady_clock::now();
    qdfn_hmat* h = new_hmat(n, prec);
    if (!h) return nullptr;
    DevBuf d1, d2, d3, dr, dc, dn, dsl, dks, dkn;
    const size_t v3 = sizeof(double) * 3 * n, b = sizeof(double) * n * n;
    int rc = QDFN_OK;
    if (d1.alloc(v3) || d2.alloc(v3) || d3.alloc(v3) || dr.alloc(sizeof(double) * n) || dc.alloc(v3) ||
        dn.alloc(v3) || dsl.alloc(v3) || dks.alloc(b) || dkn.alloc(b))
        rc = fail(QDFN_ERR_NOMEM, "qdfn_hmat_build_tri: out of device memory for the full matrices");
    const void* src[7] = {P1, P2, P3, rake, ctr, nrm, slip};
    DevBuf* dst[7] = {&d1, &d2, &d3, &dr, &dc, &dn, &dsl};
    for (int k = 0; rc == QDFN_OK && k < 7; ++k)
        if (hipMemcpy(dst[k]->p, src[k], k == 3 ? sizeof(double) * n : v3, hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(QDFN_ERR_HIP, "qdfn_hmat_build_tri: upload failed");
    if (rc == QDFN_OK) {
        ScopeMs timer(&h->build_phase_ms[0]);
        rc = assemble_tri_device(n, d1.as<double>(), d2.as<double>(), d3.as<double>(), dr.as<double>(), n,
                                 dc.as<double>(), dn.as<double>(), dsl.as<double>(), mu, lambda, dks.as<double>(),
                                 dkn.as<double>(), n, h->stream);
    }
    if (rc == QDFN_OK)
        rc = blocks_from_full(h, dks.as<double>(), dkn.as<double>(), nblk, range_SR, admissible, aca_rtol, aca_atol);
    dks.release();
    dkn.release();
    h->build_entries = n * n;
    h->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_build).count();
    if (rc == QDFN_OK) rc = hmat_layout(h);
    if (rc != QDFN_OK) {
        qdfn_hmat_destroy(h);
        return nullptr;
    }
    return h;
}

extern "C" qdfn_hmat_t* qdfn_hmat_create(int64_t n, int64_t nblk, const int64_t* range_SR, const int64_t* admissible,
                                         const int32_t* kind_S, const int64_t* rank_S, const double* const* S_a,
                                         const double* const* S_b, const int32_t* kind_N, const int64_t* rank_N,
                                         const double* const* N_a, const double* const* N_b, int32_t prec) {
    if (check_blocks(n, nblk, range_SR, admissible)) return nullptr;
    if (!kind_S || !rank_S || !S_a || !S_b || !kind_N || !rank_N || !N_a || !N_b) {
        fail(QDFN_ERR_ARG, "qdfn_hmat_create: null block arrays");
        return nullptr;
    }
    qdfn_hmat* h = new_hmat(n, prec);
    if (!h) return nullptr;
    h->blocks.resize(nblk);
    const int32_t* kinds[2] = {kind_S, kind_N};
    const int64_t* ranks[2] = {rank_S, rank_N};
    const double* const* as[2] = {S_a, N_a};
    const double* const* bs[2] = {S_b, N_b};
    for (int64_t q = 0; q < nblk; ++q) {
        HBlock& b = h->blocks[q];
        b.s0 = range_SR[4 * q] - 1;
        b.s1 = range_SR[4 * q + 1] - 1;
        b.r0 = range_SR[4 * q + 2] - 1;
        b.r1 = range_SR[4 * q + 3] - 1;
        b.adm = admissible[q] > 0 ? 1 : 0;
        const int64_t m = b.r1 - b.r0 + 1, k = b.s1 - b.s0 + 1;
        for (int w = 0; w < 2; ++w) {
            b.kind[w] = kinds[w][q];
            if (b.kind[w] == QDFN_HBLK_DENSE) {
                if (!as[w][q]) {
                    fail(QDFN_ERR_ARG, "qdfn_hmat_create: null dense block");
                    qdfn_hmat_destroy(h);
                    return nullptr;
                }
                b.rank[w] = 0;
                b.a[w].assign(as[w][q], as[w][q] + m * k);
            } else if (b.kind[w] == QDFN_HBLK_LOWRANK) {
                const int64_t r = ranks[w][q];
                if (r < 0 || (r > 0 && (!as[w][q] || !bs[w][q]))) {
                    fail(QDFN_ERR_ARG, "qdfn_hmat_create: bad low-rank block");
                    qdfn_hmat_destroy(h);
                    return nullptr;
                }
                b.rank[w] = r;
                if (r > 0) {
                    b.a[w].assign(as[w][q], as[w][q] + m * r);
                    b.b[w].assign(bs[w][q], bs[w][q] + r * k);
                }
            } else {
                fail(QDFN_ERR_ARG, "qdfn_hmat_create: unknown block kind");
                qdfn_hmat_destroy(h);
                return nullptr;
            }
        }
    }
    if (hmat_layout(h) != QDFN_OK) {
        qdfn_hmat_destroy(h);
        return nullptr;
    }
    return h;
}

extern "C" void qdfn_hmat_destroy(qdfn_hmat_t* h) {
    if (!h) return;
    hipStream_t s = h->stream;
    if (s) (void)hipStreamSynchronize(s);
    delete h;
    if (s) (void)hipStreamDestroy(s);
}

extern "C" int qdfn_hmat_info(qdfn_hmat_t* h, int64_t* n, int64_t* nblk, int64_t* device_bytes,
                              int64_t* stored_entries, int64_t* phase_bytes) {
    if (!h) return fail(QDFN_ERR_ARG, "qdfn_hmat_info: null handle");
    if (n) *n = h->n;
    if (nblk) *nblk = (int64_t)h->blocks.size();
    if (device_bytes) *device_bytes = h->phase_bytes;
    if (stored_entries) *stored_entries = h->stored_entries;
    if (phase_bytes) {
        phase_bytes[0] = h->alg_bytes1;
        phase_bytes[1] = h->alg_bytes2;
    }
    return QDFN_OK;
}

extern "C" int qdfn_hmat_bytes(qdfn_hmat_t* h, int64_t* out) {
    if (!h || !out) return fail(QDFN_ERR_ARG, "qdfn_hmat_bytes: null argument");
    const int64_t v[6] = {h->p1_entries, h->p1_other, h->p2_entries, h->p2_index, h->p2_gather, h->p2_out};
    std::memcpy(out, v, sizeof(v));
    return QDFN_OK;
}

extern "C" int qdfn_hmat_get_ranks(qdfn_hmat_t* h, int32_t* kind_S, int64_t* rank_S, int32_t* kind_N,
                                   int64_t* rank_N) {
    if (!h) return fail(QDFN_ERR_ARG, "qdfn_hmat_get_ranks: null handle");
    for (size_t q = 0; q < h->blocks.size(); ++q) {
        const HBlock& b = h->blocks[q];
        if (kind_S) kind_S[q] = b.kind[0];
        if (rank_S) rank_S[q] = b.rank[0];
        if (kind_N) kind_N[q] = b.kind[1];
        if (rank_N) rank_N[q] = b.rank[1];
    }
    return QDFN_OK;
}

// value of a 16-bit pool entry
static double from_lo(uint16_t b, bool bf16) {
    if (bf16) {
        const uint32_t w = (uint32_t)b << 16;
        float f;
        std::memcpy(&f, &w, 4);
        return f;
    }
    _Float16 f;
    std::memcpy(&f, &b, 2);
    return (double)f;
}

extern "C" int qdfn_hmat_get_block_stored(qdfn_hmat_t* h, int64_t bi, int32_t which, double* a, double* bv) {
    if (!h || bi < 0 || bi >= (int64_t)h->blocks.size() || which < 0 || which > 1)
        return fail(QDFN_ERR_ARG, "qdfn_hmat_get_block_stored: bad arguments");
    const HBlock& b = h->blocks[bi];
    const int w = which;
    const bool f64 = h->prec == QDFN_PREC_FP64, bf = h->prec == QDFN_PREC_BF16_FAR;
    auto hi = [&](double v) { return f64 ? v : (double)(float)v; };
    const int64_t m = b.r1 - b.r0 + 1, k = b.s1 - b.s0 + 1;
    if (b.kind[w] == QDFN_HBLK_DENSE) {
        // the product zeroes the shear diagonal of diagonal blocks and applies
        // K_S,ii in FP64 (kdiag), so that diagonal is not rounded
        const bool diag = w == 0 && b.s0 == b.r0 && b.s1 == b.r1;
        if (a)
            for (int64_t c = 0; c < k; ++c)
                for (int64_t i = 0; i < m; ++i)
                    a[i + c * m] = (diag && i == c) ? b.a[w][i + c * m] : hi(b.a[w][i + c * m]);
        return QDFN_OK;
    }
    const int64_t r = b.rank[w], r_hi = b.r_hi[w], r_8 = b.r_8[w];
    for (int64_t c = 0; c < r; ++c) {
        const bool lo = c >= r_hi, f8 = c >= r_8;
        const bool scaled = lo && !bf;
        const double mu = scaled ? b.fu[w][c - r_hi] : 1.0, mv = scaled ? b.fv[w][c - r_hi] : 1.0;
        const double sc = scaled ? b.fc[w][c - r_hi] : 1.0;
        // the value one stored entry v of this component stands for
        auto stored = [&](double v, double mult) {
            if (!lo) return hi(v);
            if (f8) return from_fp8(to_fp8(v * mult));
            return from_lo(bf ? to_bf16(v) : to_fp16(v * mult), bf);
        };
        if (a)
            for (int64_t i = 0; i < m; ++i) a[i + c * m] = lo ? stored(b.a[w][i + c * m], mu) * sc : hi(b.a[w][i + c * m]);
        if (bv)
            for (int64_t j = 0; j < k; ++j) bv[c + j * r] = stored(b.b[w][c + j * r], mv);
    }
    return QDFN_OK;
}

extern "C" int qdfn_hmat_get_split(qdfn_hmat_t* h, int64_t* rhi_S, int64_t* rhi_N) {
    if (!h) return fail(QDFN_ERR_ARG, "qdfn_hmat_get_split: null handle");
    for (size_t q = 0; q < h->blocks.size(); ++q) {
        if (rhi_S) rhi_S[q] = h->blocks[q].r_hi[0];
        if (rhi_N) rhi_N[q] = h->blocks[q].r_hi[1];
    }
    return QDFN_OK;
}

extern "C" int qdfn_hmat_get_split8(qdfn_hmat_t* h, int64_t* r8_S, int64_t* r8_N) {
    if (!h) return fail(QDFN_ERR_ARG, "qdfn_hmat_get_split8: null handle");
    for (size_t q = 0; q < h->blocks.size(); ++q) {
        if (r8_S) r8_S[q] = h->blocks[q].r_8[0];
        if (r8_N) r8_N[q] = h->blocks[q].r_8[1];
    }
    return QDFN_OK;
}

extern "C" int qdfn_hmat_get_block(qdfn_hmat_t* h, int64_t bi, int32_t which, double* a, double* bv) {
    if (!h || bi < 0 || bi >= (int64_t)h->blocks.size() || which < 0 || which > 1)
        return fail(QDFN_ERR_ARG, "qdfn_hmat_get_block: bad arguments");
    const HBlock& b = h->blocks[bi];
    if (a && !b.a[which].empty()) std::memcpy(a, b.a[which].data(), sizeof(double) * b.a[which].size());
    if (bv && !b.b[which].empty()) std::memcpy(bv, b.b[which].data(), sizeof(double) * b.b[which].size());
    return QDFN_OK;
}

extern "C" int qdfn_hmat_kself(qdfn_hmat_t* h, double* kself, double* diag) {
    if (!h) return fail(QDFN_ERR_ARG, "qdfn_hmat_kself: null handle");
    for (int64_t i = 0; i < h->n; ++i) {
        if (kself) kself[i] = -h->kdiag_h[i];
        if (diag) diag[i] = h->kdiag_h[i];
    }
    return QDFN_OK;
}

extern "C" int qdfn_hmat_normal_zero(qdfn_hmat_t* h) { return h ? h->normal_zero : QDFN_ERR_ARG; }

// ReduceTooStrongInteraction_Hmat (scripts/Functions_Solvers.jl:546-579): in
// every inadmissible block, an entry of the loading-form shear matrix (the
// diagonal of diagonal blocks zeroed) larger in magnitude than multiple |K_Self|
// of its receiver scales its shear and normal entries by multiple |K_Self| /
// |entry|; then the product layout is rebuilt from the changed blocks.  An
// inadmissible block stored low-rank (tail storage) is first expanded to its
// dense product U V (FP64), which the rule then scales.
extern "C" int qdfn_hmat_reduce_strong_interaction(qdfn_hmat_t* h, double multiple, int64_t* n_pairs) {
    if (!h || !(multiple > 0.0)) return fail(QDFN_ERR_ARG, "qdfn_hmat_reduce_strong_interaction: bad arguments");
    if (h->users > 0)
        return fail(QDFN_ERR_STATE, "qdfn_hmat_reduce_strong_interaction: runs exist on this operator");
    int64_t pairs = 0;
    for (HBlock& b : h->blocks) {
        if (b.adm != 0) continue;
        const int64_t m = b.r1 - b.r0 + 1, k = b.s1 - b.s0 + 1;
        for (int w = 0; w < 2; ++w) {
            if (b.kind[w] == QDFN_HBLK_DENSE) continue;
            const int64_t r = b.rank[w];
            std::vector<double> d(m * k, 0.0);
            for (int64_t c = 0; c < k; ++c)
                for (int64_t a = 0; a < r; ++a) {
                    const double v = b.b[w][a + c * r];
                    for (int64_t i = 0; i < m; ++i) d[i + c * m] += b.a[w][i + a * m] * v;
                }
            set_dense(b, w, d.data());
            b.slack[w] = 0.0;
        }
        const bool diag = b.s0 == b.r0 && b.s1 == b.r1;
        for (int64_t c = 0; c < k; ++c)
            for (int64_t i = 0; i < m; ++i) {
                if (diag && i == c) continue;  // zero in the loading form
                const double e = b.a[0][i + c * m];
                const double lim = std::fabs(-h->kdiag_h[b.r0 + i]) * multiple;
                if (lim < std::fabs(e)) {
                    const double ratio = lim / std::fabs(e);
                    b.a[0][i + c * m] = e * ratio;
                    b.a[1][i + c * m] = b.a[1][i + c * m] * ratio;
                    ++pairs;
                }
            }
    }
    if (n_pairs) *n_pairs = pairs;
    return hmat_layout(h);
}

extern "C" int qdfn_hmat_matvec(qdfn_hmat_t* h, const double* x, double* yS, double* yN, int32_t loading_form) {
    if (!h || !x) return fail(QDFN_ERR_ARG, "qdfn_hmat_matvec: null argument");
    hipStream_t s = h->stream;
    const size_t b = sizeof(double) * h->n;
    QDFN_HIP(hipMemcpyAsync(h->x.p, x, b, hipMemcpyHostToDevice, s));
    int rc = launch_hmat(h, h->x.as<double>(), CouplingGate{}, loading_form ? 1 : 0, h->yS.as<double>(),
                         h->yN.as<double>(), s, nullptr, nullptr, nullptr);
    if (rc) return rc;
    if (yS) QDFN_HIP(hipMemcpyAsync(yS, h->yS.p, b, hipMemcpyDeviceToHost, s));
    if (yN) QDFN_HIP(hipMemcpyAsync(yN, h->yN.p, b, hipMemcpyDeviceToHost, s));
    QDFN_HIP(hipStreamSynchronize(s));
    return QDFN_OK;
}

extern "C" int qdfn_hmat_time_phases(qdfn_hmat_t* h, int32_t reps, double* ms_phase1, double* ms_phase2,
                                     double* ms_total) {
    if (!h || reps <= 0) return fail(QDFN_ERR_ARG, "qdfn_hmat_time_phases: bad arguments");
    hipStream_t s = h->stream;
    QDFN_HIP(hipMemsetAsync(h->x.p, 0, sizeof(double) * h->n, s));
    hipEvent_t e[4];
    for (auto& x : e) QDFN_HIP(hipEventCreate(&x));
    double a = 0, b = 0, c = 0;
    int rc = QDFN_OK;
    for (int r = 0; r <= reps && rc == QDFN_OK; ++r) {
        rc = launch_hmat(h, h->x.as<double>(), CouplingGate{}, 1, h->yS.as<double>(), h->yN.as<double>(), s, e[0], e[1],
                         e[2]);
        if (rc) break;
        QDFN_HIP(hipEventRecord(e[3], s));
        QDFN_HIP(hipEventSynchronize(e[3]));
        if (r == 0) continue;  // warm-up
        float t01 = 0, t12 = 0, t03 = 0;
        QDFN_HIP(hipEventElapsedTime(&t01, e[0], e[1]));
        QDFN_HIP(hipEventElapsedTime(&t12, e[1], e[2]));
        QDFN_HIP(hipEventElapsedTime(&t03, e[0], e[3]));
        a += t01;
        b += t12;
        c += t03;
    }
    for (auto& x : e) (void)hipEventDestroy(x);
    if (ms_phase1) *ms_phase1 = a / reps;
    if (ms_phase2) *ms_phase2 = b / reps;
    if (ms_total) *ms_total = c / reps;
    return rc;
}

extern "C" int qdfn_hmat_time_matvec(qdfn_hmat_t* h, int32_t reps, double* avg_ms) {
    if (!h || reps <= 0 || !avg_ms) return fail(QDFN_ERR_ARG, "qdfn_hmat_time_matvec: bad arguments");
    hipStream_t s = h->stream;
    QDFN_HIP(hipMemsetAsync(h->x.p, 0, sizeof(double) * h->n, s));
    int rc = launch_hmat(h, h->x.as<double>(), CouplingGate{}, 1, h->yS.as<double>(), h->yN.as<double>(), s, nullptr,
                         nullptr, nullptr);
    if (rc) return rc;
    hipEvent_t e0, e1;
    QDFN_HIP(hipEventCreate(&e0));
    QDFN_HIP(hipEventCreate(&e1));
    QDFN_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps && rc == QDFN_OK; ++r)
        rc = launch_hmat(h, h->x.as<double>(), CouplingGate{}, 1, h->yS.as<double>(), h->yN.as<double>(), s, nullptr,
                         nullptr, nullptr);
    QDFN_HIP(hipEventRecord(e1, s));
    QDFN_HIP(hipEventSynchronize(e1));
    float ms = 0;
    QDFN_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = ms / reps;
    return rc;
}

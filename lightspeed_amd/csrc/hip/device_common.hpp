// Shared device-side definitions for the J/K kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

namespace lsjk {

constexpr int kMaxL = 2;            // per-primitive angular momentum supported by the kernels
constexpr int kMaxHerm = 35;        // Hermite functions of a d-d pair (L = 4)
constexpr int kMaxEwald = 8;        // operator terms per call
constexpr double kPi = 3.14159265358979323846;

__host__ __device__ constexpr int ncart(int L) { return (L + 1) * (L + 2) / 2; }
__host__ __device__ constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

// Cartesian component k of shell L in lexical order (xx, xy, xz, yy, yz, zz ...):
// exponent of x, y, z.
__host__ __device__ constexpr int cart_lx(int L, int k)
{
    int idx = 0;
    for (int i = 0; i <= L; i++)
        for (int j = 0; j <= i; j++, idx++)
            if (idx == k) return L - i;
    return 0;
}
__host__ __device__ constexpr int cart_ly(int L, int k)
{
    int idx = 0;
    for (int i = 0; i <= L; i++)
        for (int j = 0; j <= i; j++, idx++)
            if (idx == k) return i - j;
    return 0;
}
__host__ __device__ constexpr int cart_lz(int L, int k)
{
    int idx = 0;
    for (int i = 0; i <= L; i++)
        for (int j = 0; j <= i; j++, idx++)
            if (idx == k) return j;
    return 0;
}

// Exponent of axis k (0 = x, 1 = y, 2 = z) of Cartesian component idx of shell L.
__host__ __device__ constexpr int cart_lk(int L, int idx, int k)
{
    return k == 0 ? cart_lx(L, idx) : (k == 1 ? cart_ly(L, idx) : cart_lz(L, idx));
}

// Hermite index t of total order <= L in the nested (L, lexical) order: components.
__host__ __device__ constexpr int herm_tx(int t)
{
    int idx = 0;
    for (int L = 0; L <= 12; L++)
        for (int i = 0; i <= L; i++)
            for (int j = 0; j <= i; j++, idx++)
                if (idx == t) return L - i;
    return 0;
}
__host__ __device__ constexpr int herm_ty(int t)
{
    int idx = 0;
    for (int L = 0; L <= 12; L++)
        for (int i = 0; i <= L; i++)
            for (int j = 0; j <= i; j++, idx++)
                if (idx == t) return i - j;
    return 0;
}
__host__ __device__ constexpr int herm_tz(int t)
{
    int idx = 0;
    for (int L = 0; L <= 12; L++)
        for (int i = 0; i <= L; i++)
            for (int j = 0; j <= i; j++, idx++)
                if (idx == t) return j;
    return 0;
}
// Index of Hermite tuple (tx, ty, tz) in the same order.
__host__ __device__ constexpr int herm_index(int tx, int ty, int tz)
{
    int L = tx + ty + tz;
    int base = nherm(L - 1);
    if (L == 0) base = 0;
    // within shell L: i = L - tx, j = tz
    int i = L - tx;
    int off = i * (i + 1) / 2 + tz;
    return base + off;
}

// Compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1, fully expanded, so every
// index derived from I (Hermite / Cartesian tuples, table offsets) is a constant expression.
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Rys table in device memory: pointers set up by the library at init.
struct RysTable {
    const double* coef;      // [offset[n] + bin*n*2*NCOEF + root*2*NCOEF + {x|w}*NCOEF + k]
    const double* asym_r2;   // [n(n-1)/2 + i]
    const double* asym_w;
    int offset[10];
    int nbin[10];
    double h_inv[10];
    double Tc[10];
    // FP32 rule for the single-precision route: same bins, kRysNcoef32 coefficients per
    // polynomial, layout [offset32[n] + bin*n*2*kRysNcoef32 + root*2*kRysNcoef32 + {x|w}*kRysNcoef32 + k]
    const float* coef32;
    int offset32[10];
};

// Coefficients per polynomial of the FP32 rule: a degree-5 Chebyshev interpolant of the FP64
// polynomial on each bin reaches the FP32 rounding floor (1.2e-7 relative to F0).
constexpr int kRysNcoef32 = 6;

// Evaluate the nrys-root rule: x_i = t_i^2, w_i.  The oracle's rys_eval evaluates the same rule
// (same table, same asymptotic formula; here 1/T and T^(-1/2) come from one rsqrt).
// A non-finite T takes the asymptotic branch (no table access), and the bin is clamped, so the
// table is never read out of bounds.
template <int NRYS>
__device__ __forceinline__ void rys_roots(const RysTable& tab, double T, double* x, double* w)
{
    constexpr int NC = 12;  // LSJK_RYS_NCOEF
    if (!(T < tab.Tc[NRYS])) {
        const double s = rsqrt(T);
        const double iT = s * s;
#pragma unroll
        for (int i = 0; i < NRYS; i++) {
            x[i] = tab.asym_r2[NRYS * (NRYS - 1) / 2 + i] * iT;
            w[i] = tab.asym_w[NRYS * (NRYS - 1) / 2 + i] * s;
        }
        return;
    }
    const double sT = T * tab.h_inv[NRYS];  // 0 <= T < Tc here
    const int b = min((int)sT, tab.nbin[NRYS] - 1);
    const double u = 2.0 * (sT - b) - 1.0;
    const double* c = tab.coef + tab.offset[NRYS] + (size_t)b * NRYS * 2 * NC;
#pragma unroll
    for (int i = 0; i < NRYS; i++) {
        const double* cx = c + i * 2 * NC;
        const double* cw = cx + NC;
        double px = cx[NC - 1];
        double pw = cw[NC - 1];
#pragma unroll
        for (int k = NC - 2; k >= 0; k--) {
            px = fma(px, u, cx[k]);
            pw = fma(pw, u, cw[k]);
        }
        x[i] = px;
        w[i] = pw;
    }
}

// LDS copy of the table block of one nrys: [bin][root][x|w][12] doubles with one pad double per
// bin, so that lanes reading the same coefficient of different bins hit different LDS banks (an
// unpadded bin of 24 nrys doubles puts every bin on the same few banks).
__host__ __device__ constexpr int rys_lds_stride(int nrys) { return nrys * 2 * 12 + 1; }
__host__ __device__ constexpr size_t rys_lds_doubles(int nbin, int nrys) { return (size_t)nbin * rys_lds_stride(nrys); }

// Kernels stage the table block in LDS for nrys <= 3 (at most 27.6 KB, so several workgroups
// share a CU); for more roots they read it from global memory (rys_roots).
__host__ __device__ constexpr bool rys_staged(int nrys) { return nrys <= 3; }

// Upper limit of the dynamic LDS one workgroup of these kernels requests.
constexpr size_t kMaxLdsPerGroup = 65536;

// Roots from the LDS copy when the kernel staged one, else from the global table.
template <int NRYS>
__device__ __forceinline__ void rys_roots_any(const RysTable& tab, const double* s_coef, double T, double* x,
                                              double* w);

// Cooperative copy of the nrys block of the table into LDS, bins padded to rys_lds_stride (all
// threads of the workgroup; the caller synchronises the workgroup afterwards).
template <int NRYS>
__device__ __forceinline__ void rys_stage_lds(const RysTable& tab, double* s_coef)
{
    constexpr int B = NRYS * 2 * 12, S = rys_lds_stride(NRYS);
    const int n = tab.nbin[NRYS] * B;
    const double* src = tab.coef + tab.offset[NRYS];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int b = i / B;
        s_coef[b * S + (i - b * B)] = src[i];
    }
}

// rys_roots with the coefficients of this nrys read from LDS (staged by rys_stage_lds).
template <int NRYS>
__device__ __forceinline__ void rys_roots_lds(const RysTable& tab, const double* s_coef, double T, double* x,
                                              double* w)
{
    constexpr int NC = 12;
    if (!(T < tab.Tc[NRYS])) {
        const double s = rsqrt(T);
        const double iT = s * s;
#pragma unroll
        for (int i = 0; i < NRYS; i++) {
            x[i] = tab.asym_r2[NRYS * (NRYS - 1) / 2 + i] * iT;
            w[i] = tab.asym_w[NRYS * (NRYS - 1) / 2 + i] * s;
        }
        return;
    }
    const double sT = T * tab.h_inv[NRYS];  // 0 <= T < Tc here
    const int b = min((int)sT, tab.nbin[NRYS] - 1);
    const double u = 2.0 * (sT - b) - 1.0;
    const double* c = s_coef + (size_t)b * rys_lds_stride(NRYS);
#pragma unroll
    for (int i = 0; i < NRYS; i++) {
        const double* cx = c + i * 2 * NC;
        const double* cw = cx + NC;
        double px = cx[NC - 1];
        double pw = cw[NC - 1];
#pragma unroll
        for (int k = NC - 2; k >= 0; k--) {
            px = fma(px, u, cx[k]);
            pw = fma(pw, u, cw[k]);
        }
        x[i] = px;
        w[i] = pw;
    }
}

template <int NRYS>
__device__ __forceinline__ void rys_roots_any(const RysTable& tab, const double* s_coef, double T, double* x,
                                              double* w)
{
    if constexpr (rys_staged(NRYS))
        rys_roots_lds<NRYS>(tab, s_coef, T, x, w);
    else
        rys_roots<NRYS>(tab, T, x, w);
}

// LDS copy of the FP32 table block of one nrys: bins of nrys 2 kRysNcoef32 floats plus one pad.
__host__ __device__ constexpr int rys32_lds_stride(int nrys) { return nrys * 2 * kRysNcoef32 + 1; }

// FP32 Rys rule for the single-precision route.  T is formed in FP64 by the caller; for T >= Tc
// the asymptotic roots are formed in FP64 and rounded; otherwise the FP32 bin polynomial is read
// from `c32`: the global table (LDS = false) or an LDS copy of this nrys block with padded bins
// (LDS = true).
template <int NRYS, bool LDS>
__device__ __forceinline__ void rys_roots_f32(const RysTable& tab, const float* c32, double T, float* x, float* w)
{
    constexpr int NC = kRysNcoef32;
    if (!(T < tab.Tc[NRYS])) {
        const double s = rsqrt(T);
        const double iT = s * s;
#pragma unroll
        for (int i = 0; i < NRYS; i++) {
            x[i] = (float)(tab.asym_r2[NRYS * (NRYS - 1) / 2 + i] * iT);
            w[i] = (float)(tab.asym_w[NRYS * (NRYS - 1) / 2 + i] * s);
        }
        return;
    }
    const double sT = T * tab.h_inv[NRYS];  // 0 <= T < Tc here
    const int b = min((int)sT, tab.nbin[NRYS] - 1);
    const float u = (float)(2.0 * (sT - b) - 1.0);
    const float* c = c32 + (size_t)b * (LDS ? rys32_lds_stride(NRYS) : NRYS * 2 * NC);
#pragma unroll
    for (int i = 0; i < NRYS; i++) {
        const float* cx = c + i * 2 * NC;
        const float* cw = cx + NC;
        float px = cx[NC - 1];
        float pw = cw[NC - 1];
#pragma unroll
        for (int k = NC - 2; k >= 0; k--) {
            px = fmaf(px, u, cx[k]);
            pw = fmaf(pw, u, cw[k]);
        }
        x[i] = px;
        w[i] = pw;
    }
}

__host__ __device__ constexpr size_t rys32_lds_floats(int nbin, int nrys) { return (size_t)nbin * rys32_lds_stride(nrys); }

// Kernels of the FP32 route stage the FP32 block in LDS for nrys <= 5 (at most 14.2 KB).
__host__ __device__ constexpr bool rys32_staged(int nrys) { return nrys <= 5; }

// Cooperative copy of the FP32 nrys block into LDS, bins padded to rys32_lds_stride.
template <int NRYS>
__device__ __forceinline__ void rys32_stage_lds(const RysTable& tab, float* s_coef)
{
    constexpr int B = NRYS * 2 * kRysNcoef32, S = rys32_lds_stride(NRYS);
    const int n = tab.nbin[NRYS] * B;
    const float* src = tab.coef32 + tab.offset32[NRYS];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int b = i / B;
        s_coef[b * S + (i - b * B)] = src[i];
    }
}

// 1/sqrt(x) for finite x > 0 (sums of Gaussian exponents): the hardware estimate y refined by one
// second-order step, 1/sqrt(x) = y (1 - e)^(-1/2) ~ y (1 + e/2 + 3 e^2 / 8) with e = 1 - x y^2.
// These are the operations the device library's rsqrt performs, without its selection of
// special values (zero, infinity, denormals), so for such x the two agree bit for bit.
__device__ __forceinline__ double rsqrt_pos(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

// Load through the constant address space: a wave-uniform address becomes a scalar (SMEM) load,
// a divergent one an ordinary vector load.  Only for data no kernel of the launch writes.
template <typename T>
__device__ __forceinline__ T ld_const(const T* p)
{
    return *(const __attribute__((address_space(4))) T*)(p);
}

// Next work item of a dynamic queue for the calling wave (lane 0 takes it; all lanes get it).
__device__ __forceinline__ unsigned int wave_take(unsigned int* ctr)
{
    unsigned int t = 0;
    if ((threadIdx.x & 63) == 0) t = atomicAdd(ctr, 1u);
    return __builtin_amdgcn_readfirstlane(__shfl(t, 0));
}

// v + (v moved across lanes by the DPP control CTRL in the rows of ROW_MASK; 0 elsewhere), the
// double moved as its two 32-bit halves
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_add(double v)
{
    const long long b = __double_as_longlong(v);
    int lo, hi;
    if constexpr (ROW_MASK == 0xf) {  // every lane written from a valid source: no initial value
        lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, ROW_MASK, 0xf, false);
        hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
    } else {
        lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xf, false);
        hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
    }
    return v + __longlong_as_double((long long)(((unsigned long long)(unsigned int)hi << 32) | (unsigned int)lo));
}

// Sum of v over the 64 lanes of a wave with every lane active, complete in lane 63: DPP moves
// within 16-lane rows (swap neighbours, swap pairs, half-row mirror, row mirror: each lane then
// holds its row's sum), then row broadcasts (lane 15 into rows 1 and 3, lane 31 into rows 2 and
// 3).  No LDS traffic.
__device__ __forceinline__ double wave_sum_last(double v)
{
    v = dpp_add<0xB1, 0xf>(v);   // quad_perm [1,0,3,2]
    v = dpp_add<0x4E, 0xf>(v);   // quad_perm [2,3,0,1]
    v = dpp_add<0x141, 0xf>(v);  // row_half_mirror
    v = dpp_add<0x140, 0xf>(v);  // row_mirror
    v = dpp_add<0x142, 0xa>(v);  // row_bcast:15
    v = dpp_add<0x143, 0xc>(v);  // row_bcast:31
    return v;
}

// McMurchie-Davidson Hermite expansion coefficients for one axis:
// E[t][i][j] for i <= LA, j <= LB, t <= i + j (Helgaker, Jorgensen & Olsen eqs. 9.5.6-9.5.7):
//   E^{i+1,j}_t = E^{ij}_{t-1}/(2p) + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1}
//   E^{i,j+1}_t = E^{ij}_{t-1}/(2p) + X_PB E^{ij}_t + (t+1) E^{ij}_{t+1}
template <int LA, int LB>
struct HermE {
    double v[LA + LB + 1][LA + 1][LB + 1];
};

template <int LA, int LB>
__device__ __forceinline__ void hermite_E(double oo2p, double XPA, double XPB, HermE<LA, LB>& E)
{
#pragma unroll
    for (int t = 0; t <= LA + LB; t++)
#pragma unroll
        for (int i = 0; i <= LA; i++)
#pragma unroll
            for (int j = 0; j <= LB; j++) E.v[t][i][j] = 0.0;
    E.v[0][0][0] = 1.0;
#pragma unroll
    for (int i = 0; i < LA; i++) {
#pragma unroll
        for (int t = 0; t <= i + 1; t++) {
            double a = (t > 0) ? E.v[t - 1][i][0] * oo2p : 0.0;
            double b = (t <= i) ? XPA * E.v[t][i][0] : 0.0;
            double c = (t + 1 <= i) ? (t + 1) * E.v[t + 1][i][0] : 0.0;
            E.v[t][i + 1][0] = a + b + c;
        }
    }
#pragma unroll
    for (int i = 0; i <= LA; i++) {
#pragma unroll
        for (int j = 0; j < LB; j++) {
#pragma unroll
            for (int t = 0; t <= i + j + 1; t++) {
                double a = (t > 0) ? E.v[t - 1][i][j] * oo2p : 0.0;
                double b = (t <= i + j) ? XPB * E.v[t][i][j] : 0.0;
                double c = (t + 1 <= i + j) ? (t + 1) * E.v[t + 1][i][j] : 0.0;
                E.v[t][i][j + 1] = a + b + c;
            }
        }
    }
}

// Double-precision atomic add that maps to the hardware global FP64 atomic.
__device__ __forceinline__ void atomic_add(double* p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace lsjk

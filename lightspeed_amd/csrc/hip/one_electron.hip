// One-electron integrals on gfx950: overlap S, kinetic T and the point-charge potential V of a
// pair list, in the Cartesian representation, accumulated into the outputs.
//
// Behaviour follows IntBox::overlapCPU / kineticCPU / potentialCPU (reference
// src/cpp/intbox/cpu/overlap.cpp:14-147, kinetic.cpp:14-175, potential.cpp:14-147):
//   * S_ab = K_ab prod_axis s(i, j), s(i, j) = E^{ij}_0 sqrt(pi/p) (the reference integrates the
//     same 1-D overlaps by Gauss-Hermite quadrature);
//   * T_ab = K_ab / 2 sum_axis t(i, j) s s, t(i, j) = <d x^i | d x^j> with d x^l = l x^(l-1) - 2 e x^(l+1);
//   * V_ab = sum_A sum_terms alpha (-c_A) (2 pi / p) [sqrt(d^2)] sum_t E_t R_t(p d^2, P - A): Hermite
//     accumulators per pair, then the bra Hermite -> Cartesian transform shared with J.
// A symmetric list adds the transposed block for pairs of distinct primitives.
//
// Kernels: one_e_st<LA,LB> (thread per pair, list order), one_e_v<LA,LB> (thread per pair,
// writes the pair's Hermite accumulators); j_bra_scatter then transforms V.
#include "internal.hpp"
#include "kernels.hpp"

namespace lsjk {

template <int LA, int LB>
__global__ __launch_bounds__(256) void one_e_st(OneEArgs a, int k0, int nk)
{
    constexpr int NA = ncart(LA), NB = ncart(LB);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nk) return;
    const int k = k0 + i;
    const int ia = a.p1[k], ib = a.p2[k];
    const double eA = a.eA[ia], eB = a.eB[ib];
    const double p = eA + eB, ip = 1.0 / p;
    const double Ax = a.xA[ia], Ay = a.yA[ia], Az = a.zA[ia];
    const double Bx = a.xB[ib], By = a.yB[ib], Bz = a.zB[ib];
    const double Px = (eA * Ax + eB * Bx) * ip, Py = (eA * Ay + eB * By) * ip, Pz = (eA * Az + eB * Bz) * ip;
    // E^{ij}_0 for i <= LA + 1, j <= LB + 1 (the kinetic derivatives raise each side by one)
    HermE<LA + 1, LB + 1> Ex, Ey, Ez;
    hermite_E<LA + 1, LB + 1>(0.5 * ip, Px - Ax, Px - Bx, Ex);
    hermite_E<LA + 1, LB + 1>(0.5 * ip, Py - Ay, Py - By, Ey);
    hermite_E<LA + 1, LB + 1>(0.5 * ip, Pz - Az, Pz - Bz, Ez);
    const double sp = sqrt(kPi * ip);
    const double Kab = a.K[k];
    const bool mirror = a.symmetric && ia != ib;
    const int ca = a.cartA[ia], cb = a.cartB[ib];
    auto s1 = [&](const HermE<LA + 1, LB + 1>& E, int i1, int j1) __attribute__((always_inline)) {
        return E.v[0][i1][j1] * sp;
    };
    auto t1 = [&](const HermE<LA + 1, LB + 1>& E, int i1, int j1) __attribute__((always_inline)) {
        double v = 4.0 * eA * eB * s1(E, i1 + 1, j1 + 1);
        if (i1 > 0) v -= 2.0 * eB * i1 * s1(E, i1 - 1, j1 + 1);
        if (j1 > 0) v -= 2.0 * eA * j1 * s1(E, i1 + 1, j1 - 1);
        if (i1 > 0 && j1 > 0) v += (double)(i1 * j1) * s1(E, i1 - 1, j1 - 1);
        return v;
    };
    static_for<NA>([&](auto P_) {
        constexpr int pa = decltype(P_)::value;
        constexpr int ax = cart_lx(LA, pa), ay = cart_ly(LA, pa), az = cart_lz(LA, pa);
        static_for<NB>([&](auto Q_) {
            constexpr int qb = decltype(Q_)::value;
            constexpr int bx = cart_lx(LB, qb), by = cart_ly(LB, qb), bz = cart_lz(LB, qb);
            const double sx = s1(Ex, ax, bx), sy = s1(Ey, ay, by), sz = s1(Ez, az, bz);
            const size_t ij = (size_t)(ca + pa) * a.ld + (cb + qb);
            const size_t ji = (size_t)(cb + qb) * a.ld + (ca + pa);
            if (a.S) {
                const double v = Kab * sx * sy * sz;
                atomic_add(a.S + ij, v);
                if (mirror) atomic_add(a.S + ji, v);
            }
            if (a.T) {
                const double v = 0.5 * Kab * (t1(Ex, ax, bx) * sy * sz + sx * t1(Ey, ay, by) * sz + sx * sy * t1(Ez, az, bz));
                atomic_add(a.T + ij, v);
                if (mirror) atomic_add(a.T + ji, v);
            }
        });
    });
}

template <int LA, int LB>
__global__ __launch_bounds__(256) void one_e_v(OneEArgs a, int k0, int nk, int hstart, size_t hoff)
{
    constexpr int L = LA + LB, H = nherm(L), NR = L / 2 + 1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nk) return;
    const int k = k0 + i;
    const int ia = a.p1[k], ib = a.p2[k];
    const double eA = a.eA[ia], eB = a.eB[ib];
    const double p = eA + eB, ip = 1.0 / p;
    const double Px = (eA * a.xA[ia] + eB * a.xB[ib]) * ip;
    const double Py = (eA * a.yA[ia] + eB * a.yB[ib]) * ip;
    const double Pz = (eA * a.zA[ia] + eB * a.zB[ib]) * ip;
    double acc[H];
#pragma unroll
    for (int t = 0; t < H; t++) acc[t] = 0.0;
    for (int it = 0; it < a.ew.n; it++) {
        const double om = a.ew.omega[it];
        double pref = a.ew.alpha[it] * 2.0 * kPi * ip, d2 = 1.0;
        if (om != -1.0) {
            d2 = om * om / (p + om * om);
            pref *= sqrt(d2);
        }
        for (int A = 0; A < a.nA; A++) {
            const double X = Px - ld_const(a.xyzc + 4 * A), Y = Py - ld_const(a.xyzc + 4 * A + 1);
            const double Z = Pz - ld_const(a.xyzc + 4 * A + 2), cA = ld_const(a.xyzc + 4 * A + 3);
            double x[NR], w[NR];
            rys_roots<NR>(a.rys, p * d2 * (X * X + Y * Y + Z * Z), x, w);
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const double U = -2.0 * p * d2 * x[r];
                double Hx[L + 1], Hy[L + 1], Hz[L + 1];
                Hx[0] = -pref * w[r] * cA;
                Hy[0] = 1.0;
                Hz[0] = 1.0;
                if constexpr (L > 0) {
                    Hx[1] = U * X * Hx[0];
                    Hy[1] = U * Y;
                    Hz[1] = U * Z;
                }
#pragma unroll
                for (int u = 1; u < L; u++) {
                    Hx[u + 1] = U * (X * Hx[u] + u * Hx[u - 1]);
                    Hy[u + 1] = U * (Y * Hy[u] + u * Hy[u - 1]);
                    Hz[u + 1] = U * (Z * Hz[u] + u * Hz[u - 1]);
                }
                static_for<H>([&](auto T_) {
                    constexpr int t = decltype(T_)::value;
                    constexpr int tx = herm_tx(t), ty = herm_ty(t), tz = herm_tz(t);
                    acc[t] += Hx[tx] * Hy[ty] * Hz[tz];
                });
            }
        }
    }
    double* dst = a.vp + hoff + (size_t)(a.hpos[k] - hstart) * H;
#pragma unroll
    for (int t = 0; t < H; t++) dst[t] = acc[t];
}

#define LSJK_ONE_E_CASES(LAUNCH) \
    LAUNCH(0, 0) LAUNCH(0, 1) LAUNCH(0, 2) LAUNCH(1, 0) LAUNCH(1, 1) LAUNCH(1, 2) LAUNCH(2, 0) LAUNCH(2, 1) LAUNCH(2, 2)

void launch_one_e_st(int LA, int LB, const OneEArgs& a, int k0, int nk, hipStream_t s)
{
    if (nk <= 0) return;
    dim3 grid((nk + 255) / 256), block(256);
#define L_ST(X, Y) \
    if (LA == X && LB == Y) { hipLaunchKernelGGL((one_e_st<X, Y>), grid, block, 0, s, a, k0, nk); return; }
    LSJK_ONE_E_CASES(L_ST)
#undef L_ST
}

void launch_one_e_v(int LA, int LB, const OneEArgs& a, int k0, int nk, int hstart, size_t hoff, hipStream_t s)
{
    if (nk <= 0) return;
    dim3 grid((nk + 255) / 256), block(256);
#define L_V(X, Y)                                                                                   \
    if (LA == X && LB == Y) {                                                                       \
        hipLaunchKernelGGL((one_e_v<X, Y>), grid, block, 0, s, a, k0, nk, hstart, hoff);            \
        return;                                                                                     \
    }
    LSJK_ONE_E_CASES(L_V)
#undef L_V
}

}  // namespace lsjk

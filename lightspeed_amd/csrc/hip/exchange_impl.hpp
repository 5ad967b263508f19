// Exchange (K) quartet kernel templates, shared by the translation units that instantiate them
// (exchange_l0.hip, exchange_l1.hip, exchange_l2.hip: one per bra row angular momentum L1, so the
// 243 instantiations compile in parallel).
// Direct exchange (K) build on gfx950.
//
// K13 += sum_{2,4} (12|34) D24 over the quartets the reference evaluates
// (IntBox::exchangeCPU, reference src/cpp/intbox/cpu/exchange.cpp:17-355):
//   * pairs (1,2) and (3,4) taken from the K-invert partner lists with b >= thresp/(maxb*D24max)
//     (KBlock::build21 / build_symmetric, src/cpp/core/kvec.cpp:128-232; exchange.cpp:54-71),
//   * quartet sieve float(b12*b34) * max|D24 block| >= thresp (exchange.cpp:145-152, 186).
// The reference loops pivot pairs (2,4) and scatters each quartet's (1,3) block with atomics.  Here
// one wave owns a tile of outputs (one row-1 primitive x 16 row-3 primitives): its lanes test the
// (2,4) partner combinations of each output pair, compact the survivors into an LDS queue, evaluate
// them 64 at a time, and add each quartet's N1 x N3 block to the tile's LDS accumulator, which is
// flushed to K with one atomic add per element.  With a symmetric D24 and one list on both sides
// only row1 <= row3 is computed and the transposed block is added as well (K = K^T), the same 2x
// the reference's D24symm half-sieve (exchange.cpp:138-139, 326-336) obtains.
//
// Integrals: Rys quadrature with the 2-D integrals built by the vertical recurrence on centres
// A and C and the horizontal transfer to B and D (Rys, Dupuis & King, J. Comput. Chem. 4, 154
// (1983)), in registers, for each root x_i = t_i^2 (attenuated as t^2 d^2 for erf operators).
#pragma once
#include "internal.hpp"
#include "kernels.hpp"

namespace lsjk {

template <typename Real, int L1, int L2, int L3, int L4>
struct Axis4 {
    Real v[L1 + 1][L2 + 1][L3 + 1][L4 + 1];
};

// 2-D Rys integrals of one axis and one root: vertical recurrence on centres A and C
// (G[n][m] = (n 0 | m 0)), then the horizontal transfers to D and B.  All loops are compile-time
// expansions, so every array index is a constant and the arrays live in registers.
template <typename Real, int L1, int L2, int L3, int L4>
__device__ __forceinline__ void rys_axis(Real g00, Real C00, Real D00, Real B10, Real B01, Real B00, Real AB, Real CD,
                                         Axis4<Real, L1, L2, L3, L4>& I)
{
    constexpr int NX = L1 + L2 + 1, MX = L3 + L4 + 1;
    Real G[NX][MX];
    G[0][0] = g00;
    if constexpr (NX > 1) G[1][0] = C00 * g00;
    static_for<(NX > 2 ? NX - 2 : 0)>([&](auto N_) {
        constexpr int n = decltype(N_)::value + 1;
        G[n + 1][0] = C00 * G[n][0] + Real(n) * B10 * G[n - 1][0];
    });
    static_for<MX - 1>([&](auto M_) {
        constexpr int m = decltype(M_)::value;
        static_for<NX>([&](auto N_) {
            constexpr int n = decltype(N_)::value;
            Real v = D00 * G[n][m];
            if constexpr (m > 0) v += Real(m) * B01 * G[n][m - 1];
            if constexpr (n > 0) v += Real(n) * B00 * G[n - 1][m];
            G[n][m + 1] = v;
        });
    });
    // ket transfer: T[n][c][d] = (n 0 | c d)
    Real T[NX][MX][L4 + 1];
    static_for<NX>([&](auto N_) {
        constexpr int n = decltype(N_)::value;
        static_for<MX>([&](auto C_) {
            constexpr int c = decltype(C_)::value;
            T[n][c][0] = G[n][c];
        });
    });
    static_for<L4>([&](auto D_) {
        constexpr int d = decltype(D_)::value;
        static_for<NX>([&](auto N_) {
            constexpr int n = decltype(N_)::value;
            static_for<MX - 1 - d>([&](auto C_) {
                constexpr int c = decltype(C_)::value;
                T[n][c][d + 1] = T[n][c + 1][d] + CD * T[n][c][d];
            });
        });
    });
    // bra transfer
    static_for<L3 + 1>([&](auto C_) {
        constexpr int c = decltype(C_)::value;
        static_for<L4 + 1>([&](auto D_) {
            constexpr int d = decltype(D_)::value;
            Real U[NX][L2 + 1];
            static_for<NX>([&](auto A_) {
                constexpr int a = decltype(A_)::value;
                U[a][0] = T[a][c][d];
            });
            static_for<L2>([&](auto B_) {
                constexpr int b = decltype(B_)::value;
                static_for<NX - 1 - b>([&](auto A_) {
                    constexpr int a = decltype(A_)::value;
                    U[a][b + 1] = U[a + 1][b] + AB * U[a][b];
                });
            });
            static_for<L1 + 1>([&](auto A_) {
                constexpr int a = decltype(A_)::value;
                static_for<L2 + 1>([&](auto B_) {
                    constexpr int b = decltype(B_)::value;
                    I.v[a][b][c][d] = U[a][b];
                });
            });
        });
    });
}

// A partner entry in the arithmetic of a route: the centre P stays FP64 on both routes (P - Q is
// formed in FP64); the other fields are Real.
template <typename Real>
struct KEnt {
    double Px, Py, Pz;
    Real p, K, PAx, PAy, PAz, ABx, ABy, ABz, i2p;  // K: Kp of the packed entry
    int cart, blk;
};

// One packed FP64 entry: a 128-byte, 16-byte aligned record (p, K, P, P - A, A - B, {cart, blk}, 1/(2p),
// K view {cart, blk}); KV: take the K view ids.
template <bool KV = false>
__device__ __forceinline__ KEnt<double> load_ent(const double* __restrict__ eg, int i)
{
    const double* g = static_cast<const double*>(__builtin_assume_aligned(eg + (size_t)i * kKEntDoubles, 16));
    KEnt<double> e;
    e.p = g[0];
    e.K = g[1];
    e.Px = g[2];
    e.Py = g[3];
    e.Pz = g[4];
    e.PAx = g[5];
    e.PAy = g[6];
    e.PAz = g[7];
    e.ABx = g[8];
    e.ABy = g[9];
    e.ABz = g[10];
    const int* ids = reinterpret_cast<const int*>(g + (KV ? 13 : 11));
    e.cart = ids[0];
    e.blk = ids[1];
    e.i2p = g[12];
    return e;
}

// One packed FP32-route entry: an 80-byte record (P as 3 doubles; p, Kp, P - A, A - B, 1/(2p) as
// floats; {cart, blk}; K view {cart, blk}); KV: take the K view ids.
template <bool KV = false>
__device__ __forceinline__ KEnt<float> load_ent32(const double* __restrict__ eg32, int i)
{
    const double* g = static_cast<const double*>(__builtin_assume_aligned(eg32 + (size_t)i * kKEnt32Doubles, 16));
    KEnt<float> e;
    e.Px = g[0];
    e.Py = g[1];
    e.Pz = g[2];
    const float* f = reinterpret_cast<const float*>(g + 3);
    e.p = f[0];
    e.K = f[1];
    e.PAx = f[2];
    e.PAy = f[3];
    e.PAz = f[4];
    e.ABx = f[5];
    e.ABy = f[6];
    e.ABz = f[7];
    e.i2p = f[8];
    const int* ids = reinterpret_cast<const int*>(g + (KV ? 9 : 8));
    e.cart = ids[0];
    e.blk = ids[1];
    return e;
}

template <typename Real>
__device__ __forceinline__ Real kdiv(Real a, Real b)
{
    if constexpr (std::is_same<Real, float>::value)
        return __fdividef(a, b);
    else
        return a / b;
}

template <typename Real>
__device__ __forceinline__ Real krsqrt(Real x)
{
    if constexpr (std::is_same<Real, float>::value)
        return __frsqrt_rn(x);
    else
        return rsqrt_pos(x);
}

// One primitive quartet: acc[a][c] += sum_{b,d} (ab|cd) Db[b][d], evaluated in the route's
// arithmetic (Real) and added to the FP64 accumulator.  STAGED: the table block of this nrys was
// copied to LDS at s_tab (FP64 block for Real = double, FP32 block for Real = float); otherwise
// the global table is read.  KETD (L3 = 0 only): e34's P - A holds Q - D, and the ket's 2-D
// integrals (ab|0d) are recurred at D directly, as (ab|d0) with D in the place of C (the same
// integrals; no transfer from C, so C - D is not needed).
template <typename Real, int L1, int L2, int L3, int L4, bool STAGED, bool KETD = false>
__device__ __forceinline__ void k_quartet(const RysTable& tab, const void* s_tab, const Ewald& ew,
                                          const KEnt<Real>& e12, const KEnt<Real>& e34,
                                          const Real (&Db)[ncart(L2)][ncart(L4)], double (&acc)[ncart(L1)][ncart(L3)])
{
    constexpr int N1 = ncart(L1), N2 = ncart(L2), N3 = ncart(L3), N4 = ncart(L4);
    constexpr int NR = (L1 + L2 + L3 + L4) / 2 + 1;
    const Real p = e12.p, q = e34.p;
    const Real pq = p + q;
    const Real rpq = krsqrt(pq);
    const Real ipq = rpq * rpq;
    const Real rho = p * q * ipq;
    const double PQxd = e12.Px - e34.Px, PQyd = e12.Py - e34.Py, PQzd = e12.Pz - e34.Pz;
    const double r2 = PQxd * PQxd + PQyd * PQyd + PQzd * PQzd;
    const Real PQx = (Real)PQxd, PQy = (Real)PQyd, PQz = (Real)PQzd;
    // 2 pi^(5/2) K_AB K_CD / (p q sqrt(p + q))
    const Real base = e12.K * e34.K * rpq;
    const Real PAx = e12.PAx, PAy = e12.PAy, PAz = e12.PAz;
    const Real QCx = e34.PAx, QCy = e34.PAy, QCz = e34.PAz;
    const Real ABx = e12.ABx, ABy = e12.ABy, ABz = e12.ABz;
    const Real CDx = e34.ABx, CDy = e34.ABy, CDz = e34.ABz;
    const Real qf = q * ipq, pf = p * ipq;  // rho/p, rho/q
    const Real i2p = e12.i2p, i2q = e34.i2p;
    // FP32 quartets: the roots' sum is formed in FP32 and added to the FP64 accumulator once
    constexpr bool FP32 = std::is_same<Real, float>::value;
    Real qacc[N1][N3];
    if constexpr (FP32) {
#pragma unroll
        for (int a = 0; a < N1; a++)
#pragma unroll
            for (int c = 0; c < N3; c++) qacc[a][c] = Real(0);
    }
    // one operator term: Rys roots at T = rho d2 |PQ|^2, assembly with prefactor pref
    auto term = [&](Real pref, Real d2) __attribute__((always_inline)) {
        Real x[NR], w[NR];
        if constexpr (std::is_same<Real, float>::value)
            rys_roots_f32<NR, STAGED>(tab, STAGED ? static_cast<const float*>(s_tab) : tab.coef32 + tab.offset32[NR],
                                      (double)(rho * d2) * r2, x, w);
        else if constexpr (STAGED)
            rys_roots_lds<NR>(tab, static_cast<const double*>(s_tab), rho * d2 * r2, x, w);
        else
            rys_roots<NR>(tab, rho * d2 * r2, x, w);
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const Real u = x[i] * d2;
            const Real B00 = Real(0.5) * u * ipq;
            const Real B10 = i2p * (Real(1) - u * qf);
            const Real B01 = i2q * (Real(1) - u * pf);
            const Real cq = u * qf, cp = u * pf;
            static_assert(!KETD || L3 == 0, "KETD: no angular momentum on C");
            using AX = std::conditional_t<KETD, Axis4<Real, L1, L2, L4, 0>, Axis4<Real, L1, L2, L3, L4>>;
            AX Ix, Iy, Iz;
            if constexpr (KETD) {
                rys_axis<Real, L1, L2, L4, 0>(pref * w[i], PAx - cq * PQx, QCx + cp * PQx, B10, B01, B00, ABx, CDx, Ix);
                rys_axis<Real, L1, L2, L4, 0>(Real(1), PAy - cq * PQy, QCy + cp * PQy, B10, B01, B00, ABy, CDy, Iy);
                rys_axis<Real, L1, L2, L4, 0>(Real(1), PAz - cq * PQz, QCz + cp * PQz, B10, B01, B00, ABz, CDz, Iz);
            } else {
                rys_axis<Real, L1, L2, L3, L4>(pref * w[i], PAx - cq * PQx, QCx + cp * PQx, B10, B01, B00, ABx, CDx, Ix);
                rys_axis<Real, L1, L2, L3, L4>(Real(1), PAy - cq * PQy, QCy + cp * PQy, B10, B01, B00, ABy, CDy, Iy);
                rys_axis<Real, L1, L2, L3, L4>(Real(1), PAz - cq * PQz, QCz + cp * PQz, B10, B01, B00, ABz, CDz, Iz);
            }
            // (a b | c d) of one axis from either layout (indices are compile-time constants)
            auto at = [](const AX& I, int a_, int b_, int c_, int d_) __attribute__((always_inline)) {
                if constexpr (KETD)
                    return I.v[a_][b_][d_][0];
                else
                    return I.v[a_][b_][c_][d_];
            };
            static_for<N1>([&](auto A_) {
                constexpr int a = decltype(A_)::value;
                constexpr int ax = cart_lx(L1, a), ay = cart_ly(L1, a), az = cart_lz(L1, a);
                static_for<N3>([&](auto C_) {
                    constexpr int c = decltype(C_)::value;
                    constexpr int cx = cart_lx(L3, c), cy = cart_ly(L3, c), cz = cart_lz(L3, c);
                    Real s = Real(0);
                    static_for<N2>([&](auto B_) {
                        constexpr int b = decltype(B_)::value;
                        constexpr int bx = cart_lx(L2, b), by = cart_ly(L2, b), bz = cart_lz(L2, b);
                        static_for<N4>([&](auto D_) {
                            constexpr int d = decltype(D_)::value;
                            constexpr int dx = cart_lx(L4, d), dy = cart_ly(L4, d), dz = cart_lz(L4, d);
                            s = fma(at(Ix, ax, bx, cx, dx) * at(Iy, ay, by, cy, dy) * at(Iz, az, bz, cz, dz), Db[b][d], s);
                        });
                    });
                    if constexpr (FP32)
                        qacc[a][c] += s;
                    else
                        acc[a][c] += s;
                });
            });
        }
    };
    // one bare 1/r term (the common case): d2 = 1 and no term loop.  Taken for the light classes
    // only, where the loop's overhead is a visible share of a quartet; for the others the second
    // copy of the quartet code measured slower.
    constexpr bool kPlainBranch = L1 + L2 + L3 + L4 <= 2;
    if (kPlainBranch && ew.plain) {
        term((Real)ew.alpha[0] * base, Real(1));
    } else {
        for (int it = 0; it < ew.n; it++) {
            const double om = ew.omega[it];
            Real d2 = Real(1);
            Real pref = (Real)ew.alpha[it] * base;
            if (__double_as_longlong(om) != __double_as_longlong(-1.0)) {  // erf-attenuated term (bit test: scalar)
                const Real o = (Real)om;
                d2 = kdiv(o * o, rho + o * o);
                pref *= sqrt(d2);
            }
            term(pref, d2);
        }
    }
    if constexpr (FP32) {
#pragma unroll
        for (int a = 0; a < N1; a++)
#pragma unroll
            for (int c = 0; c < N3; c++) acc[a][c] += (double)qacc[a][c];
    }
}

// Derivatives of one primitive quartet with respect to the centres A, B, C (D by translational
// invariance, dE/dD = -(dE/dA + dE/dB + dE/dC)), contracted with D13[a][c] D24[b][d]:
//   d/dA_x (ab|cd) = 2 alpha (a+1x b|cd) - a_x (a-1x b|cd), likewise B (beta), C (gamma),
// from 2-D Rys integrals with L1, L2, L3 raised by one (reference exchange_grad.cpp).
template <int L1, int L2, int L3, int L4>
__device__ __forceinline__ void k_quartet_grad(const RysTable& tab, const Ewald& ew, const KEnt<double>& e12,
                                               const KEnt<double>& e34, double alpha, double beta, double gamma,
                                               const double (&D13)[ncart(L1)][ncart(L3)],
                                               const double (&D24)[ncart(L2)][ncart(L4)], double (&gA)[3],
                                               double (&gB)[3], double (&gC)[3])
{
    constexpr int N1 = ncart(L1), N2 = ncart(L2), N3 = ncart(L3), N4 = ncart(L4);
    constexpr int NR = (L1 + L2 + L3 + L4 + 1) / 2 + 1;
    const double p = e12.p, q = e34.p;
    const double pq = p + q;
    const double rpq = rsqrt_pos(pq);
    const double ipq = rpq * rpq;
    const double rho = p * q * ipq;
    const double PQx = e12.Px - e34.Px, PQy = e12.Py - e34.Py, PQz = e12.Pz - e34.Pz;
    const double r2 = PQx * PQx + PQy * PQy + PQz * PQz;
    const double base = e12.K * e34.K * rpq;  // 2 pi^(5/2) K_AB K_CD / (p q sqrt(p + q))
    const double qf = q * ipq, pf = p * ipq;
    const double i2p = e12.i2p, i2q = e34.i2p;
    for (int it = 0; it < ew.n; it++) {
        const double om = ew.omega[it];
        double d2 = 1.0;
        double pref = ew.alpha[it] * base;
        if (__double_as_longlong(om) != __double_as_longlong(-1.0)) {  // erf-attenuated term (bit test: scalar)
            d2 = om * om / (rho + om * om);
            pref *= sqrt(d2);
        }
        double x[NR], w[NR];
        rys_roots<NR>(tab, rho * d2 * r2, x, w);
        // roots one after another (not unrolled): keeps the derivative kernels' code size and
        // compile time bounded
#pragma unroll 1
        for (int i = 0; i < NR; i++) {
            const double u = x[i] * d2;
            const double B00 = 0.5 * u * ipq;
            const double B10 = i2p * (1.0 - u * qf);
            const double B01 = i2q * (1.0 - u * pf);
            const double cq = u * qf, cp = u * pf;
            Axis4<double, L1 + 1, L2 + 1, L3 + 1, L4> I[3];
            rys_axis<double, L1 + 1, L2 + 1, L3 + 1, L4>(pref * w[i], e12.PAx - cq * PQx, e34.PAx + cp * PQx, B10, B01,
                                                         B00, e12.ABx, e34.ABx, I[0]);
            rys_axis<double, L1 + 1, L2 + 1, L3 + 1, L4>(1.0, e12.PAy - cq * PQy, e34.PAy + cp * PQy, B10, B01, B00,
                                                         e12.ABy, e34.ABy, I[1]);
            rys_axis<double, L1 + 1, L2 + 1, L3 + 1, L4>(1.0, e12.PAz - cq * PQz, e34.PAz + cp * PQz, B10, B01, B00,
                                                         e12.ABz, e34.ABz, I[2]);
            // (a, c) run as ordinary loops (the 2-D integrals are then indexed at run time), the
            // partner components (b, d) and the three axes are expanded at compile time
#pragma unroll 1
            for (int a = 0; a < N1; a++) {
                int an[3];
                for (int k = 0; k < 3; k++) an[k] = cart_lk(L1, a, k);
#pragma unroll 1
                for (int c = 0; c < N3; c++) {
                    int cn[3];
                    for (int k = 0; k < 3; k++) cn[k] = cart_lk(L3, c, k);
                    const double d13 = D13[a][c];
                    if (d13 == 0.0) continue;
                    static_for<N2>([&](auto B_) {
                        constexpr int b = decltype(B_)::value;
                        static_for<N4>([&](auto D_) {
                            constexpr int d = decltype(D_)::value;
                            const double wgt = d13 * D24[b][d];
                            double base3[3], dA[3], dB[3], dC[3];
                            static_for<3>([&](auto K_) {
                                constexpr int k = decltype(K_)::value;
                                constexpr int ib = cart_lk(L2, b, k), id = cart_lk(L4, d, k);
                                const int ia = an[k], ic = cn[k];
                                const auto& V = I[k].v;
                                base3[k] = V[ia][ib][ic][id];
                                double va = 2.0 * alpha * V[ia + 1][ib][ic][id];
                                if (ia > 0) va -= ia * V[ia - 1][ib][ic][id];
                                double vb = 2.0 * beta * V[ia][ib + 1][ic][id];
                                if constexpr (ib > 0) vb -= ib * V[ia][ib - 1][ic][id];
                                double vc = 2.0 * gamma * V[ia][ib][ic + 1][id];
                                if (ic > 0) vc -= ic * V[ia][ib][ic - 1][id];
                                dA[k] = va;
                                dB[k] = vb;
                                dC[k] = vc;
                            });
                            gA[0] += wgt * dA[0] * base3[1] * base3[2];
                            gA[1] += wgt * base3[0] * dA[1] * base3[2];
                            gA[2] += wgt * base3[0] * base3[1] * dA[2];
                            gB[0] += wgt * dB[0] * base3[1] * base3[2];
                            gB[1] += wgt * base3[0] * dB[1] * base3[2];
                            gB[2] += wgt * base3[0] * base3[1] * dB[2];
                            gC[0] += wgt * dC[0] * base3[1] * base3[2];
                            gC[1] += wgt * base3[0] * dC[1] * base3[2];
                            gC[2] += wgt * base3[0] * base3[1] * dC[2];
                        });
                    });
                }
            }
        }
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

constexpr int kKQueue = 128;

// Exchange gradient.  One wave per tile = (row1, block of kKTile row3s) of a class (L1 L2 | L3 L4).
// For each (row1, row3) of the tile the lanes test the sieve over the candidate rectangle
// (partners that can pass with the global D maximum), compact the passing (2,4) combinations into
// an LDS queue shared by all pairs of the tile, and evaluate the quartets 64 at a time; each
// quartet's derivative contributions go to the four atoms' gradients with atomics.
template <int L1, int L2, int L3, int L4>
__global__ __launch_bounds__(64) void k_tiles_grad(KArgs a, const KUnit* __restrict__ units)
{
    __shared__ int s_e2[kKQueue], s_e4[kKQueue], s_t[kKQueue];
    const int lane = threadIdx.x;
    const KUnit unit = units[blockIdx.x];
    const int r1 = a.s12.rows[unit.i1];
    const int j3a = unit.ib * kKTile;
    const int j3b = min(j3a + kKTile, a.s34.nrows);

    const float thresp = a.thresp;
    const double dmax = __longlong_as_double((long long)a.sc->dmax_bits);
    const float thre12 = (float)((double)thresp / ((double)a.maxb34 * dmax));
    const float thre34 = (float)((double)thresp / ((double)a.maxb12 * dmax));
    // this unit's chunk of row1's partners (bound order, so the surviving ones are a prefix)
    const int a0 = a.s12.row_start[r1 * (a.s12.maxLp + 1) + L2] + unit.e2off;
    const int a1 = a0 + unit.e2cnt;
    const float b12top = a.s12.eb[a0];
    // partners of row1 that can pass with the largest possible ket bound are a prefix
    int n2max = 0;
    for (int k = lane; k < a1 - a0; k += 64) n2max += (a.s12.eb[a0 + k] >= thre12) ? 1 : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) n2max += __shfl_xor(n2max, o);

    int qn = 0;  // queue length (wave-uniform)
    // inlined at both call sites: an outlined copy would take the address of the kernel arguments
    auto evaluate = [&](int slot) __attribute__((always_inline)) {
        constexpr int N1 = ncart(L1), N2 = ncart(L2), N3 = ncart(L3), N4 = ncart(L4);
        const int e2 = s_e2[slot], e4 = s_e4[slot], tt = s_t[slot];
        const KEnt<double> E12 = load_ent(a.s12.eg, e2), E34 = load_ent(a.s34.eg, e4);
        double Db[N2][N4];
#pragma unroll
        for (int b = 0; b < N2; b++)
#pragma unroll
            for (int d = 0; d < N4; d++) Db[b][d] = a.D[(size_t)(E12.cart + b) * a.ldD + (E34.cart + d)];
        const int r3 = a.s34.rows[j3a + tt];
        const int p2 = a.eprim12[e2], p4 = a.eprim34[e4];
        const int o1 = a.s12.cart[r1], o3 = a.s34.cart[r3];
        double D13[N1][N3];
#pragma unroll
        for (int i = 0; i < N1; i++)
#pragma unroll
            for (int j = 0; j < N3; j++) D13[i][j] = a.D13[(size_t)(o1 + i) * a.ldD13 + (o3 + j)];
        double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0}, gC[3] = {0.0, 0.0, 0.0};
        k_quartet_grad<L1, L2, L3, L4>(a.rys, a.ew, E12, E34, a.exp1[r1], a.exp2[p2], a.exp3[r3], D13, Db, gA, gB, gC);
        const int at1 = a.atom1[r1], at2 = a.atom2[p2], at3 = a.atom3[r3], at4 = a.atom4[p4];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            atomic_add(a.G + 3 * at1 + k, gA[k]);
            atomic_add(a.G + 3 * at2 + k, gB[k]);
            atomic_add(a.G + 3 * at3 + k, gC[k]);
            atomic_add(a.G + 3 * at4 + k, -(gA[k] + gB[k] + gC[k]));
        }
    };

    for (int t = j3a; t < j3b && n2max > 0; t++) {
        const int r3 = a.s34.rows[t];
        const int c0 = a.s34.row_start[r3 * (a.s34.maxLp + 1) + L4];
        const int c1 = a.s34.row_start[r3 * (a.s34.maxLp + 1) + L4 + 1];
        const float b34top = a.s34.eb[c0];
        if (!(b34top >= thre34)) continue;
        int n2 = 0, n4 = 0;
        for (int k = lane; k < n2max; k += 64)
            n2 += ((double)(a.s12.eb[a0 + k] * b34top) * dmax >= (double)thresp) ? 1 : 0;
        for (int k = lane; k < c1 - c0; k += 64) {
            const float b = a.s34.eb[c0 + k];
            n4 += (b >= thre34 && (double)(b12top * b) * dmax >= (double)thresp) ? 1 : 0;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            n2 += __shfl_xor(n2, o);
            n4 += __shfl_xor(n4, o);
        }
        const int ntot = n2 * n4;
        for (int base = 0; base < ntot; base += 64) {
            const int f = base + lane;
            bool pass = false;
            int e2 = 0, e4 = 0;
            if (f < ntot) {
                const int j2 = f / n4;
                e2 = a0 + j2;
                e4 = c0 + (f - j2 * n4);
                const float b12 = a.s12.eb[e2], b34 = a.s34.eb[e4];
                if (b12 >= thre12 && b34 >= thre34) {
                    const double dblk = a.Dsh[(size_t)a.s12.eblk[e2] * a.ldDsh + a.s34.eblk[e4]];
                    const double est = (double)(b12 * b34) * dblk;
                    pass = est >= (double)thresp;
                }
            }
            const unsigned long long m = __ballot(pass);
            if (pass) {
                const int pos = qn + __popcll(m & ((1ull << lane) - 1ull));
                s_e2[pos] = e2;
                s_e4[pos] = e4;
                s_t[pos] = t - j3a;
            }
            qn += __popcll(m);
            if (qn >= 64) {
                __syncthreads();
                evaluate(lane);
                __syncthreads();
                const int rest = qn - 64;
                int m2 = 0, m4 = 0, mt = 0;
                if (lane < rest) {
                    m2 = s_e2[64 + lane];
                    m4 = s_e4[64 + lane];
                    mt = s_t[64 + lane];
                }
                __syncthreads();
                if (lane < rest) {
                    s_e2[lane] = m2;
                    s_e4[lane] = m4;
                    s_t[lane] = mt;
                }
                qn = rest;
            }
        }
    }
    __syncthreads();
    if (lane < qn) evaluate(lane);
}

// One packed FP64 entry at a wave-uniform index, through scalar loads (K view ids).
__device__ __forceinline__ KEnt<double> load_ent_uniform(const double* __restrict__ eg, int i)
{
    const double* g = eg + (size_t)i * kKEntDoubles;
    KEnt<double> e;
    e.p = ld_const(g + 0);
    e.K = ld_const(g + 1);
    e.Px = ld_const(g + 2);
    e.Py = ld_const(g + 3);
    e.Pz = ld_const(g + 4);
    e.PAx = ld_const(g + 5);
    e.PAy = ld_const(g + 6);
    e.PAz = ld_const(g + 7);
    e.ABx = ld_const(g + 8);
    e.ABy = ld_const(g + 9);
    e.ABz = ld_const(g + 10);
    const int* ids = reinterpret_cast<const int*>(g + 13);
    e.cart = ld_const(ids);
    e.blk = ld_const(ids + 1);
    e.i2p = ld_const(g + 12);
    return e;
}

// One packed FP32-route entry at a wave-uniform index, through scalar loads (K view ids).
__device__ __forceinline__ KEnt<float> load_ent32_uniform(const double* __restrict__ eg32, int i)
{
    const double* g = eg32 + (size_t)i * kKEnt32Doubles;
    KEnt<float> e;
    e.Px = ld_const(g + 0);
    e.Py = ld_const(g + 1);
    e.Pz = ld_const(g + 2);
    const float* f = reinterpret_cast<const float*>(g + 3);
    e.p = ld_const(f + 0);
    e.K = ld_const(f + 1);
    e.PAx = ld_const(f + 2);
    e.PAy = ld_const(f + 3);
    e.PAz = ld_const(f + 4);
    e.ABx = ld_const(f + 5);
    e.ABy = ld_const(f + 6);
    e.ABz = ld_const(f + 7);
    e.i2p = ld_const(f + 8);
    const int* ids = reinterpret_cast<const int*>(g + 9);
    e.cart = ld_const(ids);
    e.blk = ld_const(ids + 1);
    return e;
}

// LDS of one k_ket workgroup: the staged table blocks first, FP64 (routes with FP64 quartets,
// nrys <= 3) then FP32 (routes with FP32 quartets, nrys <= 5), then kKWaves group accumulators
// (gsize x N1 x N3 doubles each, gsize = rows per ket group).  The FP64 table at address 0 keeps a bin's coefficient offsets
// (0..8 nrys * 24 bytes) inside the immediate offset field of the paired LDS reads, so the table
// lookup needs one address per bin.
template <int L1, int L2, int L3, int L4, int ROUTE>
struct KKetLds {
    static constexpr int NR = (L1 + L2 + L3 + L4) / 2 + 1;
    static constexpr bool HAS32 = ROUTE == kRouteLo || ROUTE == kRouteMixed;
    static constexpr bool HAS64 = ROUTE != kRouteLo;
    static constexpr bool STAGED64 = HAS64 && rys_staged(NR);
    static constexpr bool STAGED32 = HAS32 && rys32_staged(NR);
    static constexpr int N13 = ncart(L1) * ncart(L3);  // accumulator doubles per group row
    static __host__ __device__ size_t bytes64(const RysTable& tab)
    {
        return STAGED64 ? rys_lds_doubles(tab.nbin[NR], NR) * sizeof(double) : 0;
    }
    // byte offset of the accumulators: after both tables, 8-byte aligned
    static __host__ __device__ size_t acc_offset(const RysTable& tab)
    {
        const size_t b = bytes64(tab) + (STAGED32 ? rys32_lds_floats(tab.nbin[NR], NR) * sizeof(float) : 0);
        return (b + 7) & ~(size_t)7;
    }
    static size_t bytes(const RysTable& tab, int gsize)
    {
        return acc_offset(tab) + (size_t)kKWaves * gsize * N13 * sizeof(double);
    }
};

// An FP64 entry rounded to the FP32 route's arithmetic (the values of its FP32 record).
__device__ __forceinline__ KEnt<float> ent_to_f32(const KEnt<double>& e)
{
    KEnt<float> f;
    f.Px = e.Px;
    f.Py = e.Py;
    f.Pz = e.Pz;
    f.p = (float)e.p;
    f.K = (float)e.K;
    f.PAx = (float)e.PAx;
    f.PAy = (float)e.PAy;
    f.PAz = (float)e.PAz;
    f.ABx = (float)e.ABx;
    f.ABy = (float)e.ABy;
    f.ABz = (float)e.ABz;
    f.i2p = (float)e.i2p;
    f.cart = e.cart;
    f.blk = e.blk;
    return f;
}

// Waves per SIMD the compiler is asked to keep room for (register budget 512 / waves), per plain
// route class: classes whose register count sits just above an occupancy step are held below it
// where that costs at most a few spilled dwords and measured faster (classes that already spill,
// or spill more, ran 15-55 % slower when held).  1 = no constraint.
constexpr int kket_class(int L1, int L2, int L3, int L4) { return ((L1 * 3 + L2) * 3 + L3) * 3 + L4; }
constexpr int kket_waves(int L1, int L2, int L3, int L4, int route)
{
    if (route != kRoutePlain) return 1;
    switch (kket_class(L1, L2, L3, L4)) {
    case kket_class(0, 0, 0, 0):
        return 8;  // 62 -> 57 VGPRs, no spill
    case kket_class(0, 0, 0, 1):
    case kket_class(0, 0, 2, 0):
    case kket_class(0, 0, 1, 1):
        return 5;  // 98-108 -> 96 VGPRs, 12-32 bytes spilled
    default:
        return 1;
    }
}

// K build over ket groups.  Each wave of the workgroup takes one unit (bra row r1, chunks of ket
// group g); for each 64-slot chunk, lane l holds ket pair (r3, 4) of slot l in registers and the
// wave walks r1's partners 2 in bound order (wave-uniform, scalar loads), stopping where even the
// chunk's largest ket bound cannot pass with the largest D block.  Each lane tests the reference's
// sieve float(b12 b34) max|D24 block| >= thresp for its (2, 4) and, where it passes, evaluates the
// quartet into FP64 register accumulators for (r1, r3) (reference exchange.cpp:145-152, 186).  On
// the mixed route (one launch) quartets with est >= thredp are evaluated in FP64 and the others in
// FP32, except that a wave step holding any FP64 quartet evaluates all of its quartets in FP64
// (never less precision than the route asks for, and one evaluation per step).  At the end of a chunk the lanes add their
// accumulators into the wave's LDS block of the group (row position t); at the end of the unit
// the block is added to K with one atomic per element, plus the transposed block in the symmetric
// mode (K = K^T) for r3 != r1.
template <int L1, int L2, int L3, int L4, int ROUTE>
__global__ __launch_bounds__(64 * kKWaves) __attribute__((amdgpu_waves_per_eu(kket_waves(L1, L2, L3, L4, ROUTE)))) void k_ket(KArgs a, KGroups kg, const KUnitG* __restrict__ units, int nunits)
{
    using Lds = KKetLds<L1, L2, L3, L4, ROUTE>;
    constexpr bool LO = ROUTE == kRouteLo;        // FP32 only
    constexpr bool SPLIT = ROUTE == kRouteMixed;  // FP64 and FP32 in one launch
    constexpr int NR = Lds::NR;
    constexpr int N1 = ncart(L1), N2 = ncart(L2), N3 = ncart(L3), N4 = ncart(L4), N13 = N1 * N3;
    extern __shared__ double s_dyn[];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double* s_tab64 = s_dyn;
    float* s_tab32 = reinterpret_cast<float*>(reinterpret_cast<char*>(s_dyn) + Lds::bytes64(a.rys));
    const int G = kg.gsize;  // rows per ket group
    const int ACC = G * N13;  // accumulator doubles per wave
    double* s_acc = reinterpret_cast<double*>(reinterpret_cast<char*>(s_dyn) + Lds::acc_offset(a.rys)) + w * ACC;
    if constexpr (Lds::STAGED64) rys_stage_lds<NR>(a.rys, s_tab64);
    if constexpr (Lds::STAGED32) rys32_stage_lds<NR>(a.rys, s_tab32);
    for (int q = lane; q < ACC; q += 64) s_acc[q] = 0.0;
    __syncthreads();
    // no workgroup barrier after this point: the waves are independent
    const int u = blockIdx.x * kKWaves + w;
    if (u >= nunits) return;
    const int i1 = ld_const(&units[u].i1), g = ld_const(&units[u].g), rank1 = ld_const(&units[u].rank1);
    const int crange = ld_const(&units[u].crange);
    const int r1 = ld_const(a.s12.rows + i1);
    int a0 = ld_const(a.s12.row_start + r1 * (a.s12.maxLp + 1) + L2);
    int a1 = ld_const(a.s12.row_start + r1 * (a.s12.maxLp + 1) + L2 + 1);
    const int prange = ld_const(&units[u].prange);
    if (prange != 0) {  // a piece of the row's partner walk
        a0 = min(a1, a0 + (prange & 0xffff));
        if (prange >> 16) a1 = min(a1, a0 + (prange >> 16));
    }
    const int s0 = ld_const(kg.grp_slot + g) + 64 * (crange & 0xffff);
    const int s1 = min(ld_const(kg.grp_slot + g + 1), s0 + 64 * (crange >> 16));

    // quartets with est at or above `thresp` here (and, on the FP32 route, below thredp) are evaluated
    const float thresp = (ROUTE == kRouteHi) ? a.thredp : a.thresp;
    const double dmax = __longlong_as_double((long long)a.sc->dmax_bits);
    // learned route: the unit's certified factor tau <= 1 (est tau bounds every quartet's
    // contribution), so the thresholds scale by 1 / tau.  It holds for quartets whose Schwarz
    // product is at least tau_smin; quartets below that fail the plain test whenever
    // thresp >= tau_smin max|D|, else the factor is not applied.
    const double tau = (a.tau && (double)thresp >= (double)a.tau_smin * dmax) ? (double)ld_const(a.tau + u) : 1.0;
    const double thr = (double)thresp / tau;
    const double thredp = (double)a.thredp / tau;
    const float thre12 = (float)(thr / ((double)a.maxb34 * dmax));
    const float thre34 = (float)(thr / ((double)a.maxb12 * dmax));
    const float b12top = ld_const(a.s12.eb + a0);

    unsigned long long nq = 0, nq32 = 0;
    unsigned int nsteps = 0;  // wave steps with an evaluated quartet (counting mode)
    bool any = false;
    for (int c0 = s0; c0 < s1; c0 += 64) {
        const float b34max = ld_const(kg.chunk_b + (c0 >> 6));
        // no quartet of this chunk can pass, and the later chunks of the group hold smaller bounds
        if (!(b34max >= thre34) || (double)(b12top * b34max) * dmax < thr) break;
        const int2 sl = kg.slot[c0 + lane];
        const int e4 = max(sl.x, 0);
        const float b34 = a.s34.eb[e4];
        const bool valid = sl.x >= 0 && kg.grp_rank[g * G + sl.y] >= rank1 && b34 >= thre34;
        const int blk4 = a.s34.eblk[e4];
        KEnt<double> E34;
        KEnt<float> E34f;
        // s-type C with a higher D: the quartets recur at D (k_quartet KETD), from Q - D = (Q - C) + (C - D)
        constexpr bool KETD = L3 == 0 && L4 > 0;
        if constexpr (LO) {
            E34f = load_ent32<true>(a.s34.eg32, e4);
            if constexpr (KETD) {
                E34f.PAx += E34f.ABx;
                E34f.PAy += E34f.ABy;
                E34f.PAz += E34f.ABz;
            }
        } else {
            E34 = load_ent<true>(a.s34.eg, e4);
            if constexpr (KETD) {
                E34.PAx += E34.ABx;
                E34.PAy += E34.ABy;
                E34.PAz += E34.ABz;
            }
            if constexpr (SPLIT) E34f = ent_to_f32(E34);
        }
        const int cart4 = LO ? E34f.cart : E34.cart;
        const double* __restrict__ dsh = a.Dsh + blk4;
        const double* __restrict__ dcol = a.D + cart4;
        double acc[N1][N3];
#pragma unroll
        for (int i = 0; i < N1; i++)
#pragma unroll
            for (int j = 0; j < N3; j++) acc[i][j] = 0.0;
        bool hit = false;
        // Bra partners below bstop cannot pass with any slot of the chunk: float(b12 b34max) dmax
        // < thr for b12 < thr / (b34max dmax) (1 - 2^-22), and below thre12 the partner fails the
        // truncation.  The partners are sorted, so the walk stops at the first one below bstop
        // (compared as bit patterns: the bounds are non-negative floats).
        const float bstop = fmaxf(thre12, (float)(thr / ((double)b34max * dmax)) * (1.0f - 4.0f * 1.1920929e-7f));
        const unsigned int bstop_bits = __builtin_amdgcn_readfirstlane(__float_as_uint(bstop));
        // the next partner's bound, block and D block maximum are loaded one step ahead
        float b12n = ld_const(a.s12.eb + a0);
        double dnext = valid ? dsh[(size_t)ld_const(a.s12.eblk + a0) * a.ldDsh] : 0.0;
        for (int k = a0; k < a1; k++) {
            const float b12 = b12n;
            if (__float_as_uint(b12) < bstop_bits) break;
            const double dblk = dnext;
            if (k + 1 < a1) {
                b12n = ld_const(a.s12.eb + k + 1);
                const int blk2n = ld_const(a.s12.eblk + k + 1);
                if (valid) dnext = dsh[(size_t)blk2n * a.ldDsh];
            }
            bool pass = false, low = LO;
            if (valid) {
                const double est = (double)(b12 * b34) * dblk;
                pass = est >= thr;
                if constexpr (LO) pass = pass && !(est >= thredp);
                if constexpr (SPLIT) low = !(est >= thredp);
            }
            const unsigned long long m = __ballot(pass);
            if (m == 0ull) continue;
            if (a.count) nsteps++;
            if constexpr (ROUTE == kRoutePlain || ROUTE == kRouteHi) {
                const KEnt<double> E12 = load_ent_uniform(a.s12.eg, k);
                if (pass) {
                    double Db[N2][N4];
#pragma unroll
                    for (int b = 0; b < N2; b++)
#pragma unroll
                        for (int d = 0; d < N4; d++) Db[b][d] = dcol[(size_t)(E12.cart + b) * a.ldD + d];
                    k_quartet<double, L1, L2, L3, L4, Lds::STAGED64, KETD>(a.rys, s_tab64, a.ew, E12, E34, Db, acc);
                    nq++;
                    hit = true;
                }
                continue;
            }
            // mixed route: a wave step with any quartet at or above thredp evaluates all its
            // quartets in FP64 (at least the precision the route asks for); the other steps in FP32
            const bool step32 = LO || (SPLIT && __ballot(pass && !low) == 0ull);
            if (!step32) {
                const KEnt<double> E12 = load_ent_uniform(a.s12.eg, k);
                if (pass) {
                    double Db[N2][N4];
#pragma unroll
                    for (int b = 0; b < N2; b++)
#pragma unroll
                        for (int d = 0; d < N4; d++) Db[b][d] = dcol[(size_t)(E12.cart + b) * a.ldD + d];
                    k_quartet<double, L1, L2, L3, L4, Lds::STAGED64, KETD>(a.rys, s_tab64, a.ew, E12, E34, Db, acc);
                }
            } else {
                const KEnt<float> E12f = load_ent32_uniform(a.s12.eg32, k);
                if (pass) {
                    float Dbf[N2][N4];
#pragma unroll
                    for (int b = 0; b < N2; b++)
#pragma unroll
                        for (int d = 0; d < N4; d++) Dbf[b][d] = (float)dcol[(size_t)(E12f.cart + b) * a.ldD + d];
                    k_quartet<float, L1, L2, L3, L4, Lds::STAGED32, KETD>(a.rys, s_tab32, a.ew, E12f, E34f, Dbf, acc);
                }
            }
            if (pass) {
                nq++;
                if (step32) nq32++;
                hit = true;
            }
        }
        if (hit) {
#pragma unroll
            for (int i = 0; i < N1; i++)
#pragma unroll
                for (int j = 0; j < N3; j++) atomicAdd(&s_acc[sl.y * N13 + i * N3 + j], acc[i][j]);
        }
        any = any || (__ballot(hit) != 0ull);
    }
    if (any) {
        // the lanes' LDS additions above are complete before any lane reads the block
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int o1 = ld_const(a.s12.cart + r1);
        for (int q = lane; q < ACC; q += 64) {
            const double v = s_acc[q];
            if (v == 0.0) continue;
            const int t = q / N13, ij = q - t * N13, i = ij / N3, j = ij - i * N3;
            const int r3 = kg.grp_rows[g * G + t];
            const int o3 = a.s34.cart[r3];
            atomic_add(a.K + (size_t)(o1 + i) * a.ldK + (o3 + j), v);
            if (a.symmetric && (r1 != r3 || a.mirror_diag)) atomic_add(a.K + (size_t)(o3 + j) * a.ldK + (o1 + i), v);
        }
    }
    if (a.count) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            nq += __shfl_xor(nq, o);
            nq32 += __shfl_xor(nq32, o);
        }
        if (lane == 0 && nq) {
            atomicAdd(&a.sc->counts[kcount_index(L1, L2, L3, L4)], nq);
            if (nq32) atomicAdd(&a.sc->counts32[kcount_index(L1, L2, L3, L4)], nq32);
            atomicAdd(&a.sc->wsteps[kcount_index(L1, L2, L3, L4)], (unsigned long long)nsteps);
            atomicAdd(&a.sc->wlanes[kcount_index(L1, L2, L3, L4)], nq);
        }
    }
}

// Training data of the learned screen: for each K work unit, the largest ratio M_q / (b12 b34)
// over the unit's quartets q with Schwarz product b12 b34 >= smin, M_q = max over the Cartesian
// components of |(ab|O|cd)| (the exact Rys evaluation of k_quartet, one component pair (b, d) at a
// time against a unit D block).  The unit's quartets are those k_ket would visit without any
// density screening: the bra row's partners (its walk piece) x the unit's slots (triangle rank
// rule included).  One wave per unit.
template <int L1, int L2, int L3, int L4>
__global__ __launch_bounds__(64 * kKWaves) void k_unit_mag(KArgs a, KGroups kg, const KUnitG* __restrict__ units,
                                                           int nunits, float smin, float* __restrict__ out)
{
    constexpr int N1 = ncart(L1), N2 = ncart(L2), N3 = ncart(L3), N4 = ncart(L4);
    const int G = kg.gsize;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int u = blockIdx.x * kKWaves + w;
    if (u >= nunits) return;
    const int i1 = ld_const(&units[u].i1), g = ld_const(&units[u].g), rank1 = ld_const(&units[u].rank1);
    const int crange = ld_const(&units[u].crange);
    const int r1 = ld_const(a.s12.rows + i1);
    int a0 = ld_const(a.s12.row_start + r1 * (a.s12.maxLp + 1) + L2);
    int a1 = ld_const(a.s12.row_start + r1 * (a.s12.maxLp + 1) + L2 + 1);
    const int prange = ld_const(&units[u].prange);
    if (prange != 0) {
        a0 = min(a1, a0 + (prange & 0xffff));
        if (prange >> 16) a1 = min(a1, a0 + (prange >> 16));
    }
    const int s0 = ld_const(kg.grp_slot + g) + 64 * (crange & 0xffff);
    const int s1 = min(ld_const(kg.grp_slot + g + 1), s0 + 64 * (crange >> 16));
    float best = 0.0f;
    for (int c0 = s0; c0 < s1; c0 += 64) {
        const int2 sl = kg.slot[c0 + lane];
        const int e4 = max(sl.x, 0);
        const bool valid = sl.x >= 0 && kg.grp_rank[g * G + sl.y] >= rank1;
        const double b34 = (double)a.s34.eb[e4];
        const KEnt<double> E34 = load_ent(a.s34.eg, e4);
        for (int k = a0; k < a1; k++) {
            const double b12 = (double)ld_const(a.s12.eb + k);
            const KEnt<double> E12 = load_ent_uniform(a.s12.eg, k);
            if (valid && b12 * b34 >= (double)smin) {
                double m = 0.0;
                for (int b = 0; b < N2; b++)
                    for (int d = 0; d < N4; d++) {
                        double Db[N2][N4];
#pragma unroll
                        for (int bb = 0; bb < N2; bb++)
#pragma unroll
                            for (int dd = 0; dd < N4; dd++) Db[bb][dd] = (bb == b && dd == d) ? 1.0 : 0.0;
                        double acc[N1][N3];
#pragma unroll
                        for (int i = 0; i < N1; i++)
#pragma unroll
                            for (int j = 0; j < N3; j++) acc[i][j] = 0.0;
                        k_quartet<double, L1, L2, L3, L4, false>(a.rys, nullptr, a.ew, E12, E34, Db, acc);
#pragma unroll
                        for (int i = 0; i < N1; i++)
#pragma unroll
                            for (int j = 0; j < N3; j++) m = fmax(m, fabs(acc[i][j]));
                    }
                best = fmaxf(best, (float)(m / (b12 * b34)));
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) best = fmaxf(best, __shfl_xor(best, o));
    if (lane == 0) out[u] = best;
}

// Launch of one K build class at row angular momentum L1 (instantiated in exchange_l<L1>.hip).
template <int L1>
void launch_k_ket_l1(int L2, int L3, int L4, int route, const KArgs& a, const KGroups& kg, const KUnitG* units,
                     int nunits, hipStream_t s)
{
    dim3 grid((unsigned)((nunits + kKWaves - 1) / kKWaves)), block(64 * kKWaves);
#define ONE(B, C, D, R)                                                                             \
    {                                                                                               \
        const size_t lds = KKetLds<L1, B, C, D, R>::bytes(a.rys, kg.gsize);                         \
        if (lds > kMaxLdsPerGroup || lds != kket_lds_bytes(a.rys, L1, B, C, D, R, kg.gsize))       \
            throw std::runtime_error("lsjk: k_ket LDS request too large or inconsistent");         \
        hipLaunchKernelGGL((k_ket<L1, B, C, D, R>), grid, block, lds, s, a, kg, units, nunits);     \
    }
#define CASE(B, C, D)                                                                               \
    if (L2 == B && L3 == C && L4 == D) {                                                            \
        if (route == kRouteMixed)                                                                   \
            ONE(B, C, D, kRouteMixed)                                                               \
        else if (route == kRoutePlain)                                                              \
            ONE(B, C, D, kRoutePlain)                                                               \
        else                                                                                        \
            throw std::runtime_error("lsjk: k_ket route not instantiated");                         \
        return;                                                                                     \
    }
#define C4(B, C) CASE(B, C, 0) CASE(B, C, 1) CASE(B, C, 2)
#define C3(B) C4(B, 0) C4(B, 1) C4(B, 2)
    C3(0) C3(1) C3(2)
#undef C3
#undef C4
#undef CASE
#undef ONE
}

// Launch of one exchange-gradient class (instantiated per (L1, L2) in exchange_grad_<L1><L2>.hip).
template <int L1, int L2>
void launch_k_grad_l12(int L3, int L4, const KArgs& a, const KUnit* units, int nunits, hipStream_t s)
{
    dim3 grid((unsigned)nunits), block(64);
#define CASE(C, D)                                                                                  \
    if (L3 == C && L4 == D) {                                                                       \
        hipLaunchKernelGGL((k_tiles_grad<L1, L2, C, D>), grid, block, 0, s, a, units);             \
        return;                                                                                     \
    }
    CASE(0, 0) CASE(0, 1) CASE(0, 2) CASE(1, 0) CASE(1, 1) CASE(1, 2) CASE(2, 0) CASE(2, 1) CASE(2, 2)
#undef CASE
}

}  // namespace lsjk

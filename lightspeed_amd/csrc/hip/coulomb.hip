// Direct Coulomb (J) build on gfx950: Hermite (McMurchie-Davidson) J-engine.
//
// Behaviour follows IntBox::coulombCPU (reference src/cpp/intbox/cpu/coulomb.cpp:14-196):
//   * ket truncation  B34' = float(B34 * max|D block|) kept if > thresp/max12
//     (PairListUtil::truncate_ket, src/cpp/core/pair_list_util.cpp:60-123),
//   * bra truncation  B12 >= thresp/max34 (truncate_bra, pair_list_util.cpp:32-58),
//   * quartet sieve   float(B12 * B34') >= thresp (coulomb.cpp:115),
//   * Hermite densities with (-1)^|t| phases (HermiteL::cart_to_herm, hermite.cpp:9-206) and
//     the transposed bra transform (herm_to_cart, hermite.cpp:207-411).
// The same set of primitive quartets is evaluated; only the order of evaluation differs.
//
// Kernels (one launch each per call, per class):
//   j_ket_prep<LA,LB>    one thread per ket pair: truncated bound, Hermite density d_u.
//   j_quartets<L12,L34>  one wave per (64 bra pairs x ket chunk) task: each lane owns one bra
//                        pair and sweeps the chunk's kets, staged through LDS; Hermite integrals
//                        R_{t+u} by Rys quadrature; jp_t += sum_u R_{t+u} d_u in registers.
//   j_bra_scatter<LA,LB> one thread per bra pair: J_ab += K sum_t E E E jp_t (symmetrised).
#include "internal.hpp"
#include "kernels.hpp"

namespace lsjk {

// ------------------------------------------------------------------ ket preparation
template <int LA, int LB>
__global__ __launch_bounds__(256) void j_ket_prep(JPrepArgs a, int k0, int nk, int hstart, size_t hoff)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), L = LA + LB, H = nherm(L);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nk) return;
    const int k = k0 + i;
    const int ia = a.p1[k], ib = a.p2[k];
    const int ca = a.cartA[ia], cb = a.cartB[ib];
    const int pos = a.hpos[k];
    double Dt[NA][NB];
    double val = 0.0;
#pragma unroll
    for (int p = 0; p < NA; p++)
#pragma unroll
        for (int q = 0; q < NB; q++) {
            double v = a.D[(size_t)(ca + p) * a.ldD + (cb + q)];
            Dt[p][q] = v;
            val = fmax(val, fabs(v));
        }
    if (a.symmetric) {
        const double scale = (ia == ib) ? 0.5 : 1.0;
#pragma unroll
        for (int p = 0; p < NA; p++)
#pragma unroll
            for (int q = 0; q < NB; q++) {
                double v = a.D[(size_t)(cb + q) * a.ldD + (ca + p)];
                val = fmax(val, fabs(v));
                Dt[p][q] = (Dt[p][q] + v) * scale;
            }
    }
    atomicMax(&a.sc->dmax_bits, (unsigned long long)__double_as_longlong(val));
    const float B2 = (float)(val * (double)a.bound[k]);
    const bool keep = B2 > a.thre_ket;
    a.kb[pos] = keep ? B2 : -1.0f;
    double* dst = a.kd + hoff + (size_t)(pos - hstart) * H;
    float* dst32 = a.kd32 ? a.kd32 + hoff + (size_t)(pos - hstart) * H : nullptr;
    if (!keep) {
#pragma unroll
        for (int t = 0; t < H; t++) dst[t] = 0.0;
        if (dst32)
#pragma unroll
            for (int t = 0; t < H; t++) dst32[t] = 0.0f;
        return;
    }
    atomicMax(&a.sc->max34_bits, __float_as_uint(B2));
    const double eA = a.eA[ia], eB = a.eB[ib];
    const double p = eA + eB, ip = 1.0 / p;
    const double Ax = a.xA[ia], Ay = a.yA[ia], Az = a.zA[ia];
    const double Bx = a.xB[ib], By = a.yB[ib], Bz = a.zB[ib];
    const double Px = (eA * Ax + eB * Bx) * ip, Py = (eA * Ay + eB * By) * ip, Pz = (eA * Az + eB * Bz) * ip;
    HermE<LA, LB> Ex, Ey, Ez;
    hermite_E<LA, LB>(0.5 * ip, Px - Ax, Px - Bx, Ex);
    hermite_E<LA, LB>(0.5 * ip, Py - Ay, Py - By, Ey);
    hermite_E<LA, LB>(0.5 * ip, Pz - Az, Pz - Bz, Ez);
    const double Kab = a.K[k];
    static_for<H>([&](auto T_) {
        constexpr int t = decltype(T_)::value;
        constexpr int tx = herm_tx(t), ty = herm_ty(t), tz = herm_tz(t);
        double s = 0.0;
        static_for<NA>([&](auto P_) {
            constexpr int p = decltype(P_)::value;
            constexpr int ax = cart_lx(LA, p), ay = cart_ly(LA, p), az = cart_lz(LA, p);
            static_for<NB>([&](auto Q_) {
                constexpr int q = decltype(Q_)::value;
                constexpr int bx = cart_lx(LB, q), by = cart_ly(LB, q), bz = cart_lz(LB, q);
                if constexpr (tx <= ax + bx && ty <= ay + by && tz <= az + bz)
                    s += Ex.v[tx][ax][bx] * Ey.v[ty][ay][by] * Ez.v[tz][az][bz] * Dt[p][q];
            });
        });
        const double v = (((tx + ty + tz) & 1) ? -Kab : Kab) * s;
        dst[t] = v;
        if (dst32) dst32[t] = (float)v;
    });
}

// ------------------------------------------------------------------ quartets
// Arithmetic of the two routes: Real = double (FP64 route) or float (FP32 route).  The FP32 route
// forms T in FP64 (from FP64 centre differences) and uses the FP32 Rys rule; everything else of
// the quartet is FP32; the caller accumulates in FP64.
template <typename Real>
__device__ __forceinline__ Real rdiv(Real a, Real b)
{
    if constexpr (std::is_same<Real, float>::value)
        return __fdividef(a, b);
    else
        return a / b;
}

template <typename Real>
__device__ __forceinline__ Real rrsqrt(Real x)
{
    if constexpr (std::is_same<Real, float>::value)
        return __frsqrt_rn(x);
    else
        return rsqrt_pos(x);
}

template <typename Real, int NR>
__device__ __forceinline__ void roots_real(const RysTable& tab, const void* s_tab, double T, Real* x, Real* w)
{
    if constexpr (std::is_same<Real, float>::value)
        rys_roots_f32<NR, true>(tab, static_cast<const float*>(s_tab), T, x, w);
    else
        rys_roots_any<NR>(tab, static_cast<const double*>(s_tab), T, x, w);
}

// Per-root one-dimensional Hermite factors of one bra-ket pair and operator term:
// H_0 = 1 (x: pref w_i), H_1 = U X, H_{n+1} = U (X H_n + n H_{n-1}), U = -2 rho d^2 x_i
// (coulomb.cpp:138-158).
template <typename Real, int L>
struct HermRoot {
    Real x[L + 1], y[L + 1], z[L + 1];
};

template <typename Real, int L>
__device__ __forceinline__ void herm_factors(Real pref_w, Real U, Real X, Real Y, Real Z, HermRoot<Real, L>& H)
{
    H.x[0] = pref_w;
    H.y[0] = Real(1);
    H.z[0] = Real(1);
    if constexpr (L > 0) {
        H.x[1] = U * X * H.x[0];
        H.y[1] = U * Y;
        H.z[1] = U * Z;
    }
#pragma unroll
    for (int u = 1; u < L; u++) {
        H.x[u + 1] = U * (X * H.x[u] + Real(u) * H.x[u - 1]);
        H.y[u + 1] = U * (Y * H.y[u] + Real(u) * H.y[u - 1]);
        H.z[u + 1] = U * (Z * H.z[u] + Real(u) * H.z[u - 1]);
    }
}

// Calls f(HermRoot) for every Rys root of every operator term of one bra-ket pair.
// c12, c34: sqrt(2 pi^(5/2)) / p of each pair, so that c12 c34 (p + q)^(-1/2) = 2 pi^(5/2) / (p q sqrt(p + q)).
template <typename Real, int L, typename F>
__device__ __forceinline__ void for_each_root(const RysTable& tab, const void* s_tab, const Ewald& ew, double p12,
                                              double c12, double p34, double c34, double X, double Y, double Z, F&& f)
{
    constexpr int NR = L / 2 + 1;
    const Real p = (Real)p12, q = (Real)p34;
    const Real psum = p + q;
    const Real rs = rrsqrt(psum);
    const Real rho = p * q * (rs * rs);
    const Real base = (Real)c12 * (Real)c34 * rs;  // 2 pi^{5/2} / (p q sqrt(p+q))
    const double r2 = X * X + Y * Y + Z * Z;
    const Real Xr = (Real)X, Yr = (Real)Y, Zr = (Real)Z;
    // one operator term: Rys roots at T = rho d2 r2, Hermite factors with prefactor pref
    auto term = [&](Real pref, Real d2) __attribute__((always_inline)) {
        const double T = (double)(rho * d2) * r2;
        Real x[NR], w[NR];
        roots_real<Real, NR>(tab, s_tab, T, x, w);
#pragma unroll
        for (int i = 0; i < NR; i++) {
            HermRoot<Real, L> H;
            herm_factors<Real, L>(pref * w[i], Real(-2) * rho * d2 * x[i], Xr, Yr, Zr, H);
            f(H);
        }
    };
    if (ew.plain) {
        // one bare 1/r term: d2 = 1, no term loop (the common case)
        term((Real)ew.alpha[0] * base, Real(1));
        return;
    }
    for (int it = 0; it < ew.n; it++) {
        const double om = ew.omega[it];
        Real d2 = Real(1);
        Real pref = (Real)ew.alpha[it] * base;
        if (__double_as_longlong(om) != __double_as_longlong(-1.0)) {  // erf-attenuated term (bit test: scalar)
            const Real o = (Real)om;
            d2 = rdiv(o * o, rho + o * o);
            pref *= sqrt(d2);
        }
        term(pref, d2);
    }
}

// Hermite Coulomb integrals R_{tuv} (t+u+v <= L) of one bra-ket pair by Rys quadrature:
// R_{tuv} = pref * sum_i w_i Hx_t Hy_u Hz_v.
template <typename Real, int L>
__device__ __forceinline__ void hermite_R(const RysTable& tab, const void* s_tab, const Ewald& ew, double p12,
                                          double c12, double p34, double c34, double X, double Y, double Z,
                                          Real (&R)[nherm(L)])
{
    constexpr int HL = nherm(L);
#pragma unroll
    for (int t = 0; t < HL; t++) R[t] = Real(0);
    for_each_root<Real, L>(tab, s_tab, ew, p12, c12, p34, c34, X, Y, Z, [&](const HermRoot<Real, L>& H) __attribute__((always_inline)) {
        static_for<HL>([&](auto T_) {
            constexpr int t = decltype(T_)::value;
            constexpr int tx = herm_tx(t), ty = herm_ty(t), tz = herm_tz(t);
            R[t] += H.x[tx] * H.y[ty] * H.z[tz];
        });
    });
}

// High angular momentum: the same contraction factorised per Rys root, so the (L+1)(L+2)(L+3)/6
// Hermite integrals are never held at once:
//   A[uy][tz] = sum_uz Hz[tz+uz] d[ux,uy,uz];  B[ty][tz] = sum_uy Hy[ty+uy] A[uy][tz];
//   jp[tx,ty,tz] += Hx[tx+ux] B[ty][tz]   (for each ux).
template <typename Real, int L12, int L34, typename Acc>
__device__ __forceinline__ void hermite_contract_factorized(const RysTable& tab, const void* s_tab, const Ewald& ew,
                                                            double p12, double c12, double p34, double c34, double X,
                                                            double Y, double Z, const Real (&d)[nherm(L34)], Acc& jp)
{
    constexpr int L = L12 + L34;
    for_each_root<Real, L>(tab, s_tab, ew, p12, c12, p34, c34, X, Y, Z, [&](const HermRoot<Real, L>& H) __attribute__((always_inline)) {
        static_for<L34 + 1>([&](auto UX_) {
            constexpr int ux = decltype(UX_)::value;
            Real A[L34 + 1 - ux][L12 + 1];
            static_for<L34 + 1 - ux>([&](auto UY_) {
                constexpr int uy = decltype(UY_)::value;
                static_for<L12 + 1>([&](auto TZ_) {
                    constexpr int tz = decltype(TZ_)::value;
                    Real s = Real(0);
                    static_for<L34 + 1 - ux - uy>([&](auto UZ_) {
                        constexpr int uz = decltype(UZ_)::value;
                        constexpr int iu = herm_index(ux, uy, uz);
                        s = fma(H.z[tz + uz], d[iu], s);
                    });
                    A[uy][tz] = s;
                });
            });
            static_for<L12 + 1>([&](auto TY_) {
                constexpr int ty = decltype(TY_)::value;
                static_for<L12 + 1 - ty>([&](auto TZ_) {
                    constexpr int tz = decltype(TZ_)::value;
                    Real s = Real(0);
                    static_for<L34 + 1 - ux>([&](auto UY_) {
                        constexpr int uy = decltype(UY_)::value;
                        s = fma(H.y[ty + uy], A[uy][tz], s);
                    });
                    static_for<L12 + 1 - ty - tz>([&](auto TX_) {
                        constexpr int tx = decltype(TX_)::value;
                        constexpr int it = herm_index(tx, ty, tz);
                        jp[it] = fma(H.x[tx + ux], s, jp[it]);
                    });
                });
            });
        });
    });
}

// LDS of one j_quartets workgroup: the FP64 table block (routes with FP64 quartets, nrys <= 3),
// then the FP32 block (routes with FP32 quartets).
template <int L12, int L34, int ROUTE>
struct JLds {
    static constexpr int NR = (L12 + L34) / 2 + 1;
    static constexpr bool HAS32 = ROUTE == kRouteLo || ROUTE == kRouteMixed;
    static constexpr bool STAGED64 = ROUTE != kRouteLo && rys_staged(NR);
    static __host__ __device__ size_t bytes64(const RysTable& tab)
    {
        return STAGED64 ? rys_lds_doubles(tab.nbin[NR], NR) * sizeof(double) : 0;
    }
    static size_t bytes(const RysTable& tab)
    {
        return bytes64(tab) + (HAS32 ? rys32_lds_floats(tab.nbin[NR], NR) * sizeof(float) : 0);
    }
};

// Persistent kernel: each wave takes (64 bra pairs x ket range) tasks from a queue until it is
// empty.  The workgroup's waves share an LDS copy of the Rys table (FP64 rule for nrys <= 3 on the
// routes with FP64 quartets, the FP32 rule on the routes with FP32 quartets).  Lanes own bra
// pairs; the kets are read at wave-uniform addresses (scalar loads), in bound order.  ROUTE
// selects the quartets by est = B12 * B34' (kRoutePlain / kRouteHi / kRouteLo), or evaluates
// the mixed route in one launch (kRouteMixed): a ket step with any quartet at or above thredp is
// evaluated in FP64, the others in FP32.
template <int L12, int L34, int ROUTE>
__global__ __launch_bounds__(256) void j_quartets(JQuartetArgs a, const JTask* __restrict__ tasks, int ntask,
                                                  unsigned int* qctr)
{
    using Lds = JLds<L12, L34, ROUTE>;
    constexpr int H12 = nherm(L12), H34 = nherm(L34), L = L12 + L34, HL = nherm(L), NR = L / 2 + 1;
    constexpr bool LO = ROUTE == kRouteLo;
    constexpr bool MIXED = ROUTE == kRouteMixed;
    extern __shared__ double s_dyn[];
    double* s_tab64 = s_dyn;
    float* s_tab32 = reinterpret_cast<float*>(reinterpret_cast<char*>(s_dyn) + Lds::bytes64(a.rys));
    if constexpr (Lds::STAGED64) rys_stage_lds<NR>(a.rys, s_tab64);
    if constexpr (Lds::HAS32) rys32_stage_lds<NR>(a.rys, s_tab32);
    if constexpr (Lds::STAGED64 || Lds::HAS32) __syncthreads();

    const int lane = threadIdx.x & 63;
    const float thresp = a.thresp;
    // quartets at or above `thre_lo` (and, on the FP32 route, below thredp) are evaluated
    const float thre_lo = (ROUTE == kRouteHi) ? a.thredp : thresp;
    const float max34 = __uint_as_float(a.sc->max34_bits);
    const float thre_bra = thresp / max34;
    const double dmax = __longlong_as_double((long long)a.sc->dmax_bits);
    unsigned long long nq = 0, nq32 = 0;
    unsigned int nsteps = 0;  // wave steps with an evaluated quartet (counting mode)

    for (;;) {
        const int t = (int)wave_take(qctr);
        if (t >= ntask) break;
        JTask task;
        task.bra0 = ld_const(&tasks[t].bra0);
        task.nbra = ld_const(&tasks[t].nbra);
        task.ket0 = ld_const(&tasks[t].ket0);
        task.ket1 = ld_const(&tasks[t].ket1);
        // largest bra bound of the tile (sorted, descending): its first pair
        const float B12max = ld_const(a.bra_b + task.bra0);
        if (!(B12max >= thre_bra)) continue;  // no bra of the tile survives truncation
        const int bra = task.bra0 + lane;
        const bool active = lane < task.nbra;
        float B12 = 0.0f;
        double p12 = 1.0, c12 = 0.0, Px = 0.0, Py = 0.0, Pz = 0.0;
        if (active) {
            B12 = a.bra_b[bra];
            p12 = a.bra_geom[bra];
            Px = a.bra_geom[a.bra_n + bra];
            Py = a.bra_geom[2 * a.bra_n + bra];
            Pz = a.bra_geom[3 * a.bra_n + bra];
            c12 = a.bra_geom[4 * a.bra_n + bra];
        }
        const bool bra_ok = active && (B12 >= thre_bra);
        double jp[H12];
#pragma unroll
        for (int u = 0; u < H12; u++) jp[u] = 0.0;

        // Early exit: B34' <= float(dmax * B34) and the static ket bounds are sorted, so no ket with
        // static bound below kstop = thre_lo / (B12max dmax) (1 - 2^-21) can pass with a bra of the
        // tile; the test compares bit patterns (non-negative floats) on the scalar unit.
        const float kstop = (float)((double)thre_lo / ((double)B12max * dmax)) * (1.0f - 4.0f * 1.1920929e-7f);
        const unsigned int kstop_bits = __builtin_amdgcn_readfirstlane(__float_as_uint(kstop));
        // the next ket's bounds are loaded one step ahead
        float bsn = ld_const(a.ket_static_b + task.ket0), kbn = ld_const(a.ket_b + task.ket0);
        for (int k = task.ket0; k < task.ket1; k++) {
            const float bs = bsn, kb = kbn;
            if (__float_as_uint(bs) < kstop_bits) break;
            if (k + 1 < task.ket1) {
                bsn = ld_const(a.ket_static_b + k + 1);
                kbn = ld_const(a.ket_b + k + 1);
            }
            const float est = B12 * kb;
            bool take = bra_ok && est >= thre_lo;
            if constexpr (LO) take = take && !(est >= a.thredp);
            bool step32 = LO;
            if constexpr (MIXED) step32 = __ballot(take && est >= a.thredp) == 0ull;
            if (a.count && __ballot(take) != 0ull) nsteps++;
            if (!take) continue;
            nq++;
            if (step32) nq32++;
            const double q = ld_const(a.ket_geom + k);
            const double c34 = ld_const(a.ket_geom + 4 * a.ket_n + k);
            const double X = Px - ld_const(a.ket_geom + a.ket_n + k);
            const double Y = Py - ld_const(a.ket_geom + 2 * a.ket_n + k);
            const double Z = Pz - ld_const(a.ket_geom + 3 * a.ket_n + k);
            const size_t dofs = (size_t)(k - a.ket_hstart) * H34;
            // one quartet in the arithmetic of `acc`'s element type, reading the ket density of
            // that precision and the matching table block
            auto contract = [&](auto& acc, const auto* dsrc, const void* tab) __attribute__((always_inline)) {
                using Real = typename std::remove_reference<decltype(acc[0])>::type;
                Real dreg[H34];
#pragma unroll
                for (int u = 0; u < H34; u++) dreg[u] = ld_const(dsrc + dofs + u);
                if constexpr (L >= 5) {
                    hermite_contract_factorized<Real, L12, L34>(a.rys, tab, a.ew, p12, c12, q, c34, X, Y, Z, dreg, acc);
                } else {
                    Real R[HL];
                    hermite_R<Real, L>(a.rys, tab, a.ew, p12, c12, q, c34, X, Y, Z, R);
                    static_for<H34>([&](auto U_) {
                        constexpr int u = decltype(U_)::value;
                        constexpr int ux = herm_tx(u), uy = herm_ty(u), uz = herm_tz(u);
                        static_for<H12>([&](auto T_) {
                            constexpr int tt = decltype(T_)::value;
                            constexpr int idx = herm_index(herm_tx(tt) + ux, herm_ty(tt) + uy, herm_tz(tt) + uz);
                            acc[tt] = fma(R[idx], dreg[u], acc[tt]);
                        });
                    });
                }
            };
            auto fp32 = [&]() __attribute__((always_inline)) {
                // FP32 quartet, FP64 accumulation
                float jq[H12];
#pragma unroll
                for (int u = 0; u < H12; u++) jq[u] = 0.0f;
                contract(jq, a.ket_d32, s_tab32);
#pragma unroll
                for (int u = 0; u < H12; u++) jp[u] += (double)jq[u];
            };
            if constexpr (LO) {
                fp32();
            } else if constexpr (MIXED) {
                if (step32)
                    fp32();
                else
                    contract(jp, a.ket_d, s_tab64);
            } else {
                contract(jp, a.ket_d, s_tab64);
            }
        }
        if (bra_ok) {
            double* dst = a.jp + (size_t)(bra - a.bra_hstart) * H12;
#pragma unroll
            for (int u = 0; u < H12; u++)
                if (jp[u] != 0.0) atomic_add(dst + u, jp[u]);
        }
    }
    if (a.count) {
        // per-wave reduction of the evaluated-quartet counts
        for (int o = 32; o > 0; o >>= 1) {
            nq += __shfl_xor(nq, o);
            nq32 += __shfl_xor(nq32, o);
        }
        if (lane == 0 && nq) {
            atomicAdd(&a.sc->counts[a.count_cls], nq);
            if (nq32) atomicAdd(&a.sc->counts32[a.count_cls], nq32);
            atomicAdd(&a.sc->wsteps[a.count_cls], (unsigned long long)nsteps);
            atomicAdd(&a.sc->wlanes[a.count_cls], nq);
        }
    }
}

// Symmetric J (one list on both sides): each integral (x|y) of bra group L12 and ket group
// L34 (L12 >= L34; y < x by Hermite position within one group) is evaluated once and used twice,
// as the reference's two quartets:
//   A. J_x += (x|y) d_y  when x passes the bra truncation and float(B_x B'_y) >= thresp;
//   B. J_y += (y|x) d_x  when y passes the bra truncation and float(B_y B'_x) >= thresp
//      (not for y = x, where it is A again),
// B' being the truncated (D-weighted) bounds.  Lanes own bra pairs x as in j_quartets; the B
// contributions of a ket step are summed over the wave and added to J_y by one lane.  ROUTE
// kRoutePlain: FP64; kRouteMixed: a ket step in which no lane's estimate reaches thredp, in either
// use, is evaluated in FP32 (table, R, contraction; FP64 accumulation), the others in FP64.
template <int L12, int L34, int ROUTE>
__global__ __launch_bounds__(256) void j_quartets_sym(JQuartetArgs a, const JTask* __restrict__ tasks, int ntask,
                                                      unsigned int* qctr)
{
    using Lds = JLds<L12, L34, ROUTE>;
    constexpr bool MIXED = ROUTE == kRouteMixed;
    constexpr int H12 = nherm(L12), H34 = nherm(L34), L = L12 + L34, HL = nherm(L), NR = L / 2 + 1;
    constexpr bool SAME = L12 == L34;
    extern __shared__ double s_dyn[];
    double* s_tab64 = s_dyn;
    float* s_tab32 = reinterpret_cast<float*>(reinterpret_cast<char*>(s_dyn) + Lds::bytes64(a.rys));
    if constexpr (Lds::STAGED64) rys_stage_lds<NR>(a.rys, s_tab64);
    if constexpr (Lds::HAS32) rys32_stage_lds<NR>(a.rys, s_tab32);
    if constexpr (Lds::STAGED64 || Lds::HAS32) __syncthreads();
    const int lane = threadIdx.x & 63;
    const float thre_lo = a.thresp;
    const float max34 = __uint_as_float(a.sc->max34_bits);
    const float thre_bra = a.thresp / max34;
    const double dmax = __longlong_as_double((long long)a.sc->dmax_bits);
    unsigned long long nqa = 0, nqb = 0, nq32a = 0, nq32b = 0;
    unsigned long long nlanes = 0;  // integrals evaluated (one serves both uses; counting mode)
    unsigned int nsteps = 0;

    for (;;) {
        const int t = (int)wave_take(qctr);
        if (t >= ntask) break;
        JTask task;
        task.bra0 = ld_const(&tasks[t].bra0);
        task.nbra = ld_const(&tasks[t].nbra);
        task.ket0 = ld_const(&tasks[t].ket0);
        task.ket1 = ld_const(&tasks[t].ket1);
        const float B12max = ld_const(a.bra_b + task.bra0);
        const int x = task.bra0 + lane;
        const bool active = lane < task.nbra;
        float Bx = 0.0f, kbx = -1.0f;
        double p12 = 1.0, c12 = 0.0, Px = 0.0, Py = 0.0, Pz = 0.0;
        double dx[H12];
#pragma unroll
        for (int u = 0; u < H12; u++) dx[u] = 0.0;
        if (active) {
            Bx = a.bra_b[x];
            kbx = a.ket_b[x];
            p12 = a.bra_geom[x];
            Px = a.bra_geom[a.bra_n + x];
            Py = a.bra_geom[2 * a.bra_n + x];
            Pz = a.bra_geom[3 * a.bra_n + x];
            c12 = a.bra_geom[4 * a.bra_n + x];
#pragma unroll
            for (int u = 0; u < H12; u++) dx[u] = a.bra_d[(size_t)(x - a.bra_hstart) * H12 + u];
        }
        const bool bra_ok = active && (Bx >= thre_bra);
        double jp[H12];
#pragma unroll
        for (int u = 0; u < H12; u++) jp[u] = 0.0;

        // As in j_quartets, no ket with static bound below kstop passes with any bra of the tile,
        // in either use: B'_x <= float(dmax B_x) too (margin 2^-20 for the two roundings).
        const float kstop = (float)((double)thre_lo / ((double)B12max * dmax)) * (1.0f - 8.0f * 1.1920929e-7f);
        const unsigned int kstop_bits = __builtin_amdgcn_readfirstlane(__float_as_uint(kstop));
        float bsn = ld_const(a.ket_static_b + task.ket0), kbn = ld_const(a.ket_b + task.ket0);
        for (int k = task.ket0; k < task.ket1; k++) {
            const float By = bsn, kby = kbn;
            if (__float_as_uint(By) < kstop_bits) break;
            if (k + 1 < task.ket1) {
                bsn = ld_const(a.ket_static_b + k + 1);
                kbn = ld_const(a.ket_b + k + 1);
            }
            const bool takeA = bra_ok && (!SAME || k <= x) && Bx * kby >= thre_lo;
            const bool takeB = active && (By >= thre_bra) && (!SAME || k < x) && By * kbx >= thre_lo;
            const unsigned long long mB = __ballot(takeB);
            if ((__ballot(takeA) | mB) == 0ull) continue;
            if (a.count) nsteps++;
            bool step32 = false;
            if constexpr (MIXED)
                step32 = __ballot((takeA && Bx * kby >= a.thredp) || (takeB && By * kbx >= a.thredp)) == 0ull;
            const double q = ld_const(a.ket_geom + k);
            const double c34 = ld_const(a.ket_geom + 4 * a.ket_n + k);
            const double X = Px - ld_const(a.ket_geom + a.ket_n + k);
            const double Y = Py - ld_const(a.ket_geom + 2 * a.ket_n + k);
            const double Z = Pz - ld_const(a.ket_geom + 3 * a.ket_n + k);
            const size_t dofs = (size_t)(k - a.ket_hstart) * H34;
            double dy[H34];
#pragma unroll
            for (int u = 0; u < H34; u++) dy[u] = ld_const(a.ket_d + dofs + u);
            double jy[H34];
#pragma unroll
            for (int u = 0; u < H34; u++) jy[u] = 0.0;
            if (takeA || takeB) {
                nqa += takeA ? 1 : 0;
                nqb += takeB ? 1 : 0;
                nlanes++;
                // one integral in the arithmetic Real, both uses; the A part lands in accA
                auto both = [&](auto* R, auto* accA, const void* tab) __attribute__((always_inline)) {
                    using Real = typename std::remove_reference<decltype(R[0])>::type;
                    hermite_R<Real, L>(a.rys, tab, a.ew, p12, c12, q, c34, X, Y, Z,
                                       *reinterpret_cast<Real(*)[HL]>(R));
                    const Real wa = takeA ? Real(1) : Real(0), wb = takeB ? Real(1) : Real(0);
                    static_for<H34>([&](auto U_) {
                        constexpr int u = decltype(U_)::value;
                        constexpr int ux = herm_tx(u), uy = herm_ty(u), uz = herm_tz(u);
                        const Real du = wa * (Real)dy[u];
                        Real s = Real(0);
                        static_for<H12>([&](auto T_) {
                            constexpr int tt = decltype(T_)::value;
                            constexpr int idx = herm_index(herm_tx(tt) + ux, herm_ty(tt) + uy, herm_tz(tt) + uz);
                            accA[tt] = fma(R[idx], du, accA[tt]);
                            // (y|x) at Q - P = -(P - Q): R_{t+u}(-X) = (-1)^{|t|+|u|} R_{t+u}(X)
                            if constexpr (((herm_tx(tt) + herm_ty(tt) + herm_tz(tt) + ux + uy + uz) & 1) != 0)
                                s = fma(-R[idx], (Real)dx[tt], s);
                            else
                                s = fma(R[idx], (Real)dx[tt], s);
                        });
                        jy[u] = (double)(wb * s);
                    });
                };
                if (MIXED && step32) {
                    float R[HL], ja[H12];
#pragma unroll
                    for (int u = 0; u < H12; u++) ja[u] = 0.0f;
                    both(R, ja, s_tab32);
#pragma unroll
                    for (int u = 0; u < H12; u++) jp[u] += (double)ja[u];
                    nq32a += takeA ? 1 : 0;
                    nq32b += takeB ? 1 : 0;
                } else {
                    double R[HL];
                    both(R, jp, s_tab64);
                }
            }
            if (mB != 0ull) {
                // sum of the B contributions over the wave (all lanes active here: the ket loop
                // and this test are wave-uniform); lane 63 adds them to J_y
#pragma unroll
                for (int u = 0; u < H34; u++) jy[u] = wave_sum_last(jy[u]);
                if (lane == 63) {
                    double* dst = a.jq + dofs;
#pragma unroll
                    for (int u = 0; u < H34; u++)
                        if (jy[u] != 0.0) atomic_add(dst + u, jy[u]);
                }
            }
        }
        if (bra_ok) {
            double* dst = a.jp + (size_t)(x - a.bra_hstart) * H12;
#pragma unroll
            for (int u = 0; u < H12; u++)
                if (jp[u] != 0.0) atomic_add(dst + u, jp[u]);
        }
    }
    if (a.count) {
        for (int o = 32; o > 0; o >>= 1) {
            nqa += __shfl_xor(nqa, o);
            nqb += __shfl_xor(nqb, o);
            nlanes += __shfl_xor(nlanes, o);
        }
        if (lane == 0) {
            if (nqa) atomicAdd(&a.sc->counts[a.count_cls], nqa);
            if (nqb) atomicAdd(&a.sc->counts[a.count_cls2], nqb);
            if (nlanes) {
                atomicAdd(&a.sc->wsteps[a.count_cls], (unsigned long long)nsteps);
                atomicAdd(&a.sc->wlanes[a.count_cls], nlanes);
            }
        }
        if constexpr (MIXED) {
            for (int o = 32; o > 0; o >>= 1) {
                nq32a += __shfl_xor(nq32a, o);
                nq32b += __shfl_xor(nq32b, o);
            }
            if (lane == 0) {
                if (nq32a) atomicAdd(&a.sc->counts32[a.count_cls], nq32a);
                if (nq32b) atomicAdd(&a.sc->counts32[a.count_cls2], nq32b);
            }
        }
    }
}

void launch_j_quartets_sym(int L12, int L34, int route, const JQuartetArgs& a, const JTask* tasks, int ntask,
                           unsigned int* qctr, hipStream_t s)
{
    if (ntask <= 0) return;
#define ONE(X, Y, R)                                                                                     \
    {                                                                                                    \
        const size_t lds = JLds<X, Y, R>::bytes(a.rys);                                                  \
        const int grid = persistent_grid((const void*)j_quartets_sym<X, Y, R>, 256, lds, ((long long)ntask + 3) / 4); \
        hipLaunchKernelGGL((j_quartets_sym<X, Y, R>), dim3(grid), dim3(256), lds, s, a, tasks, ntask, qctr); \
        return;                                                                                          \
    }
#define CASE(X, Y)                                                                                       \
    if (L12 == X && L34 == Y) {                                                                          \
        if (route == kRouteMixed) ONE(X, Y, kRouteMixed) else ONE(X, Y, kRoutePlain)                    \
    }
    CASE(0, 0) CASE(1, 0) CASE(2, 0) CASE(3, 0) CASE(4, 0)
#undef CASE
#undef ONE
    throw std::runtime_error("lsjk: j_quartets_sym class not instantiated");
}

// ------------------------------------------------------------------ gradient (bra centres)
// dE/dA_x = sum_ab Dsym_ab K_ab sum_t [2 alpha E^{a+1x,b}_tx - a_x E^{a-1x,b}_tx] E_ty E_tz V_t, the
// derivative of the primitive, d phi_a / dA_x = 2 alpha phi_{a+1x} - a_x phi_{a-1x}, expanded in
// Hermite Gaussians at P (t <= L12 + 1); likewise dE/dB with beta and b.  Dsym is the bra density
// block symmetrised as in the J bra transform (both triangles, 1/2 for a primitive with itself).
// The sum over both bra and ket roles gives the full gradient (reference coulomb_grad.cpp:19-563).
template <int LA, int LB>
__global__ __launch_bounds__(256) void j_grad_scatter(JGradArgs a, int k0, int nk, int hstart, size_t hoff)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), L = LA + LB, H = nherm(L + 1);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nk) return;
    const int k = k0 + i;
    const float max34 = __uint_as_float(a.sc->max34_bits);
    const float thre_bra = a.thresp / max34;
    if (!(a.bound[k] >= thre_bra)) return;
    const int ia = a.p1[k], ib = a.p2[k];
    const double* vp = a.vp + hoff + (size_t)(a.hpos[k] - hstart) * H;
    double V[H];
#pragma unroll
    for (int t = 0; t < H; t++) V[t] = vp[t];
    const double eA = a.eA[ia], eB = a.eB[ib];
    const double p = eA + eB, ip = 1.0 / p;
    const double Ax = a.xA[ia], Ay = a.yA[ia], Az = a.zA[ia];
    const double Bx = a.xB[ib], By = a.yB[ib], Bz = a.zB[ib];
    const double Px = (eA * Ax + eB * Bx) * ip, Py = (eA * Ay + eB * By) * ip, Pz = (eA * Az + eB * Bz) * ip;
    HermE<LA + 1, LB + 1> Ex, Ey, Ez;
    hermite_E<LA + 1, LB + 1>(0.5 * ip, Px - Ax, Px - Bx, Ex);
    hermite_E<LA + 1, LB + 1>(0.5 * ip, Py - Ay, Py - By, Ey);
    hermite_E<LA + 1, LB + 1>(0.5 * ip, Pz - Az, Pz - Bz, Ez);
    const double scale = (a.symmetric && ia == ib) ? 0.5 : 1.0;
    const int ca = a.cartA[ia], cb = a.cartB[ib];
    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
    // one-axis factor of the shifted function: sum over t of the axis, for exponents (i1, j1)
    static_for<NA>([&](auto P_) {
        constexpr int pa = decltype(P_)::value;
        constexpr int ax = cart_lx(LA, pa), ay = cart_ly(LA, pa), az = cart_lz(LA, pa);
        static_for<NB>([&](auto Q_) {
            constexpr int qb = decltype(Q_)::value;
            constexpr int bx = cart_lx(LB, qb), by = cart_ly(LB, qb), bz = cart_lz(LB, qb);
            double Dv = a.D[(size_t)(ca + pa) * a.ldD + (cb + qb)];
            if (a.symmetric) Dv = (Dv + a.D[(size_t)(cb + qb) * a.ldD + (ca + pa)]) * scale;
            if (Dv == 0.0) return;
            double dA[3] = {0.0, 0.0, 0.0}, dB[3] = {0.0, 0.0, 0.0};
            static_for<H>([&](auto T_) {
                constexpr int t = decltype(T_)::value;
                constexpr int tx = herm_tx(t), ty = herm_ty(t), tz = herm_tz(t);
                const double v = V[t];
                // plain factors (t within reach of the unshifted exponents)
                const double ex = (tx <= ax + bx) ? Ex.v[tx][ax][bx] : 0.0;
                const double ey = (ty <= ay + by) ? Ey.v[ty][ay][by] : 0.0;
                const double ez = (tz <= az + bz) ? Ez.v[tz][az][bz] : 0.0;
                // shifted factors: 2 alpha E^{a+1,b} - a E^{a-1,b} (bra A) and 2 beta E^{a,b+1} - b E^{a,b-1}
                auto shiftA = [&](const HermE<LA + 1, LB + 1>& E, int tt, int i1, int j1) __attribute__((always_inline)) {
                    double s = (tt <= i1 + 1 + j1) ? 2.0 * eA * E.v[tt][i1 + 1][j1] : 0.0;
                    if (i1 > 0 && tt <= i1 - 1 + j1) s -= i1 * E.v[tt][i1 - 1][j1];
                    return s;
                };
                auto shiftB = [&](const HermE<LA + 1, LB + 1>& E, int tt, int i1, int j1) __attribute__((always_inline)) {
                    double s = (tt <= i1 + j1 + 1) ? 2.0 * eB * E.v[tt][i1][j1 + 1] : 0.0;
                    if (j1 > 0 && tt <= i1 + j1 - 1) s -= j1 * E.v[tt][i1][j1 - 1];
                    return s;
                };
                dA[0] += shiftA(Ex, tx, ax, bx) * ey * ez * v;
                dA[1] += ex * shiftA(Ey, ty, ay, by) * ez * v;
                dA[2] += ex * ey * shiftA(Ez, tz, az, bz) * v;
                dB[0] += shiftB(Ex, tx, ax, bx) * ey * ez * v;
                dB[1] += ex * shiftB(Ey, ty, ay, by) * ez * v;
                dB[2] += ex * ey * shiftB(Ez, tz, az, bz) * v;
            });
#pragma unroll
            for (int d = 0; d < 3; d++) {
                gA[d] += Dv * dA[d];
                gB[d] += Dv * dB[d];
            }
        });
    });
    const double Kab = a.K[k];
    const int atA = a.atomA[ia], atB = a.atomB[ib];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        atomic_add(a.G + 3 * atA + d, Kab * gA[d]);
        atomic_add(a.G + 3 * atB + d, Kab * gB[d]);
    }
}

// ------------------------------------------------------------------ bra scatter
template <int LA, int LB>
__global__ __launch_bounds__(256) void j_bra_scatter(JScatterArgs a, int k0, int nk, int hstart, size_t hoff)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), L = LA + LB, H = nherm(L);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nk) return;
    const int k = k0 + i;
    if (a.screen) {
        const float max34 = __uint_as_float(a.sc->max34_bits);
        const float thre_bra = a.thresp / max34;
        if (!(a.bound[k] >= thre_bra)) return;
    }
    const int ia = a.p1[k], ib = a.p2[k];
    const int pos = a.hpos[k];
    const double* jp = a.jp + hoff + (size_t)(pos - hstart) * H;
    double jt[H];
    bool any = false;
#pragma unroll
    for (int t = 0; t < H; t++) {
        jt[t] = jp[t];
        any |= (jt[t] != 0.0);
    }
    if (!any) return;
    const double eA = a.eA[ia], eB = a.eB[ib];
    const double p = eA + eB, ip = 1.0 / p;
    const double Ax = a.xA[ia], Ay = a.yA[ia], Az = a.zA[ia];
    const double Bx = a.xB[ib], By = a.yB[ib], Bz = a.zB[ib];
    const double Px = (eA * Ax + eB * Bx) * ip, Py = (eA * Ay + eB * By) * ip, Pz = (eA * Az + eB * Bz) * ip;
    HermE<LA, LB> Ex, Ey, Ez;
    hermite_E<LA, LB>(0.5 * ip, Px - Ax, Px - Bx, Ex);
    hermite_E<LA, LB>(0.5 * ip, Py - Ay, Py - By, Ey);
    hermite_E<LA, LB>(0.5 * ip, Pz - Az, Pz - Bz, Ez);
    double Kab = a.K[k];
    if (a.symmetric && ia == ib) Kab *= 0.5;
    const int ca = a.cartA[ia], cb = a.cartB[ib];
    static_for<NA>([&](auto P_) {
        constexpr int p = decltype(P_)::value;
        constexpr int ax = cart_lx(LA, p), ay = cart_ly(LA, p), az = cart_lz(LA, p);
        static_for<NB>([&](auto Q_) {
            constexpr int q = decltype(Q_)::value;
            constexpr int bx = cart_lx(LB, q), by = cart_ly(LB, q), bz = cart_lz(LB, q);
            double s = 0.0;
            static_for<H>([&](auto T_) {
                constexpr int t = decltype(T_)::value;
                constexpr int tx = herm_tx(t), ty = herm_ty(t), tz = herm_tz(t);
                if constexpr (tx <= ax + bx && ty <= ay + by && tz <= az + bz)
                    s += Ex.v[tx][ax][bx] * Ey.v[ty][ay][by] * Ez.v[tz][az][bz] * jt[t];
            });
            s *= Kab;
            atomic_add(a.J + (size_t)(ca + p) * a.ldJ + (cb + q), s);
            if (a.symmetric) atomic_add(a.J + (size_t)(cb + q) * a.ldJ + (ca + p), s);
        });
    });
}

// ------------------------------------------------------------------ launchers
#define LSJK_L2_CASES(F) F(0, 0) F(0, 1) F(0, 2) F(1, 0) F(1, 1) F(1, 2) F(2, 0) F(2, 1) F(2, 2)

void launch_j_ket_prep(int LA, int LB, const JPrepArgs& a, int k0, int nk, int hstart, size_t hoff, hipStream_t s)
{
    if (nk <= 0) return;
    dim3 grid((nk + 255) / 256), block(256);
#define CASE(X, Y) \
    if (LA == X && LB == Y) { hipLaunchKernelGGL((j_ket_prep<X, Y>), grid, block, 0, s, a, k0, nk, hstart, hoff); return; }
    LSJK_L2_CASES(CASE)
#undef CASE
}

void launch_j_bra_scatter(int LA, int LB, const JScatterArgs& a, int k0, int nk, int hstart, size_t hoff, hipStream_t s)
{
    if (nk <= 0) return;
    dim3 grid((nk + 255) / 256), block(256);
#define CASE(X, Y) \
    if (LA == X && LB == Y) { hipLaunchKernelGGL((j_bra_scatter<X, Y>), grid, block, 0, s, a, k0, nk, hstart, hoff); return; }
    LSJK_L2_CASES(CASE)
#undef CASE
}

void launch_j_grad_scatter(int LA, int LB, const JGradArgs& a, int k0, int nk, int hstart, size_t hoff,
                           hipStream_t s)
{
    if (nk <= 0) return;
    dim3 grid((nk + 255) / 256), block(256);
#define CASE(X, Y)                                                                                  \
    if (LA == X && LB == Y) {                                                                       \
        hipLaunchKernelGGL((j_grad_scatter<X, Y>), grid, block, 0, s, a, k0, nk, hstart, hoff);     \
        return;                                                                                     \
    }
    CASE(0, 0) CASE(0, 1) CASE(0, 2) CASE(1, 0) CASE(1, 1) CASE(1, 2) CASE(2, 0) CASE(2, 1) CASE(2, 2)
#undef CASE
}

template <int L12, int L34, int ROUTE>
static void launch_jq(const JQuartetArgs& a, const JTask* tasks, int ntask, unsigned int* qctr, hipStream_t s)
{
    const size_t lds = JLds<L12, L34, ROUTE>::bytes(a.rys);
    const int grid = persistent_grid((const void*)j_quartets<L12, L34, ROUTE>, 256, lds, ((long long)ntask + 3) / 4);
    hipLaunchKernelGGL((j_quartets<L12, L34, ROUTE>), dim3(grid), dim3(256), lds, s, a, tasks, ntask, qctr);
}

void launch_j_quartets(int L12, int L34, int route, const JQuartetArgs& a, const JTask* tasks, int ntask,
                       unsigned int* qctr, hipStream_t s)
{
    if (ntask <= 0) return;
#define CASE(X, Y)                                                                     \
    if (L12 == X && L34 == Y) {                                                        \
        if (route == kRouteMixed)                                                      \
            launch_jq<X, Y, kRouteMixed>(a, tasks, ntask, qctr, s);                   \
        else                                                                           \
            launch_jq<X, Y, kRoutePlain>(a, tasks, ntask, qctr, s);                   \
        return;                                                                        \
    }
#define ROW(X) CASE(X, 0) CASE(X, 1) CASE(X, 2) CASE(X, 3) CASE(X, 4)
    ROW(0) ROW(1) ROW(2) ROW(3) ROW(4)
#undef ROW
#undef CASE
}

// Bra order L12 + 1 (gradients): FP64 route only.
void launch_j_quartets_grad(int L12p1, int L34, const JQuartetArgs& a, const JTask* tasks, int ntask,
                            unsigned int* qctr, hipStream_t s)
{
    if (ntask <= 0) return;
#define CASE(X, Y)                                                               \
    if (L12p1 == X && L34 == Y) {                                                \
        launch_jq<X, Y, kRoutePlain>(a, tasks, ntask, qctr, s);                 \
        return;                                                                  \
    }
#define ROW(X) CASE(X, 0) CASE(X, 1) CASE(X, 2) CASE(X, 3) CASE(X, 4)
    ROW(1) ROW(2) ROW(3) ROW(4) ROW(5)
#undef ROW
#undef CASE
}

}  // namespace lsjk

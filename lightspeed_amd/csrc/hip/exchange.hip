// Direct exchange (K) build on gfx950.
//
// K13 += sum_{2,4} (12|34) D24 over the quartets the reference evaluates
// (IntBox::exchangeCPU, reference src/cpp/intbox/cpu/exchange.cpp:17-355):
//   * pairs (1,2) and (3,4) taken from the K-invert partner lists with b >= thresp/(maxb*D24max)
//     (KBlock::build21 / build_symmetric, src/cpp/core/kvec.cpp:128-232; exchange.cpp:54-71),
//   * quartet sieve float(b12*b34) * max|D24 block| >= thresp (exchange.cpp:145-152, 186).
// The reference loops pivot pairs (2,4) and scatters each quartet's (1,3) block with atomics.  Here
// the loop is output-stationary: one wave owns one (1,3) primitive pair, its lanes sweep the
// (2,4) partner combinations, accumulate the N1 x N3 block in registers, and the wave reduces it
// before one atomic add per element.  With a symmetric D24 and one symmetric list on both sides
// only row1 <= row3 is computed and the transposed block is added as well (K = K^T), the same 2x
// the reference's D24symm half-sieve (exchange.cpp:138-139, 326-336) obtains.
//
// Integrals: Rys quadrature with the 2-D integrals built by the vertical recurrence on centres
// A and C and the horizontal transfer to B and D (Rys, Dupuis & King, J. Comput. Chem. 4, 154
// (1983)), in registers, for each root x_i = t_i^2 (attenuated as t^2 d^2 for erf operators).
#include "internal.hpp"
#include "kernels.hpp"

namespace lsjk {

template <int L1, int L2, int L3, int L4>
struct Axis4 {
    double v[L1 + 1][L2 + 1][L3 + 1][L4 + 1];
};

template <int L1, int L2, int L3, int L4>
__device__ __forceinline__ void rys_axis(double g00, double C00, double D00, double B10, double B01, double B00,
                                         double AB, double CD, Axis4<L1, L2, L3, L4>& I)
{
    constexpr int NX = L1 + L2 + 1, MX = L3 + L4 + 1;
    double G[NX][MX];
    G[0][0] = g00;
    if (NX > 1) G[1][0] = C00 * g00;
#pragma unroll
    for (int n = 1; n < NX - 1; n++) G[n + 1][0] = C00 * G[n][0] + n * B10 * G[n - 1][0];
#pragma unroll
    for (int m = 0; m < MX - 1; m++) {
#pragma unroll
        for (int n = 0; n < NX; n++) {
            double v = D00 * G[n][m];
            if (m > 0) v += m * B01 * G[n][m - 1];
            if (n > 0) v += n * B00 * G[n - 1][m];
            G[n][m + 1] = v;
        }
    }
    // ket transfer: T[n][c][d] = (n 0 | c d)
    double T[NX][MX][L4 + 1];
#pragma unroll
    for (int n = 0; n < NX; n++)
#pragma unroll
        for (int c = 0; c < MX; c++) T[n][c][0] = G[n][c];
#pragma unroll
    for (int d = 0; d < L4; d++)
#pragma unroll
        for (int n = 0; n < NX; n++)
#pragma unroll
            for (int c = 0; c < MX - 1 - d; c++) T[n][c][d + 1] = T[n][c + 1][d] + CD * T[n][c][d];
    // bra transfer
#pragma unroll
    for (int c = 0; c <= L3; c++)
#pragma unroll
        for (int d = 0; d <= L4; d++) {
            double U[NX][L2 + 1];
#pragma unroll
            for (int a = 0; a < NX; a++) U[a][0] = T[a][c][d];
#pragma unroll
            for (int b = 0; b < L2; b++)
#pragma unroll
                for (int a = 0; a < NX - 1 - b; a++) U[a][b + 1] = U[a + 1][b] + AB * U[a][b];
#pragma unroll
            for (int a = 0; a <= L1; a++)
#pragma unroll
                for (int b = 0; b <= L2; b++) I.v[a][b][c][d] = U[a][b];
        }
}

struct KEnt {
    double p, Px, Py, Pz, K, Bx, By, Bz;
};

__device__ __forceinline__ KEnt load_ent(const double* eg, int n, int i)
{
    KEnt e;
    e.p = eg[i];
    e.Px = eg[n + i];
    e.Py = eg[2 * n + i];
    e.Pz = eg[3 * n + i];
    e.K = eg[4 * n + i];
    e.Bx = eg[5 * n + i];
    e.By = eg[6 * n + i];
    e.Bz = eg[7 * n + i];
    return e;
}

// One primitive quartet: acc[a][c] += sum_{b,d} (ab|cd) Db[b][d].
template <int L1, int L2, int L3, int L4>
__device__ __forceinline__ void k_quartet(const RysTable& tab, const double* s_rys, const Ewald& ew, double Ax,
                                          double Ay, double Az,
                                          double Cx, double Cy, double Cz, const KEnt& e12, const KEnt& e34,
                                          const double (&Db)[ncart(L2)][ncart(L4)],
                                          double (&acc)[ncart(L1)][ncart(L3)])
{
    constexpr int N1 = ncart(L1), N2 = ncart(L2), N3 = ncart(L3), N4 = ncart(L4);
    constexpr int NR = (L1 + L2 + L3 + L4) / 2 + 1;
    const double p = e12.p, q = e34.p;
    const double pq = p + q;
    const double ipq = 1.0 / pq;
    const double rho = p * q * ipq;
    const double PQx = e12.Px - e34.Px, PQy = e12.Py - e34.Py, PQz = e12.Pz - e34.Pz;
    const double r2 = PQx * PQx + PQy * PQy + PQz * PQz;
    const double base = 2.0 * 17.493418327624862 / (p * q * sqrt(pq)) * e12.K * e34.K;
    const double PAx = e12.Px - Ax, PAy = e12.Py - Ay, PAz = e12.Pz - Az;
    const double QCx = e34.Px - Cx, QCy = e34.Py - Cy, QCz = e34.Pz - Cz;
    const double ABx = Ax - e12.Bx, ABy = Ay - e12.By, ABz = Az - e12.Bz;
    const double CDx = Cx - e34.Bx, CDy = Cy - e34.By, CDz = Cz - e34.Bz;
    const double qf = q * ipq, pf = p * ipq;  // rho/p, rho/q
    const double i2p = 0.5 / p, i2q = 0.5 / q;
    for (int it = 0; it < ew.n; it++) {
        const double om = ew.omega[it];
        double d2 = 1.0;
        double pref = ew.alpha[it] * base;
        if (om != -1.0) {
            d2 = om * om / (rho + om * om);
            pref *= sqrt(d2);
        }
        double x[NR], w[NR];
        rys_roots_any<NR>(tab, s_rys, rho * d2 * r2, x, w);
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const double u = x[i] * d2;
            const double B00 = 0.5 * u * ipq;
            const double B10 = i2p * (1.0 - u * qf);
            const double B01 = i2q * (1.0 - u * pf);
            const double cq = u * qf, cp = u * pf;
            Axis4<L1, L2, L3, L4> Ix, Iy, Iz;
            rys_axis<L1, L2, L3, L4>(pref * w[i], PAx - cq * PQx, QCx + cp * PQx, B10, B01, B00, ABx, CDx, Ix);
            rys_axis<L1, L2, L3, L4>(1.0, PAy - cq * PQy, QCy + cp * PQy, B10, B01, B00, ABy, CDy, Iy);
            rys_axis<L1, L2, L3, L4>(1.0, PAz - cq * PQz, QCz + cp * PQz, B10, B01, B00, ABz, CDz, Iz);
            static_for<N1>([&](auto A_) {
                constexpr int a = decltype(A_)::value;
                constexpr int ax = cart_lx(L1, a), ay = cart_ly(L1, a), az = cart_lz(L1, a);
                static_for<N3>([&](auto C_) {
                    constexpr int c = decltype(C_)::value;
                    constexpr int cx = cart_lx(L3, c), cy = cart_ly(L3, c), cz = cart_lz(L3, c);
                    double s = 0.0;
                    static_for<N2>([&](auto B_) {
                        constexpr int b = decltype(B_)::value;
                        constexpr int bx = cart_lx(L2, b), by = cart_ly(L2, b), bz = cart_lz(L2, b);
                        static_for<N4>([&](auto D_) {
                            constexpr int d = decltype(D_)::value;
                            constexpr int dx = cart_lx(L4, d), dy = cart_ly(L4, d), dz = cart_lz(L4, d);
                            s = fma(Ix.v[ax][bx][cx][dx] * Iy.v[ay][by][cy][dy] * Iz.v[az][bz][cz][dz], Db[b][d], s);
                        });
                    });
                    acc[a][c] += s;
                });
            });
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

// Per-wave LDS of k_tiles: the tile accumulator (kKTile x N1 x N3 doubles) and the work queue
// (3 x kKQueue ints).
__host__ __device__ constexpr size_t k_wave_lds_bytes(int N13) { return (size_t)kKTile * N13 * 8 + 3 * kKQueue * 4; }

// Persistent kernel over the tiles of one class (L1 L2 | L3 L4).  A tile is (row1, block of
// kKTile row3s); tiles are numbered row1-major and split into 8 contiguous partitions, one per
// XCD, so that the waves of one XCD (one L2) work on neighbouring rows of D and K; a wave drains
// its own partition's queue first and then helps the others.  The workgroup's waves share the
// LDS copy of the Rys table (when rys_staged(nrys)).
// For each unit (row1, row3) of a tile the lanes test the sieve over the candidate rectangle
// (partners that can pass with the global D maximum), compact the passing (2,4) combinations
// into the wave's LDS queue, and evaluate the quartets 64 at a time; each quartet's N1 x N3 block
// is added to the tile's LDS accumulator, which is flushed to K once per tile (plus the
// transposed block in the symmetric mode).
template <int L1, int L2, int L3, int L4>
__global__ __launch_bounds__(512) void k_tiles(KArgs a, int nb3, long long nunits, unsigned int* qctr)
{
    constexpr int N1 = ncart(L1), N2 = ncart(L2), N3 = ncart(L3), N4 = ncart(L4), N13 = N1 * N3;
    constexpr int NR = (L1 + L2 + L3 + L4) / 2 + 1;
    extern __shared__ double s_mem[];
    const int nw = blockDim.x >> 6, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t ntab = rys_staged(NR) ? rys_lds_doubles(a.rys.nbin[NR], NR) : 0;
    double* s_rys = s_mem;
    double* s_acc = s_mem + ntab + (size_t)wave * (kKTile * N13);
    int* s_e2 = reinterpret_cast<int*>(s_mem + ntab + (size_t)nw * (kKTile * N13)) + wave * 3 * kKQueue;
    int* s_e4 = s_e2 + kKQueue;
    int* s_t = s_e4 + kKQueue;
    if constexpr (rys_staged(NR)) {
        rys_stage_lds<NR>(a.rys, s_rys);
        __syncthreads();
    }

    const float thresp = a.thresp;
    const double dmax = __longlong_as_double((long long)a.sc->dmax_bits);
    const float thre12 = (float)((double)thresp / ((double)a.maxb34 * dmax));
    const float thre34 = (float)((double)thresp / ((double)a.maxb12 * dmax));
    unsigned long long nq = 0;
    const int xcd = blockIdx.x & 7;

    for (int sweep = 0; sweep < 8; sweep++) {
        const int part = (xcd + sweep) & 7;
        const long long lo = nunits * part / 8, hi = nunits * (part + 1) / 8;
        for (;;) {
            const long long unit = lo + (long long)wave_take(qctr + part);
            if (unit >= hi) break;
            const int i1 = (int)(unit / nb3);
            const int ib = (int)(unit - (long long)i1 * nb3);
            const int r1 = a.s12.rows[i1];
            const int j3a = ib * kKTile;
            const int j3b = min(j3a + kKTile, a.s34.nrows);
            if (a.tri && a.s34.rows[j3b - 1] < r1) continue;  // rows are in primitive order
            for (int q = lane; q < kKTile * N13; q += 64) s_acc[q] = 0.0;
            __syncwarp();

            const int a0 = a.s12.row_start[r1 * (a.s12.maxLp + 1) + L2];
            const int a1 = a.s12.row_start[r1 * (a.s12.maxLp + 1) + L2 + 1];
            const float b12top = a.s12.eb[a0];
            const double Ax = a.s12.x[r1], Ay = a.s12.y[r1], Az = a.s12.z[r1];
            // partners of row1 that can pass with the largest possible ket bound are a prefix
            int n2max = 0;
            for (int k = lane; k < a1 - a0; k += 64) n2max += (a.s12.eb[a0 + k] >= thre12) ? 1 : 0;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) n2max += __shfl_xor(n2max, o);

            int qn = 0;  // queue length (wave-uniform)
            auto evaluate = [&](int slot) {
                const int e2 = s_e2[slot], e4 = s_e4[slot], tt = s_t[slot];
                const int r3 = a.s34.rows[j3a + tt];
                const int o2 = a.s12.ecart[e2], o4 = a.s34.ecart[e4];
                double Db[N2][N4];
#pragma unroll
                for (int b = 0; b < N2; b++)
#pragma unroll
                    for (int d = 0; d < N4; d++) Db[b][d] = a.D[(size_t)(o2 + b) * a.ldD + (o4 + d)];
                const KEnt E12 = load_ent(a.s12.eg, a.s12.nent, e2);
                const KEnt E34 = load_ent(a.s34.eg, a.s34.nent, e4);
                double acc[N1][N3];
#pragma unroll
                for (int i = 0; i < N1; i++)
#pragma unroll
                    for (int j = 0; j < N3; j++) acc[i][j] = 0.0;
                k_quartet<L1, L2, L3, L4>(a.rys, s_rys, a.ew, Ax, Ay, Az, a.s34.x[r3], a.s34.y[r3], a.s34.z[r3], E12,
                                          E34, Db, acc);
#pragma unroll
                for (int i = 0; i < N1; i++)
#pragma unroll
                    for (int j = 0; j < N3; j++) atomicAdd(&s_acc[tt * N13 + i * N3 + j], acc[i][j]);
                nq++;
            };

            for (int t = j3a; t < j3b && n2max > 0; t++) {
                const int r3 = a.s34.rows[t];
                if (a.tri && r3 < r1) continue;
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
                            pass = (double)(b12 * b34) * dblk >= (double)thresp;
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
                        __syncwarp();
                        evaluate(lane);
                        __syncwarp();
                        const int rest = qn - 64;
                        int m2 = 0, m4 = 0, mt = 0;
                        if (lane < rest) {
                            m2 = s_e2[64 + lane];
                            m4 = s_e4[64 + lane];
                            mt = s_t[64 + lane];
                        }
                        __syncwarp();
                        if (lane < rest) {
                            s_e2[lane] = m2;
                            s_e4[lane] = m4;
                            s_t[lane] = mt;
                        }
                        __syncwarp();
                        qn = rest;
                    }
                }
            }
            __syncwarp();
            if (lane < qn) evaluate(lane);
            __syncwarp();
            // flush the tile accumulator
            const int o1 = a.s12.cart[r1];
            for (int q = lane; q < (j3b - j3a) * N13; q += 64) {
                const double v = s_acc[q];
                if (v == 0.0) continue;
                const int tt = q / N13, ij = q - tt * N13, i = ij / N3, j = ij - i * N3;
                const int r3 = a.s34.rows[j3a + tt];
                const int o3 = a.s34.cart[r3];
                atomic_add(a.K + (size_t)(o1 + i) * a.ldK + (o3 + j), v);
                if (a.symmetric && r1 != r3) atomic_add(a.K + (size_t)(o3 + j) * a.ldK + (o1 + i), v);
            }
            __syncwarp();
        }
    }
    if (a.count) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) nq += __shfl_xor(nq, o);
        if (lane == 0 && nq) atomicAdd(&a.sc->counts[kcount_index(L1, L2, L3, L4)], nq);
    }
}

// Block maxima of |D24| per (block2, block4) and the global maximum.
__global__ void k_blockmax(const double* D, int ldD, const int* bc2, const int* bL2, int nb2, const int* bc4,
                           const int* bL4, int nb4, double* Dsh, CallScalars* sc)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nb2 * nb4) return;
    const int i2 = (int)(i / nb4), i4 = (int)(i - (long long)i2 * nb4);
    const int o2 = bc2[i2], o4 = bc4[i4];
    const int n2 = ncart(bL2[i2]), n4 = ncart(bL4[i4]);
    double m = 0.0;
    for (int b = 0; b < n2; b++)
        for (int d = 0; d < n4; d++) m = fmax(m, fabs(D[(size_t)(o2 + b) * ldD + (o4 + d)]));
    Dsh[i] = m;
    atomicMax(&sc->dmax_bits, (unsigned long long)__double_as_longlong(m));
}

__global__ void fill_zero(unsigned int* p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

void launch_fill(void* p, size_t bytes, hipStream_t s)
{
    size_t n = bytes / 4;
    if (n == 0) return;
    int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_zero, dim3(blocks), dim3(256), 0, s, (unsigned int*)p, n);
}

// Plain copy, 16 bytes per lane per step (grid-stride); the tail below 16 bytes byte-wise.
__global__ void copy_bytes(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n16, unsigned char* dtail,
                           const unsigned char* stail, int ntail)
{
    const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t i = i0; i < n16; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
    if (i0 < (size_t)ntail) dtail[i0] = stail[i0];
}

void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    if (bytes == 0) return;
    const size_t n16 = bytes / 16;
    const int ntail = (int)(bytes % 16);
    int blocks = (int)std::max<size_t>(1, std::min<size_t>((n16 + 255) / 256, 8192));
    hipLaunchKernelGGL(copy_bytes, dim3(blocks), dim3(256), 0, s, (uint4*)dst, (const uint4*)src, n16,
                       (unsigned char*)dst + n16 * 16, (const unsigned char*)src + n16 * 16, ntail);
}

void launch_k_blockmax(const double* D, int ldD, const int* bc2, const int* bL2, int nb2, const int* bc4,
                       const int* bL4, int nb4, double* Dsh, CallScalars* sc, hipStream_t s)
{
    long long n = (long long)nb2 * nb4;
    if (n == 0) return;
    hipLaunchKernelGGL(k_blockmax, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, D, ldD, bc2, bL2, nb2, bc4,
                       bL4, nb4, Dsh, sc);
}

template <int L1, int L2, int L3, int L4>
static void launch_kt(const KArgs& a, long long nunits, unsigned int* qctr, hipStream_t s)
{
    constexpr int NR = (L1 + L2 + L3 + L4) / 2 + 1, N13 = ncart(L1) * ncart(L3);
    const int nb3 = (a.s34.nrows + kKTile - 1) / kKTile;
    const size_t tab = rys_staged(NR) ? rys_lds_doubles(a.rys.nbin[NR], NR) * sizeof(double) : 0;
    // waves per workgroup: the count that keeps the most waves resident (the table is shared by
    // the workgroup's waves; the accumulators and queues are per wave)
    int best_nw = 4, best_res = -1, best_grid = 1;
    for (int nw : {8, 4, 2, 1}) {
        const size_t lds = tab + (size_t)nw * k_wave_lds_bytes(N13);
        int grid = 0;
        try {
            grid = persistent_grid((const void*)k_tiles<L1, L2, L3, L4>, 64 * nw, lds, 1LL << 40);
        } catch (...) {
            continue;  // LDS above the per-workgroup limit at this size
        }
        if (grid * nw > best_res) {
            best_res = grid * nw;
            best_nw = nw;
            best_grid = grid;
        }
    }
    if (best_res < 0) throw std::runtime_error("lsjk: exchange kernel does not fit the device's LDS");
    const size_t lds = tab + (size_t)best_nw * k_wave_lds_bytes(N13);
    const long long need = (nunits + best_nw - 1) / best_nw;
    const int grid = (int)std::max<long long>(1, std::min<long long>(best_grid, need));
    hipLaunchKernelGGL((k_tiles<L1, L2, L3, L4>), dim3(grid), dim3(64 * best_nw), lds, s, a, nb3, nunits, qctr);
}

void launch_k_quartets(int L1, int L2, int L3, int L4, const KArgs& a, long long nunits, unsigned int* qctr,
                       hipStream_t s)
{
    if (nunits <= 0) return;
#define CASE(A, B, C, D)                                   \
    if (L1 == A && L2 == B && L3 == C && L4 == D) {        \
        launch_kt<A, B, C, D>(a, nunits, qctr, s);         \
        return;                                            \
    }
#define C4(A, B, C) CASE(A, B, C, 0) CASE(A, B, C, 1) CASE(A, B, C, 2)
#define C3(A, B) C4(A, B, 0) C4(A, B, 1) C4(A, B, 2)
#define C2(A) C3(A, 0) C3(A, 1) C3(A, 2)
    C2(0) C2(1) C2(2)
#undef C2
#undef C3
#undef C4
#undef CASE
}

}  // namespace lsjk

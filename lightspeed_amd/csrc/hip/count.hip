// Quartet counting (no integrals) under the reference's exact screening conventions, the
// denominator of the quartets/s metric:
//   * J: (bra pair, ket pair) evaluations after truncate_ket / truncate_bra with
//     float(B12 * B34') >= thresp (reference src/cpp/intbox/cpu/coulomb.cpp:38-52, 115);
//   * K: (1,2,3,4) evaluations of the K-invert loops: pivot pairs (2,4) (with the D24symm
//     half-sieve K2 <= K4), partners b >= thresp/(maxb*D24max), and
//     float(b12 * b34) * max|D24 block| >= thresp (exchange.cpp:54-71, 124-139, 186).
// Counts are per (L1, L2, L3, L4) class, indexed ((L1*5+L2)*5+L3)*5+L4.
#include "internal.hpp"
#include "kernels.hpp"

namespace lsjk {

// One thread per bra pair of class (L1, L2) (list order); scans every ket pair of the ket list.
__global__ void count_j(CountJArgs a, int k0, int nk, int L1, int L2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nk) return;
    const int k = k0 + i;
    const float thresp = a.thresp;
    const float max34 = __uint_as_float(a.sc->max34_bits);
    const float thre_bra = thresp / max34;
    const float B12 = a.bra_bound[k];
    if (!(B12 >= thre_bra)) return;
    for (int c = 0; c < a.ket_ncls; c++) {
        unsigned long long n = 0;
        for (int j = a.ket_cls_start[c]; j < a.ket_cls_start[c + 1]; j++) {
            const float B34 = a.ket_b[a.ket_hpos[j]];
            if (B12 * B34 >= thresp) n++;
        }
        if (n) atomicAdd(&a.out[((L1 * 5 + L2) * 5 + a.ket_cls_L1[c]) * 5 + a.ket_cls_L2[c]], n);
    }
}

// One thread per pivot pair (2, 4): two-pointer count over the bound-sorted partner lists.
__global__ void count_k(CountKArgs a, long long npiv)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npiv) return;
    const int i2 = (int)(t / a.n4), i4 = (int)(t - (long long)i2 * a.n4);
    if (a.half && i2 > i4) return;
    const float thresp = a.thresp;
    const double dmax = __longlong_as_double((long long)a.sc->dmax_bits);
    const float thre12 = (float)((double)thresp / ((double)a.maxb34 * dmax));
    const float thre34 = (float)((double)thresp / ((double)a.maxb12 * dmax));
    const double dblk = a.Dsh[(size_t)a.blk2[i2] * a.ldDsh + a.blk4[i4]];
    const int L2 = a.L2[i2], L4 = a.L4[i4];
    for (int La = 0; La <= a.maxLa; La++) {
        const int s12 = a.rs12[i2 * (a.maxLa + 1) + La], e12 = a.rs12[i2 * (a.maxLa + 1) + La + 1];
        if (e12 <= s12) continue;
        for (int Lc = 0; Lc <= a.maxLc; Lc++) {
            const int s34 = a.rs34[i4 * (a.maxLc + 1) + Lc], e34 = a.rs34[i4 * (a.maxLc + 1) + Lc + 1];
            if (e34 <= s34) continue;
            unsigned long long n = 0;
            int m = e34 - s34;  // valid prefix length of the ket partners for the current bra partner
            for (int q = s12; q < e12 && m > 0; q++) {
                const float b12 = a.b12[q];
                if (!(b12 >= thre12)) break;
                while (m > 0) {
                    const float b34 = a.b34[s34 + m - 1];
                    if (b34 >= thre34 && (double)(b12 * b34) * dblk >= (double)thresp) break;
                    m--;
                }
                n += m;
            }
            if (n) atomicAdd(&a.out[((La * 5 + L2) * 5 + Lc) * 5 + L4], n);
        }
    }
}

void launch_count_j(const CountJArgs& a, int k0, int nk, int L1, int L2, hipStream_t s)
{
    if (nk <= 0) return;
    hipLaunchKernelGGL(count_j, dim3((nk + 255) / 256), dim3(256), 0, s, a, k0, nk, L1, L2);
}

void launch_count_k(const CountKArgs& a, hipStream_t s)
{
    long long n = (long long)a.n2 * a.n4;
    if (n <= 0) return;
    hipLaunchKernelGGL(count_k, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, n);
}

}  // namespace lsjk

// Exchange (K) build: small helper kernels (D block maxima, fills, copies) and the class
// dispatcher; the quartet kernels are in exchange_impl.hpp (instantiated by exchange_l<L1>.hip).
#include "internal.hpp"
#include "kernels.hpp"

namespace lsjk {

template <int L1>
void launch_k_ket_l1(int L2, int L3, int L4, int route, const KArgs& a, const KGroups& kg, const KUnitG* units,
                     int nunits, hipStream_t s);
#define LSJK_KET_EXTERN(X)                                                                                   \
    extern template void launch_k_ket_l1<X>(int, int, int, int, const KArgs&, const KGroups&, const KUnitG*, int, \
                                            hipStream_t);
LSJK_KET_EXTERN(0) LSJK_KET_EXTERN(1) LSJK_KET_EXTERN(2)
#undef LSJK_KET_EXTERN

template <int L1, int L2>
void launch_k_grad_l12(int L3, int L4, const KArgs& a, const KUnit* units, int nunits, hipStream_t s);
#define LSJK_GRAD_EXTERN(X, Y) \
    extern template void launch_k_grad_l12<X, Y>(int, int, const KArgs&, const KUnit*, int, hipStream_t);
LSJK_GRAD_EXTERN(0, 0) LSJK_GRAD_EXTERN(0, 1) LSJK_GRAD_EXTERN(0, 2)
LSJK_GRAD_EXTERN(1, 0) LSJK_GRAD_EXTERN(1, 1) LSJK_GRAD_EXTERN(1, 2)
LSJK_GRAD_EXTERN(2, 0) LSJK_GRAD_EXTERN(2, 1) LSJK_GRAD_EXTERN(2, 2)
#undef LSJK_GRAD_EXTERN

// Block maxima of |D24| per (block2, block4), Dsh[i2 * ldDsh + i4], and the global maximum.
// Dense: every (i2, i4), thread per pair.  Banded (blo != nullptr, a batch basis of independent
// molecules): only the pairs of one molecule, i4 in [blo[i2], bhi[i2]); thread (i2, j) takes
// i4 = blo[i2] + j, j < band (the largest molecule's block count).
__global__ void k_blockmax(const double* D, int ldD, const int* bc2, const int* bL2, int nb2, const int* bc4,
                           const int* bL4, int nb4, double* Dsh, int ldDsh, const int* blo, const int* bhi, int band,
                           CallScalars* sc)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int w = blo ? band : nb4;
    if (i >= (long long)nb2 * w) return;
    const int i2 = (int)(i / w);
    int i4 = (int)(i - (long long)i2 * w);
    if (blo) {
        i4 += blo[i2];
        if (i4 >= bhi[i2]) return;
    }
    const int o2 = bc2[i2], o4 = bc4[i4];
    const int n2 = ncart(bL2[i2]), n4 = ncart(bL4[i4]);
    double m = 0.0;
    for (int b = 0; b < n2; b++)
        for (int d = 0; d < n4; d++) m = fmax(m, fabs(D[(size_t)(o2 + b) * ldD + (o4 + d)]));
    Dsh[(size_t)i2 * ldDsh + i4] = m;
    atomicMax(&sc->dmax_bits, (unsigned long long)__double_as_longlong(m));
}

// out[pr[i] * ldo + pc[j]] = in[i * ldi + j] for i < nr, j < nc (pr, pc permutations; null: the
// identity on that side).  One thread per element, rows of `in` read coalesced.
__global__ void k_permute(const double* __restrict__ in, int ldi, int nr, int nc, const int* __restrict__ pr,
                          const int* __restrict__ pc, double* __restrict__ out, int ldo)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nr * nc) return;
    const int i = (int)(t / nc), j = (int)(t - (long long)i * nc);
    const int oi = pr ? pr[i] : i, oj = pc ? pc[j] : j;
    out[(size_t)oi * ldo + oj] = in[(size_t)i * ldi + j];
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
                       const int* bL4, int nb4, double* Dsh, int ldDsh, const int* blo, const int* bhi, int band,
                       CallScalars* sc, hipStream_t s)
{
    long long n = (long long)nb2 * (blo ? band : nb4);
    if (n == 0) return;
    hipLaunchKernelGGL(k_blockmax, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, D, ldD, bc2, bL2, nb2, bc4,
                       bL4, nb4, Dsh, ldDsh, blo, bhi, band, sc);
}

void launch_permute(const double* in, int ldi, int nr, int nc, const int* pr, const int* pc, double* out, int ldo,
                    hipStream_t s)
{
    const long long n = (long long)nr * nc;
    if (n == 0) return;
    hipLaunchKernelGGL(k_permute, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, ldi, nr, nc, pr, pc, out,
                       ldo);
}

void launch_k_quartets(int L1, int L2, int L3, int L4, int route, const KArgs& a, const KUnit* units, int nunits,
                       hipStream_t s)
{
    if (nunits <= 0) return;
    if (route != kRouteGrad) throw std::runtime_error("lsjk: tile kernels serve the gradient route only");
#define G(X, Y) \
    if (L1 == X && L2 == Y) return launch_k_grad_l12<X, Y>(L3, L4, a, units, nunits, s);
    G(0, 0) G(0, 1) G(0, 2) G(1, 0) G(1, 1) G(1, 2) G(2, 0) G(2, 1) G(2, 2)
#undef G
}

void launch_k_ket(int L1, int L2, int L3, int L4, int route, const KArgs& a, const KGroups& kg, const KUnitG* units,
                  int nunits, hipStream_t s)
{
    if (nunits <= 0) return;
    if (L1 == 0)
        launch_k_ket_l1<0>(L2, L3, L4, route, a, kg, units, nunits, s);
    else if (L1 == 1)
        launch_k_ket_l1<1>(L2, L3, L4, route, a, kg, units, nunits, s);
    else
        launch_k_ket_l1<2>(L2, L3, L4, route, a, kg, units, nunits, s);
}

}  // namespace lsjk

// C-ABI of liblsjk: session, device-resident basis / pair-list layouts, and the orchestration of
// the J and K kernels on the library stream.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <tuple>
#include <numeric>
#include <set>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/lsjk.h"
#include "internal.hpp"
#include "kernels.hpp"

namespace {
#include "../common/rys_table.inc"
}

struct lsjk_basis : public lsjk::Basis {};
struct lsjk_pairlist : public lsjk::PairList {};

// Device scratch of lsjk_coulomb_noalloc: its per-call J buffers (truncated ket bounds, ket
// Hermite densities and their FP32 copy, bra accumulators) are carved from this array instead of
// the session workspace, so that the call allocates nothing.  The runtime places one copy on each
// device at module load and keeps it until the process ends.  Host code never passes the name
// itself to a kernel (on the host it is a stand-in, not a device address): lsjk_init looks up the
// device address once with hipGetSymbolAddress and every use goes through Session::scratch.
__device__ __align__(256) unsigned char scratch[LSJK_NOALLOC_SCRATCH_BYTES];

namespace {

struct LsjkError : public std::runtime_error {
    int code;
    LsjkError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                       \
    do {                                                                                                \
        hipError_t e_ = (x);                                                                            \
        if (e_ != hipSuccess)                                                                           \
            throw LsjkError(LSJK_E_HIP, std::string("HIP error '") + hipGetErrorString(e_) + "' at " #x); \
    } while (0)

thread_local std::string g_err;
std::recursive_mutex g_mu;

struct Session {
    bool init = false;
    int device = 0;
    hipStream_t stream = nullptr;
    double* d_rys_coef = nullptr;
    double* d_rys_r2 = nullptr;
    double* d_rys_w = nullptr;
    float* d_rys_coef32 = nullptr;
    lsjk::RysTable rys{};
    lsjk::CallScalars* d_sc = nullptr;
    void* ws = nullptr;
    size_t ws_bytes = 0;
    void* scratch = nullptr;  // device address of `scratch` on the session's device
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool counting = false;
    uint64_t counts[1024] = {0};
    uint64_t counts32[1024] = {0};
    uint64_t wsteps[1024] = {0};
    uint64_t wlanes[1024] = {0};
    double ms[2] = {0.0, 0.0};
    hipStream_t side[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr;
    hipEvent_t join[4] = {nullptr, nullptr, nullptr, nullptr};
    int nside = 4;  // side streams in use (lsjk_set_streams / environment LSJK_STREAMS = 1..4; 1 serialises)
    int num_cu = 1;         // compute units of the device
    size_t max_lds = 65536; // LDS bytes per workgroup
    bool profiling = false;
    std::vector<hipEvent_t> kev;      // event pool for per-launch timing
    std::vector<int> kid;             // launch ids of the last call (count index)
    std::vector<double> kms;          // their durations (ms)
};
Session S;

constexpr int kSide = 4;

template <typename T>
T* dev_upload(const std::vector<T>& v)
{
    T* p = nullptr;
    size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
    HIPCHK(hipMalloc((void**)&p, bytes));
    if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
}

template <typename T>
void dev_free(T*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}


// FP32 Rys rule derived from the FP64 table: on each bin, the degree-(kRysNcoef32-1) Chebyshev
// interpolant (at the Chebyshev points of the first kind) of each FP64 polynomial, converted to
// monomial coefficients in the bin variable u and rounded to FP32.  Blocks are padded to a
// multiple of 4 floats so every block starts 16-byte aligned.
void build_rys_f32(std::vector<float>& out, int* offset32)
{
    constexpr int NC = LSJK_RYS_NCOEF, M = lsjk::kRysNcoef32;
    // monomial coefficients of the Chebyshev polynomials T_0..T_{M-1}
    double Tm[M][M] = {};
    Tm[0][0] = 1.0;
    if (M > 1) Tm[1][1] = 1.0;
    for (int k = 2; k < M; k++)
        for (int j = 0; j < M; j++) Tm[k][j] = (j > 0 ? 2.0 * Tm[k - 1][j - 1] : 0.0) - Tm[k - 2][j];
    double node[M];
    for (int j = 0; j < M; j++) node[j] = std::cos(M_PI * (j + 0.5) / M);
    out.clear();
    offset32[0] = 0;
    for (int n = 1; n <= 9; n++) {
        while (out.size() % 4) out.push_back(0.0f);
        offset32[n] = (int)out.size();
        for (int b = 0; b < lsjk_rys_nbin[n]; b++)
            for (int i = 0; i < n; i++)
                for (int xw = 0; xw < 2; xw++) {
                    const double* c = lsjk_rys_coef + lsjk_rys_offset[n] + ((size_t)b * n + i) * 2 * NC + xw * NC;
                    double f[M];
                    for (int j = 0; j < M; j++) {
                        double v = c[NC - 1];
                        for (int k = NC - 2; k >= 0; k--) v = v * node[j] + c[k];
                        f[j] = v;
                    }
                    double mono[M] = {};
                    for (int k = 0; k < M; k++) {
                        double ck = 0.0;
                        for (int j = 0; j < M; j++) ck += f[j] * std::cos(M_PI * k * (j + 0.5) / M);
                        ck *= (k == 0 ? 1.0 : 2.0) / M;
                        for (int q = 0; q < M; q++) mono[q] += ck * Tm[k][q];
                    }
                    for (int q = 0; q < M; q++) out.push_back((float)mono[q]);
                }
    }
    while (out.size() % 4) out.push_back(0.0f);
}

void require_init()
{
    if (!S.init) throw LsjkError(LSJK_E_NOINIT, "lsjk: lsjk_init has not been called");
}

// Elements of a Cartesian matrix over (basis a, basis b) and of the block-maximum matrix: dense
// rows x columns, or for a batch basis the span of its molecules' diagonal blocks at stride ld.
size_t mat_elems(const lsjk::Basis* a, const lsjk::Basis* b)
{
    if (a->ld) return a->ncart == 0 ? 1 : (size_t)(a->ncart - 1) * (size_t)(a->ld + 1) + 1;
    return (size_t)a->ncart * b->ncart;
}
size_t blk_elems(const lsjk::Basis* a, const lsjk::Basis* b)
{
    if (a->ldblk) return a->nblk == 0 ? 1 : (size_t)(a->nblk - 1) * (size_t)(a->ldblk + 1) + 1;
    return (size_t)a->nblk * b->nblk;
}

// Bump allocator over one grown workspace buffer (256-byte aligned pieces).
struct Workspace {
    std::vector<std::pair<void**, size_t>> req;
    template <typename T>
    void add(T** p, size_t count) { req.push_back({(void**)p, std::max<size_t>(1, count) * sizeof(T)}); }
    // ext: carve the pieces from this fixed buffer of ext_bytes (nothing is allocated; a request
    // that does not fit raises); nullptr: from the session workspace, grown as needed.
    void commit(void* ext = nullptr, size_t ext_bytes = 0, const char* who = nullptr)
    {
        size_t total = 0;
        for (auto& r : req) total += (r.second + 255) & ~size_t(255);
        if (ext) {
            if (total > ext_bytes)
                throw LsjkError(LSJK_E_UNSUPPORTED, std::string(who ? who : "lsjk") + ": the call needs " +
                                                        std::to_string(total) + " bytes of scratch, more than the " +
                                                        std::to_string(ext_bytes) + " bytes of the fixed scratch array");
            size_t off = 0;
            for (auto& r : req) {
                *r.first = (char*)ext + off;
                off += (r.second + 255) & ~size_t(255);
            }
            return;
        }
        if (total > S.ws_bytes) {
            if (S.ws) HIPCHK(hipFree(S.ws));
            S.ws = nullptr;
            S.ws_bytes = 0;
            HIPCHK(hipMalloc(&S.ws, total));
            S.ws_bytes = total;
        }
        size_t off = 0;
        for (auto& r : req) {
            *r.first = (char*)S.ws + off;
            off += (r.second + 255) & ~size_t(255);
        }
    }
};

int ret_error(const std::exception& e, int code)
{
    g_err = e.what();
    return code;
}

template <typename F>
int guarded(F&& f)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    try {
        f();
        g_err.clear();
        return LSJK_OK;
    } catch (const LsjkError& e) {
        return ret_error(e, e.code);
    } catch (const std::exception& e) {
        return ret_error(e, LSJK_E_INVALID);
    }
}

void check_ewald(int nterm, const double* scales, const double* omegas, lsjk::Ewald& ew)
{
    if (nterm < 0 || (nterm > 0 && (!scales || !omegas)))
        throw LsjkError(LSJK_E_INVALID, "lsjk: invalid Ewald operator arrays");
    ew.n = 0;
    for (int i = 0; i < nterm; i++) {
        // terms with a zero scale or zero omega contribute nothing (coulomb.cpp:70-73)
        if (scales[i] == 0.0 || omegas[i] == 0.0) continue;
        if (ew.n == lsjk::kMaxEwald) throw LsjkError(LSJK_E_UNSUPPORTED, "lsjk: more than 8 Ewald terms");
        ew.alpha[ew.n] = scales[i];
        ew.omega[ew.n] = omegas[i];
        ew.n++;
    }
    ew.plain = ew.n == 1 && ew.omega[0] == -1.0;
}

// ------------------------------------------------------------------ pair-list layouts

void build_pairlist(lsjk_pairlist* pl)
{
    using namespace lsjk;
    const Basis* b1 = pl->b1;
    const Basis* b2 = pl->b2;
    const int n = pl->npair;
    // classes: contiguous (L1, L2) runs, bound descending inside each
    pl->cls_L1.clear();
    pl->cls_L2.clear();
    pl->cls_start.clear();
    for (int k = 0; k < n; k++) {
        int L1 = b1->h_L[pl->h_p1[k]], L2 = b2->h_L[pl->h_p2[k]];
        if (k == 0 || L1 != pl->cls_L1.back() || L2 != pl->cls_L2.back()) {
            for (size_t c = 0; c < pl->cls_L1.size(); c++)
                if (pl->cls_L1[c] == L1 && pl->cls_L2[c] == L2)
                    throw LsjkError(LSJK_E_INVALID, "lsjk_pairlist_create: pairs are not grouped by (L1, L2) class");
            pl->cls_L1.push_back(L1);
            pl->cls_L2.push_back(L2);
            pl->cls_start.push_back(k);
        } else if (pl->h_bound[k] > pl->h_bound[k - 1] &&
                   (pl->nmol == 0 || pl->h_mol[pl->h_p1[k]] == pl->h_mol[pl->h_p1[k - 1]])) {
            throw LsjkError(LSJK_E_INVALID, "lsjk_pairlist_create: bounds not sorted descending within a class");
        }
    }
    pl->cls_start.push_back(n);
    pl->maxbound = 0.0f;
    for (int k = 0; k < n; k++) pl->maxbound = std::max(pl->maxbound, pl->h_bound[k]);

    // pair prefactors K = c1 c2 exp(-e1 e2 / p R^2)
    std::vector<double> K(n);
    for (int k = 0; k < n; k++) {
        int i = pl->h_p1[k], j = pl->h_p2[k];
        double e1 = b1->h_e[i], e2 = b2->h_e[j];
        double dx = b1->h_xyz[3 * i] - b2->h_xyz[3 * j];
        double dy = b1->h_xyz[3 * i + 1] - b2->h_xyz[3 * j + 1];
        double dz = b1->h_xyz[3 * i + 2] - b2->h_xyz[3 * j + 2];
        K[k] = b1->h_c[i] * b2->h_c[j] * std::exp(-e1 * e2 / (e1 + e2) * (dx * dx + dy * dy + dz * dz));
    }

    // Hermite order
    std::vector<int> order;
    order.reserve(n);
    std::vector<int> hpos(n);
    size_t hoff = 0;
    for (int L = 0; L <= 2 * kMaxL; L++) {
        pl->hstart[L] = (int)order.size();
        pl->hoff[L] = hoff;
        std::vector<int> grp;
        for (int k = 0; k < n; k++)
            if (b1->h_L[pl->h_p1[k]] + b2->h_L[pl->h_p2[k]] == L) grp.push_back(k);
        if (pl->nmol > 0) {
            // batch list: molecule-major, bound-sorted inside each molecule
            std::stable_sort(grp.begin(), grp.end(), [&](int a, int b) {
                const int ma = pl->h_mol[pl->h_p1[a]], mb = pl->h_mol[pl->h_p1[b]];
                return ma != mb ? ma < mb : pl->h_bound[a] > pl->h_bound[b];
            });
            auto& hm = pl->h_hmol[L];
            hm.assign(pl->nmol + 1, 0);
            for (int k : grp) hm[pl->h_mol[pl->h_p1[k]] + 1]++;
            hm[0] = (int)order.size();
            for (int m = 0; m < pl->nmol; m++) hm[m + 1] += hm[m];
        } else {
            std::stable_sort(grp.begin(), grp.end(), [&](int a, int b) { return pl->h_bound[a] > pl->h_bound[b]; });
        }
        for (int k : grp) {
            hpos[k] = (int)order.size();
            order.push_back(k);
        }
        hoff += grp.size() * (size_t)nherm(L);
    }
    pl->hstart[2 * kMaxL + 1] = (int)order.size();
    pl->hoff[2 * kMaxL + 1] = hoff;
    std::vector<double> hgeom(5 * (size_t)n);
    pl->h_hb.assign(n, 0.0f);
    for (int h = 0; h < n; h++) {
        int k = order[h];
        int i = pl->h_p1[k], j = pl->h_p2[k];
        double e1 = b1->h_e[i], e2 = b2->h_e[j], p = e1 + e2, ip = 1.0 / p;
        hgeom[h] = p;
        for (int d = 0; d < 3; d++) hgeom[(size_t)(d + 1) * n + h] = (e1 * b1->h_xyz[3 * i + d] + e2 * b2->h_xyz[3 * j + d]) * ip;
        hgeom[(size_t)4 * n + h] = lsjk::kKPrefRoot * ip;
        pl->h_hb[h] = pl->h_bound[k];
    }

    // K-invert partner lists: rows = basis-1 primitives
    struct Ent {
        int row, part;
        float b;
    };
    std::vector<Ent> ents;
    ents.reserve(pl->symmetric ? 2 * (size_t)n : (size_t)n);
    for (int k = 0; k < n; k++) {
        ents.push_back({pl->h_p1[k], pl->h_p2[k], pl->h_bound[k]});
        if (pl->symmetric && pl->h_p1[k] != pl->h_p2[k]) ents.push_back({pl->h_p2[k], pl->h_p1[k], pl->h_bound[k]});
    }
    const int maxLp = b2->maxL;
    std::sort(ents.begin(), ents.end(), [&](const Ent& a, const Ent& b) {
        if (a.row != b.row) return a.row < b.row;
        int la = b2->h_L[a.part], lb = b2->h_L[b.part];
        if (la != lb) return la < lb;
        if (a.b != b.b) return a.b > b.b;
        return a.part < b.part;
    });
    pl->nrow = b1->nprim;
    pl->kmaxL2 = maxLp;
    pl->h_row_start.assign((size_t)pl->nrow * (maxLp + 1) + 1, 0);
    {
        std::vector<int> cnt((size_t)pl->nrow * (maxLp + 1), 0);
        for (auto& e : ents) cnt[(size_t)e.row * (maxLp + 1) + b2->h_L[e.part]]++;
        for (size_t s = 0; s < cnt.size(); s++) pl->h_row_start[s + 1] = pl->h_row_start[s] + cnt[s];
    }
    const int ne = (int)ents.size();
    pl->nent = ne;
    std::vector<float> eb(ne);
    std::vector<int> eprim(ne), ecart(ne), eblk(ne), keblk(ne);
    // packed entries (kKEntDoubles per entry, see KEnt): p, K, P, P - A, A - B, {cart, blk}, 1/(2p),
    // {cart, blk} in the partner basis's K view order
    std::vector<double> eg((size_t)kKEntDoubles * (size_t)std::max(ne, 1), 0.0);
    std::vector<double> eg32((size_t)kKEnt32Doubles * (size_t)std::max(ne, 1), 0.0);
    for (int q = 0; q < ne; q++) {
        const Ent& e = ents[q];
        int r = e.row, s = e.part;
        double er = b1->h_e[r], es = b2->h_e[s], p = er + es, ip = 1.0 / p;
        double R2 = 0.0;
        for (int d = 0; d < 3; d++) {
            double dd = b1->h_xyz[3 * r + d] - b2->h_xyz[3 * s + d];
            R2 += dd * dd;
        }
        eb[q] = e.b;
        eprim[q] = s;
        ecart[q] = b2->h_cart[s];
        eblk[q] = b2->h_blk[s];
        double* g = eg.data() + (size_t)kKEntDoubles * q;
        g[0] = p;
        // K_AB sqrt(2 pi^(5/2)) / p: the product of the two sides' values times (p + q)^(-1/2) is the
        // quartet prefactor 2 pi^(5/2) K_AB K_CD / (p q sqrt(p + q))
        g[1] = b1->h_c[r] * b2->h_c[s] * std::exp(-er * es * ip * R2) * kKPrefRoot * ip;
        for (int d = 0; d < 3; d++) {
            const double A = b1->h_xyz[3 * r + d], B = b2->h_xyz[3 * s + d];
            const double P = (er * A + es * B) * ip;
            g[2 + d] = P;
            g[5 + d] = P - A;
            g[8 + d] = A - B;
        }
        const int ids[2] = {ecart[q], eblk[q]};
        std::memcpy(&g[11], ids, sizeof(ids));
        g[12] = 0.5 * ip;
        keblk[q] = b2->h_kblk[eblk[q]];
        const int kids[2] = {b2->h_kcart[ecart[q]], keblk[q]};
        std::memcpy(&g[13], kids, sizeof(kids));
        // FP32-route copy: P in FP64, the rest rounded to FP32
        double* h = eg32.data() + (size_t)kKEnt32Doubles * q;
        h[0] = g[2];
        h[1] = g[3];
        h[2] = g[4];
        const float f[10] = {(float)g[0], (float)g[1], (float)g[5], (float)g[6], (float)g[7],
                             (float)g[8], (float)g[9], (float)g[10], (float)g[12], 0.0f};
        std::memcpy(&h[3], f, sizeof(f));
        std::memcpy(&h[8], ids, sizeof(ids));
        std::memcpy(&h[9], kids, sizeof(kids));
    }
    for (int L = 0; L <= kMaxL; L++)
        for (int Lp = 0; Lp <= kMaxL; Lp++) pl->h_rows[L][Lp].clear();
    for (int r = 0; r < pl->nrow; r++)
        for (int Lp = 0; Lp <= maxLp; Lp++)
            if (pl->h_row_start[(size_t)r * (maxLp + 1) + Lp + 1] > pl->h_row_start[(size_t)r * (maxLp + 1) + Lp])
                pl->h_rows[b1->h_L[r]][Lp].push_back(r);
    // triangle enumerations (a <= b) for the symmetric-mode launches with equal row momenta; the
    // symmetric mode depends on the list being used on both sides with a symmetric D24, not on the
    // list itself being symmetric, so every list gets them (for every Lp <= kMaxL)
    std::vector<int> tri_start[kMaxL + 1][kMaxL + 1][kMaxL + 1];
    {
        for (int L = 0; L <= kMaxL; L++)
            for (int La = 0; La <= kMaxL; La++)
                for (int Lb = 0; Lb <= kMaxL; Lb++) {
                    const auto& A = pl->h_rows[L][La];
                    const auto& B = pl->h_rows[L][Lb];
                    auto& pre = pl->h_tri_pre[L][La][Lb];
                    auto& st = tri_start[L][La][Lb];
                    pre.assign(A.size() + 1, 0);
                    st.assign(A.size(), 0);
                    for (size_t i = 0; i < A.size(); i++) {
                        st[i] = (int)(std::lower_bound(B.begin(), B.end(), A[i]) - B.begin());
                        pre[i + 1] = pre[i] + (long long)(B.size() - st[i]);
                    }
                }
    }

    // uploads
    pl->d_p1 = dev_upload(pl->h_p1);
    pl->d_p2 = dev_upload(pl->h_p2);
    pl->d_bound = dev_upload(pl->h_bound);
    pl->d_hpos = dev_upload(hpos);
    pl->d_K = dev_upload(K);
    pl->d_hb = dev_upload(pl->h_hb);
    pl->d_hgeom = dev_upload(hgeom);
    pl->d_row_start = dev_upload(pl->h_row_start);
    pl->d_ebound = dev_upload(eb);
    pl->h_eb = eb;
    pl->h_eprim = eprim;
    pl->d_eprim = dev_upload(eprim);
    pl->d_ecart = dev_upload(ecart);
    pl->d_eblk = dev_upload(eblk);
    pl->d_keblk = dev_upload(keblk);
    pl->d_egeom = dev_upload(eg);
    pl->d_egeom32 = dev_upload(eg32);
    for (int L = 0; L <= kMaxL; L++)
        for (int Lp = 0; Lp <= kMaxL; Lp++) {
            pl->d_rows[L][Lp] = dev_upload(pl->h_rows[L][Lp]);
            for (int Lb = 0; Lb <= kMaxL; Lb++) {
                    pl->d_tri_pre[L][Lp][Lb] = dev_upload(pl->h_tri_pre[L][Lp][Lb]);
                    pl->d_tri_start[L][Lp][Lb] = dev_upload(tri_start[L][Lp][Lb]);
                }
        }
    pl->d_cls_start = dev_upload(pl->cls_start);
    pl->d_cls_L1 = dev_upload(pl->cls_L1);
    pl->d_cls_L2 = dev_upload(pl->cls_L2);
    if (!pl->symmetric) {
        // pivot view: rows = basis-2 primitives, partners = basis-1 primitives (KBlock::build21)
        std::vector<Ent> pv(n);
        for (int k = 0; k < n; k++) pv[k] = {pl->h_p2[k], pl->h_p1[k], pl->h_bound[k]};
        const int maxLa = b1->maxL;
        std::sort(pv.begin(), pv.end(), [&](const Ent& a, const Ent& b) {
            if (a.row != b.row) return a.row < b.row;
            int la = b1->h_L[a.part], lb = b1->h_L[b.part];
            if (la != lb) return la < lb;
            if (a.b != b.b) return a.b > b.b;
            return a.part < b.part;
        });
        std::vector<int> rs((size_t)b2->nprim * (maxLa + 1) + 1, 0);
        std::vector<int> cnt((size_t)b2->nprim * (maxLa + 1), 0);
        for (auto& e : pv) cnt[(size_t)e.row * (maxLa + 1) + b1->h_L[e.part]]++;
        for (size_t q = 0; q < cnt.size(); q++) rs[q + 1] = rs[q] + cnt[q];
        std::vector<float> pb(n);
        for (int k = 0; k < n; k++) pb[k] = pv[k].b;
        pl->d_piv_row_start = dev_upload(rs);
        pl->d_piv_b = dev_upload(pb);
    }
}

void free_pairlist(lsjk_pairlist* pl)
{
    dev_free(pl->d_p1);
    dev_free(pl->d_p2);
    dev_free(pl->d_bound);
    dev_free(pl->d_hpos);
    dev_free(pl->d_K);
    dev_free(pl->d_hb);
    dev_free(pl->d_hgeom);
    dev_free(pl->d_row_start);
    dev_free(pl->d_ebound);
    dev_free(pl->d_eprim);
    dev_free(pl->d_ecart);
    dev_free(pl->d_eblk);
    dev_free(pl->d_keblk);
    dev_free(pl->d_egeom);
    dev_free(pl->d_egeom32);
    for (int L = 0; L <= lsjk::kMaxL; L++)
        for (int Lp = 0; Lp <= lsjk::kMaxL; Lp++) {
            dev_free(pl->d_rows[L][Lp]);
            for (int Lb = 0; Lb <= lsjk::kMaxL; Lb++) {
                dev_free(pl->d_tri_pre[L][Lp][Lb]);
                dev_free(pl->d_tri_start[L][Lp][Lb]);
            }
        }
    dev_free(pl->d_piv_row_start);
    dev_free(pl->d_piv_b);
    dev_free(pl->d_cls_start);
    dev_free(pl->d_cls_L1);
    dev_free(pl->d_cls_L2);
    for (auto& kv : pl->jtasks) dev_free(kv.second.d_tasks);
    pl->jtasks.clear();
    dev_free(pl->jtasks_sym.d_tasks);
    pl->jtasks_sym = lsjk::JTaskList();
    for (auto& kv : pl->kunits) dev_free(kv.second.first);
    pl->kunits.clear();
    pl->kcost.clear();
    for (auto& kv : pl->kunits_g) dev_free(kv.second.first);
    pl->kunits_g.clear();
    pl->kcost_g.clear();
    for (auto& kv : pl->kgroups) {
        dev_free(kv.second.d_slot);
        dev_free(kv.second.d_grp_slot);
        dev_free(kv.second.d_grp_rows);
        dev_free(kv.second.d_grp_rank);
        dev_free(kv.second.d_chunk_b);
    }
    pl->kgroups.clear();
}

// Live pair lists.  The caches on a list keyed by another list's uid (J task lists, K units and
// their costs, mirror-class choices) are dropped when that other list is destroyed, so a
// long-lived list paired with many short-lived ones does not keep their device buffers.
std::set<lsjk_pairlist*> g_lists;

template <typename M, typename F>
void erase_if_key(M& m, F&& pred, bool free_dev)
{
    for (auto it = m.begin(); it != m.end();) {
        if (pred(it->first)) {
            if constexpr (std::is_same<typename M::mapped_type, std::pair<void*, int>>::value)
                if (free_dev) dev_free(it->second.first);
            it = m.erase(it);
        } else {
            ++it;
        }
    }
}

void drop_caches_for(lsjk_pairlist* pl, unsigned long long uid)
{
    auto it = pl->jtasks.find(uid);
    if (it != pl->jtasks.end()) {
        dev_free(it->second.d_tasks);
        pl->jtasks.erase(it);
    }
    auto by_uid = [uid](const auto& k) { return k.first == uid; };
    erase_if_key(pl->kunits, by_uid, true);
    erase_if_key(pl->kcost, by_uid, false);
    erase_if_key(pl->kunits_g, by_uid, true);
    pl->kunits_gen++;
    erase_if_key(pl->kcost_g, by_uid, false);
    erase_if_key(pl->korient, by_uid, false);
}

// J task list for (bra list, ket list): per (L12, L34) group pair, tiles of 64 bra pairs x chunks of
// ket pairs (at most 1024, sized so a group pair yields about kJTasksTarget tasks), heaviest
// (largest bounds) first.
constexpr int kJChunkMax = 1024;
constexpr int kJTasksTarget = 16384;

// Ket chunk of a group pair with `nket` ket pairs and `ntile` bra tiles: normally between 64 and
// kJChunkMax kets, sized for about kJTasksTarget tasks.  A group pair that would then have fewer
// than 256 tasks (heavy classes with few kets and few bra tiles: a handful of long-running waves)
// takes shorter chunks, down to nket / 16 kets: each task adds its tile's bra potentials with
// atomics, so more tasks per tile would contend on the same addresses.  For total angular
// momentum L >= 6 a quartet costs so much more than the atomics that the floor is nket / 64.
// A ket chunk scaled by an autotuned factor (16 .. 4096 kets).
long long scaled_chunk(long long c, float scale)
{
    if (scale == 1.0f) return c;
    return std::min<long long>(4096, std::max<long long>(16, (long long)std::llround((double)c * scale)));
}

long long j_chunk(long long nket, long long ntile, int L)
{
    long long c = (nket * ntile + kJTasksTarget - 1) / kJTasksTarget;
    c = std::min<long long>(kJChunkMax, std::max<long long>(64, (c + 63) / 64 * 64));
    const long long per_tile = L >= 6 ? 64 : 16;
    const long long few = std::max<long long>({(nket * ntile + 255) / 256, (nket + per_tile - 1) / per_tile, 1});
    return std::min(c, few);
}

// K work units of one class.  A unit is (row1, block of kKTile row3s, chunk of at most kKChunk2 of
// row1's partners); its cost estimate is (chunk partners) x (partners of the block's row3s) from
// the bound-screened lists alone (D only removes quartets, it never adds).  In the triangle mode
// row3s below row1 are not counted and units without work are dropped.  Heavy units are
// dispatched first (so that long units do not finish after the rest of the launch), in an order
// that keeps each XCD on neighbouring rows (see below).
std::pair<const lsjk::KUnit*, int> get_kunits(const lsjk_pairlist* pl12c, const lsjk_pairlist* pl34, int L1, int L2,
                                              int L3, int L4, bool tri, double* cost_out)
{
    using lsjk::KUnit;
    lsjk_pairlist* pl12 = const_cast<lsjk_pairlist*>(pl12c);
    const int cls = ((((L1 * 3 + L2) * 3 + L3) * 3 + L4) << 1) | (tri ? 1 : 0);
    const auto key = std::make_pair(pl34->uid, cls);
    auto it = pl12->kunits.find(key);
    if (it == pl12->kunits.end()) {
        const auto& rows1 = pl12->h_rows[L1][L2];
        const auto& rows3 = pl34->h_rows[L3][L4];
        const int m12 = pl12->kmaxL2 + 1, m34 = pl34->kmaxL2 + 1;
        std::vector<long long> n4(rows3.size());
        for (size_t j = 0; j < rows3.size(); j++) {
            const size_t r3 = (size_t)rows3[j];
            n4[j] = pl34->h_row_start[r3 * m34 + L4 + 1] - pl34->h_row_start[r3 * m34 + L4];
        }
        const int nb3 = (int)((rows3.size() + lsjk::kKTile - 1) / lsjk::kKTile);
        std::vector<std::pair<long long, KUnit>> v;
        double total = 0.0;
        for (size_t i1 = 0; i1 < rows1.size(); i1++) {
            const size_t r1 = (size_t)rows1[i1];
            const int n2 = pl12->h_row_start[r1 * m12 + L2 + 1] - pl12->h_row_start[r1 * m12 + L2];
            for (int ib = 0; ib < nb3; ib++) {
                long long sum4 = 0;
                const size_t j0 = (size_t)ib * lsjk::kKTile, j1 = std::min(rows3.size(), j0 + lsjk::kKTile);
                for (size_t j = j0; j < j1; j++)
                    if (!tri || rows3[j] >= (int)r1) sum4 += n4[j];
                if (sum4 == 0) continue;
                for (int e = 0; e < n2; e += lsjk::kKChunk2) {
                    const int cnt = std::min(lsjk::kKChunk2, n2 - e);
                    v.push_back({(long long)cnt * sum4, KUnit{(int)i1, ib, e, cnt}});
                    total += (double)cnt * (double)sum4;
                }
            }
        }
        // Dispatch order: coarse cost buckets (powers of two), heaviest bucket first; inside a bucket
        // the units stay in row order (v is built row-major) and are split into 8 contiguous ranges
        // that are interleaved, so that with the round-robin workgroup-to-XCD dispatch each XCD (one
        // L2) works through neighbouring rows of D and K.
        auto bucket = [](long long c) { return c > 0 ? 63 - __builtin_clzll((unsigned long long)c) : 0; };
        std::stable_sort(v.begin(), v.end(), [&](const std::pair<long long, KUnit>& a,
                                                 const std::pair<long long, KUnit>& b) {
            return bucket(a.first) > bucket(b.first);
        });
        std::vector<KUnit> u;
        u.reserve(v.size());
        for (size_t s0 = 0; s0 < v.size();) {
            size_t s1 = s0;
            while (s1 < v.size() && bucket(v[s1].first) == bucket(v[s0].first)) s1++;
            const size_t n = s1 - s0, chunk = (n + 7) / 8;
            for (size_t k = 0; k < chunk; k++)
                for (size_t x = 0; x < 8; x++) {
                    const size_t src = s0 + x * chunk + k;
                    if (x * chunk + k < n) u.push_back(v[src].second);
                }
            s0 = s1;
        }
        if (u.size() > (size_t)std::numeric_limits<int>::max())
            throw LsjkError(LSJK_E_UNSUPPORTED, "lsjk: too many exchange work units");
        KUnit* d = u.empty() ? nullptr : dev_upload(u);
        it = pl12->kunits.insert({key, {(void*)d, (int)u.size()}}).first;
        pl12->kcost[key] = total;
    }
    if (cost_out) *cost_out = pl12->kcost[key];
    return {static_cast<const KUnit*>(it->second.first), it->second.second};
}

// Spatial rank of every basis-1 primitive of a list: primitives sorted by the Morton code of
// their centre (10 bits per axis over the bounding box), ties by index.
const std::vector<int>& get_mrank(lsjk_pairlist* pl)
{
    if (!pl->h_mrank.empty() || pl->b1->nprim == 0) return pl->h_mrank;
    const lsjk::Basis* b = pl->b1;
    const int n = b->nprim;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], b->h_xyz[3 * i + d]);
            hi[d] = std::max(hi[d], b->h_xyz[3 * i + d]);
        }
    auto spread = [](unsigned int v) {  // 10 bits -> every third bit of 30
        unsigned long long x = v & 1023u, r = 0;
        for (int k = 0; k < 10; k++) r |= ((x >> k) & 1ull) << (3 * k);
        return r;
    };
    std::vector<std::pair<unsigned long long, int>> key(n);
    for (int i = 0; i < n; i++) {
        unsigned long long m = 0;
        for (int d = 0; d < 3; d++) {
            const double ext = hi[d] - lo[d];
            const unsigned int q = ext > 0.0 ? (unsigned int)std::min(1023.0, (b->h_xyz[3 * i + d] - lo[d]) / ext * 1023.0) : 0u;
            m |= spread(q) << d;
        }
        // batch list: molecule-major (the molecule in the bits above the 30-bit Morton code)
        if (pl->nmol > 0) m |= (unsigned long long)pl->h_mol[i] << 30;
        key[i] = {m, i};
    }
    std::sort(key.begin(), key.end());
    pl->h_mrank.assign(n, 0);
    for (int r = 0; r < n; r++) pl->h_mrank[key[r].second] = r;
    return pl->h_mrank;
}

// Ket groups of the (row L, partner Lp) class of a list (see KGroups): the class's rows in spatial
// order, gsize per group, each group's partner entries merged by bound (descending, ties by
// entry index) and padded to a multiple of 64 slots.
// Rows per ket group of (row L, partner L) of a list: the autotuned value or kgroup_size.
int kgroup_for(const lsjk_pairlist* pl, int L, int Lp)
{
    auto it = pl->kgsize.find(L * 8 + Lp);
    return it != pl->kgsize.end() ? it->second : lsjk::kgroup_size(L, Lp);
}

const lsjk_pairlist::KGroupData& get_kgroups(lsjk_pairlist* pl, int L, int Lp, int gsize)
{
    const int key = (L * 8 + Lp) * 256 + gsize;
    auto it = pl->kgroups.find(key);
    if (it != pl->kgroups.end()) return it->second;
    lsjk_pairlist::KGroupData G;
    G.gsize = gsize;
    const auto& rows = pl->h_rows[L][Lp];
    const auto& mrank = get_mrank(pl);
    const int nr = (int)rows.size();
    G.order.resize(nr);
    for (int i = 0; i < nr; i++) G.order[i] = i;
    std::sort(G.order.begin(), G.order.end(), [&](int a, int b) { return mrank[rows[a]] < mrank[rows[b]]; });
    // groups: gsize consecutive rows of the spatial order; a batch list starts a new group at
    // every molecule (a group's rows are of one molecule)
    std::vector<int> gstart;  // first position (in G.order) of each group
    {
        int cur = -1, fill = 0;
        for (int i = 0; i < nr; i++) {
            const int mol = pl->nmol > 0 ? pl->h_mol[rows[G.order[i]]] : 0;
            if (gstart.empty() || fill == gsize || mol != cur) {
                gstart.push_back(i);
                fill = 0;
                cur = mol;
            }
            fill++;
        }
        gstart.push_back(nr);
    }
    G.ngroup = (int)gstart.size() - 1;
    if (pl->nmol > 0) {
        G.mol_grp.assign(pl->nmol + 1, 0);
        for (int g = 0; g < G.ngroup; g++) G.mol_grp[pl->h_mol[rows[G.order[gstart[g]]]] + 1]++;
        for (int q = 0; q < pl->nmol; q++) G.mol_grp[q + 1] += G.mol_grp[q];
    }
    const int m = pl->kmaxL2 + 1;
    std::vector<int2> slot;
    std::vector<int> grp_rows((size_t)G.ngroup * gsize, -1);
    G.grp_rank.assign((size_t)G.ngroup * gsize, std::numeric_limits<int>::max());
    G.grp_slot.assign(G.ngroup + 1, 0);
    for (int g = 0; g < G.ngroup; g++) {
        std::vector<int2> v;
        for (int t = 0; t < gsize && gstart[g] + t < gstart[g + 1]; t++) {
            const int r = rows[G.order[gstart[g] + t]];
            grp_rows[(size_t)g * gsize + t] = r;
            G.grp_rank[(size_t)g * gsize + t] = mrank[r];
            for (int e = pl->h_row_start[(size_t)r * m + Lp]; e < pl->h_row_start[(size_t)r * m + Lp + 1]; e++)
                v.push_back(make_int2(e, t));
        }
        std::stable_sort(v.begin(), v.end(), [&](const int2& a, const int2& b) { return pl->h_eb[a.x] > pl->h_eb[b.x]; });
        while (v.size() % 64) v.push_back(make_int2(-1, 0));
        for (size_t c = 0; c < v.size(); c += 64) G.chunk_b.push_back(pl->h_eb[v[c].x]);
        slot.insert(slot.end(), v.begin(), v.end());
        G.grp_slot[g + 1] = (int)slot.size();
    }
    G.d_slot = dev_upload(slot);
    G.d_grp_slot = dev_upload(G.grp_slot);
    G.d_grp_rows = dev_upload(grp_rows);
    G.slot = slot;
    G.grp_rows = grp_rows;
    G.d_grp_rank = dev_upload(G.grp_rank);
    G.d_chunk_b = dev_upload(G.chunk_b);
    return pl->kgroups.emplace(key, std::move(G)).first->second;
}

// Work units per K class below which the groups' chunk ranges are split.
constexpr long long kKUnitsTargetDefault = 12288;
// Unit tiles inside an XCD's range (get_kunits_g): bra rows x ket groups.  Measured on (H2O)256 /
// 6-31G* (K 247 -> 199 ms serialised against cost order alone) and (H2O)64 / 6-31G** (no loss);
// 64 x 2 .. whole band x 2 measured within 2 % of each other.
constexpr int kKTileRows = 256;
constexpr int kKTileGroups = 2;
constexpr int kcount_index_host(int L1, int L2, int L3, int L4) { return 256 + ((L1 * 5 + L2) * 5 + L3) * 5 + L4; }
// thresp values whose K unit lists are cached per pair of lists (older ones are dropped)
constexpr int kKThrespCached = 4;

// K build units of one class: (bra row, ket group) pairs.  Rows are taken in spatial order and,
// in the triangle mode, groups lying wholly before the bra row in that order are dropped.  The
// cost estimate of a unit is the number of (bra partner, ket chunk) steps that pass the bound
// test with |D| <= 1 (an ordering heuristic only; the kernel screens with the actual D).  Units
// are dispatched in coarse cost buckets, heaviest first; inside a bucket in (spatial row, group)
// order, split into 8 contiguous ranges that are interleaved by workgroup, so that with the
// round-robin workgroup-to-XCD dispatch each XCD (one L2) works through neighbouring bra rows, in
// tiles of rows x groups (below).  The tiles pay off with the K view copy of D (neighbouring
// shells in neighbouring columns); in the basis order they measured no faster.
std::pair<const lsjk::KUnitG*, int> get_kunits_g(const lsjk_pairlist* pl12c, const lsjk_pairlist* pl34c, int L1,
                                                 int L2, int L3, int L4, bool tri, float thresp, double* cost_out)
{
    using lsjk::KUnitG;
    lsjk_pairlist* pl12 = const_cast<lsjk_pairlist*>(pl12c);
    lsjk_pairlist* pl34 = const_cast<lsjk_pairlist*>(pl34c);
    const int gsize = kgroup_for(pl34, L3, L4);
    unsigned int tb;
    std::memcpy(&tb, &thresp, sizeof(tb));
    // key: rows per ket group (bits 40..47), thresp bits (8..39), class and triangle flag
    const long long cls = ((long long)gsize << 40) | ((long long)tb << 8) |
                          (((((L1 * 3 + L2) * 3 + L3) * 3 + L4) << 1) | (tri ? 1 : 0));
    const auto key = std::make_pair(pl34->uid, cls);
    auto it = pl12->kunits_g.find(key);
    if (it == pl12->kunits_g.end()) {
        // the unit lists depend on thresp: keep those of at most kKThrespCached values per ket list
        {
            std::set<long long> seen;
            for (const auto& kv : pl12->kunits_g)
                if (kv.first.first == pl34->uid) seen.insert((kv.first.second >> 8) & 0xffffffffLL);
            if (!seen.count((cls >> 8) & 0xffffffffLL) && (int)seen.size() >= kKThrespCached) {
                auto same = [&](const std::pair<unsigned long long, long long>& k) { return k.first == pl34->uid; };
                erase_if_key(pl12->kunits_g, same, true);
                erase_if_key(pl12->kcost_g, same, false);
                pl12->kunits_gen++;
            }
        }
        const auto& rows1 = pl12->h_rows[L1][L2];
        const auto& G34 = get_kgroups(pl34, L3, L4, gsize);
        const auto& G12 = get_kgroups(pl12, L1, L2, kgroup_for(pl12, L1, L2));  // its row order only
        // the class's unit-split target: autotuned (this list on both sides) or the default
        long long kKUnitsTarget = kKUnitsTargetDefault;
        if (pl12 == pl34) {
            auto ks = pl12->ksplit.find(kcount_index_host(L1, L2, L3, L4));
            if (ks != pl12->ksplit.end()) kKUnitsTarget = ks->second;
        }
        const auto& mrank = get_mrank(pl12);
        const int m12 = pl12->kmaxL2 + 1;
        // chunk bounds per group, and the smallest rank of each group (triangle mode)
        std::vector<std::pair<long long, KUnitG>> v;
        double total = 0.0;
        for (int i1 : G12.order) {
            const int r1 = rows1[i1];
            const int a0 = pl12->h_row_start[(size_t)r1 * m12 + L2], a1 = pl12->h_row_start[(size_t)r1 * m12 + L2 + 1];
            // bounds are sorted descending: the partners with b12 >= x are a prefix
            auto npass = [&](double x) {
                int lo = a0, hi = a1;
                while (lo < hi) {
                    const int mid = (lo + hi) / 2;
                    if ((double)pl12->h_eb[mid] >= x) lo = mid + 1; else hi = mid;
                }
                return (long long)(lo - a0);
            };
            const int rank1 = tri ? mrank[r1] : -1;
            // a batch list pairs a row only with the groups of its own molecule
            const int g0 = pl34->nmol > 0 ? G34.mol_grp[pl12->h_mol[r1]] : 0;
            const int g1 = pl34->nmol > 0 ? G34.mol_grp[pl12->h_mol[r1] + 1] : G34.ngroup;
            for (int g = g0; g < g1; g++) {
                if (tri) {
                    int rmax = -1;
                    for (int t = 0; t < gsize; t++) {
                        const int rk = G34.grp_rank[(size_t)g * gsize + t];
                        if (rk != std::numeric_limits<int>::max()) rmax = std::max(rmax, rk);
                    }
                    if (rmax < rank1) continue;
                }
                long long c = 0;
                for (int s = G34.grp_slot[g]; s < G34.grp_slot[g + 1]; s += 64)
                    c += npass((double)thresp / (double)G34.chunk_b[s / 64]);
                v.push_back({c, KUnitG{i1, g, rank1, 0}});
                total += (double)c;
            }
        }
        // classes with few (row, group) pairs: split the groups' chunk ranges so that the launch
        // has enough waves to fill the device
        long long nchunk = 0;
        for (const auto& x : v) nchunk += (G34.grp_slot[x.second.g + 1] - G34.grp_slot[x.second.g]) / 64;
        const int per = (int)std::max<long long>(1, std::min<long long>(
            (long long)(v.size() >= kKUnitsTarget ? 1 << 15 : (nchunk + kKUnitsTarget - 1) / kKUnitsTarget), 1 << 15));
        {
            std::vector<std::pair<long long, KUnitG>> w;
            w.reserve(v.size());
            for (const auto& x : v) {
                const int g = x.second.g;
                const int nch = (G34.grp_slot[g + 1] - G34.grp_slot[g]) / 64;
                if (nch >= (1 << 15)) throw LsjkError(LSJK_E_UNSUPPORTED, "lsjk: exchange ket group too large");
                const int r1 = rows1[x.second.i1];
                const int a0 = pl12->h_row_start[(size_t)r1 * m12 + L2], a1 = pl12->h_row_start[(size_t)r1 * m12 + L2 + 1];
                // equal pieces of at most `per` chunks
                const int npiece = (nch + per - 1) / per, step = (nch + npiece - 1) / npiece;
                for (int c0 = 0; c0 < nch; c0 += step) {
                    const int nc = std::min(step, nch - c0);
                    long long c = 0;
                    if (step < nch)
                        for (int k = 0; k < nc; k++) {
                            const double x0 = (double)thresp / (double)G34.chunk_b[G34.grp_slot[g] / 64 + c0 + k];
                            c += (long long)(std::partition_point(pl12->h_eb.begin() + a0, pl12->h_eb.begin() + a1,
                                                                  [&](float b) { return (double)b >= x0; }) -
                                             (pl12->h_eb.begin() + a0));
                        }
                    else
                        c = x.first;
                    KUnitG uu = x.second;
                    uu.crange = c0 | (nc << 16);
                    w.push_back({c, uu});
                }
            }
            v.swap(w);
        }
        // still few units (heavy classes: few rows, costly quartets): split each unit's walk over
        // its bra row's partners into pieces, so that the launch has enough waves; each piece adds
        // its part of the same K block
        if (v.size() < (size_t)kKUnitsTarget / 4) {
            long long walk_total = 0;
            std::vector<int> walk(v.size());
            for (size_t q = 0; q < v.size(); q++) {
                const KUnitG& x = v[q].second;
                const int r1 = rows1[x.i1];
                const int a0 = pl12->h_row_start[(size_t)r1 * m12 + L2], a1 = pl12->h_row_start[(size_t)r1 * m12 + L2 + 1];
                // partners that can pass with the unit's first (largest-bound) chunk
                const double x0 = (double)thresp / (double)G34.chunk_b[G34.grp_slot[x.g] / 64 + (x.crange & 0xffff)];
                walk[q] = (a1 - a0 >= 0xffff) ? 0 : (int)(std::partition_point(pl12->h_eb.begin() + a0, pl12->h_eb.begin() + a1,
                                                                                [&](float b) { return (double)b >= x0; }) -
                                                           (pl12->h_eb.begin() + a0));
                walk_total += walk[q];
            }
            const long long piece = std::max<long long>(1, (walk_total + kKUnitsTarget / 4 - 1) / (kKUnitsTarget / 4));
            std::vector<std::pair<long long, KUnitG>> w;
            for (size_t q = 0; q < v.size(); q++) {
                const int n = walk[q];
                if (n <= piece) {
                    w.push_back(v[q]);
                    continue;
                }
                const int np = (int)((n + piece - 1) / piece), step = (n + np - 1) / np;
                for (int p0 = 0; p0 < n; p0 += step) {
                    KUnitG uu = v[q].second;
                    // the last piece runs to the end of the row's partners (the stop test ends it)
                    uu.prange = p0 | ((p0 + step < n ? step : 0) << 16);
                    w.push_back({std::max<long long>(1, v[q].first / np), uu});
                }
            }
            v.swap(w);
        }
        auto bucket = [](long long c) { return c > 0 ? 64 - __builtin_clzll((unsigned long long)c) : 0; };
        std::stable_sort(v.begin(), v.end(), [&](const std::pair<long long, KUnitG>& a,
                                                 const std::pair<long long, KUnitG>& b) {
            return bucket(a.first) > bucket(b.first);
        });
        // Workgroup w (units kKWaves w .. kKWaves w + kKWaves - 1) runs on XCD w % 8.  Each bucket is
        // cut into 8 contiguous ranges (bands of neighbouring bra rows); workgroups take kKWaves
        // consecutive units of range w % 8.  Inside a range the units go in tiles of kKTileRows
        // rows x kKTileGroups ket groups (spatial order both ways), so that an XCD's L2 holds the
        // tile's block of the K view D: rows = the tile's bra partners, columns = the groups'
        // partners.  Inside a tile the units are sorted by cost, so the waves of one workgroup
        // (which holds its slots until its last wave ends) cost about the same.  Units that do not
        // fill a whole workgroup of their range go at the bucket's end.
        constexpr size_t W = lsjk::kKWaves;
        std::vector<int> pos1(rows1.size(), 0);  // spatial position of each bra row
        for (size_t q = 0; q < G12.order.size(); q++) pos1[G12.order[q]] = (int)q;
        std::vector<KUnitG> u;
        u.reserve(v.size());
        for (size_t s0 = 0; s0 < v.size();) {
            size_t s1 = s0;
            while (s1 < v.size() && bucket(v[s1].first) == bucket(v[s0].first)) s1++;
            const size_t n = s1 - s0, per = ((n + 7) / 8 + W - 1) / W * W;
            for (size_t x = 0; x < 8; x++) {
                const size_t b = std::min(n, x * per), e = std::min(n, (x + 1) * per);
                std::stable_sort(v.begin() + s0 + b, v.begin() + s0 + e,
                                 [&](const std::pair<long long, KUnitG>& a, const std::pair<long long, KUnitG>& c) {
                                     const int ra = pos1[a.second.i1] / kKTileRows, rc = pos1[c.second.i1] / kKTileRows;
                                     if (ra != rc) return ra < rc;
                                     const int ga = a.second.g / kKTileGroups, gc = c.second.g / kKTileGroups;
                                     if (ga != gc) return ga < gc;
                                     if (a.first != c.first) return a.first > c.first;
                                     if (pos1[a.second.i1] != pos1[c.second.i1]) return pos1[a.second.i1] < pos1[c.second.i1];
                                     return a.second.g < c.second.g;
                                 });
            }
            std::vector<KUnitG> rest;
            for (size_t k = 0; k < per; k += W)
                for (size_t x = 0; x < 8; x++) {
                    const size_t b = x * per + k, e = std::min(std::min(n, (x + 1) * per), b + W);
                    if (b >= e) continue;
                    for (size_t q = b; q < e; q++) (e - b == W ? u : rest).push_back(v[s0 + q].second);
                }
            u.insert(u.end(), rest.begin(), rest.end());
            s0 = s1;
        }
        if (u.size() > (size_t)std::numeric_limits<int>::max() / lsjk::kKWaves)
            throw LsjkError(LSJK_E_UNSUPPORTED, "lsjk: too many exchange work units");
        KUnitG* d = u.empty() ? nullptr : dev_upload(u);
        it = pl12->kunits_g.insert({key, {(void*)d, (int)u.size()}}).first;
        pl12->kcost_g[key] = total;
    }
    if (cost_out) *cost_out = pl12->kcost_g[key];
    return {static_cast<const KUnitG*>(it->second.first), it->second.second};
}

lsjk::KGroups kgroups_args(const lsjk_pairlist::KGroupData& G)
{
    lsjk::KGroups k;
    k.gsize = G.gsize;
    k.slot = G.d_slot;
    k.grp_slot = G.d_grp_slot;
    k.grp_rows = G.d_grp_rows;
    k.grp_rank = G.d_grp_rank;
    k.chunk_b = G.d_chunk_b;
    return k;
}

lsjk::JTaskList upload_jtasks(const std::vector<lsjk::JTask>& tasks, const lsjk::JTaskList& seg)
{
    lsjk::JTaskList out = seg;
    out.ntask = (int)tasks.size();
    out.d_tasks = reinterpret_cast<int*>(dev_upload(tasks));
    return out;
}

// J tasks of one group pair of a batch list (the list on both sides): per molecule, tiles of 64 of
// its bra pairs x chunks of its own ket pairs, never another molecule's.  The ket chunk is sized for
// the whole group pair as for one list (about kJTasksTarget tasks over all molecules, 64..1024 kets)
// and clipped to each molecule's kets.  sym: the symmetric kernel's triangle (within one group a
// tile's kets stop at its last bra pair).
void append_batch_jtasks(const lsjk_pairlist* pl, int L12, int L34, bool sym, std::vector<lsjk::JTask>& tasks)
{
    const auto& hb = pl->h_hmol[L12];
    const auto& hk = pl->h_hmol[L34];
    double work = 0.0;
    for (int m = 0; m < pl->nmol; m++)
        work += (double)((hb[m + 1] - hb[m] + 63) / 64) * (double)(hk[m + 1] - hk[m]);
    long long chunk = (long long)std::ceil(work / kJTasksTarget);
    chunk = std::min<long long>(kJChunkMax, std::max<long long>(64, (chunk + 63) / 64 * 64));
    chunk = scaled_chunk(chunk, pl->jscale[sym ? 1 : 0][L12][L34]);
    for (int m = 0; m < pl->nmol; m++) {
        const int b0 = hb[m], b1 = hb[m + 1], k0 = hk[m], k1 = hk[m + 1];
        if (b1 <= b0 || k1 <= k0) continue;
        for (int b = b0; b < b1; b += 64) {
            const int kend = (sym && L12 == L34) ? std::min(k1, b + 64) : k1;
            for (int k = k0; k < kend; k += (int)chunk)
                tasks.push_back({b, std::min(64, b1 - b), k, std::min(k + (int)chunk, kend)});
        }
    }
}

// The J task list of (bra list, ket list), built and uploaded at first use and cached on the bra
// list under the ket list's uid.  `build` false: only look it up (nullptr if absent).
const lsjk::JTaskList* get_jtasks(const lsjk_pairlist* pl12c, const lsjk_pairlist* pl34, bool build = true)
{
    lsjk_pairlist* pl12 = const_cast<lsjk_pairlist*>(pl12c);
    auto it = pl12->jtasks.find(pl34->uid);
    if (it != pl12->jtasks.end()) return &it->second;
    if (!build) return nullptr;
    lsjk::JTaskList seg;
    std::vector<lsjk::JTask> tasks;
    for (int L12 = 0; L12 <= 4; L12++)
        for (int L34 = 0; L34 <= 4; L34++) {
            seg.off[L12][L34] = (int)tasks.size();
            int b0 = pl12->hstart[L12], b1 = pl12->hstart[L12 + 1];
            int k0 = pl34->hstart[L34], k1 = pl34->hstart[L34 + 1];
            if (b1 <= b0 || k1 <= k0) continue;
            if (pl12->nmol > 0) {
                append_batch_jtasks(pl12, L12, L34, false, tasks);
            } else {
                const long long ntile = (b1 - b0 + 63) / 64;
                const long long chunk = scaled_chunk(j_chunk(k1 - k0, ntile, L12 + L34),
                                                     pl12 == pl34 ? pl12->jscale[0][L12][L34] : 1.0f);
                for (int b = b0; b < b1; b += 64)
                    for (int k = k0; k < k1; k += (int)chunk)
                        tasks.push_back({b, std::min(64, b1 - b), k, std::min(k + (int)chunk, k1)});
            }
            seg.cnt[L12][L34] = (int)tasks.size() - seg.off[L12][L34];
        }
    return &pl12->jtasks.emplace(pl34->uid, upload_jtasks(tasks, seg)).first->second;
}

// Task list of the symmetric J kernel (one list on both sides): for the group pairs js_enabled
// admits, tiles of 64 bra pairs x ket chunks as in get_jtasks; within one group a tile's kets stop
// at its last bra pair (the triangle y <= x).  Built with the list.
void build_jtasks_sym(lsjk_pairlist* pl)
{
    lsjk::JTaskList seg;
    std::vector<lsjk::JTask> tasks;
    for (int L12 = 0; L12 <= 4; L12++)
        for (int L34 = 0; L34 <= 4; L34++) {
            seg.off[L12][L34] = (int)tasks.size();
            if (!lsjk::js_enabled(L12, L34)) continue;
            const int b0 = pl->hstart[L12], b1 = pl->hstart[L12 + 1];
            const int k0 = pl->hstart[L34], k1 = pl->hstart[L34 + 1];
            if (b1 <= b0 || k1 <= k0) continue;
            if (pl->nmol > 0) {
                append_batch_jtasks(pl, L12, L34, true, tasks);
                seg.cnt[L12][L34] = (int)tasks.size() - seg.off[L12][L34];
                continue;
            }
            const long long ntile = (b1 - b0 + 63) / 64;
            const long long chunk = scaled_chunk(j_chunk((k1 - k0) / (L12 == L34 ? 2 : 1), ntile, L12 + L34),
                                                 pl->jscale[1][L12][L34]);
            for (int b = b0; b < b1; b += 64) {
                const int kend = (L12 == L34) ? std::min(k1, b + 64) : k1;
                for (int k = k0; k < kend; k += (int)chunk)
                    tasks.push_back({b, std::min(64, b1 - b), k, std::min(k + (int)chunk, kend)});
            }
            seg.cnt[L12][L34] = (int)tasks.size() - seg.off[L12][L34];
        }
    pl->jtasks_sym = upload_jtasks(tasks, seg);
}

// Fork/join of the class launches over the side streams: launches are dispatched heaviest first,
// round-robin, so small classes run concurrently instead of leaving the device idle.
struct Fanout {
    int next = 0;
    void begin()
    {
        HIPCHK(hipEventRecord(S.fork, S.stream));
        for (int i = 0; i < kSide; i++) HIPCHK(hipStreamWaitEvent(S.side[i], S.fork, 0));
    }
    hipStream_t pick() { return S.side[(next++) % S.nside]; }
    void end()
    {
        for (int i = 0; i < kSide; i++) {
            HIPCHK(hipEventRecord(S.join[i], S.side[i]));
            HIPCHK(hipStreamWaitEvent(S.stream, S.join[i], 0));
        }
    }
};

// Per-launch timing: bracket a launch with an event pair from the pool when profiling.
struct LaunchTimer {
    int n = 0;
    bool on = S.profiling;
    hipStream_t cur = nullptr;
    void begin(int id, hipStream_t st)
    {
        cur = st;
        if (!on) return;
        if ((size_t)(2 * n + 2) > S.kev.size()) {
            size_t old = S.kev.size();
            S.kev.resize(old + 64);
            for (size_t i = old; i < S.kev.size(); i++) HIPCHK(hipEventCreate(&S.kev[i]));
        }
        S.kid.push_back(id);
        HIPCHK(hipEventRecord(S.kev[2 * n], cur));
    }
    void end()
    {
        if (!on) return;
        HIPCHK(hipEventRecord(S.kev[2 * n + 1], cur));
        n++;
    }
};

void finish_call(bool count, const lsjk::CallScalars* sc)
{
    HIPCHK(hipStreamSynchronize(S.stream));
    S.kms.assign(S.kid.size(), 0.0);
    for (size_t i = 0; i < S.kid.size(); i++) {
        float m = 0.0f;
        HIPCHK(hipEventElapsedTime(&m, S.kev[2 * i], S.kev[2 * i + 1]));
        S.kms[i] = m;
    }
    float m0 = 0.0f, m1 = 0.0f;
    HIPCHK(hipEventElapsedTime(&m0, S.ev[0], S.ev[3]));
    HIPCHK(hipEventElapsedTime(&m1, S.ev[1], S.ev[2]));
    S.ms[0] = m0;
    S.ms[1] = m1;
    if (count) {
        lsjk::CallScalars h;
        HIPCHK(hipMemcpy(&h, sc, sizeof(h), hipMemcpyDeviceToHost));
        std::memcpy(S.counts, h.counts, sizeof(S.counts));
        std::memcpy(S.counts32, h.counts32, sizeof(S.counts32));
        std::memcpy(S.wsteps, h.wsteps, sizeof(S.wsteps));
        std::memcpy(S.wlanes, h.wlanes, sizeof(S.wlanes));
    } else {
        std::memset(S.counts, 0, sizeof(S.counts));
        std::memset(S.counts32, 0, sizeof(S.counts32));
        std::memset(S.wsteps, 0, sizeof(S.wsteps));
        std::memset(S.wlanes, 0, sizeof(S.wlanes));
    }
}

}  // namespace

int lsjk::persistent_grid(const void* kernel, int block, size_t lds, long long nwork_groups)
{
    // Dynamic LDS per workgroup stays within 64 KB whatever the device reports: a launch that
    // requested more did not make progress on this stack.
    if (lds > std::min(S.max_lds, lsjk::kMaxLdsPerGroup))
        throw LsjkError(LSJK_E_UNSUPPORTED, "lsjk: kernel needs more LDS per workgroup than allowed");
    // called with the session lock held
    static std::map<std::tuple<const void*, int, size_t>, int> cache;
    const auto key = std::make_tuple(kernel, block, lds);
    auto it = cache.find(key);
    int per_cu = 0;
    if (it == cache.end()) {
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds));
        cache[key] = per_cu;
    } else {
        per_cu = it->second;
    }
    if (per_cu <= 0) throw LsjkError(LSJK_E_UNSUPPORTED, "lsjk: kernel cannot be resident at this block size");
    const long long g = (long long)per_cu * S.num_cu;
    return (int)std::max<long long>(1, std::min<long long>(g, nwork_groups));
}

// ------------------------------------------------------------------ learned screen
//
// A screen holds the trained network (lsjk_mlp_* layout), the normalisation of its inputs, and per
// K class the certified margin delta (log scale; NaN: the class is not certified).  For a work
// unit u of a certified class the factor tau_u = min(1, exp(f(x_u) + delta)) bounds, for every
// quartet of the unit with Schwarz product >= smin, the ratio of its largest integral component to
// its Schwarz product.  The factors are computed once per unit list and kept on the screen.
namespace {
struct LearnedScreen {
    int ni = 0;
    float* d_params = nullptr;
    std::vector<float> mean, scale;
    int nbin = 1;
    std::vector<float> fedge;    // 625 x (nbin - 1): bin edges of the network output per class
    std::vector<double> delta;   // 625 x nbin: margin per class and bin (NaN: not certified)
    float smin = 0.0f;
    std::map<std::tuple<unsigned long long, unsigned long long, long long>, std::pair<float*, int>> tau;
    std::map<std::tuple<unsigned long long, unsigned long long, long long>, unsigned long long> tau_gen;
};
}  // namespace

struct lsjk_screen : public LearnedScreen {};

namespace {

// Features of the K work units of one class, LSJK_SCREEN_NFEAT per unit, from the lists alone:
//   L1, L2, L3, L4;
//   ln(exponent of the bra row), ln(smallest exponent among the bra row's partners of the unit),
//   ln(smallest exponent among the ket group's rows), ln(smallest exponent among the unit's ket
//   partner entries);
//   |bra row - centroid of the ket group's rows|, the group's radius about its centroid,
//   largest bra pair extent |A - B| among the unit's bra partners, largest ket pair extent among the
//   unit's slots (bohr);
//   ln(largest bra bound of the unit), ln(largest ket bound of the unit's first chunk);
//   ln of the (ss|ss) ratio to its Schwarz product for the unit's most diffuse pairs at the closest
//   approach (centroid distance less the group radius and both pair extents) and at the centroid
//   distance.
// sqrt(2) (pq)^(1/4) (p + q)^(-1/2) F0(rho R^2), rho = pq / (p + q): (ss|ss) over its Schwarz product
// for pair exponents p, q at centre distance R (F0(T) = sqrt(pi / T) erf(sqrt T) / 2).
double ss_ratio(double p, double q, double R)
{
    const double rho = p * q / (p + q), T = rho * R * R;
    const double F0 = T < 1e-12 ? 1.0 - T / 3.0 : 0.5 * std::sqrt(M_PI / T) * std::erf(std::sqrt(T));
    return std::sqrt(2.0) * std::pow(p * q, 0.25) / std::sqrt(p + q) * F0;
}

void unit_features(const lsjk_pairlist* pl12, const lsjk_pairlist* pl34c, int L1, int L2, int L3, int L4,
                   const std::vector<lsjk::KUnitG>& units, std::vector<float>& out)
{
    constexpr int NF = LSJK_SCREEN_NFEAT;
    lsjk_pairlist* pl34 = const_cast<lsjk_pairlist*>(pl34c);
    const auto& G = get_kgroups(pl34, L3, L4, kgroup_for(pl34, L3, L4));
    const int gsize = G.gsize;
    const lsjk::Basis *b1 = pl12->b1, *b2 = pl12->b2, *b3 = pl34->b1, *b4 = pl34->b2;
    const auto& rows1 = pl12->h_rows[L1][L2];
    const int m12 = pl12->kmaxL2 + 1;
    auto dist = [](const double* a, const double* b) {
        return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
    };
    // per group: centroid, radius, smallest row exponent
    std::vector<double> gc((size_t)G.ngroup * 3, 0.0), grad(G.ngroup, 0.0), gemin(G.ngroup, 1e300);
    for (int g = 0; g < G.ngroup; g++) {
        int n = 0;
        for (int t = 0; t < gsize; t++) {
            const int r = G.grp_rows[(size_t)g * gsize + t];
            if (r < 0) continue;
            for (int d = 0; d < 3; d++) gc[3 * g + d] += b3->h_xyz[3 * r + d];
            gemin[g] = std::min(gemin[g], b3->h_e[r]);
            n++;
        }
        for (int d = 0; d < 3; d++) gc[3 * g + d] /= std::max(n, 1);
        for (int t = 0; t < gsize; t++) {
            const int r = G.grp_rows[(size_t)g * gsize + t];
            if (r >= 0) grad[g] = std::max(grad[g], dist(&b3->h_xyz[3 * r], &gc[3 * g]));
        }
    }
    // per 64-slot chunk: smallest partner exponent, largest pair extent
    const size_t nchunk = G.slot.size() / 64;
    std::vector<double> cemin(nchunk, 1e300), cext(nchunk, 0.0);
    for (size_t c = 0; c < nchunk; c++) {
        const int g = (int)(std::upper_bound(G.grp_slot.begin(), G.grp_slot.end(), (int)(c * 64)) - G.grp_slot.begin()) - 1;
        for (int l = 0; l < 64; l++) {
            const int2 sl = G.slot[c * 64 + l];
            if (sl.x < 0) continue;
            const int r3 = G.grp_rows[(size_t)g * gsize + sl.y];
            const int p4 = pl34->h_eprim[sl.x];
            cemin[c] = std::min(cemin[c], b4->h_e[p4]);
            cext[c] = std::max(cext[c], dist(&b3->h_xyz[3 * r3], &b4->h_xyz[3 * p4]));
        }
    }
    out.assign(units.size() * NF, 0.0f);
    for (size_t q = 0; q < units.size(); q++) {
        const lsjk::KUnitG& u = units[q];
        const int r1 = rows1[u.i1];
        int a0 = pl12->h_row_start[(size_t)r1 * m12 + L2], a1 = pl12->h_row_start[(size_t)r1 * m12 + L2 + 1];
        if (u.prange != 0) {
            a0 = std::min(a1, a0 + (u.prange & 0xffff));
            if (u.prange >> 16) a1 = std::min(a1, a0 + (u.prange >> 16));
        }
        const double* A = &b1->h_xyz[3 * r1];
        double emin2 = 1e300, ext12 = 0.0;
        for (int k = a0; k < a1; k++) {
            const int p2 = pl12->h_eprim[k];
            emin2 = std::min(emin2, b2->h_e[p2]);
            ext12 = std::max(ext12, dist(A, &b2->h_xyz[3 * p2]));
        }
        const size_t c0 = (size_t)G.grp_slot[u.g] / 64 + (u.crange & 0xffff);
        const size_t c1 = std::min((size_t)G.grp_slot[u.g + 1] / 64, c0 + (size_t)(u.crange >> 16));
        double emin4 = 1e300, ext34 = 0.0;
        for (size_t c = c0; c < c1; c++) {
            emin4 = std::min(emin4, cemin[c]);
            ext34 = std::max(ext34, cext[c]);
        }
        auto lg = [](double v) { return (float)std::log(std::max(v, 1e-300)); };
        float* f = &out[q * NF];
        f[0] = (float)L1;
        f[1] = (float)L2;
        f[2] = (float)L3;
        f[3] = (float)L4;
        f[4] = lg(b1->h_e[r1]);
        f[5] = lg(a1 > a0 ? emin2 : b1->h_e[r1]);
        f[6] = lg(gemin[u.g]);
        f[7] = lg(c1 > c0 ? emin4 : gemin[u.g]);
        f[8] = (float)dist(A, &gc[3 * u.g]);
        f[9] = (float)grad[u.g];
        f[10] = (float)ext12;
        f[11] = (float)ext34;
        f[12] = lg(a1 > a0 ? (double)pl12->h_eb[a0] : 1e-300);
        f[13] = lg(c0 < nchunk ? (double)G.chunk_b[c0] : 1e-300);
        // s-type baseline: the exact (ss|ss) ratio to its Schwarz product,
        // sqrt(2) (pq)^(1/4) (p + q)^(-1/2) F0(rho R^2), for the unit's most diffuse pairs at the
        // closest approach the geometry allows and at the centroid distance
        const double p = b1->h_e[r1] + (a1 > a0 ? emin2 : b1->h_e[r1]);
        const double qq = gemin[u.g] + (c1 > c0 ? emin4 : gemin[u.g]);
        const double R = f[8], Rmin = std::max(0.0, R - f[9] - f[10] - f[11]);
        f[14] = lg(ss_ratio(p, qq, Rmin));
        f[15] = lg(ss_ratio(p, qq, R));
    }
}

std::vector<lsjk::KUnitG> download_units(const lsjk::KUnitG* d, int n)
{
    std::vector<lsjk::KUnitG> h((size_t)std::max(n, 0));
    if (n > 0) HIPCHK(hipMemcpy(h.data(), d, (size_t)n * sizeof(lsjk::KUnitG), hipMemcpyDeviceToHost));
    return h;
}

// The screen's factors for the units of one class (computed and cached at first use): network
// over the normalised unit features, tau = min(1, exp(f + delta)).
const float* screen_tau(lsjk_screen* sc, const lsjk_pairlist* pl12, const lsjk_pairlist* pl34, int L1, int L2,
                        int L3, int L4, bool tri, float thresp, const lsjk::KUnitG* units, int nunits)
{
    unsigned int tb;
    std::memcpy(&tb, &thresp, sizeof(tb));
    const long long cls = ((long long)tb << 8) | (((((L1 * 3 + L2) * 3 + L3) * 3 + L4) << 1) | (tri ? 1 : 0));
    const auto key = std::make_tuple(pl12->uid, pl34->uid, cls);
    auto it = sc->tau.find(key);
    if (it != sc->tau.end() && it->second.second == nunits && sc->tau_gen[key] == pl12->kunits_gen)
        return it->second.first;
    const int ci = ((L1 * 5 + L2) * 5 + L3) * 5 + L4;
    std::vector<float> feat;
    unit_features(pl12, pl34, L1, L2, L3, L4, download_units(units, nunits), feat);
    constexpr int NF = LSJK_SCREEN_NFEAT;
    std::vector<float> x((size_t)nunits * sc->ni, 0.0f);
    for (int u = 0; u < nunits; u++)
        for (int k = 0; k < sc->ni && k < NF; k++)
            x[(size_t)u * sc->ni + k] = (feat[(size_t)u * NF + k] - sc->mean[k]) / sc->scale[k];
    float* dx = dev_upload(x);
    float* df = nullptr;
    std::vector<float> f(nunits);
    try {
        HIPCHK(hipMalloc((void**)&df, std::max(nunits, 1) * sizeof(float)));
        lsjk::launch_mlp_forward(sc->d_params, sc->ni, nunits, dx, df, S.stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(S.stream));
        if (nunits) HIPCHK(hipMemcpy(f.data(), df, nunits * sizeof(float), hipMemcpyDeviceToHost));
    } catch (...) {
        dev_free(dx);
        dev_free(df);
        throw;
    }
    dev_free(dx);
    dev_free(df);
    std::vector<float> tau(nunits);
    for (int u = 0; u < nunits; u++) {
        // the margin of the output's bin (edges ascending); an uncertified bin keeps tau = 1
        int b = 0;
        while (b + 1 < sc->nbin && f[u] >= sc->fedge[(size_t)ci * (sc->nbin - 1) + b]) b++;
        const double d = sc->delta[(size_t)ci * sc->nbin + b];
        tau[u] = (d == d) ? (float)std::min(1.0, std::max(1e-30, std::exp((double)f[u] + d))) : 1.0f;
    }
    if (it != sc->tau.end()) {
        dev_free(it->second.first);
        sc->tau.erase(it);
    }
    float* d = dev_upload(tau);
    sc->tau[key] = {d, nunits};
    sc->tau_gen[key] = pl12->kunits_gen;
    return d;
}

}  // namespace

// How a build call runs.  ws: a fixed workspace of ws_bytes for the per-call buffers (nullptr:
// the session workspace, grown as needed).  noalloc: allocate no device memory (the J task list
// must exist already; lsjk_coulomb_noalloc).  graph: only enqueue the launches on the library
// stream (it is being captured into a hipGraph): no timing events, no synchronisation, no
// mirror-class timing, no per-launch timers.
struct CallOpts {
    void* ws = nullptr;
    size_t ws_bytes = 0;
    bool noalloc = false;
    bool graph = false;
    const char* who = nullptr;
    lsjk_screen* screen = nullptr;  // exchange: the learned route (certified classes)
    bool tune_orient = false;       // exchange: time both orientations of each mirror-class pair
    std::map<int, double>* orient_ms = nullptr;     // with tune_orient: each timed class's ms
    const std::map<int, int>* orient_set = nullptr; // exchange: store and use these orientations
};

void commit_ws(Workspace& wsp, const CallOpts& o, const char* who)
{
    if (o.ws) {
        wsp.commit(o.ws, o.ws_bytes, who);
    } else {
        if (o.noalloc || o.graph) throw LsjkError(LSJK_E_INVALID, std::string(who) + ": no fixed workspace given");
        wsp.commit();
    }
}

// lsjk_coulomb: workspace, task lists, the launches on the library stream.
static void coulomb_call(const lsjk_pairlist* pl12, const lsjk_pairlist* pl34, const double* D34_dev, int nterm,
                         const double* scales, const double* omegas, float thresp, float thredp, double* J12_dev,
                         const CallOpts& o = CallOpts())
{
        using namespace lsjk;
        const bool noalloc = o.noalloc;
        const char* who = o.who ? o.who : "lsjk_coulomb";
        require_init();
        if (!pl12 || !pl34 || !D34_dev || !J12_dev) throw LsjkError(LSJK_E_INVALID, std::string(who) + ": null argument");
        // mixed precision (intbox.hpp:736-742) when thredp > thresp: FP64 at or above thredp,
        // FP32 between thresp and thredp
        const bool mixed = thredp > thresp;
        Ewald ew;
        check_ewald(nterm, scales, omegas, ew);
        hipStream_t st = S.stream;
        const Basis *b1 = pl12->b1, *b2 = pl12->b2, *b3 = pl34->b1, *b4 = pl34->b2;
        // workspace
        float* kb = nullptr;
        double* kd = nullptr;
        double* jp = nullptr;
        Workspace wsp;
        wsp.add(&kb, (size_t)pl34->npair);
        wsp.add(&kd, pl34->hoff[2 * kMaxL + 1]);
        wsp.add(&jp, pl12->hoff[2 * kMaxL + 1]);
        float* kd32 = nullptr;
        if (mixed) wsp.add(&kd32, pl34->hoff[2 * kMaxL + 1]);
        CallScalars* sc = S.d_sc;
        commit_ws(wsp, o, who);
        // one list on both sides (FP64 route): the symmetric kernel takes the group pairs
        // js_enabled admits, each integral once for both J blocks
        const bool jsym = (pl12 == pl34);
        const JTaskList* segp = get_jtasks(pl12, pl34, !noalloc && !o.graph);
        if (!segp)
            throw LsjkError(LSJK_E_INVALID, std::string(who) + ": no J task list for this pair of lists (one is built "
                                            "when a list is created, for that list on both sides, and by the first "
                                            "lsjk_coulomb call of any other pair)");
        const JTaskList& seg = *segp;
        const JTaskList& sseg = pl12->jtasks_sym;

        const bool timed = !o.graph;
        if (timed) HIPCHK(hipEventRecord(S.ev[0], st));
        launch_fill(sc, sizeof(CallScalars), st);
        launch_fill(jp, std::max<size_t>(1, pl12->hoff[2 * kMaxL + 1]) * sizeof(double), st);
        const float max12 = pl12->maxbound;
        const float thre_ket = thresp / max12;

        JPrepArgs pa{};
        pa.p1 = pl34->d_p1;
        pa.p2 = pl34->d_p2;
        pa.bound = pl34->d_bound;
        pa.hpos = pl34->d_hpos;
        pa.K = pl34->d_K;
        pa.eA = b3->d_e; pa.xA = b3->d_x; pa.yA = b3->d_y; pa.zA = b3->d_z; pa.cartA = b3->d_cart;
        pa.eB = b4->d_e; pa.xB = b4->d_x; pa.yB = b4->d_y; pa.zB = b4->d_z; pa.cartB = b4->d_cart;
        pa.D = D34_dev;
        pa.ldD = mat_ld(b4);
        pa.symmetric = pl34->symmetric;
        pa.thre_ket = thre_ket;
        pa.kb = kb;
        pa.kd = kd;
        pa.kd32 = kd32;
        pa.sc = sc;
        for (size_t c = 0; c + 1 < pl34->cls_start.size(); c++) {
            int L = pl34->cls_L1[c] + pl34->cls_L2[c];
            launch_j_ket_prep(pl34->cls_L1[c], pl34->cls_L2[c], pa, pl34->cls_start[c],
                              pl34->cls_start[c + 1] - pl34->cls_start[c], pl34->hstart[L], pl34->hoff[L], st);
        }
        HIPCHK(hipGetLastError());

        if (timed) HIPCHK(hipEventRecord(S.ev[1], st));
        S.kid.clear();
        LaunchTimer lt;
        if (noalloc || o.graph) lt.on = false;  // the per-launch event pool would grow
        if (ew.n > 0) {
            Fanout fo;
            fo.begin();
            std::vector<std::pair<int, int>> order;
            for (int L12 = 4; L12 >= 0; L12--)
                for (int L34 = 4; L34 >= 0; L34--) order.push_back({L12, L34});
            std::stable_sort(order.begin(), order.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
                return a.first + a.second > b.first + b.second;
            });
            for (auto& o : order) {
                {
                    const int L12 = o.first, L34 = o.second;
                    const bool sym = jsym && js_enabled(L12, L34);
                    if (jsym && !sym && js_enabled(L34, L12)) continue;  // served by (L34, L12)
                    if ((sym ? sseg.cnt : seg.cnt)[L12][L34] == 0) continue;
                    JQuartetArgs qa{};
                    qa.rys = S.rys;
                    qa.ew = ew;
                    qa.thresp = thresp;
                    qa.count = S.counting ? 1 : 0;
                    qa.count_cls = jcount_index(L12, L34);
                    qa.bra_b = pl12->d_hb;
                    qa.bra_geom = pl12->d_hgeom;
                    qa.bra_n = pl12->npair;
                    qa.bra_hstart = pl12->hstart[L12];
                    qa.jp = jp + pl12->hoff[L12];
                    qa.ket_b = kb;
                    qa.ket_static_b = pl34->d_hb;
                    qa.ket_geom = pl34->d_hgeom;
                    qa.ket_n = pl34->npair;
                    qa.ket_hstart = pl34->hstart[L34];
                    qa.ket_d = kd + pl34->hoff[L34];
                    qa.ket_d32 = kd32 ? kd32 + pl34->hoff[L34] : nullptr;
                    qa.thredp = thredp;
                    qa.sc = sc;
                    int idx = jcount_index(L12, L34);
                    // the mixed route's kernel, or the FP64 one where the autotuner found it faster
                    auto jroute = [&](int lid) {
                        const bool plain = pl12 == pl34 && (pl12->force_plain || pl12->plain_on_mixed.count(lid));
                        return mixed && !plain ? kRouteMixed : kRoutePlain;
                    };
                    hipStream_t ss = fo.pick();
                    if (sym) {
                        // the bra group's densities and the ket group's accumulators, for the
                        // second use; launch id: bit 24 set, the other class's index + 1 in bits
                        // 12..23 when it differs
                        qa.bra_d = kd + pl12->hoff[L12];
                        qa.jq = jp + pl12->hoff[L34];
                        qa.count_cls2 = jcount_index(L34, L12);
                        const JTask* tk = (const JTask*)sseg.d_tasks + sseg.off[L12][L34];
                        const int lid = idx | (1 << 24) | (L12 != L34 ? (jcount_index(L34, L12) + 1) << 12 : 0);
                        lt.begin(lid, ss);
                        launch_j_quartets_sym(L12, L34, jroute(lid), qa, tk, sseg.cnt[L12][L34], &sc->queue[idx], ss);
                        lt.end();
                    } else {
                        const JTask* tk = (const JTask*)seg.d_tasks + seg.off[L12][L34];
                        lt.begin(idx, ss);
                        launch_j_quartets(L12, L34, jroute(idx), qa, tk, seg.cnt[L12][L34], &sc->queue[idx], ss);
                        lt.end();
                    }
                }
            }
            fo.end();
            HIPCHK(hipGetLastError());
        }
        if (timed) HIPCHK(hipEventRecord(S.ev[2], st));

        JScatterArgs sa{};
        sa.p1 = pl12->d_p1;
        sa.p2 = pl12->d_p2;
        sa.bound = pl12->d_bound;
        sa.hpos = pl12->d_hpos;
        sa.K = pl12->d_K;
        sa.eA = b1->d_e; sa.xA = b1->d_x; sa.yA = b1->d_y; sa.zA = b1->d_z; sa.cartA = b1->d_cart;
        sa.eB = b2->d_e; sa.xB = b2->d_x; sa.yB = b2->d_y; sa.zB = b2->d_z; sa.cartB = b2->d_cart;
        sa.jp = jp;
        sa.J = J12_dev;
        sa.ldJ = mat_ld(b2);
        sa.symmetric = pl12->symmetric;
        sa.screen = 1;
        sa.thresp = thresp;
        sa.sc = sc;
        if (ew.n > 0) {
            for (size_t c = 0; c + 1 < pl12->cls_start.size(); c++) {
                int L = pl12->cls_L1[c] + pl12->cls_L2[c];
                launch_j_bra_scatter(pl12->cls_L1[c], pl12->cls_L2[c], sa, pl12->cls_start[c],
                                     pl12->cls_start[c + 1] - pl12->cls_start[c], pl12->hstart[L], pl12->hoff[L], st);
            }
            HIPCHK(hipGetLastError());
        }
        if (timed) {
            HIPCHK(hipEventRecord(S.ev[3], st));
            finish_call(S.counting, sc);
        }
}


namespace {
void basis_free(lsjk_basis* b)
{
    if (!b) return;
    dev_free(b->d_atom);
    dev_free(b->d_c);
    dev_free(b->d_e);
    dev_free(b->d_x);
    dev_free(b->d_y);
    dev_free(b->d_z);
    dev_free(b->d_L);
    dev_free(b->d_cart);
    dev_free(b->d_blk);
    dev_free(b->d_blk_cart);
    dev_free(b->d_blk_L);
    dev_free(b->d_blk_lo);
    dev_free(b->d_blk_hi);
    dev_free(b->d_kcart);
    dev_free(b->d_kblk);
    delete b;
}

// K view order of a basis's blocks (Basis::kperm): by angular momentum, then by the Morton code
// of the shell centre (10 bits per axis over the basis's bounding box), then by block index; the
// components of a block stay contiguous.  The identity when kview is false or when the blocks do
// not tile [0, ncart).
void set_kview(lsjk::Basis* b, bool kview)
{
    const int nblk = b->nblk, ncart = b->ncart;
    b->h_kblk.resize(nblk);
    b->h_kcart.resize(ncart);
    for (int k = 0; k < nblk; k++) b->h_kblk[k] = k;
    for (int i = 0; i < ncart; i++) b->h_kcart[i] = i;
    b->kperm = false;
    if (!kview || nblk == 0) return;
    std::vector<int> cover(ncart, 0);
    for (int k = 0; k < nblk; k++)
        for (int c = 0; c < lsjk::ncart(b->h_blk_L[k]); c++) cover[b->h_blk_cart[k] + c]++;
    for (int i = 0; i < ncart; i++)
        if (cover[i] != 1) return;
    std::vector<int> first(nblk, -1);
    for (int i = 0; i < b->nprim; i++)
        if (first[b->h_blk[i]] < 0) first[b->h_blk[i]] = i;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < b->nprim; i++)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], b->h_xyz[3 * (size_t)i + d]);
            hi[d] = std::max(hi[d], b->h_xyz[3 * (size_t)i + d]);
        }
    auto spread = [](unsigned v) {  // 10 bits -> every third bit
        unsigned long long x = v & 1023u, r = 0;
        for (int t = 0; t < 10; t++) r |= ((x >> t) & 1ull) << (3 * t);
        return r;
    };
    std::vector<unsigned long long> code(nblk);
    for (int k = 0; k < nblk; k++) {
        unsigned long long m = 0;
        for (int d = 0; d < 3; d++) {
            const double w = hi[d] - lo[d];
            const double t = w > 0.0 ? (b->h_xyz[3 * (size_t)first[k] + d] - lo[d]) / w : 0.0;
            m |= spread((unsigned)std::min(1023.0, std::max(0.0, t * 1023.0))) << d;
        }
        code[k] = m;
    }
    std::vector<int> order(nblk);
    for (int k = 0; k < nblk; k++) order[k] = k;
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        if (b->h_blk_L[x] != b->h_blk_L[y]) return b->h_blk_L[x] < b->h_blk_L[y];
        if (code[x] != code[y]) return code[x] < code[y];
        return x < y;
    });
    int pos = 0;
    for (int r = 0; r < nblk; r++) {
        const int k = order[r];
        b->h_kblk[k] = r;
        for (int c = 0; c < lsjk::ncart(b->h_blk_L[k]); c++) b->h_kcart[b->h_blk_cart[k] + c] = pos + c;
        pos += lsjk::ncart(b->h_blk_L[k]);
    }
    for (int i = 0; i < ncart; i++) b->kperm = b->kperm || b->h_kcart[i] != i;
    b->d_kcart = dev_upload(b->h_kcart);
    b->d_kblk = dev_upload(b->h_kblk);
}

// A device basis from primitive arrays (lsjk_basis_create; also the combined basis of a batch).
lsjk_basis* basis_new(int nprim, const double* c, const double* e, const double* xyz, const int* L, const int* cart,
                      int ncart, bool kview = true)
{
    lsjk_basis* b = new lsjk_basis();
    try {
        b->nprim = nprim;
        b->ncart = ncart;
        b->h_c.assign(c, c + nprim);
        b->h_e.assign(e, e + nprim);
        b->h_xyz.assign(xyz, xyz + 3 * (size_t)nprim);
        b->h_L.assign(L, L + nprim);
        b->h_cart.assign(cart, cart + nprim);
        b->maxL = 0;
        for (int i = 0; i < nprim; i++) {
            if (L[i] < 0 || L[i] > lsjk::kMaxL)
                throw LsjkError(LSJK_E_UNSUPPORTED, "lsjk_basis_create: angular momentum above d is not supported");
            if (cart[i] < 0 || cart[i] + lsjk::ncart(L[i]) > ncart)
                throw LsjkError(LSJK_E_INVALID, "lsjk_basis_create: Cartesian offset out of range");
            b->maxL = std::max(b->maxL, L[i]);
        }
        // blocks: distinct Cartesian offsets (one per shell)
        b->h_blk.resize(nprim);
        for (int i = 0; i < nprim; i++) {
            if (i == 0 || cart[i] != cart[i - 1]) {
                b->h_blk_cart.push_back(cart[i]);
                b->h_blk_L.push_back(L[i]);
            } else if (L[i] != L[i - 1]) {
                throw LsjkError(LSJK_E_INVALID, "lsjk_basis_create: primitives sharing a Cartesian offset differ in L");
            }
            b->h_blk[i] = (int)b->h_blk_cart.size() - 1;
        }
        b->nblk = (int)b->h_blk_cart.size();
        std::vector<double> x(nprim), y(nprim), z(nprim);
        for (int i = 0; i < nprim; i++) {
            x[i] = xyz[3 * i];
            y[i] = xyz[3 * i + 1];
            z[i] = xyz[3 * i + 2];
        }
        b->d_c = dev_upload(b->h_c);
        b->d_e = dev_upload(b->h_e);
        b->d_x = dev_upload(x);
        b->d_y = dev_upload(y);
        b->d_z = dev_upload(z);
        b->d_L = dev_upload(b->h_L);
        b->d_cart = dev_upload(b->h_cart);
        b->d_blk = dev_upload(b->h_blk);
        b->d_blk_cart = dev_upload(b->h_blk_cart);
        b->d_blk_L = dev_upload(b->h_blk_L);
        set_kview(b, kview);
    } catch (...) {
        basis_free(b);
        throw;
    }
    return b;
}

}  // namespace

// lsjk_exchange: workspace, unit lists, mirror-class choices, the launches on the library stream.
static void exchange_call(const lsjk_pairlist* pl12, const lsjk_pairlist* pl34, const double* D24_dev, int D24symm,
                          int nterm, const double* scales, const double* omegas, float thresp, float thredp,
                          double* K13_dev, const CallOpts& o = CallOpts())
{
        using namespace lsjk;
        const char* who = o.who ? o.who : "lsjk_exchange";
        require_init();
        if (!pl12 || !pl34 || !D24_dev || !K13_dev) throw LsjkError(LSJK_E_INVALID, "lsjk_exchange: null argument");
        // mixed precision when thredp > thresp (intbox.hpp:787-793)
        const bool mixed = thredp > thresp;
        Ewald ew;
        check_ewald(nterm, scales, omegas, ew);
        hipStream_t st = S.stream;
        const Basis *b1 = pl12->b1, *b2 = pl12->b2, *b3 = pl34->b1, *b4 = pl34->b2;
        const bool symmode = D24symm && (pl12 == pl34);
        // mirror-class choices of the symmetric mode (see below); a scratch K while they are timed
        unsigned tb;
        std::memcpy(&tb, &thresp, sizeof(tb));
        // keyed by route as well: the two routes' kernels differ in speed
        const auto okey = std::make_pair(pl34->uid, ((unsigned long long)tb << 2) | (mixed ? 1ull : 0ull) |
                                                        (o.screen ? 2ull : 0ull));
        auto& ocache = const_cast<lsjk_pairlist*>(pl12)->korient;
        // the mirror-class orientations are timed only when the autotuner asks (o.tune_orient); a
        // first call otherwise takes the deterministic cost-model choice below
        const bool tune = symmode && ew.n > 0 && o.tune_orient && !o.orient_set && !o.graph && !o.noalloc;
        std::map<int, int> tuned;  // filled by the tuning below, stored only once it is complete
        const std::map<int, int>* orient = nullptr;
        if (symmode && o.orient_set) {
            orient = &(ocache[okey] = *o.orient_set);
        } else if (symmode && !tune) {
            auto it = ocache.find(okey);
            // a learned-route call without choices of its own takes its route's tuned ones
            if (it == ocache.end() && o.screen) it = ocache.find(std::make_pair(okey.first, okey.second & ~2ull));
            if (it != ocache.end()) orient = &it->second;
        }
        double* Dsh = nullptr;
        double* Ks = nullptr;
        // K view copies of D24 and of its block maxima (Basis::kperm): the kernel gathers from them
        double* Dk = nullptr;
        double* Dshk = nullptr;
        const bool kview = b2->kperm || b4->kperm;
        Workspace wsp;
        wsp.add(&Dsh, blk_elems(b2, b4));
        if (kview) {
            wsp.add(&Dk, mat_elems(b2, b4));
            wsp.add(&Dshk, blk_elems(b2, b4));
        }
        if (tune) wsp.add(&Ks, mat_elems(b1, b3));
        commit_ws(wsp, o, who);
        const bool timed = !o.graph;

        if (timed) HIPCHK(hipEventRecord(S.ev[0], st));
        launch_fill(S.d_sc, sizeof(CallScalars), st);
        launch_k_blockmax(D24_dev, mat_ld(b4), b2->d_blk_cart, b2->d_blk_L, b2->nblk, b4->d_blk_cart, b4->d_blk_L,
                          b4->nblk, Dsh, blk_ld(b4), b2->d_blk_lo, b2->d_blk_hi, b2->ldblk, S.d_sc, st);
        if (kview) {
            launch_permute(D24_dev, mat_ld(b4), b2->ncart, b4->ncart, b2->kperm ? b2->d_kcart : nullptr,
                           b4->kperm ? b4->d_kcart : nullptr, Dk, mat_ld(b4), st);
            launch_permute(Dsh, blk_ld(b4), b2->nblk, b4->nblk, b2->kperm ? b2->d_kblk : nullptr,
                           b4->kperm ? b4->d_kblk : nullptr, Dshk, blk_ld(b4), st);
        } else {
            Dk = const_cast<double*>(D24_dev);
            Dshk = Dsh;
        }
        HIPCHK(hipGetLastError());
        if (timed) HIPCHK(hipEventRecord(S.ev[1], st));
        S.kid.clear();
        LaunchTimer lt;
        if (o.noalloc || o.graph) lt.on = false;
        if (ew.n > 0) {
            Fanout fo;
            struct KL {
                int L1, L2, L3, L4;
                const KUnitG* units;
                int nunits;
                double cost;
                KArgs ka;
                KGroups kg;
                int route;
            };
            // The launch of one class: tri = the triangle rule (row1 rank <= row3 rank, equal row
            // momenta of one list); K = the output.  False when the class has no work.
            auto make_kl = [&](int L1, int L2, int L3, int L4, bool tri, double* K, KL& out) {
                const auto& rows1 = pl12->h_rows[L1][L2];
                const auto& rows3 = pl34->h_rows[L3][L4];
                if (rows1.empty() || rows3.empty()) return false;
                double cost = 0.0;
                const auto units = get_kunits_g(pl12, pl34, L1, L2, L3, L4, tri, thresp, &cost);
                if (units.second == 0) return false;
                const KGroups kg = kgroups_args(get_kgroups(const_cast<lsjk_pairlist*>(pl34), L3, L4,
                                                            kgroup_for(pl34, L3, L4)));
                KArgs ka{};
                ka.rys = S.rys;
                ka.ew = ew;
                ka.thresp = thresp;
                ka.thredp = thredp;
                ka.count = S.counting ? 1 : 0;
                ka.symmetric = symmode ? 1 : 0;
                ka.tri = tri ? 1 : 0;
                const lsjk_pairlist* sides[2] = {pl12, pl34};
                KSide* ks[2] = {&ka.s12, &ka.s34};
                const int Ls[2] = {L1, L3}, Lps[2] = {L2, L4};
                for (int q = 0; q < 2; q++) {
                    const lsjk_pairlist* pl = sides[q];
                    KSide& sd = *ks[q];
                    sd.row_start = pl->d_row_start;
                    sd.maxLp = pl->kmaxL2;
                    sd.eb = pl->d_ebound;
                    sd.ecart = pl->d_ecart;
                    sd.eblk = pl->d_keblk;  // the K view order of Dshk
                    sd.eg = pl->d_egeom;
                    sd.eg32 = pl->d_egeom32;
                    sd.nent = pl->nent;
                    sd.rows = pl->d_rows[Ls[q]][Lps[q]];
                    sd.nrows = (int)pl->h_rows[Ls[q]][Lps[q]].size();
                    sd.x = pl->b1->d_x;
                    sd.y = pl->b1->d_y;
                    sd.z = pl->b1->d_z;
                    sd.cart = pl->b1->d_cart;
                }
                // k_ket gathers D and the block maxima at the entries' K view ids
                ka.D = Dk;
                ka.ldD = mat_ld(b4);
                ka.Dsh = Dshk;
                ka.ldDsh = blk_ld(b4);
                ka.K = K;
                ka.ldK = mat_ld(b3);
                ka.maxb12 = pl12->maxbound;
                ka.maxb34 = pl34->maxbound;
                ka.sc = S.d_sc;
                // class cost: candidate pairs x FLOPs of one quartet of the class
                const int Lt = L1 + L2 + L3 + L4;
                const double Pi = (double)ncart(L1) * ncart(L2) * ncart(L3) * ncart(L4);
                const double G = (double)(L1 + L2 + 1) * (L3 + L4 + 1);
                const double F = 40.0 + (Lt / 2 + 1) * (30.0 + 9.0 * G + 3.0 * Pi) + 2.0 * Pi;
                int route = mixed ? kRouteMixed : kRoutePlain;
                // the FP64 kernel where the autotuner found it faster on the mixed route
                if (mixed && pl12 == pl34 &&
                    (pl12->force_plain || pl12->plain_on_mixed.count(kcount_index(L1, L2, L3, L4))))
                    route = kRoutePlain;
                if (o.screen) {
                    // learned route: certified classes (s and p functions) take their units'
                    // factors; every other quartet is evaluated in FP64 under the plain test
                    const int ci = ((L1 * 5 + L2) * 5 + L3) * 5 + L4;
                    bool cert = false;
                    for (int b = 0; b < o.screen->nbin; b++) {
                        const double dl = o.screen->delta[(size_t)ci * o.screen->nbin + b];
                        cert = cert || dl == dl;
                    }
                    if (cert && L1 <= 1 && L2 <= 1 && L3 <= 1 && L4 <= 1) {
                        ka.tau = screen_tau(o.screen, pl12, pl34, L1, L2, L3, L4, tri, thresp, units.first,
                                            units.second);
                        ka.tau_smin = o.screen->smin;
                    } else {
                        route = kRoutePlain;
                    }
                }
                out = KL{L1, L2, L3, L4, units.first, units.second, cost * F, ka, kg, route};
                return true;
            };
            // Symmetric mode (K = K^T): a class and its mirror (L3 L4 | L1 L2) hold the same quartets
            // with bra and ket swapped, so one of them, over all row pairs and with the transposed
            // block, gives both.  Which one runs faster depends on the class shapes and the lists.
            // By default the orientation with fewer estimated wave steps is kept (the unit lists'
            // cost: (bra partner, ket chunk) steps that pass the bound test with |D| <= 1), a
            // deterministic choice; the autotuner times both (twice each, on the library stream,
            // into a scratch K) and keeps the faster.  A class that is its own mirror takes the row
            // triangle instead.
            if (symmode && !tune && !orient && ew.n > 0 && !o.graph) {
                std::map<int, int> pick;
                for (int L1 = 0; L1 <= b1->maxL; L1++)
                    for (int L2 = 0; L2 <= b2->maxL; L2++)
                        for (int L3 = 0; L3 <= b3->maxL; L3++)
                            for (int L4 = 0; L4 <= b4->maxL; L4++) {
                                const int c = kcount_index(L1, L2, L3, L4), m = kcount_index(L3, L4, L1, L2);
                                if (c >= m) continue;
                                auto steps = [&](int A, int B, int C, int D_) {
                                    if (pl12->h_rows[A][B].empty() || pl34->h_rows[C][D_].empty()) return 0.0;
                                    double cost = 0.0;
                                    get_kunits_g(pl12, pl34, A, B, C, D_, false, thresp, &cost);
                                    return cost;
                                };
                                const double sc = steps(L1, L2, L3, L4), sm = steps(L3, L4, L1, L2);
                                pick[c] = pick[m] = (sc <= sm) ? c : m;
                            }
                orient = &(ocache[okey] = pick);
            }
            if (tune) {
                // events and the counting flag are restored whatever happens in between
                struct Guard {
                    hipEvent_t e0 = nullptr, e1 = nullptr;
                    bool counting = S.counting;
                    ~Guard()
                    {
                        if (e0) (void)hipEventDestroy(e0);
                        if (e1) (void)hipEventDestroy(e1);
                        S.counting = counting;
                    }
                } g;
                HIPCHK(hipEventCreate(&g.e0));
                HIPCHK(hipEventCreate(&g.e1));
                S.counting = false;
                auto time_class = [&](int L1, int L2, int L3, int L4) {
                    KL k;
                    if (!make_kl(L1, L2, L3, L4, false, Ks, k)) return 0.0f;
                    k.ka.count = 0;
                    HIPCHK(hipEventRecord(g.e0, st));
                    launch_k_ket(L1, L2, L3, L4, k.route, k.ka, k.kg, k.units, k.nunits, st);
                    HIPCHK(hipEventRecord(g.e1, st));
                    HIPCHK(hipEventSynchronize(g.e1));
                    float ms = 0.0f;
                    HIPCHK(hipEventElapsedTime(&ms, g.e0, g.e1));
                    return ms;
                };
                for (int L1 = 0; L1 <= b1->maxL; L1++)
                    for (int L2 = 0; L2 <= b2->maxL; L2++)
                        for (int L3 = 0; L3 <= b3->maxL; L3++)
                            for (int L4 = 0; L4 <= b4->maxL; L4++) {
                                const int c = kcount_index(L1, L2, L3, L4), m = kcount_index(L3, L4, L1, L2);
                                if (c >= m) continue;  // each pair once; self-mirrors need no choice
                                // twice each, alternating; the faster of each pair of runs counts
                                float tc = time_class(L1, L2, L3, L4), tm = time_class(L3, L4, L1, L2);
                                tc = std::min(tc, time_class(L1, L2, L3, L4));
                                tm = std::min(tm, time_class(L3, L4, L1, L2));
                                tuned[c] = tuned[m] = (tc <= tm) ? c : m;
                                if (o.orient_ms) {
                                    (*o.orient_ms)[c] = tc;
                                    (*o.orient_ms)[m] = tm;
                                }
                            }
                orient = &(ocache[okey] = tuned);
            }
            std::vector<KL> kls;
            for (int L1 = 0; L1 <= b1->maxL; L1++)
                for (int L3 = 0; L3 <= b3->maxL; L3++)
                    for (int L2 = 0; L2 <= b2->maxL; L2++)
                        for (int L4 = 0; L4 <= b4->maxL; L4++) {
                            bool tri = false;
                            if (symmode) {
                                const int c = kcount_index(L1, L2, L3, L4), m = kcount_index(L3, L4, L1, L2);
                                if (c == m) {
                                    tri = true;
                                } else {
                                    // the stored choice, else the class of the lower index (both
                                    // classes of a pair never run together)
                                    const auto it = orient ? orient->find(c) : std::map<int, int>::const_iterator();
                                    const int pick = (orient && it != orient->end()) ? it->second : std::min(c, m);
                                    if (pick != c) continue;
                                }
                            }
                            KL k;
                            if (!make_kl(L1, L2, L3, L4, tri, K13_dev, k)) continue;
                            // evaluated for the mirror class too: its diagonal blocks are the transposes
                            k.ka.mirror_diag = (symmode && !tri) ? 1 : 0;
                            kls.push_back(k);
                        }
            // heaviest classes first (estimated cost)
            std::stable_sort(kls.begin(), kls.end(), [](const KL& a, const KL& b) { return a.cost > b.cost; });
            fo.begin();
            for (auto& k : kls) {
                const int idx = kcount_index(k.L1, k.L2, k.L3, k.L4);
                hipStream_t ss = fo.pick();
                lt.begin(idx, ss);
                launch_k_ket(k.L1, k.L2, k.L3, k.L4, k.route, k.ka, k.kg, k.units, k.nunits, ss);
                lt.end();
            }
            fo.end();
            HIPCHK(hipGetLastError());
        }
        if (timed) {
            HIPCHK(hipEventRecord(S.ev[2], st));
            HIPCHK(hipEventRecord(S.ev[3], st));
            finish_call(S.counting, S.d_sc);
        }
}

// ------------------------------------------------------------------ autotuner
//
// Launch parameters that the kernels take at run time, timed on a list on both sides with a given
// density and route: per J group pair the ket chunk of a task (scale of the default chunk: 1/4 ..
// 4), per K class the work-unit split target (how finely few-unit classes are cut: 3072 .. 49152
// units).  Class launches are serialised on one stream and timed with HIP events (lsjk's per-launch
// profiling), each candidate twice (the faster counts); the fastest candidate is kept per group
// pair / class.  Group size, waves per workgroup and register budgets are compile-time template
// parameters of the kernels and are not searched here.
namespace {
constexpr float kJScales[] = {0.25f, 0.5f, 1.0f, 2.0f, 4.0f};
constexpr long long kKSplits[] = {3072, 12288, 49152};

void rebuild_jtasks(lsjk_pairlist* pl)
{
    auto it = pl->jtasks.find(pl->uid);
    if (it != pl->jtasks.end()) {
        dev_free(it->second.d_tasks);
        pl->jtasks.erase(it);
    }
    dev_free(pl->jtasks_sym.d_tasks);
    pl->jtasks_sym = lsjk::JTaskList();
    get_jtasks(pl, pl);
    build_jtasks_sym(pl);
}

void drop_own_kunits(lsjk_pairlist* pl)
{
    auto self = [&](const std::pair<unsigned long long, long long>& k) { return k.first == pl->uid; };
    erase_if_key(pl->kunits_g, self, true);
    erase_if_key(pl->kcost_g, self, false);
    pl->kunits_gen++;
}

void autotune(lsjk_pairlist* pl, const double* D_dev, int nterm, const double* scales, const double* omegas,
              float thresp, float thredp)
{
    using namespace lsjk;
    if (!pl->symmetric || pl->b1 != pl->b2) throw LsjkError(LSJK_E_INVALID, "lsjk_autotune: a symmetric list over one basis");
    const size_t n = mat_elems(pl->b1, pl->b1);
    double *dJ = nullptr, *dK = nullptr;
    struct Restore {
        int nside = S.nside;
        bool profiling = S.profiling, counting = S.counting;
        double** a;
        double** b;
        ~Restore()
        {
            S.nside = nside;
            S.profiling = profiling;
            S.counting = counting;
            dev_free(*a);
            dev_free(*b);
        }
    } restore{S.nside, S.profiling, S.counting, &dJ, &dK};
    HIPCHK(hipMalloc((void**)&dJ, n * sizeof(double)));
    HIPCHK(hipMalloc((void**)&dK, n * sizeof(double)));
    S.counting = false;
    struct ForcePlain {
        lsjk_pairlist* pl;
        ~ForcePlain() { pl->force_plain = false; }
    } force_plain_reset{pl};
    const bool mixed = thredp > thresp;
    if (mixed) pl->plain_on_mixed.clear();
    // the stored mirror-class orientations of this route (the other route's stay)
    auto drop_orient = [&]() {
        auto route = [&](const std::pair<unsigned long long, unsigned long long>& k) {
            return k.first == pl->uid && (k.second & 3ull) == (mixed ? 1ull : 0ull);
        };
        erase_if_key(pl->korient, route, false);
    };
    // the mixed route: its mirror-class orientations timed (both, twice each) and stored; then warm
    // ket groups, task and unit lists of the defaults (the FP64 route times its orientations
    // together with the ket group rows, below)
    drop_orient();
    if (mixed) {
        CallOpts ot;
        ot.tune_orient = true;
        exchange_call(pl, pl, D_dev, 1, nterm, scales, omegas, thresp, thredp, dK, ot);
    }
    coulomb_call(pl, pl, D_dev, nterm, scales, omegas, thresp, thredp, dJ);
    exchange_call(pl, pl, D_dev, 1, nterm, scales, omegas, thresp, thredp, dK);
    S.nside = 1;
    S.profiling = true;
    auto launches = [&](bool J) {
        std::map<int, double> ms;
        for (int rep = 0; rep < 2; rep++) {
            if (J)
                coulomb_call(pl, pl, D_dev, nterm, scales, omegas, thresp, thredp, dJ);
            else
                exchange_call(pl, pl, D_dev, 1, nterm, scales, omegas, thresp, thredp, dK);
            for (size_t i = 0; i < S.kid.size(); i++) {
                auto it = ms.find(S.kid[i]);
                ms[S.kid[i]] = it == ms.end() ? S.kms[i] : std::min(it->second, S.kms[i]);
            }
        }
        return ms;
    };
    if (mixed) {
        // The mixed route: per launch, the mixed kernel against the FP64 kernel (the same quartets,
        // all in FP64), on the launch shapes the FP64 route was tuned to; the FP64 one is kept
        // where it is faster (kind 3 in the table: value 1 = FP64 kernel)
        pl->tuned.erase(std::remove_if(pl->tuned.begin(), pl->tuned.end(),
                                       [](const lsjk_pairlist::Tuned& t) { return t.kind == 3; }),
                        pl->tuned.end());
        std::map<int, double> mj = launches(true), mk = launches(false);
        pl->force_plain = true;
        std::map<int, double> pj = launches(true), pk = launches(false);
        pl->force_plain = false;
        const std::pair<std::map<int, double>*, std::map<int, double>*> sides[2] = {{&mj, &pj}, {&mk, &pk}};
        for (const auto& pr : sides) {
            for (auto& kv : *pr.first) {
                auto it = pr.second->find(kv.first);
                if (it == pr.second->end()) continue;
                const bool plain = it->second < kv.second;
                if (plain) pl->plain_on_mixed.insert(kv.first);
                pl->tuned.push_back({3, kv.first, plain ? 1.0 : 0.0, kv.second, plain ? it->second : kv.second});
            }
        }
        return;
    }
    pl->tuned.clear();
    // J: one candidate scale for every group pair at a time
    std::map<int, std::map<float, double>> jt;  // launch id -> scale -> ms
    for (float c : kJScales) {
        for (int q = 0; q < 2; q++)
            for (int a = 0; a < 5; a++)
                for (int b = 0; b < 5; b++) pl->jscale[q][a][b] = c;
        rebuild_jtasks(pl);
        for (auto& kv : launches(true)) jt[kv.first][c] = kv.second;
    }
    for (int q = 0; q < 2; q++)
        for (int a = 0; a < 5; a++)
            for (int b = 0; b < 5; b++) pl->jscale[q][a][b] = 1.0f;
    for (auto& kv : jt) {
        const int id = kv.first;
        const bool sym = (id >> 24) & 1;
        const int idx = id & 0xfff, L12 = idx / 5, L34 = idx % 5;
        if (!kv.second.count(1.0f)) continue;
        float best = 1.0f;
        for (auto& c : kv.second)
            if (c.second < kv.second[best]) best = c.first;
        pl->jscale[sym ? 1 : 0][L12][L34] = best;
        pl->tuned.push_back({sym ? 1 : 0, idx, (double)best, kv.second[1.0f], kv.second[best]});
    }
    rebuild_jtasks(pl);
    // K: rows per ket group and the mirror-class orientations, together.  Rows per group: 16, 32
    // or 64 per (row L, partner L) of the ket list, where every class over that ket pair fits the
    // LDS limit on both routes.  For each size, every class is timed: those of a mirror pair in
    // both orientations (twice each, the faster kept), the others (their own mirror) in the
    // regular launches.  A ket pair keeps the size with the least summed time of its classes (both
    // orientations counted), and each mirror pair then keeps its faster orientation at the kept
    // sizes (an orientation's ket pair can differ from its mirror's, so the sizes come first).
    {
        const int maxL = pl->b1->maxL;
        auto fits = [&](int c, int L3, int L4) {
            for (int L1 = 0; L1 <= maxL; L1++)
                for (int L2 = 0; L2 <= maxL; L2++)
                    for (int route : {kRoutePlain, kRouteMixed})
                        if (kket_lds_bytes(S.rys, L1, L2, L3, L4, route, c) > kMaxLdsPerGroup) return false;
            return true;
        };
        auto ket_of = [](int k) { return ((k - 256) / 5) % 5 * 8 + (k - 256) % 5; };  // L3 * 8 + L4
        std::map<int, std::map<int, double>> ct;  // class counter index -> rows per group -> ms
        for (int c : {16, 32, 64}) {
            pl->kgsize.clear();
            for (int L3 = 0; L3 <= maxL; L3++)
                for (int L4 = 0; L4 <= maxL; L4++)
                    if (fits(c, L3, L4)) pl->kgsize[L3 * 8 + L4] = c;
            drop_own_kunits(pl);
            drop_orient();
            std::map<int, double> om;
            CallOpts ot;
            ot.tune_orient = true;
            ot.orient_ms = &om;
            exchange_call(pl, pl, D_dev, 1, nterm, scales, omegas, thresp, thredp, dK, ot);
            for (auto& kv : launches(false))
                if (!om.count(kv.first)) om[kv.first] = kv.second;
            for (auto& kv : om)
                if (pl->kgsize.count(ket_of(kv.first))) ct[kv.first][c] = kv.second;
        }
        // per ket pair: the summed time of its classes at each size that every one of them has
        std::map<int, std::map<int, double>> gt;
        for (auto& kv : ct)
            for (auto& t : kv.second) gt[ket_of(kv.first)][t.first] += t.second;
        for (auto& kv : ct)
            for (auto it = gt[ket_of(kv.first)].begin(); it != gt[ket_of(kv.first)].end();)
                it = kv.second.count(it->first) ? std::next(it) : gt[ket_of(kv.first)].erase(it);
        pl->kgsize.clear();
        for (auto& kv : gt) {
            if (kv.second.empty()) continue;
            const int L3 = kv.first / 8, L4 = kv.first % 8, def = kgroup_size(L3, L4);
            int best = kv.second.count(def) ? def : kv.second.begin()->first;
            for (auto& c : kv.second)
                if (c.second < kv.second[best]) best = c.first;
            if (best != def) pl->kgsize[kv.first] = best;
            if (kv.second.count(def))
                pl->tuned.push_back({4, L3 * 5 + L4, (double)best, kv.second[def], kv.second[best]});
        }
        // each mirror pair: the faster orientation at the kept sizes
        std::map<int, int> pick;
        auto at = [&](int k) {
            const auto it = ct.find(k);
            if (it == ct.end()) return 1e300;
            const int key = ket_of(k);
            const auto jt = it->second.find(kgroup_for(pl, key / 8, key % 8));
            return jt != it->second.end() ? jt->second : 1e300;
        };
        for (auto& kv : ct) {
            const int c = kv.first, q = c - 256;
            const int m = 256 + (((q / 5) % 5 * 5 + q % 5) * 5 + q / 125) * 5 + (q / 25) % 5;
            if (c >= m) continue;
            pick[c] = pick[m] = at(c) <= at(m) ? c : m;
        }
        drop_own_kunits(pl);
        drop_orient();
        CallOpts os;
        os.orient_set = &pick;
        exchange_call(pl, pl, D_dev, 1, nterm, scales, omegas, thresp, thredp, dK, os);
    }
    // K: one candidate split target for every class at a time
    std::map<int, std::map<long long, double>> kt;  // class counter index -> target -> ms
    for (long long c : kKSplits) {
        for (int i = 256; i < 256 + 625; i++) pl->ksplit[i] = c;
        drop_own_kunits(pl);
        for (auto& kv : launches(false)) kt[kv.first][c] = kv.second;
    }
    pl->ksplit.clear();
    for (auto& kv : kt) {
        if (!kv.second.count(kKUnitsTargetDefault)) continue;
        long long best = kKUnitsTargetDefault;
        for (auto& c : kv.second)
            if (c.second < kv.second[best]) best = c.first;
        if (best != kKUnitsTargetDefault) pl->ksplit[kv.first] = best;
        pl->tuned.push_back({2, kv.first, (double)best, kv.second[kKUnitsTargetDefault], kv.second[best]});
    }
    drop_own_kunits(pl);
}
}  // namespace

namespace {
// A device pair list (lsjk_pairlist_create; also the combined list of a batch, nmol > 0 with the
// molecule of every basis-1 primitive in mol).
lsjk_pairlist* pairlist_new(const lsjk_basis* b1, const lsjk_basis* b2, int symmetric, int npair, const int* p1,
                            const int* p2, const float* bound, int nmol, const int* mol)
{
    if (symmetric && (b1->nprim != b2->nprim || b1->ncart != b2->ncart))
        throw LsjkError(LSJK_E_INVALID, "lsjk_pairlist_create: symmetric list over bases of different size");
    for (int k = 0; k < npair; k++) {
        if (p1[k] < 0 || p1[k] >= b1->nprim || p2[k] < 0 || p2[k] >= b2->nprim)
            throw LsjkError(LSJK_E_INVALID, "lsjk_pairlist_create: primitive index out of range");
        if (symmetric && p2[k] > p1[k])
            throw LsjkError(LSJK_E_INVALID, "lsjk_pairlist_create: symmetric list must have p2 <= p1");
        if (nmol > 0 && mol[p1[k]] != mol[p2[k]])
            throw LsjkError(LSJK_E_INVALID, "lsjk_batch_create: a pair joins two molecules");
    }
    lsjk_pairlist* pl = new lsjk_pairlist();
    static unsigned long long next_uid = 0;
    pl->uid = ++next_uid;
    for (int q = 0; q < 2; q++)
        for (int a = 0; a < 5; a++)
            for (int b = 0; b < 5; b++) pl->jscale[q][a][b] = 1.0f;
    try {
        pl->b1 = b1;
        pl->b2 = b2;
        pl->symmetric = symmetric ? 1 : 0;
        pl->npair = npair;
        pl->h_p1.assign(p1, p1 + npair);
        pl->h_p2.assign(p2, p2 + npair);
        pl->h_bound.assign(bound, bound + npair);
        if (nmol > 0) {
            pl->nmol = nmol;
            pl->h_mol.assign(mol, mol + b1->nprim);
        }
        build_pairlist(pl);
        // the J task lists of the list on both sides (the common use), so that a call with it
        // on both sides allocates nothing
        get_jtasks(pl, pl);
        build_jtasks_sym(pl);
    } catch (...) {
        free_pairlist(pl);
        delete pl;
        throw;
    }
    g_lists.insert(pl);
    return pl;
}

void pairlist_delete(lsjk_pairlist* pl)
{
    if (!pl) return;
    g_lists.erase(pl);
    for (lsjk_pairlist* o : g_lists) drop_caches_for(o, pl->uid);
    free_pairlist(pl);
    delete pl;
}

// ------------------------------------------------------------------ batches of molecules
//
// A batch holds many independent molecules as one basis (the molecules' primitives one after
// another, Cartesian numbering continued) and one symmetric pair list whose pairs never join two
// molecules.  J tasks and K units only pair a molecule with itself, so every class is one launch
// over all molecules.  Matrices hold only the molecules' diagonal blocks: element (i, j) of the
// global Cartesian numbering at i * ld + j with ld = the largest molecule (blocks do not overlap).
// The first lsjk_batch_jk of a (D, J, K, thresholds, operator) combination runs eagerly (building
// the caches and timing the mirror classes) and then captures the same launches into a hipGraph;
// later calls replay the graph.
struct BatchGraph {
    const double* D = nullptr;
    double *J = nullptr, *K = nullptr;
    float thresp = 0.0f, thredp = 0.0f;
    lsjk::Ewald ew{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool same(const double* d, const double* j, const double* k, float tp, float td, const lsjk::Ewald& e) const
    {
        if (d != D || j != J || k != K || tp != thresp || td != thredp || e.n != ew.n) return false;
        for (int i = 0; i < e.n; i++)
            if (e.alpha[i] != ew.alpha[i] || e.omega[i] != ew.omega[i]) return false;
        return true;
    }
};
}  // namespace

struct lsjk_batch {
    int nmol = 0;
    lsjk_basis* basis = nullptr;
    lsjk_pairlist* list = nullptr;
    std::vector<int> cart0, ncart;  // per molecule: first global Cartesian index, count
    int ld = 0;
    size_t nelem = 0;
    void* ws = nullptr;  // fixed workspace of the captured graphs
    size_t ws_bytes = 0;
    std::vector<BatchGraph> graphs;  // most recent last
};

namespace {

constexpr size_t kBatchGraphsKept = 4;

void batch_drop_graph(BatchGraph& g)
{
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
}

void batch_delete(lsjk_batch* b)
{
    if (!b) return;
    for (auto& g : b->graphs) batch_drop_graph(g);
    b->graphs.clear();
    if (b->ws) (void)hipFree(b->ws);
    pairlist_delete(b->list);
    basis_free(b->basis);
    delete b;
}

lsjk_batch* batch_new(int nmol, const lsjk_pairlist* const* lists)
{
    using namespace lsjk;
    for (int m = 0; m < nmol; m++) {
        const lsjk_pairlist* pl = lists[m];
        if (!pl) throw LsjkError(LSJK_E_INVALID, "lsjk_batch_create: null pair list");
        if (!pl->symmetric || pl->b1 != pl->b2)
            throw LsjkError(LSJK_E_INVALID, "lsjk_batch_create: each molecule needs a symmetric list over one basis");
        if (pl->b1->ld) throw LsjkError(LSJK_E_INVALID, "lsjk_batch_create: a list of a batch basis");
    }
    std::vector<double> c, e, xyz;
    std::vector<int> L, cart, mol;
    std::vector<int> prim0(nmol + 1, 0);
    lsjk_batch* b = new lsjk_batch();
    try {
        b->nmol = nmol;
        b->cart0.resize(nmol);
        b->ncart.resize(nmol);
        int nc = 0;
        for (int m = 0; m < nmol; m++) {
            const Basis* bm = lists[m]->b1;
            b->cart0[m] = nc;
            b->ncart[m] = bm->ncart;
            prim0[m] = (int)L.size();
            c.insert(c.end(), bm->h_c.begin(), bm->h_c.end());
            e.insert(e.end(), bm->h_e.begin(), bm->h_e.end());
            xyz.insert(xyz.end(), bm->h_xyz.begin(), bm->h_xyz.end());
            L.insert(L.end(), bm->h_L.begin(), bm->h_L.end());
            for (int i = 0; i < bm->nprim; i++) {
                cart.push_back(bm->h_cart[i] + nc);
                mol.push_back(m);
            }
            nc += bm->ncart;
            b->ld = std::max(b->ld, bm->ncart);
        }
        prim0[nmol] = (int)L.size();
        if ((double)std::max(nc, 1) * (double)(b->ld + 1) > 4.0e18)
            throw LsjkError(LSJK_E_UNSUPPORTED, "lsjk_batch_create: batch too large");
        b->basis = basis_new((int)L.size(), c.data(), e.data(), xyz.data(), L.data(), cart.data(), nc, false);
        Basis* bb = b->basis;
        bb->ld = std::max(b->ld, 1);
        bb->h_mol = mol;
        // per block: its molecule's block range; ldblk = the most blocks of one molecule
        std::vector<int> mb0(nmol + 1, bb->nblk);
        for (int i = (int)mol.size() - 1; i >= 0; i--) mb0[mol[i]] = bb->h_blk[i];
        for (int m = nmol - 1; m >= 0; m--)
            if (prim0[m] == prim0[m + 1]) mb0[m] = mb0[m + 1];
        bb->h_blk_lo.resize(bb->nblk);
        bb->h_blk_hi.resize(bb->nblk);
        for (int i = 0; i < (int)mol.size(); i++) {
            const int k = bb->h_blk[i], m = mol[i];
            bb->h_blk_lo[k] = mb0[m];
            bb->h_blk_hi[k] = mb0[m + 1];
        }
        for (int m = 0; m < nmol; m++) bb->ldblk = std::max(bb->ldblk, mb0[m + 1] - mb0[m]);
        bb->ldblk = std::max(bb->ldblk, 1);
        bb->d_blk_lo = dev_upload(bb->h_blk_lo);
        bb->d_blk_hi = dev_upload(bb->h_blk_hi);
        b->nelem = mat_elems(bb, bb);
        // pairs: class by class ((L1, L2) in L1-major order), molecule by molecule, each molecule's
        // pairs of the class in its own (bound-sorted) order
        std::vector<int> p1, p2;
        std::vector<float> bd;
        for (int L1 = 0; L1 <= kMaxL; L1++)
            for (int L2 = 0; L2 <= kMaxL; L2++)
                for (int m = 0; m < nmol; m++) {
                    const lsjk_pairlist* pl = lists[m];
                    const Basis* bm = pl->b1;
                    for (int k = 0; k < pl->npair; k++) {
                        if (bm->h_L[pl->h_p1[k]] != L1 || bm->h_L[pl->h_p2[k]] != L2) continue;
                        p1.push_back(pl->h_p1[k] + prim0[m]);
                        p2.push_back(pl->h_p2[k] + prim0[m]);
                        bd.push_back(pl->h_bound[k]);
                    }
                }
        b->list = pairlist_new(b->basis, b->basis, 1, (int)p1.size(), p1.data(), p2.data(), bd.data(), nmol, mol.data());
    } catch (...) {
        batch_delete(b);
        throw;
    }
    return b;
}

// Bytes of the per-call buffers of one J + K build of a batch (as coulomb_call / exchange_call
// carve them from a workspace): kb, kd, jp (and kd32 on the mixed route); Dsh.
size_t batch_ws_bytes(const lsjk_batch* b, bool mixed, size_t* jbytes)
{
    auto r = [](size_t n) { return (std::max<size_t>(1, n) + 255) & ~size_t(255); };
    const lsjk_pairlist* pl = b->list;
    const size_t h = pl->hoff[2 * lsjk::kMaxL + 1];
    size_t j = r((size_t)pl->npair * sizeof(float)) + r(h * sizeof(double)) + r(h * sizeof(double));
    if (mixed) j += r(h * sizeof(float));
    *jbytes = j;
    return j + r(blk_elems(b->basis, b->basis) * sizeof(double));
}

void batch_jk(lsjk_batch* b, const double* D_dev, int nterm, const double* scales, const double* omegas, float thresp,
              float thredp, double* J_dev, double* K_dev)
{
    using namespace lsjk;
    require_init();
    if (!b || !D_dev || (!J_dev && !K_dev)) throw LsjkError(LSJK_E_INVALID, "lsjk_batch_jk: null argument");
    Ewald ew;
    check_ewald(nterm, scales, omegas, ew);
    const lsjk_pairlist* pl = b->list;
    for (size_t i = 0; i < b->graphs.size(); i++) {
        if (!b->graphs[i].same(D_dev, J_dev, K_dev, thresp, thredp, ew)) continue;
        // replay
        BatchGraph g = b->graphs[i];
        b->graphs.erase(b->graphs.begin() + i);
        b->graphs.push_back(g);
        HIPCHK(hipEventRecord(S.ev[0], S.stream));
        HIPCHK(hipGraphLaunch(g.exec, S.stream));
        HIPCHK(hipEventRecord(S.ev[3], S.stream));
        HIPCHK(hipStreamSynchronize(S.stream));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, S.ev[0], S.ev[3]));
        S.ms[0] = S.ms[1] = ms;
        S.kid.clear();
        S.kms.clear();
        std::memset(S.counts, 0, sizeof(S.counts));
        std::memset(S.counts32, 0, sizeof(S.counts32));
        std::memset(S.wsteps, 0, sizeof(S.wsteps));
        std::memset(S.wlanes, 0, sizeof(S.wlanes));
        return;
    }
    // eager run: the results of this call, and every cache the launches need (task and unit lists,
    // ket groups, mirror-class choices, occupancy)
    double ms = 0.0;
    if (J_dev) {
        coulomb_call(pl, pl, D_dev, nterm, scales, omegas, thresp, thredp, J_dev);
        ms += S.ms[0];
    }
    if (K_dev) {
        exchange_call(pl, pl, D_dev, 1, nterm, scales, omegas, thresp, thredp, K_dev);
        ms += S.ms[0];
    }
    // then capture the same launches into a graph, for the next calls with these arguments
    size_t jbytes = 0;
    const size_t need = batch_ws_bytes(b, thredp > thresp, &jbytes);
    if (need > b->ws_bytes) {
        for (auto& g : b->graphs) batch_drop_graph(g);  // they use the old workspace
        b->graphs.clear();
        if (b->ws) HIPCHK(hipFree(b->ws));
        b->ws = nullptr;
        b->ws_bytes = 0;
        HIPCHK(hipMalloc(&b->ws, need));
        b->ws_bytes = need;
    }
    if (b->graphs.size() >= kBatchGraphsKept) {
        batch_drop_graph(b->graphs.front());
        b->graphs.erase(b->graphs.begin());
    }
    BatchGraph g;
    g.D = D_dev;
    g.J = J_dev;
    g.K = K_dev;
    g.thresp = thresp;
    g.thredp = thredp;
    g.ew = ew;
    HIPCHK(hipStreamBeginCapture(S.stream, hipStreamCaptureModeRelaxed));
    try {
        CallOpts o;
        o.graph = true;
        o.who = "lsjk_batch_jk";
        if (J_dev) {
            o.ws = b->ws;
            o.ws_bytes = jbytes;
            coulomb_call(pl, pl, D_dev, nterm, scales, omegas, thresp, thredp, J_dev, o);
        }
        if (K_dev) {
            o.ws = (char*)b->ws + jbytes;
            o.ws_bytes = b->ws_bytes - jbytes;
            exchange_call(pl, pl, D_dev, 1, nterm, scales, omegas, thresp, thredp, K_dev, o);
        }
    } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(S.stream, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
    }
    HIPCHK(hipStreamEndCapture(S.stream, &g.graph));
    if (hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0) != hipSuccess) {
        batch_drop_graph(g);
        throw LsjkError(LSJK_E_HIP, "lsjk_batch_jk: graph instantiation failed");
    }
    b->graphs.push_back(g);
    S.ms[0] = S.ms[1] = ms;
}

}  // namespace

// ====================================================================== C-ABI

extern "C" {

const char* lsjk_last_error(void) { return g_err.c_str(); }

int lsjk_device_count(int* count)
{
    return guarded([&] {
        if (!count) throw LsjkError(LSJK_E_INVALID, "lsjk_device_count: null pointer");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *count = n;
    });
}

int lsjk_init(int device)
{
    return guarded([&] {
        if (S.init) {
            if (S.device != device) throw LsjkError(LSJK_E_INVALID, "lsjk_init: session already open on another device");
            return;
        }
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw LsjkError(LSJK_E_INVALID, "lsjk_init: no such device");
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
        for (int i = 0; i < kSide; i++) {
            HIPCHK(hipStreamCreateWithFlags(&S.side[i], hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&S.join[i], hipEventDisableTiming));
        }
        HIPCHK(hipEventCreateWithFlags(&S.fork, hipEventDisableTiming));
        if (const char* env = std::getenv("LSJK_STREAMS")) {
            int v = std::atoi(env);
            S.nside = std::min(kSide, std::max(1, v));
        }
        HIPCHK(hipDeviceGetAttribute(&S.num_cu, hipDeviceAttributeMultiprocessorCount, device));
        int lds = 0;
        HIPCHK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
        S.max_lds = (size_t)lds;
        std::vector<double> coef(lsjk_rys_coef, lsjk_rys_coef + sizeof(lsjk_rys_coef) / sizeof(double));
        std::vector<double> r2(lsjk_rys_asym_r2, lsjk_rys_asym_r2 + sizeof(lsjk_rys_asym_r2) / sizeof(double));
        std::vector<double> w(lsjk_rys_asym_w, lsjk_rys_asym_w + sizeof(lsjk_rys_asym_w) / sizeof(double));
        S.d_rys_coef = dev_upload(coef);
        S.d_rys_r2 = dev_upload(r2);
        S.d_rys_w = dev_upload(w);
        S.rys.coef = S.d_rys_coef;
        S.rys.asym_r2 = S.d_rys_r2;
        S.rys.asym_w = S.d_rys_w;
        for (int n = 0; n < 10; n++) {
            S.rys.offset[n] = lsjk_rys_offset[n];
            S.rys.nbin[n] = lsjk_rys_nbin[n];
            S.rys.h_inv[n] = (n > 0) ? 1.0 / lsjk_rys_h[n] : 0.0;
            S.rys.Tc[n] = lsjk_rys_Tc[n];
        }
        {
            std::vector<float> c32;
            build_rys_f32(c32, S.rys.offset32);
            S.d_rys_coef32 = dev_upload(c32);
            S.rys.coef32 = S.d_rys_coef32;
        }
        HIPCHK(hipMalloc((void**)&S.d_sc, sizeof(lsjk::CallScalars)));
        for (auto& e : S.ev) HIPCHK(hipEventCreate(&e));
        // device address of the no-alloc scratch array on this device (the device is current)
        S.scratch = nullptr;
        HIPCHK(hipGetSymbolAddress(&S.scratch, HIP_SYMBOL(scratch)));
        if (!S.scratch) throw LsjkError(LSJK_E_HIP, "lsjk_init: no device address for the scratch array");

        S.device = device;
        S.init = true;
    });
}

int lsjk_finalize(void)
{
    return guarded([&] {
        if (!S.init) return;
        (void)hipStreamSynchronize(S.stream);
        dev_free(S.d_rys_coef);
        dev_free(S.d_rys_r2);
        dev_free(S.d_rys_w);
        dev_free(S.d_rys_coef32);
        dev_free(S.d_sc);
        if (S.ws) (void)hipFree(S.ws);
        S.ws = nullptr;
        S.ws_bytes = 0;
        for (auto& e : S.ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : S.kev) (void)hipEventDestroy(e);
        for (int i = 0; i < kSide; i++) {
            (void)hipEventDestroy(S.join[i]);
            (void)hipStreamDestroy(S.side[i]);
        }
        (void)hipEventDestroy(S.fork);
        (void)hipStreamDestroy(S.stream);
        S = Session();
    });
}

int lsjk_basis_create(int nprim, const double* c, const double* e, const double* xyz, const int* L, const int* cart,
                      int ncart, lsjk_basis** out)
{
    return guarded([&] {
        require_init();
        if (!out || nprim < 0 || (nprim > 0 && (!c || !e || !xyz || !L || !cart)))
            throw LsjkError(LSJK_E_INVALID, "lsjk_basis_create: null argument");
        *out = basis_new(nprim, c, e, xyz, L, cart, ncart);
    });
}

int lsjk_basis_set_atoms(lsjk_basis* b, int natom, const int* atom)
{
    return guarded([&] {
        require_init();
        if (!b || natom < 0 || (b->nprim > 0 && !atom)) throw LsjkError(LSJK_E_INVALID, "lsjk_basis_set_atoms: null argument");
        std::vector<int> a(atom, atom + b->nprim);
        for (int v : a)
            if (v < 0 || v >= natom) throw LsjkError(LSJK_E_INVALID, "lsjk_basis_set_atoms: atom index out of range");
        dev_free(b->d_atom);
        b->d_atom = dev_upload(a);
        b->natom = natom;
    });
}

int lsjk_basis_destroy(lsjk_basis* b)
{
    return guarded([&] { basis_free(b); });
}

int lsjk_pairlist_create(const lsjk_basis* b1, const lsjk_basis* b2, int symmetric, int npair, const int* p1,
                         const int* p2, const float* bound, lsjk_pairlist** out)
{
    return guarded([&] {
        require_init();
        if (!b1 || !b2 || !out || npair < 0 || (npair > 0 && (!p1 || !p2 || !bound)))
            throw LsjkError(LSJK_E_INVALID, "lsjk_pairlist_create: null argument");
        *out = pairlist_new(b1, b2, symmetric, npair, p1, p2, bound, 0, nullptr);
    });
}

int lsjk_pairlist_destroy(lsjk_pairlist* pl)
{
    return guarded([&] { pairlist_delete(pl); });
}

int lsjk_batch_create(int nmol, const lsjk_pairlist* const* lists, lsjk_batch** out)
{
    return guarded([&] {
        require_init();
        if (!out || nmol <= 0 || !lists) throw LsjkError(LSJK_E_INVALID, "lsjk_batch_create: null argument");
        *out = batch_new(nmol, lists);
    });
}

int lsjk_batch_destroy(lsjk_batch* b)
{
    return guarded([&] { batch_delete(b); });
}

int lsjk_batch_layout(const lsjk_batch* b, int* ld, size_t* nelem, size_t* offsets)
{
    return guarded([&] {
        if (!b || !ld || !nelem) throw LsjkError(LSJK_E_INVALID, "lsjk_batch_layout: null argument");
        *ld = b->ld;
        *nelem = b->nelem;
        if (offsets)
            for (int m = 0; m < b->nmol; m++) offsets[m] = (size_t)b->cart0[m] * (size_t)(b->ld + 1);
    });
}

int lsjk_batch_jk(lsjk_batch* b, const double* D_dev, int nterm, const double* scales, const double* omegas,
                  float thresp, float thredp, double* J_dev, double* K_dev)
{
    return guarded([&] { batch_jk(b, D_dev, nterm, scales, omegas, thresp, thredp, J_dev, K_dev); });
}

int lsjk_coulomb(const lsjk_pairlist* pl12, const lsjk_pairlist* pl34, const double* D34_dev, int nterm,
                 const double* scales, const double* omegas, float thresp, float thredp, double* J12_dev)
{
    return guarded([&] { coulomb_call(pl12, pl34, D34_dev, nterm, scales, omegas, thresp, thredp, J12_dev); });
}

int lsjk_coulomb_noalloc(const lsjk_pairlist* pl12, const lsjk_pairlist* pl34, const double* D34_dev, int nterm,
                         const double* scales, const double* omegas, float thresp, float thredp, double* J12_dev)
{
    return guarded([&] {
        require_init();
        if (!S.scratch) throw LsjkError(LSJK_E_NOINIT, "lsjk_coulomb_noalloc: scratch array address unknown");
        CallOpts o;
        o.ws = S.scratch;
        o.ws_bytes = LSJK_NOALLOC_SCRATCH_BYTES;
        o.noalloc = true;
        o.who = "lsjk_coulomb_noalloc";
        coulomb_call(pl12, pl34, D34_dev, nterm, scales, omegas, thresp, thredp, J12_dev, o);
    });
}

int lsjk_exchange(const lsjk_pairlist* pl12, const lsjk_pairlist* pl34, const double* D24_dev, int D24symm,
                  int nterm, const double* scales, const double* omegas, float thresp, float thredp, double* K13_dev)
{
    return guarded([&] {
        exchange_call(pl12, pl34, D24_dev, D24symm, nterm, scales, omegas, thresp, thredp, K13_dev);
    });
}

int lsjk_unit_features(const lsjk_pairlist* pl, int L1, int L2, int L3, int L4, float thresp, int* nunits, float* feat)
{
    return guarded([&] {
        require_init();
        if (!pl || !nunits || L1 < 0 || L2 < 0 || L3 < 0 || L4 < 0 || L1 > 2 || L2 > 2 || L3 > 2 || L4 > 2)
            throw LsjkError(LSJK_E_INVALID, "lsjk_unit_features: bad argument");
        *nunits = 0;
        if (pl->h_rows[L1][L2].empty() || pl->h_rows[L3][L4].empty()) return;
        const auto units = get_kunits_g(pl, pl, L1, L2, L3, L4, false, thresp, nullptr);
        *nunits = units.second;
        if (!feat) return;
        std::vector<float> f;
        unit_features(pl, pl, L1, L2, L3, L4, download_units(units.first, units.second), f);
        std::memcpy(feat, f.data(), f.size() * sizeof(float));
    });
}

int lsjk_unit_magnitudes(const lsjk_pairlist* pl, int L1, int L2, int L3, int L4, float thresp, float smin, float* mag)
{
    return guarded([&] {
        using namespace lsjk;
        require_init();
        if (!pl || !mag || L1 < 0 || L2 < 0 || L3 < 0 || L4 < 0 || L1 > 1 || L2 > 1 || L3 > 1 || L4 > 1)
            throw LsjkError(LSJK_E_INVALID, "lsjk_unit_magnitudes: classes with s and p functions only");
        if (pl->h_rows[L1][L2].empty() || pl->h_rows[L3][L4].empty()) return;
        const auto units = get_kunits_g(pl, pl, L1, L2, L3, L4, false, thresp, nullptr);
        if (units.second == 0) return;
        const KGroups kg = kgroups_args(get_kgroups(const_cast<lsjk_pairlist*>(pl), L3, L4, kgroup_for(pl, L3, L4)));
        KArgs ka{};
        ka.rys = S.rys;
        ka.ew.n = 1;
        ka.ew.plain = 1;
        ka.ew.alpha[0] = 1.0;
        ka.ew.omega[0] = -1.0;
        KSide* ks[2] = {&ka.s12, &ka.s34};
        const int Ls[2] = {L1, L3}, Lps[2] = {L2, L4};
        for (int q = 0; q < 2; q++) {
            KSide& sd = *ks[q];
            sd.row_start = pl->d_row_start;
            sd.maxLp = pl->kmaxL2;
            sd.eb = pl->d_ebound;
            sd.ecart = pl->d_ecart;
            sd.eblk = pl->d_eblk;
            sd.eg = pl->d_egeom;
            sd.eg32 = pl->d_egeom32;
            sd.nent = pl->nent;
            sd.rows = pl->d_rows[Ls[q]][Lps[q]];
            sd.nrows = (int)pl->h_rows[Ls[q]][Lps[q]].size();
            sd.x = pl->b1->d_x;
            sd.y = pl->b1->d_y;
            sd.z = pl->b1->d_z;
            sd.cart = pl->b1->d_cart;
        }
        float* d = nullptr;
        HIPCHK(hipMalloc((void**)&d, (size_t)units.second * sizeof(float)));
        try {
            launch_k_unit_mag_sp(L1, L2, L3, L4, ka, kg, units.first, units.second, smin, d, S.stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(S.stream));
            HIPCHK(hipMemcpy(mag, d, (size_t)units.second * sizeof(float), hipMemcpyDeviceToHost));
        } catch (...) {
            dev_free(d);
            throw;
        }
        dev_free(d);
    });
}

int lsjk_screen_create(int ni, const float* params, const float* mean, const float* scale, int nbin, const float* fedge,
                       const double* delta, float smin, lsjk_screen** out)
{
    return guarded([&] {
        require_init();
        if (!out || !params || !mean || !scale || !delta || ni < 1 || ni > lsjk::kMlpMaxIn || ni > LSJK_SCREEN_NFEAT ||
            nbin < 1 || nbin > 64 || (nbin > 1 && !fedge))
            throw LsjkError(LSJK_E_INVALID, "lsjk_screen_create: bad argument");
        for (int k = 0; k < ni; k++)
            if (!(scale[k] > 0.0f)) throw LsjkError(LSJK_E_INVALID, "lsjk_screen_create: input scales must be positive");
        lsjk_screen* sc = new lsjk_screen();
        try {
            sc->ni = ni;
            std::vector<float> p(params, params + lsjk::mlp_nparam(ni));
            sc->d_params = dev_upload(p);
            sc->mean.assign(mean, mean + ni);
            sc->scale.assign(scale, scale + ni);
            sc->nbin = nbin;
            sc->delta.assign(delta, delta + 625 * (size_t)nbin);
            if (nbin > 1) sc->fedge.assign(fedge, fedge + 625 * (size_t)(nbin - 1));
            sc->smin = smin;
        } catch (...) {
            dev_free(sc->d_params);
            delete sc;
            throw;
        }
        *out = sc;
    });
}

int lsjk_screen_destroy(lsjk_screen* sc)
{
    return guarded([&] {
        if (!sc) return;
        for (auto& kv : sc->tau) dev_free(kv.second.first);
        dev_free(sc->d_params);
        delete sc;
    });
}

int lsjk_exchange_screened(const lsjk_pairlist* pl12, const lsjk_pairlist* pl34, const double* D24_dev, int D24symm,
                           int nterm, const double* scales, const double* omegas, float thresp, float thredp,
                           lsjk_screen* screen, double* K13_dev)
{
    return guarded([&] {
        if (!screen) throw LsjkError(LSJK_E_INVALID, "lsjk_exchange_screened: null screen");
        CallOpts o;
        o.screen = screen;
        o.who = "lsjk_exchange_screened";
        exchange_call(pl12, pl34, D24_dev, D24symm, nterm, scales, omegas, thresp, thredp, K13_dev, o);
    });
}

int lsjk_boys(int nboys, int npoint, const double* T_dev, double* F_dev)
{
    return guarded([&] {
        require_init();
        if (nboys < 0 || nboys > lsjk::kBoysMax || npoint < 0 || (npoint > 0 && (!T_dev || !F_dev)))
            throw LsjkError(LSJK_E_INVALID, "lsjk_boys: bad argument (nboys must be 0..16)");
        lsjk::launch_boys(nboys, npoint, T_dev, F_dev, S.stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(S.stream));
    });
}

int lsjk_autotune(lsjk_pairlist* pl, const double* D_dev, int nterm, const double* scales, const double* omegas,
                  float thresp, float thredp)
{
    return guarded([&] {
        require_init();
        if (!pl || !D_dev) throw LsjkError(LSJK_E_INVALID, "lsjk_autotune: null argument");
        autotune(pl, D_dev, nterm, scales, omegas, thresp, thredp);
    });
}

int lsjk_set_kgroup_rows(lsjk_pairlist* pl, int L, int Lp, int rows)
{
    return guarded([&] {
        using namespace lsjk;
        require_init();
        if (!pl || L < 0 || L > kMaxL || Lp < 0 || Lp > kMaxL)
            throw LsjkError(LSJK_E_INVALID, "lsjk_set_kgroup_rows: bad list or angular momentum");
        if (rows == 0) {
            pl->kgsize.erase(L * 8 + Lp);
        } else {
            if (rows != 16 && rows != 32 && rows != 64)
                throw LsjkError(LSJK_E_INVALID, "lsjk_set_kgroup_rows: rows must be 16, 32 or 64");
            for (int L1 = 0; L1 <= kMaxL; L1++)
                for (int L2 = 0; L2 <= kMaxL; L2++)
                    for (int route : {kRoutePlain, kRouteMixed})
                        if (kket_lds_bytes(S.rys, L1, L2, L, Lp, route, rows) > kMaxLdsPerGroup)
                            throw LsjkError(LSJK_E_INVALID, "lsjk_set_kgroup_rows: the classes over this ket pair "
                                                            "would exceed the LDS limit");
            pl->kgsize[L * 8 + Lp] = rows;
        }
        drop_own_kunits(pl);
    });
}

int lsjk_tuned_params(const lsjk_pairlist* pl, int capacity, int* n, int* kind, int* idx, double* value,
                      double* ms_default, double* ms_tuned)
{
    return guarded([&] {
        if (!pl || !n) throw LsjkError(LSJK_E_INVALID, "lsjk_tuned_params: null argument");
        *n = (int)pl->tuned.size();
        for (int i = 0; i < std::min(capacity, *n); i++) {
            const auto& t = pl->tuned[i];
            if (kind) kind[i] = t.kind;
            if (idx) idx[i] = t.idx;
            if (value) value[i] = t.value;
            if (ms_default) ms_default[i] = t.ms_default;
            if (ms_tuned) ms_tuned[i] = t.ms_tuned;
        }
    });
}

int lsjk_batch_autotune(lsjk_batch* b, const double* D_dev, int nterm, const double* scales, const double* omegas,
                        float thresp, float thredp)
{
    return guarded([&] {
        require_init();
        if (!b || !D_dev) throw LsjkError(LSJK_E_INVALID, "lsjk_batch_autotune: null argument");
        for (auto& g : b->graphs) batch_drop_graph(g);  // they hold the old task and unit lists
        b->graphs.clear();
        autotune(b->list, D_dev, nterm, scales, omegas, thresp, thredp);
    });
}

int lsjk_batch_pairlist(const lsjk_batch* b, const lsjk_pairlist** out)
{
    return guarded([&] {
        if (!b || !out) throw LsjkError(LSJK_E_INVALID, "lsjk_batch_pairlist: null argument");
        *out = b->list;
    });
}

int lsjk_mlp_nparam(int ni, int* nparam)
{
    return guarded([&] {
        if (!nparam || ni < 1 || ni > lsjk::kMlpMaxIn) throw LsjkError(LSJK_E_INVALID, "lsjk_mlp_nparam: bad input width");
        *nparam = lsjk::mlp_nparam(ni);
    });
}

int lsjk_mlp_forward(int ni, const float* params_dev, int n, const float* X_dev, float* out_dev)
{
    return guarded([&] {
        require_init();
        if (ni < 1 || ni > lsjk::kMlpMaxIn || n < 0 || !params_dev || (n > 0 && (!X_dev || !out_dev)))
            throw LsjkError(LSJK_E_INVALID, "lsjk_mlp_forward: bad argument");
        lsjk::launch_mlp_forward(params_dev, ni, n, X_dev, out_dev, S.stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(S.stream));
    });
}

int lsjk_mlp_grad(int ni, const float* params_dev, int nb, const int* idx_dev, const float* X_dev, const float* y_dev,
                  const float* sw_dev, float wu, float* grad_dev, float* loss_dev)
{
    return guarded([&] {
        require_init();
        if (ni < 1 || ni > lsjk::kMlpMaxIn || nb < 0 || !params_dev || !grad_dev || !loss_dev ||
            (nb > 0 && (!X_dev || !y_dev)))
            throw LsjkError(LSJK_E_INVALID, "lsjk_mlp_grad: bad argument");
        lsjk::launch_mlp_grad(params_dev, ni, nb, idx_dev, X_dev, y_dev, sw_dev, wu, grad_dev, loss_dev, S.stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(S.stream));
    });
}

int lsjk_mlp_adam(int nparam, float* params_dev, const float* grad_dev, float* m_dev, float* v_dev, float lr,
                  float beta1, float beta2, float eps, int t)
{
    return guarded([&] {
        require_init();
        if (nparam <= 0 || !params_dev || !grad_dev || !m_dev || !v_dev || t < 1)
            throw LsjkError(LSJK_E_INVALID, "lsjk_mlp_adam: bad argument");
        lsjk::launch_mlp_adam(nparam, params_dev, grad_dev, m_dev, v_dev, lr, beta1, beta2, eps, t, S.stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(S.stream));
    });
}

int lsjk_coulomb_grad(const lsjk_pairlist* pl, const double* Dbra_dev, const double* Dket_dev, int nterm,
                      const double* scales, const double* omegas, float thresp, float thredp, double* G_dev)
{
    (void)thredp;  // FP64 route
    return guarded([&] {
        using namespace lsjk;
        require_init();
        if (!pl || !Dbra_dev || !Dket_dev || !G_dev) throw LsjkError(LSJK_E_INVALID, "lsjk_coulomb_grad: null argument");
        if (!pl->symmetric || pl->b1 != pl->b2)
            throw LsjkError(LSJK_E_INVALID, "lsjk_coulomb_grad: the pair list must be symmetric");
        if (!pl->b1->d_atom) throw LsjkError(LSJK_E_INVALID, "lsjk_coulomb_grad: basis atoms not set");
        Ewald ew;
        check_ewald(nterm, scales, omegas, ew);
        hipStream_t st = S.stream;
        const Basis* b = pl->b1;
        // potential buffer at bra order L + 1: group L holds npair(L) x H(L + 1) doubles
        size_t hoff2[2 * kMaxL + 2] = {0};
        for (int L = 0; L <= 2 * kMaxL; L++)
            hoff2[L + 1] = hoff2[L] + (size_t)(pl->hstart[L + 1] - pl->hstart[L]) * nherm(L + 1);
        float* kb = nullptr;
        double* kd = nullptr;
        double* vp = nullptr;
        Workspace wsp;
        wsp.add(&kb, (size_t)pl->npair);
        wsp.add(&kd, pl->hoff[2 * kMaxL + 1]);
        wsp.add(&vp, hoff2[2 * kMaxL + 1]);
        wsp.commit();
        const JTaskList& seg = *get_jtasks(pl, pl);
        launch_fill(S.d_sc, sizeof(CallScalars), st);
        launch_fill(vp, std::max<size_t>(1, hoff2[2 * kMaxL + 1]) * sizeof(double), st);
        JPrepArgs pa{};
        pa.p1 = pl->d_p1;
        pa.p2 = pl->d_p2;
        pa.bound = pl->d_bound;
        pa.hpos = pl->d_hpos;
        pa.K = pl->d_K;
        pa.eA = b->d_e; pa.xA = b->d_x; pa.yA = b->d_y; pa.zA = b->d_z; pa.cartA = b->d_cart;
        pa.eB = b->d_e; pa.xB = b->d_x; pa.yB = b->d_y; pa.zB = b->d_z; pa.cartB = b->d_cart;
        pa.D = Dket_dev;
        pa.ldD = b->ncart;
        pa.symmetric = 1;
        pa.thre_ket = thresp / pl->maxbound;
        pa.kb = kb;
        pa.kd = kd;
        pa.sc = S.d_sc;
        for (size_t c = 0; c + 1 < pl->cls_start.size(); c++) {
            int L = pl->cls_L1[c] + pl->cls_L2[c];
            launch_j_ket_prep(pl->cls_L1[c], pl->cls_L2[c], pa, pl->cls_start[c], pl->cls_start[c + 1] - pl->cls_start[c],
                              pl->hstart[L], pl->hoff[L], st);
        }
        if (ew.n > 0) {
            for (int L12 = 0; L12 <= 2 * kMaxL; L12++)
                for (int L34 = 0; L34 <= 2 * kMaxL; L34++) {
                    if (seg.cnt[L12][L34] == 0) continue;
                    JQuartetArgs qa{};
                    qa.rys = S.rys;
                    qa.ew = ew;
                    qa.thresp = thresp;
                    qa.thredp = thresp;
                    qa.bra_b = pl->d_hb;
                    qa.bra_geom = pl->d_hgeom;
                    qa.bra_n = pl->npair;
                    qa.bra_hstart = pl->hstart[L12];
                    qa.jp = vp + hoff2[L12];
                    qa.ket_b = kb;
                    qa.ket_static_b = pl->d_hb;
                    qa.ket_geom = pl->d_hgeom;
                    qa.ket_n = pl->npair;
                    qa.ket_hstart = pl->hstart[L34];
                    qa.ket_d = kd + pl->hoff[L34];
                    qa.sc = S.d_sc;
                    launch_j_quartets_grad(L12 + 1, L34, qa, (const JTask*)seg.d_tasks + seg.off[L12][L34],
                                           seg.cnt[L12][L34], &S.d_sc->queue[jcount_index(L12, L34)], st);
                }
            JGradArgs ga{};
            ga.p1 = pl->d_p1;
            ga.p2 = pl->d_p2;
            ga.bound = pl->d_bound;
            ga.hpos = pl->d_hpos;
            ga.K = pl->d_K;
            ga.eA = b->d_e; ga.xA = b->d_x; ga.yA = b->d_y; ga.zA = b->d_z; ga.cartA = b->d_cart; ga.atomA = b->d_atom;
            ga.eB = b->d_e; ga.xB = b->d_x; ga.yB = b->d_y; ga.zB = b->d_z; ga.cartB = b->d_cart; ga.atomB = b->d_atom;
            ga.vp = vp;
            ga.D = Dbra_dev;
            ga.ldD = b->ncart;
            ga.symmetric = 1;
            ga.thresp = thresp;
            ga.G = G_dev;
            ga.sc = S.d_sc;
            for (size_t c = 0; c + 1 < pl->cls_start.size(); c++) {
                int L = pl->cls_L1[c] + pl->cls_L2[c];
                launch_j_grad_scatter(pl->cls_L1[c], pl->cls_L2[c], ga, pl->cls_start[c],
                                      pl->cls_start[c + 1] - pl->cls_start[c], pl->hstart[L], hoff2[L], st);
            }
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    });
}

int lsjk_exchange_grad(const lsjk_pairlist* pl, const double* D13_dev, const double* D24_dev, int nterm,
                       const double* scales, const double* omegas, float thresp, float thredp, double* G_dev)
{
    (void)thredp;  // FP64 route
    return guarded([&] {
        using namespace lsjk;
        require_init();
        if (!pl || !D13_dev || !D24_dev || !G_dev) throw LsjkError(LSJK_E_INVALID, "lsjk_exchange_grad: null argument");
        if (!pl->symmetric || pl->b1 != pl->b2)
            throw LsjkError(LSJK_E_INVALID, "lsjk_exchange_grad: the pair list must be symmetric");
        const Basis* b = pl->b1;
        if (!b->d_atom) throw LsjkError(LSJK_E_INVALID, "lsjk_exchange_grad: basis atoms not set");
        Ewald ew;
        check_ewald(nterm, scales, omegas, ew);
        hipStream_t st = S.stream;
        double* Dsh = nullptr;
        Workspace wsp;
        wsp.add(&Dsh, (size_t)b->nblk * b->nblk);
        wsp.commit();
        launch_fill(S.d_sc, sizeof(CallScalars), st);
        launch_k_blockmax(D24_dev, b->ncart, b->d_blk_cart, b->d_blk_L, b->nblk, b->d_blk_cart, b->d_blk_L, b->nblk, Dsh,
                          b->nblk, nullptr, nullptr, 0, S.d_sc, st);
        if (ew.n > 0) {
            for (int L1 = 0; L1 <= b->maxL; L1++)
                for (int L2 = 0; L2 <= b->maxL; L2++)
                    for (int L3 = 0; L3 <= b->maxL; L3++)
                        for (int L4 = 0; L4 <= b->maxL; L4++) {
                            if (pl->h_rows[L1][L2].empty() || pl->h_rows[L3][L4].empty()) continue;
                            const auto units = get_kunits(pl, pl, L1, L2, L3, L4, false, nullptr);
                            if (units.second == 0) continue;
                            KArgs ka{};
                            ka.rys = S.rys;
                            ka.ew = ew;
                            ka.thresp = thresp;
                            ka.thredp = thresp;
                            KSide* ks[2] = {&ka.s12, &ka.s34};
                            const int Ls[2] = {L1, L3}, Lps[2] = {L2, L4};
                            for (int q = 0; q < 2; q++) {
                                KSide& s = *ks[q];
                                s.row_start = pl->d_row_start;
                                s.maxLp = pl->kmaxL2;
                                s.eb = pl->d_ebound;
                                s.ecart = pl->d_ecart;
                                s.eblk = pl->d_eblk;
                                s.eg = pl->d_egeom;
                                s.eg32 = pl->d_egeom32;
                                s.nent = pl->nent;
                                s.rows = pl->d_rows[Ls[q]][Lps[q]];
                                s.nrows = (int)pl->h_rows[Ls[q]][Lps[q]].size();
                                s.cart = b->d_cart;
                            }
                            ka.D = D24_dev;
                            ka.ldD = b->ncart;
                            ka.Dsh = Dsh;
                            ka.ldDsh = b->nblk;
                            ka.maxb12 = pl->maxbound;
                            ka.maxb34 = pl->maxbound;
                            ka.sc = S.d_sc;
                            ka.D13 = D13_dev;
                            ka.ldD13 = b->ncart;
                            ka.G = G_dev;
                            ka.exp1 = ka.exp2 = ka.exp3 = ka.exp4 = b->d_e;
                            ka.atom1 = ka.atom2 = ka.atom3 = ka.atom4 = b->d_atom;
                            ka.eprim12 = ka.eprim34 = pl->d_eprim;
                            launch_k_quartets(L1, L2, L3, L4, kRouteGrad, ka, units.first, units.second, st);
                        }
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    });
}

// One-electron integrals of one pair list (accumulated, Cartesian): S and/or T.
static void one_e_st(const lsjk_pairlist* pl, double* Sout, double* Tout)
{
    using namespace lsjk;
    require_init();
    if (!pl) throw LsjkError(LSJK_E_INVALID, "lsjk: null pair list");
    OneEArgs a{};
    const Basis *b1 = pl->b1, *b2 = pl->b2;
    a.p1 = pl->d_p1;
    a.p2 = pl->d_p2;
    a.hpos = pl->d_hpos;
    a.K = pl->d_K;
    a.eA = b1->d_e; a.xA = b1->d_x; a.yA = b1->d_y; a.zA = b1->d_z; a.cartA = b1->d_cart;
    a.eB = b2->d_e; a.xB = b2->d_x; a.yB = b2->d_y; a.zB = b2->d_z; a.cartB = b2->d_cart;
    a.symmetric = pl->symmetric;
    a.S = Sout;
    a.T = Tout;
    a.ld = b2->ncart;
    for (size_t c = 0; c + 1 < pl->cls_start.size(); c++)
        launch_one_e_st(pl->cls_L1[c], pl->cls_L2[c], a, pl->cls_start[c], pl->cls_start[c + 1] - pl->cls_start[c],
                        S.stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(S.stream));
}

int lsjk_overlap(const lsjk_pairlist* pl, double* S_dev)
{
    return guarded([&] {
        if (!S_dev) throw LsjkError(LSJK_E_INVALID, "lsjk_overlap: null output");
        one_e_st(pl, S_dev, nullptr);
    });
}

int lsjk_kinetic(const lsjk_pairlist* pl, double* T_dev)
{
    return guarded([&] {
        if (!T_dev) throw LsjkError(LSJK_E_INVALID, "lsjk_kinetic: null output");
        one_e_st(pl, nullptr, T_dev);
    });
}

int lsjk_potential(const lsjk_pairlist* pl, int nterm, const double* scales, const double* omegas, int ncharge,
                   const double* xyzc, double* V_dev)
{
    return guarded([&] {
        using namespace lsjk;
        require_init();
        if (!pl || !V_dev || (ncharge > 0 && !xyzc) || ncharge < 0)
            throw LsjkError(LSJK_E_INVALID, "lsjk_potential: null or invalid argument");
        Ewald ew;
        check_ewald(nterm, scales, omegas, ew);
        if (ew.n == 0 || ncharge == 0) return;
        hipStream_t st = S.stream;
        const Basis *b1 = pl->b1, *b2 = pl->b2;
        double* vp = nullptr;
        double* dxyzc = nullptr;
        Workspace wsp;
        wsp.add(&vp, pl->hoff[2 * kMaxL + 1]);
        wsp.add(&dxyzc, (size_t)4 * ncharge);
        wsp.commit();
        HIPCHK(hipMemcpyAsync(dxyzc, xyzc, sizeof(double) * 4 * ncharge, hipMemcpyHostToDevice, st));
        OneEArgs a{};
        a.p1 = pl->d_p1;
        a.p2 = pl->d_p2;
        a.hpos = pl->d_hpos;
        a.K = pl->d_K;
        a.eA = b1->d_e; a.xA = b1->d_x; a.yA = b1->d_y; a.zA = b1->d_z; a.cartA = b1->d_cart;
        a.eB = b2->d_e; a.xB = b2->d_x; a.yB = b2->d_y; a.zB = b2->d_z; a.cartB = b2->d_cart;
        a.symmetric = pl->symmetric;
        a.rys = S.rys;
        a.ew = ew;
        a.xyzc = dxyzc;
        a.nA = ncharge;
        a.vp = vp;
        for (size_t c = 0; c + 1 < pl->cls_start.size(); c++) {
            const int L = pl->cls_L1[c] + pl->cls_L2[c];
            launch_one_e_v(pl->cls_L1[c], pl->cls_L2[c], a, pl->cls_start[c], pl->cls_start[c + 1] - pl->cls_start[c],
                           pl->hstart[L], pl->hoff[L], st);
        }
        JScatterArgs sa{};
        sa.p1 = pl->d_p1;
        sa.p2 = pl->d_p2;
        sa.bound = pl->d_bound;
        sa.hpos = pl->d_hpos;
        sa.K = pl->d_K;
        sa.eA = b1->d_e; sa.xA = b1->d_x; sa.yA = b1->d_y; sa.zA = b1->d_z; sa.cartA = b1->d_cart;
        sa.eB = b2->d_e; sa.xB = b2->d_x; sa.yB = b2->d_y; sa.zB = b2->d_z; sa.cartB = b2->d_cart;
        sa.jp = vp;
        sa.J = V_dev;
        sa.ldJ = b2->ncart;
        sa.symmetric = pl->symmetric;
        sa.screen = 0;
        sa.sc = S.d_sc;
        for (size_t c = 0; c + 1 < pl->cls_start.size(); c++) {
            const int L = pl->cls_L1[c] + pl->cls_L2[c];
            launch_j_bra_scatter(pl->cls_L1[c], pl->cls_L2[c], sa, pl->cls_start[c],
                                 pl->cls_start[c + 1] - pl->cls_start[c], pl->hstart[L], pl->hoff[L], st);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    });
}

int lsjk_count_quartets(const lsjk_pairlist* pl12, const lsjk_pairlist* pl34, const double* D34_dev,
                        const double* D24_dev, int D24symm, float thresp, uint64_t* countsJ, uint64_t* countsK)
{
    return guarded([&] {
        using namespace lsjk;
        require_init();
        if (!pl12 || !pl34 || !countsJ || !countsK || (!D34_dev && !D24_dev))
            throw LsjkError(LSJK_E_INVALID, "lsjk_count_quartets: null argument");
        hipStream_t st = S.stream;
        const Basis *b2 = pl12->b2, *b3 = pl34->b1, *b4 = pl34->b2;
        float* kb = nullptr;
        double* kd = nullptr;
        double* Dsh = nullptr;
        unsigned long long* out = nullptr;
        Workspace wsp;
        wsp.add(&kb, (size_t)pl34->npair);
        wsp.add(&kd, pl34->hoff[2 * kMaxL + 1]);
        wsp.add(&Dsh, (size_t)b2->nblk * b4->nblk);
        wsp.add(&out, 2 * 625);
        wsp.commit();
        launch_fill(out, 2 * 625 * sizeof(unsigned long long), st);
        std::memset(countsJ, 0, 625 * sizeof(uint64_t));
        std::memset(countsK, 0, 625 * sizeof(uint64_t));
        if (D34_dev) {
            launch_fill(S.d_sc, sizeof(CallScalars), st);
            JPrepArgs pa{};
            pa.p1 = pl34->d_p1; pa.p2 = pl34->d_p2; pa.bound = pl34->d_bound; pa.hpos = pl34->d_hpos; pa.K = pl34->d_K;
            pa.eA = b3->d_e; pa.xA = b3->d_x; pa.yA = b3->d_y; pa.zA = b3->d_z; pa.cartA = b3->d_cart;
            pa.eB = b4->d_e; pa.xB = b4->d_x; pa.yB = b4->d_y; pa.zB = b4->d_z; pa.cartB = b4->d_cart;
            pa.D = D34_dev; pa.ldD = b4->ncart; pa.symmetric = pl34->symmetric;
            pa.thre_ket = thresp / pl12->maxbound;
            pa.kb = kb; pa.kd = kd; pa.sc = S.d_sc;
            for (size_t c = 0; c + 1 < pl34->cls_start.size(); c++) {
                int L = pl34->cls_L1[c] + pl34->cls_L2[c];
                launch_j_ket_prep(pl34->cls_L1[c], pl34->cls_L2[c], pa, pl34->cls_start[c],
                                  pl34->cls_start[c + 1] - pl34->cls_start[c], pl34->hstart[L], pl34->hoff[L], st);
            }
            CountJArgs ca{};
            ca.thresp = thresp;
            ca.bra_bound = pl12->d_bound;
            ca.ket_b = kb;
            ca.ket_hpos = pl34->d_hpos;
            ca.ket_cls_start = pl34->d_cls_start;
            ca.ket_cls_L1 = pl34->d_cls_L1;
            ca.ket_cls_L2 = pl34->d_cls_L2;
            ca.ket_ncls = (int)pl34->cls_L1.size();
            ca.out = out;
            ca.sc = S.d_sc;
            for (size_t c = 0; c + 1 < pl12->cls_start.size(); c++)
                launch_count_j(ca, pl12->cls_start[c], pl12->cls_start[c + 1] - pl12->cls_start[c], pl12->cls_L1[c],
                               pl12->cls_L2[c], st);
            HIPCHK(hipGetLastError());
        }
        if (D24_dev) {
            launch_fill(S.d_sc, sizeof(CallScalars), st);
            launch_k_blockmax(D24_dev, b4->ncart, b2->d_blk_cart, b2->d_blk_L, b2->nblk, b4->d_blk_cart, b4->d_blk_L,
                              b4->nblk, Dsh, b4->nblk, nullptr, nullptr, 0, S.d_sc, st);
            CountKArgs ka{};
            ka.thresp = thresp;
            ka.half = (D24symm && pl12 == pl34) ? 1 : 0;
            ka.n2 = b2->nprim;
            ka.n4 = b4->nprim;
            ka.rs12 = pl12->symmetric ? pl12->d_row_start : pl12->d_piv_row_start;
            ka.b12 = pl12->symmetric ? pl12->d_ebound : pl12->d_piv_b;
            ka.maxLa = pl12->b1->maxL;
            ka.rs34 = pl34->symmetric ? pl34->d_row_start : pl34->d_piv_row_start;
            ka.b34 = pl34->symmetric ? pl34->d_ebound : pl34->d_piv_b;
            ka.maxLc = pl34->b1->maxL;
            ka.L2 = b2->d_L;
            ka.L4 = b4->d_L;
            ka.blk2 = b2->d_blk;
            ka.blk4 = b4->d_blk;
            ka.Dsh = Dsh;
            ka.ldDsh = b4->nblk;
            ka.maxb12 = pl12->maxbound;
            ka.maxb34 = pl34->maxbound;
            ka.out = out + 625;
            ka.sc = S.d_sc;
            launch_count_k(ka, st);
            HIPCHK(hipGetLastError());
        }
        std::vector<unsigned long long> h(2 * 625);
        HIPCHK(hipMemcpyAsync(h.data(), out, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int i = 0; i < 625; i++) {
            countsJ[i] = h[i];
            countsK[i] = h[625 + i];
        }
    });
}

int lsjk_set_profiling(int enable)
{
    return guarded([&] { S.profiling = enable != 0; });
}

int lsjk_last_launches(int capacity, int* n, int* ids, double* ms)
{
    return guarded([&] {
        if (!n) throw LsjkError(LSJK_E_INVALID, "lsjk_last_launches: null pointer");
        *n = (int)S.kms.size();
        for (int i = 0; i < *n && i < capacity; i++) {
            if (ids) ids[i] = S.kid[i];
            if (ms) ms[i] = S.kms[i];
        }
    });
}

int lsjk_malloc(void** ptr, size_t bytes)
{
    return guarded([&] {
        require_init();
        if (!ptr) throw LsjkError(LSJK_E_INVALID, "lsjk_malloc: null pointer");
        HIPCHK(hipMalloc(ptr, std::max<size_t>(bytes, 1)));
    });
}
int lsjk_free(void* ptr)
{
    return guarded([&] {
        if (ptr) HIPCHK(hipFree(ptr));
    });
}
int lsjk_memcpy_htod(void* dst, const void* src, size_t bytes)
{
    return guarded([&] {
        require_init();
        if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S.stream));
        HIPCHK(hipStreamSynchronize(S.stream));
    });
}
int lsjk_memcpy_dtoh(void* dst, const void* src, size_t bytes)
{
    return guarded([&] {
        require_init();
        if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S.stream));
        HIPCHK(hipStreamSynchronize(S.stream));
    });
}
int lsjk_memset_zero(void* dst, size_t bytes)
{
    return guarded([&] {
        require_init();
        lsjk::launch_fill(dst, bytes, S.stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(S.stream));
    });
}
int lsjk_memcpy_dtod(void* dst, const void* src, size_t bytes)
{
    return guarded([&] {
        require_init();
        if (bytes && (!dst || !src)) throw LsjkError(LSJK_E_INVALID, "lsjk_memcpy_dtod: null pointer");
        lsjk::launch_copy(dst, src, bytes, S.stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(S.stream));
    });
}
int lsjk_synchronize(void)
{
    return guarded([&] {
        require_init();
        HIPCHK(hipStreamSynchronize(S.stream));
    });
}
int lsjk_set_streams(int n)
{
    return guarded([&] {
        require_init();
        if (n < 1 || n > kSide) throw LsjkError(LSJK_E_INVALID, "lsjk_set_streams: n must be 1..4");
        S.nside = n;
    });
}
int lsjk_set_counting(int enable)
{
    return guarded([&] { S.counting = enable != 0; });
}
int lsjk_last_counts(uint64_t* counts)
{
    return guarded([&] {
        if (!counts) throw LsjkError(LSJK_E_INVALID, "lsjk_last_counts: null pointer");
        std::memcpy(counts, S.counts, sizeof(S.counts));
    });
}
int lsjk_last_counts_fp32(uint64_t* counts)
{
    return guarded([&] {
        if (!counts) throw LsjkError(LSJK_E_INVALID, "lsjk_last_counts_fp32: null pointer");
        std::memcpy(counts, S.counts32, sizeof(S.counts32));
    });
}
int lsjk_last_wave_steps(uint64_t* steps, uint64_t* lanes)
{
    return guarded([&] {
        if (!steps || !lanes) throw LsjkError(LSJK_E_INVALID, "lsjk_last_wave_steps: null pointer");
        std::memcpy(steps, S.wsteps, sizeof(S.wsteps));
        std::memcpy(lanes, S.wlanes, sizeof(S.wlanes));
    });
}
int lsjk_last_timing(double* ms)
{
    return guarded([&] {
        if (!ms) throw LsjkError(LSJK_E_INVALID, "lsjk_last_timing: null pointer");
        ms[0] = S.ms[0];
        ms[1] = S.ms[1];
    });
}

}  // extern "C"

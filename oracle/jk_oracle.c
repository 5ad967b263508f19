/*
 * jk_oracle.c -- TEST INFRASTRUCTURE ONLY (the parity checker, never the product path).
 *
 * A plain C + OpenMP restatement of the reference's FP64 CPU direct Coulomb/exchange build,
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to check the HIP
 * path.  Every function follows the reference file:line named in its comment.  The Rys and
 * Gauss-Hermite rules come from the project's generated tables (lightspeed_amd/tables/), which
 * are pinned independently by tests/test_quadrature.py (Boys moments, closed forms, limits).
 *
 * Pinning (see tests/test_oracle.py):
 *   - pair lists, J and K against the reference's fixtures prop_h2o_* (tol 1e-9),
 *     ewald_h2o_* (tol 1e-12) and mix_* (tol 1e-12), symmetric and non-symmetric lists;
 *   - (ss|ss) closed form, linearity, symmetry, translation invariance.
 *
 * Layouts: all matrices are row-major FP64 in the Cartesian representation of each basis
 * (ncart1 x ncart2 ...), Cartesian order lexical (xx, xy, xz, yy, yz, zz), as in
 * src/lightspeed/am.hpp:48-68.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../lightspeed_amd/csrc/common/rys_table.inc"
#include "../lightspeed_amd/csrc/common/gh_table.inc"

#define OB_MAXCLS 64
#define OB_MAXL 4

typedef struct {
    int nprim;
    int ncart;
    const double* c;
    const double* e;
    const double* xyz; /* nprim x 3 */
    const int* L;
    const int* cart; /* cartIdx of each primitive */
} ob_basis;

typedef struct {
    int symmetric;
    int npair;
    int* p1; /* primitive index into basis1 */
    int* p2; /* primitive index into basis2 */
    float* bound;
    int ncls;
    int cls_L1[OB_MAXCLS];
    int cls_L2[OB_MAXCLS];
    int cls_start[OB_MAXCLS + 1];
} ob_pairlist;

static inline int ncart_of(int L) { return (L + 1) * (L + 2) / 2; }

/* ---------------------------------------------------------------- Rys table evaluation */
/* Evaluate the n-root rule at T: x_i = t_i^2, w_i (table layout documented in rys_table.inc).
 * Same role as Rys::interpolate_t/_w, src/lightspeed/rys.hpp:56-153. */
static inline void rys_eval(int n, double T, double* x, double* w)
{
    if (T >= lsjk_rys_Tc[n]) {
        const double* r2 = lsjk_rys_asym_r2 + n * (n - 1) / 2;
        const double* W = lsjk_rys_asym_w + n * (n - 1) / 2;
        double iT = 1.0 / T;
        double s = sqrt(iT);
        for (int i = 0; i < n; i++) {
            x[i] = r2[i] * iT;
            w[i] = W[i] * s;
        }
        return;
    }
    double sT = T / lsjk_rys_h[n];
    int b = (int)sT;
    double u = 2.0 * (sT - b) - 1.0;
    const double* c = lsjk_rys_coef + lsjk_rys_offset[n] + (size_t)b * n * 2 * LSJK_RYS_NCOEF;
    for (int i = 0; i < n; i++) {
        const double* cx = c + (size_t)i * 2 * LSJK_RYS_NCOEF;
        const double* cw = cx + LSJK_RYS_NCOEF;
        double px = cx[LSJK_RYS_DEG];
        double pw = cw[LSJK_RYS_DEG];
        for (int k = LSJK_RYS_DEG - 1; k >= 0; k--) {
            px = px * u + cx[k];
            pw = pw * u + cw[k];
        }
        x[i] = px;
        w[i] = pw;
    }
}

void oracle_rys(int n, int npoint, const double* T, double* x, double* w)
{
    for (int p = 0; p < npoint; p++) rys_eval(n, T[p], x + (size_t)p * n, w + (size_t)p * n);
}

/* ---------------------------------------------------------------- pair lists */

/* Pair::build_schwarz, src/cpp/core/pair_list.cpp:15-27 (same operation order). */
static float schwarz_bound(const ob_basis* b1, int i, const ob_basis* b2, int j)
{
    const double* r1 = b1->xyz + 3 * i;
    const double* r2 = b2->xyz + 3 * j;
    double R2 = 0.0;
    R2 += pow(r1[0] - r2[0], 2);
    R2 += pow(r1[1] - r2[1], 2);
    R2 += pow(r1[2] - r2[2], 2);
    double e1 = b1->e[i], e2 = b2->e[j];
    return (float)(fabs(b1->c[i] * b2->c[j]) * exp(-e1 * e2 / (e1 + e2) * R2) * pow(2.0, 0.25) *
                   pow(M_PI / (e1 + e2), 1.25));
}

/* HashedGrid::significant_box, src/lightspeed/gridbox.hpp:132-150. */
static int significant_box(int kx, int ky, int kz, double R, double xA, double yA, double zA, double Rc)
{
    double C1x = kx * R, C2x = (kx + 1) * R;
    double C1y = ky * R, C2y = (ky + 1) * R;
    double C1z = kz * R, C2z = (kz + 1) * R;
    double D2 = pow(Rc, 2);
    D2 -= (xA < C1x) * pow(xA - C1x, 2);
    D2 -= (xA > C2x) * pow(xA - C2x, 2);
    D2 -= (yA < C1y) * pow(yA - C1y, 2);
    D2 -= (yA > C2y) * pow(yA - C2y, 2);
    D2 -= (zA < C1z) * pow(zA - C1z, 2);
    D2 -= (zA > C2z) * pow(zA - C2z, 2);
    return D2 >= 0.0;
}

typedef struct {
    int p1, p2;
    float b;
} ob_pair;

/* PairListUtil::compare_prim_pairs, src/cpp/core/pair_list_util.cpp:12-20 (bound desc, p1 asc,
 * p2 asc) as a qsort comparator. */
static int cmp_pairs(const void* va, const void* vb)
{
    const ob_pair* a = (const ob_pair*)va;
    const ob_pair* b = (const ob_pair*)vb;
    if (a->b != b->b) return (a->b > b->b) ? -1 : 1;
    if (a->p1 != b->p1) return (a->p1 < b->p1) ? -1 : 1;
    if (a->p2 != b->p2) return (a->p2 < b->p2) ? -1 : 1;
    return 0;
}

static int max_L(const ob_basis* b)
{
    int m = 0;
    for (int i = 0; i < b->nprim; i++)
        if (b->L[i] > m) m = b->L[i];
    return m;
}

/* Fill pl from a list of pairs, grouping by (L1, L2) class (L1-major, empty classes dropped)
 * and sorting each class with compare_prim_pairs (pair_list.cpp:152-173). */
static void group_pairs(ob_pairlist* pl, ob_pair* pairs, size_t n, const ob_basis* b1, const ob_basis* b2,
                        int symmetric, int keep_empty_classes, const int* cls_L1, const int* cls_L2, int ncls_in)
{
    int ml1 = max_L(b1), ml2 = max_L(b2);
    int nL = (ml1 + 1) * (ml2 + 1);
    size_t* cnt = (size_t*)calloc(nL, sizeof(size_t));
    for (size_t k = 0; k < n; k++) cnt[b1->L[pairs[k].p1] * (ml2 + 1) + b2->L[pairs[k].p2]]++;
    pl->symmetric = symmetric;
    pl->npair = (int)n;
    pl->p1 = (int*)malloc(sizeof(int) * (n ? n : 1));
    pl->p2 = (int*)malloc(sizeof(int) * (n ? n : 1));
    pl->bound = (float*)malloc(sizeof(float) * (n ? n : 1));
    ob_pair* sorted = (ob_pair*)malloc(sizeof(ob_pair) * (n ? n : 1));
    size_t off = 0;
    int ncls = 0;
    for (int L1 = 0; L1 <= ml1; L1++) {
        for (int L2 = 0; L2 <= ml2; L2++) {
            int l = L1 * (ml2 + 1) + L2;
            int present = cnt[l] > 0;
            if (keep_empty_classes) {
                present = 0;
                for (int q = 0; q < ncls_in; q++)
                    if (cls_L1[q] == L1 && cls_L2[q] == L2) present = 1;
            }
            if (!present) continue;
            size_t m = 0;
            for (size_t k = 0; k < n; k++)
                if (b1->L[pairs[k].p1] == L1 && b2->L[pairs[k].p2] == L2) sorted[off + m++] = pairs[k];
            qsort(sorted + off, m, sizeof(ob_pair), cmp_pairs);
            pl->cls_L1[ncls] = L1;
            pl->cls_L2[ncls] = L2;
            pl->cls_start[ncls] = (int)off;
            ncls++;
            off += m;
        }
    }
    pl->cls_start[ncls] = (int)off;
    pl->ncls = ncls;
    for (size_t k = 0; k < n; k++) {
        pl->p1[k] = sorted[k].p1;
        pl->p2[k] = sorted[k].p2;
        pl->bound[k] = sorted[k].b;
    }
    free(sorted);
    free(cnt);
}

/* PairList::build_schwarz, src/cpp/core/pair_list.cpp:29-177 (hashed candidate search with box
 * R = 5 bohr and reach Rmax, keep bound > threpq, p2 <= p1 if symmetric). */
int oracle_pairlist_build(const ob_basis* b1, const ob_basis* b2, int symmetric, float threpq, ob_pairlist* pl)
{
    if (symmetric && b1->nprim != b2->nprim) return -1;
    double Rmax = 0.0;
    const ob_basis* bs[2] = {b1, b2};
    for (int q = 0; q < 2; q++) {
        for (int P = 0; P < bs[q]->nprim; P++) {
            double c = bs[q]->c[P], e = bs[q]->e[P];
            double R2 = -2.0 / e * log(threpq / (c * c * (pow(2.0, 0.25) * pow(M_PI / (e + e), 1.25))));
            if (R2 < 0.0) continue;
            double R = sqrt(R2);
            if (R > Rmax) Rmax = R;
        }
    }
    const double Rval = 5.0;
    const double Rinv = 1.0 / Rval;
    int nstep = (int)ceil(Rmax / Rval);
    int n1 = b1->nprim, n2 = b2->nprim;
    int* kq = (int*)malloc(sizeof(int) * 3 * (n2 ? n2 : 1));
    for (int Q = 0; Q < n2; Q++)
        for (int d = 0; d < 3; d++) kq[3 * Q + d] = (int)floor(b2->xyz[3 * Q + d] * Rinv);
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    ob_pair** tp = (ob_pair**)calloc(nthr, sizeof(ob_pair*));
    size_t* tn = (size_t*)calloc(nthr, sizeof(size_t));
    size_t* tc = (size_t*)calloc(nthr, sizeof(size_t));
#pragma omp parallel for schedule(dynamic)
    for (int P = 0; P < n1; P++) {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        const double* rP = b1->xyz + 3 * P;
        int kp[3];
        for (int d = 0; d < 3; d++) kp[d] = (int)floor(rP[d] * Rinv);
        for (int Q = 0; Q < n2; Q++) {
            if (symmetric && Q > P) continue;
            const int* k2 = kq + 3 * Q;
            if (abs(k2[0] - kp[0]) > nstep || abs(k2[1] - kp[1]) > nstep || abs(k2[2] - kp[2]) > nstep) continue;
            if (!significant_box(k2[0], k2[1], k2[2], Rval, rP[0], rP[1], rP[2], Rmax)) continue;
            float b = schwarz_bound(b1, P, b2, Q);
            if (b > threpq) {
                if (tn[t] == tc[t]) {
                    tc[t] = tc[t] ? 2 * tc[t] : 1024;
                    tp[t] = (ob_pair*)realloc(tp[t], sizeof(ob_pair) * tc[t]);
                }
                tp[t][tn[t]].p1 = P;
                tp[t][tn[t]].p2 = Q;
                tp[t][tn[t]].b = b;
                tn[t]++;
            }
        }
    }
    size_t ntot = 0;
    for (int t = 0; t < nthr; t++) ntot += tn[t];
    ob_pair* all = (ob_pair*)malloc(sizeof(ob_pair) * (ntot ? ntot : 1));
    size_t o = 0;
    for (int t = 0; t < nthr; t++) {
        if (tn[t]) memcpy(all + o, tp[t], sizeof(ob_pair) * tn[t]);
        o += tn[t];
        free(tp[t]);
    }
    free(tp);
    free(tn);
    free(tc);
    free(kq);
    group_pairs(pl, all, ntot, b1, b2, symmetric, 0, NULL, NULL, 0);
    free(all);
    return 0;
}

void oracle_pairlist_free(ob_pairlist* pl)
{
    free(pl->p1);
    free(pl->p2);
    free(pl->bound);
    pl->p1 = pl->p2 = NULL;
    pl->bound = NULL;
}

/* PairListUtil::max_bound, src/cpp/core/pair_list_util.cpp:22-31. */
static double max_bound(const ob_pairlist* pl)
{
    double v = 0.0;
    for (int c = 0; c < pl->ncls; c++) {
        double f = (pl->cls_start[c + 1] > pl->cls_start[c]) ? (double)pl->bound[pl->cls_start[c]] : 0.0;
        if (f > v) v = f;
    }
    return v;
}

/* PairListUtil::truncate_bra, src/cpp/core/pair_list_util.cpp:32-58: keep each class's prefix
 * with bound >= thre. */
static void truncate_bra(const ob_pairlist* in, float thre, ob_pairlist* out)
{
    *out = *in;
    out->p1 = (int*)malloc(sizeof(int) * (in->npair ? in->npair : 1));
    out->p2 = (int*)malloc(sizeof(int) * (in->npair ? in->npair : 1));
    out->bound = (float*)malloc(sizeof(float) * (in->npair ? in->npair : 1));
    int o = 0;
    for (int c = 0; c < in->ncls; c++) {
        out->cls_start[c] = o;
        for (int k = in->cls_start[c]; k < in->cls_start[c + 1]; k++) {
            if (in->bound[k] < thre) break;
            out->p1[o] = in->p1[k];
            out->p2[o] = in->p2[k];
            out->bound[o] = in->bound[k];
            o++;
        }
    }
    out->cls_start[in->ncls] = o;
    out->npair = o;
}

/* PairListUtil::truncate_ket, src/cpp/core/pair_list_util.cpp:60-123: bound <- float(B *
 * max|D block|) (both triangles if symmetric), keep > thre, re-sort each class. */
static void truncate_ket(const ob_pairlist* in, const ob_basis* b1, const ob_basis* b2, const double* D,
                         float thre, ob_pairlist* out)
{
    *out = *in;
    int n = in->npair;
    out->p1 = (int*)malloc(sizeof(int) * (n ? n : 1));
    out->p2 = (int*)malloc(sizeof(int) * (n ? n : 1));
    out->bound = (float*)malloc(sizeof(float) * (n ? n : 1));
    ob_pair* tmp = (ob_pair*)malloc(sizeof(ob_pair) * (n ? n : 1));
    size_t n2 = b2->ncart;
    int o = 0;
    for (int c = 0; c < in->ncls; c++) {
        out->cls_start[c] = o;
        int o0 = o;
        for (int k = in->cls_start[c]; k < in->cls_start[c + 1]; k++) {
            float B = in->bound[k];
            int i1 = in->p1[k], i2 = in->p2[k];
            int c1 = ncart_of(b1->L[i1]), c2 = ncart_of(b2->L[i2]);
            int a1 = b1->cart[i1], a2 = b2->cart[i2];
            double val = 0.0;
            for (int q1 = 0; q1 < c1; q1++)
                for (int q2 = 0; q2 < c2; q2++) val = fmax(val, fabs(D[(q1 + a1) * n2 + (q2 + a2)]));
            if (in->symmetric) {
                for (int q1 = 0; q1 < c1; q1++)
                    for (int q2 = 0; q2 < c2; q2++) val = fmax(val, fabs(D[(q2 + a2) * n2 + (q1 + a1)]));
            }
            float B2 = (float)(val * B);
            if (B2 > thre) {
                tmp[o].p1 = i1;
                tmp[o].p2 = i2;
                tmp[o].b = B2;
                o++;
            }
        }
        qsort(tmp + o0, o - o0, sizeof(ob_pair), cmp_pairs);
    }
    out->cls_start[in->ncls] = o;
    out->npair = o;
    for (int k = 0; k < o; k++) {
        out->p1[k] = tmp[k].p1;
        out->p2[k] = tmp[k].p2;
        out->bound[k] = tmp[k].b;
    }
    free(tmp);
}

/* ---------------------------------------------------------------- Hermite (J) */

/* McMurchie-Davidson expansion coefficients E[d][t][l1][l2] of one primitive pair by the exact
 * (L1+L2+1)-point Gauss-Hermite quadrature, src/cpp/core/hermite.cpp:87-134. */
static void hermite_E(int L1, int L2, double e1, const double* r1, double e2, const double* r2, double* E)
{
    int V1 = L1 + 1, V2 = L2 + 1, V = L1 + L2 + 1;
    const double* ts = lsjk_gh_x + V * (V - 1) / 2;
    const double* ws = lsjk_gh_w + V * (V - 1) / 2;
    double p = e1 + e2;
    double pm1 = 1.0 / p;
    double pm12 = pow(p, -0.5);
    memset(E, 0, sizeof(double) * 3 * V * V1 * V2);
    for (int d = 0; d < 3; d++) {
        double xPA = e2 * (r2[d] - r1[d]) * pm1;
        double xPB = e1 * (r1[d] - r2[d]) * pm1;
        double* Ex = E + d * V * V1 * V2;
        for (int t = 0; t < V; t++) {
            double xgh = ts[t] * pm12;
            double ugh = ws[t] * pm12;
            double chi[2 * OB_MAXL + 1];
            chi[0] = ugh * sqrt(p / M_PI);
            if (V > 1) chi[1] = xgh * chi[0];
            for (int k = 1; k < V - 1; k++) chi[k + 1] = 1.0 / (k + 1) * (xgh * chi[k] - 0.5 * pm1 * chi[k - 1]);
            double X1[OB_MAXL + 1], X2[OB_MAXL + 1];
            X1[0] = 1.0;
            for (int l = 1; l < V1; l++) X1[l] = X1[l - 1] * (xgh + xPA);
            X2[0] = 1.0;
            for (int l = 1; l < V2; l++) X2[l] = X2[l - 1] * (xgh + xPB);
            for (int Lh = 0; Lh < V; Lh++)
                for (int l1 = 0; l1 < V1; l1++)
                    for (int l2 = 0; l2 < V2; l2++) {
                        if (Lh > l1 + l2) continue;
                        Ex[Lh * V1 * V2 + l1 * V2 + l2] += chi[Lh] * X1[l1] * X2[l2];
                    }
        }
    }
}

static inline int hermite_count(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

/* Hermite index t -> (tx, ty, tz), HermiteUtil::angl ordering (hermite_util.cpp:6-20). */
static void hermite_tuples(int L, int* tx, int* ty, int* tz)
{
    int idx = 0;
    for (int Lh = 0; Lh <= L; Lh++)
        for (int i = 0; i <= Lh; i++)
            for (int j = 0; j <= i; j++) {
                tx[idx] = Lh - i;
                ty[idx] = i - j;
                tz[idx] = j;
                idx++;
            }
}

/* HermiteL::cart_to_herm (phases) / herm_to_cart (no phases), src/cpp/core/hermite.cpp:9-206 and
 * 207-411.  dir = +1: dh[pair][t] += sum K E E E Dcart (symmetrised, 0.5 on diagonal pairs);
 * dir = -1: Dcart += sum K E E E dh (symmetrised scatter). */
static void hermite_transform(const ob_pairlist* pl, int cls, const ob_basis* b1, const ob_basis* b2,
                              double* dh, double* Dc, int dir, int phases)
{
    int L1 = pl->cls_L1[cls], L2 = pl->cls_L2[cls];
    int L = L1 + L2, V1 = L1 + 1, V2 = L2 + 1, V = L + 1;
    int N1 = ncart_of(L1), N2 = ncart_of(L2);
    int H = hermite_count(L);
    size_t nc2 = b2->ncart;
    int symm = pl->symmetric;
    int s0 = pl->cls_start[cls], s1 = pl->cls_start[cls + 1];
#pragma omp parallel for schedule(dynamic)
    for (int k = s0; k < s1; k++) {
        int i1 = pl->p1[k], i2 = pl->p2[k];
        const double* r1 = b1->xyz + 3 * i1;
        const double* r2 = b2->xyz + 3 * i2;
        double e1 = b1->e[i1], e2 = b2->e[i2];
        double pm1 = 1.0 / (e1 + e2);
        double Kab = b1->c[i1] * b2->c[i2] *
                     exp(-e1 * e2 * pm1 *
                         ((r1[0] - r2[0]) * (r1[0] - r2[0]) + (r1[1] - r2[1]) * (r1[1] - r2[1]) +
                          (r1[2] - r2[2]) * (r1[2] - r2[2])));
        double E[3 * (2 * OB_MAXL + 1) * (OB_MAXL + 1) * (OB_MAXL + 1)];
        hermite_E(L1, L2, e1, r1, e2, r2, E);
        const double* Ex = E;
        const double* Ey = E + V * V1 * V2;
        const double* Ez = E + 2 * V * V1 * V2;
        double Dt[225];
        int off1 = b1->cart[i1], off2 = b2->cart[i2];
        double scale = (symm && i1 == i2) ? 0.5 : 1.0;
        /* the reference's symmetric/diagonal test compares primIdx; here the two bases of a
         * symmetric list are the same basis, so primitive indices compare directly */
        if (dir > 0) {
            for (int p = 0; p < N1; p++)
                for (int q = 0; q < N2; q++) {
                    double v = Dc[(p + off1) * nc2 + (q + off2)];
                    if (symm) {
                        v += Dc[(q + off2) * nc2 + (p + off1)];
                        v *= scale;
                    }
                    Dt[p * N2 + q] = v;
                }
        }
        double* dt = dh + (size_t)k * H;
        int i1c = 0;
        for (int a = 0; a <= L1; a++) {
            int l1 = L1 - a;
            for (int b = 0; b <= a; b++, i1c++) {
                int m1 = a - b, n1 = b;
                int i2c = 0;
                for (int a2 = 0; a2 <= L2; a2++) {
                    int l2 = L2 - a2;
                    for (int b2 = 0; b2 <= a2; b2++, i2c++) {
                        int m2 = a2 - b2, n2 = b2;
                        double acc = 0.0;
                        double Dval = (dir > 0) ? Dt[i1c * N2 + i2c] : 0.0;
                        int ih = 0;
                        for (int Lh = 0; Lh < V; Lh++)
                            for (int ii = 0; ii <= Lh; ii++) {
                                int lh = Lh - ii;
                                for (int jj = 0; jj <= ii; jj++, ih++) {
                                    int mh = ii - jj, nh = jj;
                                    double sign = 1.0;
                                    if (phases) sign = ((lh + mh + nh) % 2) ? -1.0 : 1.0;
                                    double f = sign * Kab * Ex[lh * V1 * V2 + l1 * V2 + l2] *
                                               Ey[mh * V1 * V2 + m1 * V2 + m2] * Ez[nh * V1 * V2 + n1 * V2 + n2];
                                    if (dir > 0)
                                        dt[ih] += f * Dval;
                                    else
                                        acc += f * dt[ih];
                                }
                            }
                        if (dir < 0) Dt[i1c * N2 + i2c] = acc;
                    }
                }
            }
        }
        if (dir < 0) {
            for (int p = 0; p < N1; p++)
                for (int q = 0; q < N2; q++) {
                    double v = Dt[p * N2 + q];
                    if (symm) {
                        v *= scale;
#pragma omp atomic
                        Dc[(p + off1) * nc2 + (q + off2)] += v;
#pragma omp atomic
                        Dc[(q + off2) * nc2 + (p + off1)] += v;
                    } else {
#pragma omp atomic
                        Dc[(p + off1) * nc2 + (q + off2)] += v;
                    }
                }
        }
    }
}

/* IntBox::coulombCPU, src/cpp/intbox/cpu/coulomb.cpp:14-196, in the Cartesian representation.
 * J12 (ncart1 x ncart2) is accumulated.  counts (optional) receives the number of evaluated
 * (bra pair, ket pair) primitive quartets per (L1,L2,L3,L4) class, indexed
 * ((L1*5+L2)*5+L3)*5+L4. */
int oracle_coulomb(const ob_basis* b1, const ob_basis* b2, const ob_basis* b3, const ob_basis* b4,
                   const ob_pairlist* pl12, const ob_pairlist* pl34, const double* D34, int nterm,
                   const double* alphas, const double* omegas, float thresp, double* J12, int64_t* counts)
{
    float max12 = (float)max_bound(pl12);
    ob_pairlist ket, bra;
    truncate_ket(pl34, b3, b4, D34, thresp / max12, &ket);
    float max34 = (float)max_bound(&ket);
    truncate_bra(pl12, thresp / max34, &bra);

    /* Hermite data (coulomb.cpp:55-59) */
    size_t off12[OB_MAXCLS + 1], off34[OB_MAXCLS + 1];
    off12[0] = 0;
    for (int c = 0; c < bra.ncls; c++)
        off12[c + 1] = off12[c] + (size_t)(bra.cls_start[c + 1] - bra.cls_start[c]) *
                                      hermite_count(bra.cls_L1[c] + bra.cls_L2[c]);
    off34[0] = 0;
    for (int c = 0; c < ket.ncls; c++)
        off34[c + 1] = off34[c] + (size_t)(ket.cls_start[c + 1] - ket.cls_start[c]) *
                                      hermite_count(ket.cls_L1[c] + ket.cls_L2[c]);
    double* h12 = (double*)calloc(off12[bra.ncls] + 1, sizeof(double));
    double* h34 = (double*)calloc(off34[ket.ncls] + 1, sizeof(double));
    for (int c = 0; c < ket.ncls; c++) {
        /* hermite_transform indexes its buffer by the absolute pair index k */
        double* base = h34 + off34[c] - (size_t)ket.cls_start[c] * hermite_count(ket.cls_L1[c] + ket.cls_L2[c]);
        hermite_transform(&ket, c, b3, b4, base, (double*)D34, +1, 1);
    }

    for (int it = 0; it < nterm; it++) {
        double alpha = alphas[it], omega = omegas[it];
        if (alpha == 0.0 || omega == 0.0) continue;
        for (int c12 = 0; c12 < bra.ncls; c12++) {
            int L12 = bra.cls_L1[c12] + bra.cls_L2[c12];
            int H12 = hermite_count(L12);
            int s12 = bra.cls_start[c12], e12 = bra.cls_start[c12 + 1];
            for (int c34 = 0; c34 < ket.ncls; c34++) {
                int L34 = ket.cls_L1[c34] + ket.cls_L2[c34];
                int H34 = hermite_count(L34);
                int s34 = ket.cls_start[c34], e34 = ket.cls_start[c34 + 1];
                int L = L12 + L34;
                int nrys = L / 2 + 1;
                int tx[165], ty[165], tz[165];
                hermite_tuples(L12 > L34 ? L12 : L34, tx, ty, tz);
                int64_t nq = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : nq)
                for (int k12 = s12; k12 < e12; k12++) {
                    float B12 = bra.bound[k12];
                    int i1 = bra.p1[k12], i2 = bra.p2[k12];
                    double e1 = b1->e[i1], e2 = b2->e[i2];
                    double p12 = e1 + e2;
                    double P[3];
                    for (int d = 0; d < 3; d++) P[d] = (e1 * b1->xyz[3 * i1 + d] + e2 * b2->xyz[3 * i2 + d]) * (1.0 / p12);
                    double jp[165];
                    memset(jp, 0, sizeof(jp));
                    for (int k34 = s34; k34 < e34; k34++) {
                        float B34 = ket.bound[k34];
                        if (B12 * B34 < thresp) break;
                        nq++;
                        int i3 = ket.p1[k34], i4 = ket.p2[k34];
                        double e3 = b3->e[i3], e4 = b4->e[i4];
                        double p34 = e3 + e4;
                        double Q[3];
                        for (int d = 0; d < 3; d++)
                            Q[d] = (e3 * b3->xyz[3 * i3 + d] + e4 * b4->xyz[3 * i4 + d]) * (1.0 / p34);
                        const double* dp = h34 + off34[c34] + (size_t)(k34 - s34) * H34;
                        double pref = alpha * 2.0 * pow(M_PI, 2.5) / (p12 * p34 * sqrt(p12 + p34));
                        double rho = p12 * p34 / (p12 + p34);
                        double d2 = 1.0;
                        if (omega != -1.0) {
                            d2 = omega * omega / (rho + omega * omega);
                            pref *= sqrt(d2);
                        }
                        double xPQ = P[0] - Q[0], yPQ = P[1] - Q[1], zPQ = P[2] - Q[2];
                        double T = rho * d2 * (xPQ * xPQ + yPQ * yPQ + zPQ * zPQ);
                        double rx[9], rw[9];
                        rys_eval(nrys, T, rx, rw);
                        for (int r = 0; r < nrys; r++) {
                            double U = -2.0 * rho * d2 * rx[r];
                            double Hx[2 * 2 * OB_MAXL + 1], Hy[2 * 2 * OB_MAXL + 1], Hz[2 * 2 * OB_MAXL + 1];
                            Hx[0] = pref * rw[r];
                            Hy[0] = 1.0;
                            Hz[0] = 1.0;
                            if (L > 0) {
                                Hx[1] = U * xPQ * Hx[0];
                                Hy[1] = U * yPQ;
                                Hz[1] = U * zPQ;
                            }
                            for (int u = 1; u < L; u++) {
                                Hx[u + 1] = U * (xPQ * Hx[u] + u * Hx[u - 1]);
                                Hy[u + 1] = U * (yPQ * Hy[u] + u * Hy[u - 1]);
                                Hz[u + 1] = U * (zPQ * Hz[u] + u * Hz[u - 1]);
                            }
                            for (int t = 0; t < H12; t++)
                                for (int u = 0; u < H34; u++)
                                    jp[t] += Hx[tx[t] + tx[u]] * Hy[ty[t] + ty[u]] * Hz[tz[t] + tz[u]] * dp[u];
                        }
                    }
                    double* dst = h12 + off12[c12] + (size_t)(k12 - s12) * H12;
                    for (int t = 0; t < H12; t++) dst[t] += jp[t];
                }
                if (counts) {
                    int idx = ((bra.cls_L1[c12] * 5 + bra.cls_L2[c12]) * 5 + ket.cls_L1[c34]) * 5 + ket.cls_L2[c34];
                    counts[idx] += nq;
                }
            }
        }
    }
    for (int c = 0; c < bra.ncls; c++) {
        double* base = h12 + off12[c] - (size_t)bra.cls_start[c] * hermite_count(bra.cls_L1[c] + bra.cls_L2[c]);
        hermite_transform(&bra, c, b1, b2, base, J12, -1, 0);
    }
    free(h12);
    free(h34);
    oracle_pairlist_free(&ket);
    oracle_pairlist_free(&bra);
    return 0;
}

/* ---------------------------------------------------------------- K-invert (K) */

typedef struct {
    int pivot;    /* primitive index of the pivot (the reference's KVecL::prim1) */
    int start;    /* first partner in the block's partner arrays */
    int n;        /* partner count */
} ob_kvec;

typedef struct {
    int Lpiv, Lpart;
    int nvec;
    ob_kvec* vecs;
    int* part;    /* partner primitive indices */
    float* bound; /* pair bounds, descending per kvec */
} ob_kblockl;

/* kblock12_comparator, src/cpp/core/kvec.cpp:17-25 (pivot asc, bound desc, partner asc). */
static int cmp_kvec(const void* va, const void* vb)
{
    const ob_pair* a = (const ob_pair*)va;
    const ob_pair* b = (const ob_pair*)vb;
    if (a->p1 != b->p1) return (a->p1 < b->p1) ? -1 : 1;
    if (a->b != b->b) return (a->b > b->b) ? -1 : 1;
    if (a->p2 != b->p2) return (a->p2 < b->p2) ? -1 : 1;
    return 0;
}

/* KBlock::build_kblockl, src/cpp/core/kvec.cpp:31-97: pairs given as (pivot, partner, bound). */
static void build_kblockl(int Lpiv, int Lpart, ob_pair* pairs, int n, ob_kblockl* kb)
{
    qsort(pairs, n, sizeof(ob_pair), cmp_kvec);
    kb->Lpiv = Lpiv;
    kb->Lpart = Lpart;
    kb->part = (int*)malloc(sizeof(int) * (n ? n : 1));
    kb->bound = (float*)malloc(sizeof(float) * (n ? n : 1));
    kb->vecs = (ob_kvec*)malloc(sizeof(ob_kvec) * (n ? n : 1));
    int nv = 0;
    for (int k = 0; k < n; k++) {
        if (k == 0 || pairs[k].p1 != pairs[k - 1].p1) {
            kb->vecs[nv].pivot = pairs[k].p1;
            kb->vecs[nv].start = k;
            kb->vecs[nv].n = 0;
            nv++;
        }
        kb->vecs[nv - 1].n++;
        kb->part[k] = pairs[k].p2;
        kb->bound[k] = pairs[k].b;
    }
    kb->nvec = nv;
}

/* KBlock::build21 / build_symmetric, src/cpp/core/kvec.cpp:128-232.  Produces blocks whose
 * pivots are "2" primitives of a (12) list (basis2 of the list) and partners "1" primitives.
 * Returns the number of blocks written to kbs. */
static int build21(const ob_pairlist* pl, float thre, ob_kblockl* kbs)
{
    int nb = 0;
    ob_pair* tmp = (ob_pair*)malloc(sizeof(ob_pair) * (2 * (size_t)pl->npair + 1));
    if (!pl->symmetric) {
        /* classes ordered by (L2, L1) */
        int order[OB_MAXCLS];
        for (int c = 0; c < pl->ncls; c++) order[c] = c;
        for (int a = 0; a < pl->ncls; a++)
            for (int b = a + 1; b < pl->ncls; b++) {
                int ca = order[a], cb = order[b];
                int ka = pl->cls_L2[ca] * 100 + pl->cls_L1[ca], kb = pl->cls_L2[cb] * 100 + pl->cls_L1[cb];
                if (kb < ka || (kb == ka && cb < ca)) {
                    order[a] = cb;
                    order[b] = ca;
                }
            }
        for (int q = 0; q < pl->ncls; q++) {
            int c = order[q];
            int n = 0;
            for (int k = pl->cls_start[c]; k < pl->cls_start[c + 1]; k++) {
                if (pl->bound[k] >= thre) {
                    tmp[n].p1 = pl->p2[k]; /* pivot: the list's 2nd primitive */
                    tmp[n].p2 = pl->p1[k];
                    tmp[n].b = pl->bound[k];
                    n++;
                }
            }
            build_kblockl(pl->cls_L2[c], pl->cls_L1[c], tmp, n, &kbs[nb++]);
        }
    } else {
        /* all (A,B) and (B,A) combos, in lexicographic order */
        int present[OB_MAXL + 1][OB_MAXL + 1];
        memset(present, 0, sizeof(present));
        for (int c = 0; c < pl->ncls; c++) {
            present[pl->cls_L1[c]][pl->cls_L2[c]] = 1;
            present[pl->cls_L2[c]][pl->cls_L1[c]] = 1;
        }
        for (int A = 0; A <= OB_MAXL; A++)
            for (int B = 0; B <= OB_MAXL; B++) {
                if (!present[A][B]) continue;
                int n = 0;
                for (int c = 0; c < pl->ncls; c++) {
                    if (pl->cls_L1[c] != A || pl->cls_L2[c] != B) continue;
                    for (int k = pl->cls_start[c]; k < pl->cls_start[c + 1]; k++)
                        if (pl->bound[k] >= thre) {
                            tmp[n].p1 = pl->p1[k];
                            tmp[n].p2 = pl->p2[k];
                            tmp[n].b = pl->bound[k];
                            n++;
                        }
                }
                for (int c = 0; c < pl->ncls; c++) {
                    if (pl->cls_L2[c] != A || pl->cls_L1[c] != B) continue;
                    for (int k = pl->cls_start[c]; k < pl->cls_start[c + 1]; k++) {
                        if (pl->p1[k] == pl->p2[k]) continue;
                        if (pl->bound[k] >= thre) {
                            tmp[n].p1 = pl->p2[k];
                            tmp[n].p2 = pl->p1[k];
                            tmp[n].b = pl->bound[k];
                            n++;
                        }
                    }
                }
                build_kblockl(A, B, tmp, n, &kbs[nb++]);
            }
    }
    free(tmp);
    return nb;
}

static void free_kblocks(ob_kblockl* kbs, int n)
{
    for (int i = 0; i < n; i++) {
        free(kbs[i].vecs);
        free(kbs[i].part);
        free(kbs[i].bound);
    }
}

/* IntBox::exchangeCPU, src/cpp/intbox/cpu/exchange.cpp:17-355, in the Cartesian representation.
 * K13 (ncart1 x ncart3) accumulated from D24 (ncart2 x ncart4).  In a symmetric list the
 * "1" basis and the "2" basis of the list are the same basis (b1 == b2 and b3 == b4).
 * D24symm must already be the effective flag (the caller checks basis equivalence as
 * exchange.cpp:73-81 does).  same_list: pl12 and pl34 are the same object. */
int oracle_exchange(const ob_basis* b1, const ob_basis* b2, const ob_basis* b3, const ob_basis* b4,
                    const ob_pairlist* pl12, const ob_pairlist* pl34, int same_list, const double* D24,
                    int D24symm, int nterm, const double* alphas, const double* omegas, float thresp,
                    double* K13, int64_t* counts)
{
    size_t nc3 = b3->ncart, nc4 = b4->ncart;
    double D24max = 0.0;
    for (size_t i = 0; i < (size_t)b2->ncart * nc4; i++) D24max = fmax(D24max, fabs(D24[i]));
    ob_kblockl kb12[OB_MAXCLS * 2], kb34[OB_MAXCLS * 2];
    int nkb12 = build21(pl12, (float)(thresp / (max_bound(pl34) * D24max)), kb12);
    int nkb34;
    ob_kblockl* k34 = kb34;
    if (same_list) {
        k34 = kb12;
        nkb34 = nkb12;
    } else {
        nkb34 = build21(pl34, (float)(thresp / (max_bound(pl12) * D24max)), kb34);
    }
    /* In the kblocks built from pl12 the pivots are basis-2 primitives and partners basis-1. */
    for (int it = 0; it < nterm; it++) {
        double alpha = alphas[it], omega = omegas[it];
        if (alpha == 0.0 || omega == 0.0) continue;
        for (int q12 = 0; q12 < nkb12; q12++) {
            const ob_kblockl* A = &kb12[q12];
            for (int q34 = 0; q34 < nkb34; q34++) {
                const ob_kblockl* B = &k34[q34];
                int L1 = A->Lpart, L2 = A->Lpiv, L3 = B->Lpart, L4 = B->Lpiv;
                int V1 = L1 + 1, V2 = L2 + 1, V3 = L3 + 1, V4 = L4 + 1;
                int nrys = (L1 + L2 + L3 + L4) / 2 + 1;
                const double* tgh = lsjk_gh_x + nrys * (nrys - 1) / 2;
                const double* wgh = lsjk_gh_w + nrys * (nrys - 1) / 2;
                int N1 = ncart_of(L1), N2 = ncart_of(L2), N3 = ncart_of(L3), N4 = ncart_of(L4);
                int V1234 = V1 * V2 * V3 * V4;
                int64_t nq = 0;
                int64_t n24 = (int64_t)A->nvec * B->nvec;
#pragma omp parallel for schedule(dynamic) reduction(+ : nq)
                for (int64_t ind24 = 0; ind24 < n24; ind24++) {
                    const ob_kvec* kv12 = &A->vecs[ind24 / B->nvec];
                    const ob_kvec* kv34 = &B->vecs[ind24 % B->nvec];
                    int K2 = kv12->pivot, K4 = kv34->pivot;
                    if (D24symm && K2 > K4) continue;
                    int flip24 = (D24symm && K2 != K4);
                    int o2 = b2->cart[K2], o4 = b4->cart[K4];
                    double D24V[225];
                    double D24Vmax = 0.0;
                    for (int p2 = 0; p2 < N2; p2++)
                        for (int p4 = 0; p4 < N4; p4++) {
                            double Dval = D24[(size_t)(p2 + o2) * nc4 + (p4 + o4)];
                            D24V[p2 * N4 + p4] = Dval;
                            D24Vmax = fmax(D24Vmax, fabs(Dval));
                        }
                    double e2 = b2->e[K2], e4 = b4->e[K4];
                    const double* x2 = b2->xyz + 3 * K2;
                    const double* x4 = b4->xyz + 3 * K4;
                    for (int i1 = 0; i1 < kv12->n; i1++) {
                        float bound12 = A->bound[kv12->start + i1];
                        int P1 = A->part[kv12->start + i1];
                        double e1 = b1->e[P1];
                        const double* x1 = b1->xyz + 3 * P1;
                        /* K12 as geom2[0] of KVecL (kvec.cpp:68-72): pivot first */
                        double K12 = b2->c[K2] * b1->c[P1] *
                                     exp(-e2 * e1 / (e2 + e1) *
                                         (pow(x2[0] - x1[0], 2) + pow(x2[1] - x1[1], 2) + pow(x2[2] - x1[2], 2)));
                        double pref12 = alpha * K12 * 2.0 * pow(M_PI, -0.5);
                        double e12 = e1 + e2;
                        double e12_m1 = 1.0 / e12;
                        int did_work = 0;
                        for (int i3 = 0; i3 < kv34->n; i3++) {
                            float bound34 = B->bound[kv34->start + i3];
                            if (bound12 * bound34 * D24Vmax < thresp) break;
                            did_work = 1;
                            nq++;
                            int P3 = B->part[kv34->start + i3];
                            double e3 = b3->e[P3];
                            const double* x3 = b3->xyz + 3 * P3;
                            double K34 = b4->c[K4] * b3->c[P3] *
                                         exp(-e4 * e3 / (e4 + e3) *
                                             (pow(x4[0] - x3[0], 2) + pow(x4[1] - x3[1], 2) + pow(x4[2] - x3[2], 2)));
                            double pref1234 = pref12 * K34;
                            double e34 = e3 + e4;
                            double e34_m1 = 1.0 / e34;
                            pref1234 *= e12_m1 * e34_m1 * pow(e12 + e34, -0.5);
                            double xP[3], xQ[3], rPQ_2 = 0.0;
                            for (int d = 0; d < 3; d++) {
                                xP[d] = (e1 * x1[d] + e2 * x2[d]) * e12_m1;
                                xQ[d] = (e3 * x3[d] + e4 * x4[d]) * e34_m1;
                                rPQ_2 += pow(xP[d] - xQ[d], 2);
                            }
                            double rho = e12 * e34 / (e12 + e34);
                            double d2 = 1.0;
                            if (omega != -1.0) {
                                d2 = omega * omega / (rho + omega * omega);
                                pref1234 *= sqrt(d2);
                            }
                            double T = rho * d2 * rPQ_2;
                            double K13V[225];
                            memset(K13V, 0, sizeof(K13V));
                            double rx[9], rw[9];
                            rys_eval(nrys, T, rx, rw);
                            for (int r = 0; r < nrys; r++) {
                                double t2 = rx[r] * d2;
                                double w = rw[r];
                                double J2[3][(OB_MAXL + 1) * (OB_MAXL + 1) * (OB_MAXL + 1) * (OB_MAXL + 1)];
                                memset(J2, 0, sizeof(J2));
                                double gamma = rho * t2 / (1.0 - t2);
                                double delta_m1 = (e34 - e12) / (2.0 * gamma);
                                double Qval = delta_m1 * pow(1.0 + delta_m1 * delta_m1, -0.5);
                                double c2val = 0.5 + 0.5 * Qval;
                                double s2val = 0.5 - 0.5 * Qval;
                                double csval = 0.5 * pow(1.0 + delta_m1 * delta_m1, -0.5);
                                double cval = sqrt(c2val);
                                double sval = sqrt(s2val);
                                double a1 = c2val * (e12 + gamma) + s2val * (e34 + gamma) - 2.0 * csval * gamma;
                                double a2 = s2val * (e12 + gamma) + c2val * (e34 + gamma) + 2.0 * csval * gamma;
                                double a1_m12 = pow(a1, -0.5);
                                double a2_m12 = pow(a2, -0.5);
                                double X11 = cval * a1_m12;
                                double X12 = -sval * a2_m12;
                                double X21 = sval * a1_m12;
                                double X22 = cval * a2_m12;
                                double xA[3], xB[3];
                                for (int d = 0; d < 3; d++) {
                                    xA[d] = xP[d] + rho * t2 * e12_m1 * (xQ[d] - xP[d]);
                                    xB[d] = xQ[d] + rho * t2 * e34_m1 * (xP[d] - xQ[d]);
                                }
                                for (int v1 = 0; v1 < nrys; v1++)
                                    for (int v2 = 0; v2 < nrys; v2++) {
                                        double zeta1 = X11 * tgh[v1] + X12 * tgh[v2];
                                        double zeta2 = X21 * tgh[v1] + X22 * tgh[v2];
                                        double wgh12 = wgh[v1] * wgh[v2];
                                        for (int d = 0; d < 3; d++) {
                                            double* J2x = J2[d];
                                            double x1l = 1.0;
                                            for (int l1 = 0; l1 < V1; l1++) {
                                                double x2l = 1.0;
                                                for (int l2 = 0; l2 < V2; l2++) {
                                                    double x3l = 1.0;
                                                    for (int l3 = 0; l3 < V3; l3++) {
                                                        double x4l = 1.0;
                                                        for (int l4 = 0; l4 < V4; l4++) {
                                                            (*J2x++) += wgh12 * x1l * x2l * x3l * x4l;
                                                            x4l *= zeta2 + xB[d] - x4[d];
                                                        }
                                                        x3l *= zeta2 + xB[d] - x3[d];
                                                    }
                                                    x2l *= zeta1 + xA[d] - x2[d];
                                                }
                                                x1l *= zeta1 + xA[d] - x1[d];
                                            }
                                        }
                                    }
                                (void)V1234;
                                int idx1 = 0;
                                for (int a1i = 0; a1i <= L1; a1i++) {
                                    int l1 = L1 - a1i;
                                    for (int b1i = 0; b1i <= a1i; b1i++, idx1++) {
                                        int m1 = a1i - b1i, n1 = b1i;
                                        int idx2 = 0;
                                        for (int a2i = 0; a2i <= L2; a2i++) {
                                            int l2 = L2 - a2i;
                                            for (int b2i = 0; b2i <= a2i; b2i++, idx2++) {
                                                int m2 = a2i - b2i, n2 = b2i;
                                                int idx3 = 0;
                                                for (int a3i = 0; a3i <= L3; a3i++) {
                                                    int l3 = L3 - a3i;
                                                    for (int b3i = 0; b3i <= a3i; b3i++, idx3++) {
                                                        int m3 = a3i - b3i, n3 = b3i;
                                                        int idx4 = 0;
                                                        for (int a4i = 0; a4i <= L4; a4i++) {
                                                            int l4 = L4 - a4i;
                                                            for (int b4i = 0; b4i <= a4i; b4i++, idx4++) {
                                                                int m4 = a4i - b4i, n4 = b4i;
                                                                double I = pref1234 * w *
                                                                           J2[0][l1 * V2 * V3 * V4 + l2 * V3 * V4 + l3 * V4 + l4] *
                                                                           J2[1][m1 * V2 * V3 * V4 + m2 * V3 * V4 + m3 * V4 + m4] *
                                                                           J2[2][n1 * V2 * V3 * V4 + n2 * V3 * V4 + n3 * V4 + n4];
                                                                K13V[idx1 * N3 + idx3] += I * D24V[idx2 * N4 + idx4];
                                                            }
                                                        }
                                                    }
                                                }
                                            }
                                        }
                                    }
                                }
                            }
                            int o1 = b1->cart[P1], o3 = b3->cart[P3];
                            for (int p1 = 0; p1 < N1; p1++)
                                for (int p3 = 0; p3 < N3; p3++) {
#pragma omp atomic
                                    K13[(size_t)(p1 + o1) * nc3 + (p3 + o3)] += K13V[p1 * N3 + p3];
                                    if (flip24) {
#pragma omp atomic
                                        K13[(size_t)(p3 + o3) * nc3 + (p1 + o1)] += K13V[p1 * N3 + p3];
                                    }
                                }
                        }
                        if (!did_work) break;
                    }
                }
                if (counts) counts[((L1 * 5 + L2) * 5 + L3) * 5 + L4] += nq;
            }
        }
    }
    free_kblocks(kb12, nkb12);
    if (!same_list) free_kblocks(kb34, nkb34);
    return 0;
}

/* ---------------------------------------------------------------- quartet magnitudes */

/* All Cartesian components (ab|O|cd) of one primitive quartet, I[((a*N2 + b)*N3 + c)*N4 + d]: the
 * per-quartet evaluation of IntBox::exchangeCPU (src/cpp/intbox/cpu/exchange.cpp:191-324: the
 * prefactor, the Rys roots, the rotated nrys x nrys Gauss-Hermite 2-D integrals and their
 * assembly), without the contraction with D24. */
static void quartet_tensor(const ob_basis* b1, int P1, const ob_basis* b2, int P2, const ob_basis* b3, int P3,
                           const ob_basis* b4, int P4, int nterm, const double* alphas, const double* omegas,
                           double* I)
{
    int L1 = b1->L[P1], L2 = b2->L[P2], L3 = b3->L[P3], L4 = b4->L[P4];
    int V1 = L1 + 1, V2 = L2 + 1, V3 = L3 + 1, V4 = L4 + 1;
    int N1 = ncart_of(L1), N2 = ncart_of(L2), N3 = ncart_of(L3), N4 = ncart_of(L4);
    int nrys = (L1 + L2 + L3 + L4) / 2 + 1;
    const double* tgh = lsjk_gh_x + nrys * (nrys - 1) / 2;
    const double* wgh = lsjk_gh_w + nrys * (nrys - 1) / 2;
    memset(I, 0, sizeof(double) * N1 * N2 * N3 * N4);
    double e1 = b1->e[P1], e2 = b2->e[P2], e3 = b3->e[P3], e4 = b4->e[P4];
    const double *x1 = b1->xyz + 3 * P1, *x2 = b2->xyz + 3 * P2, *x3 = b3->xyz + 3 * P3, *x4 = b4->xyz + 3 * P4;
    double r12 = 0.0, r34 = 0.0;
    for (int d = 0; d < 3; d++) {
        r12 += pow(x1[d] - x2[d], 2);
        r34 += pow(x3[d] - x4[d], 2);
    }
    double K12 = b1->c[P1] * b2->c[P2] * exp(-e1 * e2 / (e1 + e2) * r12);
    double K34 = b3->c[P3] * b4->c[P4] * exp(-e3 * e4 / (e3 + e4) * r34);
    double e12 = e1 + e2, e34 = e3 + e4, e12_m1 = 1.0 / e12, e34_m1 = 1.0 / e34;
    double xP[3], xQ[3], rPQ_2 = 0.0;
    for (int d = 0; d < 3; d++) {
        xP[d] = (e1 * x1[d] + e2 * x2[d]) * e12_m1;
        xQ[d] = (e3 * x3[d] + e4 * x4[d]) * e34_m1;
        rPQ_2 += pow(xP[d] - xQ[d], 2);
    }
    double rho = e12 * e34 / (e12 + e34);
    for (int it = 0; it < nterm; it++) {
        double alpha = alphas[it], omega = omegas[it];
        if (alpha == 0.0 || omega == 0.0) continue;
        double pref = alpha * K12 * K34 * 2.0 * pow(M_PI, -0.5) * e12_m1 * e34_m1 * pow(e12 + e34, -0.5);
        double d2 = 1.0;
        if (omega != -1.0) {
            d2 = omega * omega / (rho + omega * omega);
            pref *= sqrt(d2);
        }
        double rx[9], rw[9];
        rys_eval(nrys, rho * d2 * rPQ_2, rx, rw);
        for (int r = 0; r < nrys; r++) {
            double t2 = rx[r] * d2, w = rw[r];
            double J2[3][(OB_MAXL + 1) * (OB_MAXL + 1) * (OB_MAXL + 1) * (OB_MAXL + 1)];
            memset(J2, 0, sizeof(J2));
            double gamma = rho * t2 / (1.0 - t2);
            double delta_m1 = (e34 - e12) / (2.0 * gamma);
            double Qv = delta_m1 * pow(1.0 + delta_m1 * delta_m1, -0.5);
            double c2 = 0.5 + 0.5 * Qv, s2 = 0.5 - 0.5 * Qv, cs = 0.5 * pow(1.0 + delta_m1 * delta_m1, -0.5);
            double cv = sqrt(c2), sv = sqrt(s2);
            double a1 = c2 * (e12 + gamma) + s2 * (e34 + gamma) - 2.0 * cs * gamma;
            double a2 = s2 * (e12 + gamma) + c2 * (e34 + gamma) + 2.0 * cs * gamma;
            double X11 = cv * pow(a1, -0.5), X12 = -sv * pow(a2, -0.5);
            double X21 = sv * pow(a1, -0.5), X22 = cv * pow(a2, -0.5);
            double xA[3], xB[3];
            for (int d = 0; d < 3; d++) {
                xA[d] = xP[d] + rho * t2 * e12_m1 * (xQ[d] - xP[d]);
                xB[d] = xQ[d] + rho * t2 * e34_m1 * (xP[d] - xQ[d]);
            }
            for (int v1 = 0; v1 < nrys; v1++)
                for (int v2 = 0; v2 < nrys; v2++) {
                    double z1 = X11 * tgh[v1] + X12 * tgh[v2];
                    double z2 = X21 * tgh[v1] + X22 * tgh[v2];
                    double wv = wgh[v1] * wgh[v2];
                    for (int d = 0; d < 3; d++) {
                        double* out = J2[d];
                        double p1 = 1.0;
                        for (int l1 = 0; l1 < V1; l1++) {
                            double p2 = 1.0;
                            for (int l2 = 0; l2 < V2; l2++) {
                                double p3 = 1.0;
                                for (int l3 = 0; l3 < V3; l3++) {
                                    double p4 = 1.0;
                                    for (int l4 = 0; l4 < V4; l4++) {
                                        (*out++) += wv * p1 * p2 * p3 * p4;
                                        p4 *= z2 + xB[d] - x4[d];
                                    }
                                    p3 *= z2 + xB[d] - x3[d];
                                }
                                p2 *= z1 + xA[d] - x2[d];
                            }
                            p1 *= z1 + xA[d] - x1[d];
                        }
                    }
                }
            /* Cartesian component (a, b, c, d) of each centre: exponents (lx, ly, lz) in lexical order */
            int lx[4][15], ly[4][15], lz[4][15];
            const int Ls[4] = {L1, L2, L3, L4};
            for (int k = 0; k < 4; k++) {
                int idx = 0;
                for (int ai = 0; ai <= Ls[k]; ai++)
                    for (int bi = 0; bi <= ai; bi++, idx++) {
                        lx[k][idx] = Ls[k] - ai;
                        ly[k][idx] = ai - bi;
                        lz[k][idx] = bi;
                    }
            }
            for (int a = 0; a < N1; a++)
                for (int b = 0; b < N2; b++)
                    for (int c = 0; c < N3; c++)
                        for (int d = 0; d < N4; d++) {
                            int ix = ((lx[0][a] * V2 + lx[1][b]) * V3 + lx[2][c]) * V4 + lx[3][d];
                            int iy = ((ly[0][a] * V2 + ly[1][b]) * V3 + ly[2][c]) * V4 + ly[3][d];
                            int iz = ((lz[0][a] * V2 + lz[1][b]) * V3 + lz[2][c]) * V4 + lz[3][d];
                            I[((a * N2 + b) * N3 + c) * N4 + d] += pref * w * J2[0][ix] * J2[1][iy] * J2[2][iz];
                        }
        }
    }
}

/* Largest |component| of each of n primitive quartets (P1[k] P2[k] | P3[k] P4[k]): the magnitude
 * the learned screen is trained to bound (data generation checks the GPU against this). */
int oracle_quartet_max(const ob_basis* b1, const ob_basis* b2, const ob_basis* b3, const ob_basis* b4, int64_t n,
                       const int* P1, const int* P2, const int* P3, const int* P4, int nterm, const double* alphas,
                       const double* omegas, double* out)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < n; k++) {
        double I[1296];
        quartet_tensor(b1, P1[k], b2, P2[k], b3, P3[k], b4, P4[k], nterm, alphas, omegas, I);
        int N = ncart_of(b1->L[P1[k]]) * ncart_of(b2->L[P2[k]]) * ncart_of(b3->L[P3[k]]) * ncart_of(b4->L[P4[k]]);
        double m = 0.0;
        for (int i = 0; i < N; i++) m = fmax(m, fabs(I[i]));
        out[k] = m;
    }
    return 0;
}

/* The full component tensors of n quartets, out[k * 1296 + ...] in the quartet_tensor layout
 * (pinning of quartet_tensor against the oracle's J). */
int oracle_quartet_tensor(const ob_basis* b1, const ob_basis* b2, const ob_basis* b3, const ob_basis* b4, int64_t n,
                          const int* P1, const int* P2, const int* P3, const int* P4, int nterm,
                          const double* alphas, const double* omegas, double* out)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t k = 0; k < n; k++)
        quartet_tensor(b1, P1[k], b2, P2[k], b3, P3[k], b4, P4[k], nterm, alphas, omegas, out + k * 1296);
    return 0;
}

/* ---------------------------------------------------------------- one-electron (S, T, V) */

/* IntBox::overlapCPU and IntBox::kineticCPU (src/cpp/intbox/cpu/overlap.cpp:14-147,
 * src/cpp/intbox/cpu/kinetic.cpp:14-175), Cartesian representation, accumulated.  1-D integrals by
 * the Gauss-Hermite rule with (L1+L2)/2+1 (S) and (L1+L2+2)/2+1 (T) nodes in x = t/sqrt(p) about
 * the pair centre; T = 1/2 sum_axis <d a|d b> with d x^l = l x^(l-1) - 2 e x^(l+1).  A symmetric
 * list adds the transposed block for pairs of distinct primitives.  S or T may be NULL. */
int oracle_overlap_kinetic(const ob_basis* b1, const ob_basis* b2, const ob_pairlist* pl, double* S, double* T)
{
    size_t nc2 = b2->ncart;
    for (int cls = 0; cls < pl->ncls; cls++) {
        int L1 = pl->cls_L1[cls], L2 = pl->cls_L2[cls];
        int V1 = L1 + 1, V2 = L2 + 1;
        int ngs = (L1 + L2) / 2 + 1, ngt = (L1 + L2 + 2) / 2 + 1;
        int s0 = pl->cls_start[cls], s1 = pl->cls_start[cls + 1];
#pragma omp parallel for schedule(dynamic)
        for (int k = s0; k < s1; k++) {
            int i1 = pl->p1[k], i2 = pl->p2[k];
            const double* r1 = b1->xyz + 3 * i1;
            const double* r2 = b2->xyz + 3 * i2;
            double e1 = b1->e[i1], e2 = b2->e[i2], p = e1 + e2, pm12 = pow(p, -0.5);
            double R2 = 0.0, px1[3], px2[3];
            for (int d = 0; d < 3; d++) {
                double P = (e1 * r1[d] + e2 * r2[d]) / p;
                px1[d] = P - r1[d];
                px2[d] = P - r2[d];
                R2 += (r1[d] - r2[d]) * (r1[d] - r2[d]);
            }
            double K12 = b1->c[i1] * b2->c[i2] * exp(-e1 * e2 / p * R2);
            double S1[3][(OB_MAXL + 1) * (OB_MAXL + 1)], T1[3][(OB_MAXL + 1) * (OB_MAXL + 1)];
            memset(S1, 0, sizeof(S1));
            memset(T1, 0, sizeof(T1));
            for (int pass = 0; pass < 2; pass++) {
                int ng = pass ? ngt : ngs;
                const double* tg = lsjk_gh_x + ng * (ng - 1) / 2;
                const double* wg = lsjk_gh_w + ng * (ng - 1) / 2;
                for (int v = 0; v < ng; v++) {
                    double xe = tg[v] * pm12, we = wg[v] * pm12;
                    for (int d = 0; d < 3; d++) {
                        double x1 = xe + px1[d], x2 = xe + px2[d];
                        for (int i = 0; i < V1; i++)
                            for (int j = 0; j < V2; j++) {
                                if (pass == 0) {
                                    S1[d][i * V2 + j] += we * pow(x1, i) * pow(x2, j);
                                } else {
                                    double d1 = (i > 0 ? i * pow(x1, i - 1) : 0.0) - 2.0 * e1 * pow(x1, i + 1);
                                    double d2 = (j > 0 ? j * pow(x2, j - 1) : 0.0) - 2.0 * e2 * pow(x2, j + 1);
                                    T1[d][i * V2 + j] += we * d1 * d2;
                                }
                            }
                    }
                }
            }
            int off1 = b1->cart[i1], off2 = b2->cart[i2];
            int c1 = 0;
            for (int a = 0; a <= L1; a++)
                for (int b = 0; b <= a; b++, c1++) {
                    int l1 = L1 - a, m1 = a - b, n1 = b;
                    int c2 = 0;
                    for (int a2 = 0; a2 <= L2; a2++)
                        for (int b2 = 0; b2 <= a2; b2++, c2++) {
                            int l2 = L2 - a2, m2 = a2 - b2, n2 = b2;
                            double sx = S1[0][l1 * V2 + l2], sy = S1[1][m1 * V2 + m2], sz = S1[2][n1 * V2 + n2];
                            double sv = K12 * sx * sy * sz;
                            double tv = K12 * 0.5 *
                                        (T1[0][l1 * V2 + l2] * sy * sz + sx * T1[1][m1 * V2 + m2] * sz +
                                         sx * sy * T1[2][n1 * V2 + n2]);
                            size_t ij = (size_t)(c1 + off1) * nc2 + (c2 + off2);
                            size_t ji = (size_t)(c2 + off2) * nc2 + (c1 + off1);
                            int mirror = pl->symmetric && i1 != i2;
                            if (S) {
#pragma omp atomic
                                S[ij] += sv;
                                if (mirror) {
#pragma omp atomic
                                    S[ji] += sv;
                                }
                            }
                            if (T) {
#pragma omp atomic
                                T[ij] += tv;
                                if (mirror) {
#pragma omp atomic
                                    T[ji] += tv;
                                }
                            }
                        }
                }
        }
    }
    return 0;
}

/* IntBox::potentialCPU (src/cpp/intbox/cpu/potential.cpp:14-147): V += sum_A int a b O(r_1A) (-c_A)
 * for point charges xyzc (nA x 4: x, y, z, c), Cartesian, accumulated.  Per pair the Hermite
 * integrals R_t(p d^2, P - A) by Rys quadrature with pref = alpha 2 pi / p (x sqrt(d^2) for erf
 * operators, d^2 = omega^2/(p + omega^2)), then the bra Hermite -> Cartesian transform. */
int oracle_potential(const ob_basis* b1, const ob_basis* b2, const ob_pairlist* pl, int nterm, const double* scales,
                     const double* omegas, int nA, const double* xyzc, double* V)
{
    for (int cls = 0; cls < pl->ncls; cls++) {
        int L = pl->cls_L1[cls] + pl->cls_L2[cls];
        int H = hermite_count(L), nrys = L / 2 + 1;
        int tx[35], ty[35], tz[35];
        hermite_tuples(L, tx, ty, tz);
        int s0 = pl->cls_start[cls], s1 = pl->cls_start[cls + 1];
        double* dh = (double*)calloc((size_t)(pl->npair > 0 ? pl->npair : 1) * H, sizeof(double));
#pragma omp parallel for schedule(dynamic)
        for (int k = s0; k < s1; k++) {
            int i1 = pl->p1[k], i2 = pl->p2[k];
            const double* r1 = b1->xyz + 3 * i1;
            const double* r2 = b2->xyz + 3 * i2;
            double e1 = b1->e[i1], e2 = b2->e[i2], p = e1 + e2;
            double P[3];
            for (int d = 0; d < 3; d++) P[d] = (e1 * r1[d] + e2 * r2[d]) / p;
            double* acc = dh + (size_t)k * H;
            for (int it = 0; it < nterm; it++) {
                double alpha = scales[it], omega = omegas[it];
                if (alpha == 0.0 || omega == 0.0) continue;
                double pref = alpha * 2.0 * M_PI / p, d2 = 1.0;
                if (omega != -1.0) {
                    d2 = omega * omega / (p + omega * omega);
                    pref *= sqrt(d2);
                }
                for (int A = 0; A < nA; A++) {
                    double X = P[0] - xyzc[4 * A], Y = P[1] - xyzc[4 * A + 1], Z = P[2] - xyzc[4 * A + 2];
                    double cA = xyzc[4 * A + 3];
                    double Tv = p * d2 * (X * X + Y * Y + Z * Z);
                    double xr[9], wr[9];
                    rys_eval(nrys, Tv, xr, wr);
                    for (int i = 0; i < nrys; i++) {
                        double U = -2.0 * p * d2 * xr[i];
                        double Hx[2 * OB_MAXL + 1], Hy[2 * OB_MAXL + 1], Hz[2 * OB_MAXL + 1];
                        Hx[0] = -pref * wr[i] * cA;
                        Hy[0] = 1.0;
                        Hz[0] = 1.0;
                        if (L > 0) {
                            Hx[1] = U * X * Hx[0];
                            Hy[1] = U * Y;
                            Hz[1] = U * Z;
                        }
                        for (int u = 1; u < L; u++) {
                            Hx[u + 1] = U * (X * Hx[u] + u * Hx[u - 1]);
                            Hy[u + 1] = U * (Y * Hy[u] + u * Hy[u - 1]);
                            Hz[u + 1] = U * (Z * Hz[u] + u * Hz[u - 1]);
                        }
                        for (int t = 0; t < H; t++) acc[t] += Hx[tx[t]] * Hy[ty[t]] * Hz[tz[t]];
                    }
                }
            }
        }
        hermite_transform(pl, cls, b1, b2, dh, V, -1, 0);
        free(dh);
    }
    return 0;
}

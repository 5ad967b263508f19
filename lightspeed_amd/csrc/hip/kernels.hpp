// Kernel argument blocks and launcher declarations (library-internal).
#pragma once
#include "internal.hpp"

namespace lsjk {

// Index of the quartet counters in CallScalars::counts.
__host__ __device__ constexpr int jcount_index(int L12, int L34) { return L12 * 5 + L34; }
__host__ __device__ constexpr int kcount_index(int L1, int L2, int L3, int L4)
{
    return 256 + ((L1 * 5 + L2) * 5 + L3) * 5 + L4;
}

// Precision routes of a quartet kernel launch (the reference's documented mixed precision,
// intbox.hpp:736-742: FP64 when est >= thredp, FP32 when thresp <= est < thredp, else skipped).
constexpr int kRoutePlain = 0;  // est >= thresp, FP64 (thredp <= thresp)
constexpr int kRouteHi = 1;     // est >= thredp, FP64
constexpr int kRouteLo = 2;     // thresp <= est < thredp, FP32 arithmetic, FP64 accumulation
constexpr int kRouteGrad = 3;   // exchange gradient: est >= thresp, FP64 derivative integrals
constexpr int kRouteMixed = 4;  // K build: est >= thredp in FP64 and thresp <= est < thredp in FP32, one launch

struct JPrepArgs {
    const int* p1;
    const int* p2;
    const float* bound;
    const int* hpos;
    const double* K;
    const double *eA, *xA, *yA, *zA;
    const int* cartA;
    const double *eB, *xB, *yB, *zB;
    const int* cartB;
    const double* D;
    int ldD;
    int symmetric;
    float thre_ket;
    float* kb;
    double* kd;
    float* kd32;  // FP32 copy of the densities for the FP32 route (nullptr: not needed)
    CallScalars* sc;
};

struct JQuartetArgs {
    RysTable rys;
    Ewald ew;
    float thresp;
    int count;
    int count_cls;
    // bra (Hermite order of the bra list)
    const float* bra_b;
    const double* bra_geom;  // SoA p, Px, Py, Pz, sqrt(2 pi^(5/2))/p, stride bra_n
    int bra_n;
    int bra_hstart;
    double* jp;  // accumulators of the bra group
    // ket (Hermite order of the ket list)
    const float* ket_b;         // truncated bounds (per call)
    const float* ket_static_b;  // Schwarz bounds
    const double* ket_geom;
    int ket_n;
    int ket_hstart;
    const double* ket_d;  // Hermite densities of the ket group
    const float* ket_d32; // their FP32 copy (FP32 route)
    float thredp;
    CallScalars* sc;
    // symmetric J (one list on both sides): the bra group's Hermite densities and the ket group's
    // accumulators, for the second use of each integral (the ket pair as bra)
    const double* bra_d;
    double* jq;
    int count_cls2;
};

struct JScatterArgs {
    const int* p1;
    const int* p2;
    const float* bound;
    const int* hpos;
    const double* K;
    const double *eA, *xA, *yA, *zA;
    const int* cartA;
    const double *eB, *xB, *yB, *zB;
    const int* cartB;
    const double* jp;
    double* J;
    int ldJ;
    int symmetric;
    int screen;  // 1: only bra pairs that survive truncate_bra (J); 0: every pair (potential)
    float thresp;
    CallScalars* sc;
};

// Coulomb gradient, bra-centre part: per bra pair, dE/dA and dE/dB from the potential V_t of the
// ket density at bra Hermite order L12 + 1 and the symmetrised bra density block.
struct JGradArgs {
    const int* p1;
    const int* p2;
    const float* bound;
    const int* hpos;
    const double* K;
    const double *eA, *xA, *yA, *zA;
    const int *cartA, *atomA;
    const double *eB, *xB, *yB, *zB;
    const int *cartB, *atomB;
    const double* vp;  // V_t, Hermite order, H(L12+1) per pair of group L12
    const double* D;   // bra density (Cartesian)
    int ldD;
    int symmetric;
    float thresp;
    double* G;  // natom x 3
    CallScalars* sc;
};
void launch_j_grad_scatter(int LA, int LB, const JGradArgs& a, int k0, int nk, int hstart, size_t hoff,
                           hipStream_t s);

// One-electron integrals (overlap, kinetic, potential) of one pair list.
struct OneEArgs {
    const int* p1;
    const int* p2;
    const int* hpos;
    const double* K;
    const double *eA, *xA, *yA, *zA;
    const int* cartA;
    const double *eB, *xB, *yB, *zB;
    const int* cartB;
    int symmetric;
    double* S;  // ncart1 x ncart2 outputs (nullptr: not wanted)
    double* T;
    int ld;
    // potential: point charges and operator
    RysTable rys;
    Ewald ew;
    const double* xyzc;  // nA x 4
    int nA;
    double* vp;  // Hermite accumulators (Hermite order, same layout as the J bra buffer)
};
void launch_one_e_st(int LA, int LB, const OneEArgs& a, int k0, int nk, hipStream_t s);
void launch_one_e_v(int LA, int LB, const OneEArgs& a, int k0, int nk, int hstart, size_t hoff, hipStream_t s);

// K gradient: one wave per tile = (row1, kKTile consecutive row3s) of fixed angular momenta.
constexpr int kKTile = 16;
// Row-1 partners per gradient work unit: a tile's partner range is split into chunks of this size
// so that no unit is much heavier than the rest.
constexpr int kKChunk2 = 64;
// A K gradient work unit: row1 = rows1[i1], row3 block ib (rows3[ib*kKTile ...]), row1's partners
// [e2off, e2off + e2cnt) of the class's partner angular momentum.
struct KUnit {
    int i1, ib, e2off, e2cnt;
};

// K build: ket groups.  The rows of one (row L, partner L) class of the ket list are put in
// spatial (Morton) order and cut into groups of kgroup_size(L, Lp) consecutive rows; the partner
// entries of a group's rows are merged, sorted by bound (descending) and padded to a multiple of
// 64 slots.  A wave holds 64 slots (one ket pair per lane) in registers and streams the bra row's
// partners through scalar loads.  Rows with few partners per row (p or d partners) leave the
// bounds of a 16-row group's chunks spread wide, so that few lanes pass the sieve together;
// larger groups narrow each chunk's bound range: 64 rows for s rows with p or d partners, 32 for
// p rows with d partners (measured up to 42 % faster), 16 otherwise (larger groups measured
// slower there: longer units, and fewer ket rows of the bra row's triangle per group).
constexpr int kgroup_size(int L, int Lp) { return (L == 0 && Lp > 0) ? 64 : ((L == 1 && Lp == 2) ? 32 : 16); }
// These are the defaults: lsjk_autotune picks 16, 32 or 64 rows per (row L, partner L) of a list
// (KGroups::gsize), where every class over that ket pair fits the LDS limit (kket_lds_bytes).
constexpr int kKWaves = 4;  // waves (independent work units) per workgroup
struct KGroups {
    int gsize;             // rows per group
    const int2* slot;      // [nslot] {entry index or -1 (padding), position of the row in the group}
    const int* grp_slot;   // [ngroup + 1] first slot of each group
    const int* grp_rows;   // [ngroup * group size] row primitive at each position (-1: none)
    const int* grp_rank;   // [ngroup * group size] spatial rank of that row (INT_MAX: none)
    const float* chunk_b;  // [nslot / 64] largest bound of each 64-slot chunk
};
// A K work unit: bra row rows1[i1] against the 64-slot chunks [c0, c0 + nc) of ket group g
// (crange = c0 | nc << 16, chunks counted from the group's first slot).  rank1: in the triangle
// mode (one list, symmetric D24) only ket rows of spatial rank >= rank1 (the bra row's) take
// part; -1 otherwise.  prange = p0 | np << 16: the unit walks the row's partners from p0 on, np
// of them (np = 0: to the end); 0 = the whole walk.
struct KUnitG {
    int i1, g, rank1, crange, prange;
};
// A packed partner entry (row primitive A, partner primitive B), 128 bytes: p, Kp, P (3),
// P - A (3), A - B (3), {the partner's Cartesian offset, its block id} (two ints), 1/(2p), the
// same two ids in the partner basis's K view order (Basis::h_kcart, h_kblk), 2 zero pads.
// Kp = K_AB sqrt(2 pi^(5/2)) / p with K_AB = c_A c_B exp(-a b / p |A - B|^2).
constexpr int kKEntDoubles = 16;
// sqrt(2 pi^(5/2))
constexpr double kKPrefRoot = 5.914967172795613;
// The FP32-route copy, 80 bytes: P (3 doubles), p, Kp, P - A, A - B, 1/(2p), 0 (10 floats),
// {cart, blk}, {cart, blk} in the K view order.
constexpr int kKEnt32Doubles = 10;
struct KSide {
    const int* row_start;  // [nrow * (maxLp+1) + 1]
    int maxLp;
    const float* eb;       // entry bound
    const int* ecart;      // partner Cartesian offset
    const int* eblk;       // partner block id
    const double* eg;      // packed entries, kKEntDoubles doubles each (see KEnt)
    const double* eg32;    // FP32-route entries, kKEnt32Doubles doubles each
    int nent;
    const int* rows;       // rows of this class with partners of the class's partner L (primitive indices)
    int nrows;
    const double *x, *y, *z;  // row basis coordinates
    const int* cart;          // row basis Cartesian offsets
};

struct KArgs {
    RysTable rys;
    Ewald ew;
    float thresp;
    float thredp;
    int count;
    int symmetric;  // compute only row1 <= row3 and add the transposed block
    int mirror_diag;  // symmetric mode, a class evaluated for its mirror class as well: the
                      // transposed block is added on the diagonal (row1 == row3) too
    int tri;        // units are the (row1 <= row3) pairs given by tri_pre / tri_start
    const long long* tri_pre;
    const int* tri_start;
    KSide s12, s34;
    const double* D;     // D24, ncart2 x ncart4
    int ldD;
    const double* Dsh;   // block maxima, nblk2 x nblk4
    int ldDsh;
    double* K;           // K13, ncart1 x ncart3
    int ldK;
    float maxb12, maxb34;  // max_bound of the two lists
    // learned route (nullptr: none): per unit the certified factor tau in (0, 1] of the Schwarz
    // estimate, applied while thresp >= tau_smin max|D24|
    const float* tau;
    float tau_smin;
    CallScalars* sc;
    // exchange gradient (kRouteGrad): G (natom x 3) += d/dR sum D13_ac D24_bd (ab|cd)
    const double* D13;
    int ldD13;
    double* G;
    const double *exp1, *exp2, *exp3, *exp4;  // primitive exponents of the four bases
    const int *atom1, *atom2, *atom3, *atom4; // primitive atoms
    const int *eprim12, *eprim34;             // partner primitive of each entry
};

struct CountJArgs {
    float thresp;
    const float* bra_bound;   // bra list bounds (list order)
    const float* ket_b;       // truncated ket bounds (Hermite order)
    const int* ket_hpos;      // list order -> Hermite order
    const int* ket_cls_start; // ket classes (list order)
    const int* ket_cls_L1;
    const int* ket_cls_L2;
    int ket_ncls;
    unsigned long long* out;
    CallScalars* sc;
};

struct CountKArgs {
    float thresp;
    int half;                 // D24symm half-sieve (same list): pivots i2 <= i4 only
    int n2, n4;               // pivot primitives of the two lists
    const int* rs12;          // pivot-view segment starts [n2 * (maxLa+1) + 1]
    const float* b12;         // pivot-view partner bounds
    int maxLa;
    const int* rs34;
    const float* b34;
    int maxLc;
    const int *L2, *L4, *blk2, *blk4;  // pivot angular momenta and block ids
    const double* Dsh;
    int ldDsh;
    float maxb12, maxb34;
    unsigned long long* out;
    CallScalars* sc;
};

void launch_count_j(const CountJArgs& a, int k0, int nk, int L1, int L2, hipStream_t s);
void launch_count_k(const CountKArgs& a, hipStream_t s);

void launch_j_ket_prep(int LA, int LB, const JPrepArgs& a, int k0, int nk, int hstart, size_t hoff, hipStream_t s);
void launch_j_bra_scatter(int LA, int LB, const JScatterArgs& a, int k0, int nk, int hstart, size_t hoff,
                          hipStream_t s);
// Group pairs (L12 >= L34) of the symmetric J kernel: with one list on both sides it evaluates each
// bra-ket pair of the two groups once for both J blocks (the triangle within one group).  Used
// where measured faster than the two one-way launches: s-type kets (one Hermite component to
// reduce over the wave per ket); with p or d kets the per-ket reduction costs what it saves
// (measured again with the DPP reduction: (2, 1) 1.21 ms against 0.60 + 0.55 ms one-way).
__host__ __device__ constexpr bool js_enabled(int L12, int L34) { return L12 >= L34 && L34 == 0; }
void launch_j_quartets_sym(int L12, int L34, int route, const JQuartetArgs& a, const JTask* tasks, int ntask,
                           unsigned int* qctr, hipStream_t s);
void launch_j_quartets(int L12, int L34, int route, const JQuartetArgs& a, const JTask* tasks, int ntask,
                       unsigned int* qctr, hipStream_t s);
void launch_j_quartets_grad(int L12p1, int L34, const JQuartetArgs& a, const JTask* tasks, int ntask,
                            unsigned int* qctr, hipStream_t s);

// Work-queue counters of one call (CallScalars::queue, zeroed with the call's scalars): J classes
// at jcount_index.

// Grid of a persistent kernel: as many workgroups as fit on the device at once (occupancy at
// this block size and LDS), but no more than the work needs.  Raises if the LDS request exceeds
// the device's per-workgroup limit.
int persistent_grid(const void* kernel, int block, size_t lds, long long nwork_groups);

// blk_lo / blk_hi (nullptr: dense): per block 2, the range of blocks 4 of its molecule (batch
// basis); band = the largest such range.
// Dynamic LDS of one k_ket workgroup (as KKetLds lays it out): the staged Rys table blocks (FP64
// for routes with FP64 quartets and nrys <= 3, FP32 for routes with FP32 quartets and nrys <= 5),
// 8-byte aligned, then kKWaves accumulators of gsize x N1 x N3 doubles.
inline size_t kket_lds_bytes(const RysTable& tab, int L1, int L2, int L3, int L4, int route, int gsize)
{
    const int NR = (L1 + L2 + L3 + L4) / 2 + 1;
    const bool has32 = route == kRouteLo || route == kRouteMixed, has64 = route != kRouteLo;
    size_t b = 0;
    if (has64 && rys_staged(NR)) b += rys_lds_doubles(tab.nbin[NR], NR) * sizeof(double);
    if (has32 && rys32_staged(NR)) b += rys32_lds_floats(tab.nbin[NR], NR) * sizeof(float);
    b = (b + 7) & ~(size_t)7;
    return b + (size_t)kKWaves * gsize * ncart(L1) * ncart(L3) * sizeof(double);
}

// out[pr[i] * ldo + pc[j]] = in[i * ldi + j] (the K view copies of D and of its block maxima).
void launch_permute(const double* in, int ldi, int nr, int nc, const int* pr, const int* pc, double* out, int ldo,
                    hipStream_t s);
void launch_k_blockmax(const double* D, int ldD, const int* blk_cart2, const int* blk_L2, int nblk2,
                       const int* blk_cart4, const int* blk_L4, int nblk4, double* Dsh, int ldDsh, const int* blk_lo,
                       const int* blk_hi, int band, CallScalars* sc, hipStream_t s);
// Exchange gradient classes (route kRouteGrad) over KUnit tiles.
void launch_k_quartets(int L1, int L2, int L3, int L4, int route, const KArgs& a, const KUnit* units, int nunits,
                       hipStream_t s);
// Learned-screen training data: per K unit, the largest M_q / (b12 b34) over its quartets with
// b12 b34 >= smin (classes with s and p functions only).
void launch_k_unit_mag_sp(int L1, int L2, int L3, int L4, const KArgs& a, const KGroups& kg, const KUnitG* units,
                          int nunits, float smin, float* out, hipStream_t s);
// K build classes (routes plain, FP64 part, FP32 part) over ket-group units.
void launch_k_ket(int L1, int L2, int L3, int L4, int route, const KArgs& a, const KGroups& g, const KUnitG* units,
                  int nunits, hipStream_t s);
void launch_fill(void* p, size_t bytes, hipStream_t s);

// Boys functions F_0..F_nboys (nboys <= kBoysMax) of npoint arguments T, F row-major (npoint x
// (nboys + 1)), FP64 (boys.hip).
constexpr int kBoysMax = 16;
void launch_boys(int nboys, int npoint, const double* T, double* F, hipStream_t s);

// The learned screen's network (mlp.hip): ni <= kMlpMaxIn inputs, 3 hidden layers of kMlpHidden
// ReLU units, one linear output, FP32; kMlpTile samples per workgroup.
constexpr int kMlpHidden = 64;
constexpr int kMlpMaxIn = 32;
constexpr int kMlpTile = 64;
int mlp_nparam(int ni);
void launch_mlp_forward(const float* params, int ni, int n, const float* X, float* out, hipStream_t s);
// grad += d/dparams of (1/nb) sum_s sw_s r_s^2 (times wu where r_s < 0), r_s = f(X[idx[s]]) - y[idx[s]]
// (idx null: s itself; sw null: 1); loss[0] += that loss.
void launch_mlp_grad(const float* params, int ni, int nb, const int* idx, const float* X, const float* y,
                     const float* sw, float wu, float* grad, float* loss, hipStream_t s);
void launch_mlp_adam(int np, float* params, const float* grad, float* m, float* v, float lr, float b1, float b2,
                     float eps, int t, hipStream_t s);
void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);

}  // namespace lsjk

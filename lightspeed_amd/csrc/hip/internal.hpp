// Library-internal data structures of liblsjk (device-resident basis / pair-list data and the
// per-call kernel argument blocks).  Not part of the C-ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "device_common.hpp"

namespace lsjk {

// ---------------------------------------------------------------- basis
// Primitive-level SoA copy of a basis.  "Blocks" are the distinct Cartesian function blocks
// (one per contracted shell): primitives of one shell share a block.
struct Basis {
    int nprim = 0, ncart = 0, nblk = 0, maxL = 0;
    std::vector<double> h_c, h_e, h_xyz;
    std::vector<int> h_L, h_cart, h_blk, h_blk_cart, h_blk_L;
    double *d_c = nullptr, *d_e = nullptr, *d_x = nullptr, *d_y = nullptr, *d_z = nullptr;
    int *d_L = nullptr, *d_cart = nullptr, *d_blk = nullptr, *d_blk_cart = nullptr, *d_blk_L = nullptr;
    int natom = 0;            // atoms of the basis (lsjk_basis_set_atoms; 0 = not set)
    int* d_atom = nullptr;    // atom index of each primitive
    // Batch basis (independent molecules, lsjk_batch_create): matrices over it hold only each
    // molecule's diagonal block, element (i, j) of the global Cartesian (or block) numbering at
    // i * ld + j (or i * ldblk + j for block maxima); ld >= the largest molecule, so the blocks do
    // not overlap.  0: plain ncart x ncart (nblk x nblk) matrices.
    int ld = 0, ldblk = 0;
    std::vector<int> h_mol;           // molecule of each primitive (batch basis)
    std::vector<int> h_blk_lo, h_blk_hi;  // per block: its molecule's block range (batch basis)
    int *d_blk_lo = nullptr, *d_blk_hi = nullptr;
    // K view order: the exchange kernel reads D and its block maxima from copies whose blocks are
    // ordered by angular momentum, then spatially (Morton order of the shell centres), so that the
    // partners of a group of nearby rows are neighbouring columns.  h_kcart[i] = position of
    // Cartesian function i, h_kblk[k] = position of block k; kperm = false: the identity (batch
    // bases, whose matrices are banded per molecule).
    bool kperm = false;
    std::vector<int> h_kcart, h_kblk;
    int *d_kcart = nullptr, *d_kblk = nullptr;
};

// Row stride of the Cartesian matrices and of the block-maximum matrices over a basis.
inline int mat_ld(const Basis* b) { return b->ld ? b->ld : b->ncart; }
inline int blk_ld(const Basis* b) { return b->ldblk ? b->ldblk : b->nblk; }

// ---------------------------------------------------------------- pair list
// A J task list on the device: tasks of group pair (L12, L34) at [off, off + cnt).
struct JTaskList {
    int* d_tasks = nullptr;
    int ntask = 0;
    int off[5][5] = {{0}};
    int cnt[5][5] = {{0}};
};

// J view ("Hermite order"): pairs grouped by total angular momentum L = L1 + L2 and sorted by
// Schwarz bound (descending) inside each group.  The per-call ket Hermite densities and bra
// accumulators live in this order.
// K view ("partner lists"): for each row primitive r (a basis-1 primitive of the list) and each
// partner angular momentum, the partner entries (basis-2 primitives) sorted by bound, descending;
// a symmetric list contributes both orientations (the diagonal once).
struct PairList {
    const Basis* b1 = nullptr;
    const Basis* b2 = nullptr;
    int symmetric = 0;
    int npair = 0;
    float maxbound = 0.0f;  // PairListUtil::max_bound
    // original list order: classes (L1, L2)
    std::vector<int> cls_L1, cls_L2, cls_start;
    std::vector<int> h_p1, h_p2;
    std::vector<float> h_bound;
    int *d_p1 = nullptr, *d_p2 = nullptr, *d_hpos = nullptr;
    float* d_bound = nullptr;
    double* d_K = nullptr;  // c1 c2 exp(-e1 e2/p R^2) per pair (list order)
    // J view
    int hstart[2 * kMaxL + 2] = {0};  // Hermite-order start of total-L group
    size_t hoff[2 * kMaxL + 2] = {0}; // offset (in doubles) of group in the H-strided buffers
    float* d_hb = nullptr;             // static bound, Hermite order
    double* d_hgeom = nullptr;         // p, Px, Py, Pz, sqrt(2 pi^(5/2))/p (SoA, 5 * npair), Hermite order
    std::vector<float> h_hb;
    // K view
    int nrow = 0;                      // = b1->nprim
    int kmaxL2 = 0;                    // max partner L
    std::vector<int> h_row_start;      // [nrow * (kmaxL2+1) + 1]
    int* d_row_start = nullptr;
    int nent = 0;
    float* d_ebound = nullptr;
    int* d_eprim = nullptr;   // partner primitive
    int* d_ecart = nullptr;   // partner Cartesian offset
    int* d_eblk = nullptr;    // partner block id
    int* d_keblk = nullptr;   // partner block id in the K view order (Basis::h_kblk)
    double* d_egeom = nullptr; // packed entries (12 doubles each: p, K, P, P-A, A-B, {cart, blk})
    double* d_egeom32 = nullptr; // FP32-route entries (64 bytes each: P in FP64, 8 floats, {cart, blk})
    // pivot view for counting (rows = basis-2 primitives, partners = basis-1 primitives, bounds
    // only); a symmetric list uses the row view above
    int* d_piv_row_start = nullptr;
    float* d_piv_b = nullptr;
    int* d_cls_start = nullptr;
    int* d_cls_L1 = nullptr;
    int* d_cls_L2 = nullptr;
    // active rows per angular momentum (rows with at least one partner), in primitive order
    // rows with at least one partner of angular momentum Lp, per (row L, Lp), in primitive order
    std::vector<int> h_rows[kMaxL + 1][kMaxL + 1];
    int* d_rows[kMaxL + 1][kMaxL + 1] = {};
    // triangle enumeration for a symmetric list: units (a in rows[L][La], b in rows[L][Lb], a <= b);
    // tri_start[i] = first index into rows[L][Lb] with b >= a_i, tri_pre[i] = units before row i
    std::vector<long long> h_tri_pre[kMaxL + 1][kMaxL + 1][kMaxL + 1];
    long long* d_tri_pre[kMaxL + 1][kMaxL + 1][kMaxL + 1] = {};
    int* d_tri_start[kMaxL + 1][kMaxL + 1][kMaxL + 1] = {};
    // batch list (lsjk_batch_create): pairs of one molecule only; nmol molecules, h_mol[p] the
    // molecule of basis-1 primitive p; in each Hermite group the molecules' pairs are contiguous,
    // molecule m at [h_hmol[L][m], h_hmol[L][m + 1]), each sorted by bound.  Quartets are formed
    // within one molecule only (J tasks and K units never pair two molecules).
    int nmol = 0;
    std::vector<int> h_mol;
    std::vector<int> h_hmol[2 * kMaxL + 1];
    // autotuned launch parameters of this list on both sides (lsjk_autotune): J ket-chunk scale per
    // group pair of the one-way task list [0] and of the symmetric kernel's list [1] (1: the
    // default chunk), K unit-split target per class counter index (absent: the default)
    float jscale[2][5][5];
    std::map<int, long long> ksplit;
    // mixed route (thredp > thresp), this list on both sides: the launches (J and K launch ids, as
    // timed by lsjk_autotune) that run the FP64 kernel instead, because it measured faster (the
    // same quartets, all in FP64); force_plain: every launch, while the autotuner times them
    std::set<int> plain_on_mixed;
    bool force_plain = false;
    // generation of the K unit lists (raised whenever cached unit lists are dropped), so that data
    // cached per unit list elsewhere can tell a rebuilt list from the old one
    unsigned long long kunits_gen = 0;
    // the tuned choices, for reporting: kind (0 J one-way, 1 J symmetric, 2 K), counter index, value,
    // ms with the default and with the choice
    struct Tuned {
        int kind, idx;
        double value, ms_default, ms_tuned;
    };
    std::vector<Tuned> tuned;
    // process-unique id (never reused, unlike the object's address)
    unsigned long long uid = 0;
    // cached J task lists keyed by the ket list's uid
    std::map<unsigned long long, JTaskList> jtasks;
    // the symmetric J kernel's task list (this list on both sides), built with the list
    JTaskList jtasks_sym;
    // cached K gradient unit lists (device pointer, count) keyed by (ket list uid, class, triangle flag)
    std::map<std::pair<unsigned long long, int>, std::pair<void*, int>> kunits;
    // cost of the class's units (sum over units of partner-count products), same key
    std::map<std::pair<unsigned long long, int>, double> kcost;
    // host copies of the entry bounds and partner primitives (K view order)
    std::vector<float> h_eb;
    std::vector<int> h_eprim;
    // spatial (Morton) rank of each basis-1 primitive: the order of the ket groups, and the total
    // order of the triangle mode; built on first use
    std::vector<int> h_mrank;
    // K ket groups per (row L, partner L) (kernels.hpp KGroups), built on first use
    struct KGroupData {
        int gsize = 16;  // rows per group
        int2* d_slot = nullptr;
        int* d_grp_slot = nullptr;
        int* d_grp_rows = nullptr;
        int* d_grp_rank = nullptr;
        float* d_chunk_b = nullptr;
        int ngroup = 0;
        std::vector<int> grp_slot;  // host copies
        std::vector<int> grp_rank;
        std::vector<float> chunk_b;
        std::vector<int> order;     // indices into h_rows[L][Lp] in group order
        std::vector<int> mol_grp;   // batch list: groups of molecule m at [mol_grp[m], mol_grp[m + 1])
        std::vector<int2> slot;     // host copies of the slots and group rows
        std::vector<int> grp_rows;
    };
    std::map<int, KGroupData> kgroups;  // keyed by ((row L) * 8 + partner L) * 256 + rows per group
    // autotuned rows per ket group, per (row L) * 8 + partner L of this list as the ket list
    // (absent: kgroup_size)
    std::map<int, int> kgsize;
    // cached K build unit lists (device pointer, count) keyed by (ket list uid, class and
    // triangle flag, thresp bits), and their estimated cost
    std::map<std::pair<unsigned long long, long long>, std::pair<void*, int>> kunits_g;
    std::map<std::pair<unsigned long long, long long>, double> kcost_g;
    // symmetric-mode K: for each pair of mirror classes (L1 L2 | L3 L4) / (L3 L4 | L1 L2), the one
    // that evaluates both (with the transposed block), chosen by timing both at first use; keyed by
    // (ket list uid, thresp bits << 1 | mixed route), class counter index -> chosen class counter
    // index.  Every class pair with work has an entry once the map is stored.
    std::map<std::pair<unsigned long long, unsigned long long>, std::map<int, int>> korient;
};

// ---------------------------------------------------------------- per-call argument blocks

struct Ewald {
    int n = 0;
    int plain = 0;  // one bare 1/r term (omega = -1): the kernels take a branch without the term loop
    double alpha[kMaxEwald];
    double omega[kMaxEwald];
};

// Device scalars produced and consumed on the stream within one call.
struct CallScalars {
    unsigned int max34_bits;     // max truncated ket bound (float bits)
    unsigned int pad;
    unsigned long long dmax_bits;  // max |D| (double bits)
    unsigned long long counts[1024]; // evaluated primitive quartets (see jcount_index / kcount_index)
    unsigned long long counts32[1024]; // the part of them evaluated on the FP32 route
    unsigned long long wsteps[1024];   // wave steps that evaluated at least one integral
    unsigned long long wlanes[1024];   // lanes that evaluated an integral, summed over those steps
    unsigned int queue[1024];        // work-queue counters of the persistent kernels
};

struct JTask {
    int bra0, nbra, ket0, ket1;
};

}  // namespace lsjk

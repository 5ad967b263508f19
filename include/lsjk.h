/*
 * lsjk.h -- C-ABI of liblsjk, the MI355X (gfx950) direct Coulomb/exchange build.
 *
 * This is the boundary the host C++ IntBox::coulomb / IntBox::exchange dispatch into, in the
 * place of the reference's TeraChem seam (reference src/cpp/intbox/intbox.cpp:213-329 ->
 * src/cpp/intbox/tc/{coulomb,exchange}.cpp -> TCIntBox -> IBox* C API,
 * src/cpp/intbox/tc/tcintbox/tcintbox.cpp:20-415).  Plain pointers and sizes only.
 *
 * Conventions
 *   - All matrices are row-major FP64 in the Cartesian representation of their bases, Cartesian
 *     components in lexical order (xx, xy, xz, yy, yz, zz) as in reference
 *     src/lightspeed/am.hpp:48-68.  Pure <-> Cartesian transforms stay on the caller's side, as
 *     in reference coulombCPU (src/cpp/intbox/cpu/coulomb.cpp:26-35, 186-193).
 *   - Matrix arguments named *_dev are device pointers (HBM) on the library's device; the build
 *     functions ACCUMULATE into their output (J += ..., K += ...), like IBoxJFockGen.
 *   - Every function returns 0 on success and a nonzero LSJK_E* code on failure; the message is
 *     available from lsjk_last_error() (per thread).  Nothing is printed.
 *   - One process-wide session (like TCIntBox, tcintbox.cpp:18-27); calls are serialised
 *     internally and run on the library's own HIP stream; each build call returns after its
 *     kernels completed.
 */
#ifndef LSJK_H
#define LSJK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSJK_OK 0
#define LSJK_E_INVALID 1   /* bad argument (shape, angular momentum, null pointer, ...) */
#define LSJK_E_HIP 2       /* HIP runtime error */
#define LSJK_E_NOINIT 3    /* lsjk_init not called */
#define LSJK_E_UNSUPPORTED 4

typedef struct lsjk_basis lsjk_basis;
typedef struct lsjk_pairlist lsjk_pairlist;

/* Replaces IBoxStartup(ngpu, devIds, pinnedBytes, align) (reference tcintbox.cpp:44-50):
 * select the device, create the stream, upload the Rys quadrature table. */
int lsjk_init(int device);

/* Replaces the TCIntBox teardown (~TCIntBox / IBoxShutdown, reference tcintbox.cpp:52-55, run
 * from the interpreter exit hook src/python/exit_hooks.cpp:14-18): release every device
 * resource of the session. */
int lsjk_finalize(void);

/* Last error message of the calling thread ("" if none). */
const char* lsjk_last_error(void);

/* Number of visible GPUs (0 if the HIP runtime finds none); needs no session.  The
 * counterpart of the device enumeration in ResourceList::build (reference
 * src/cpp/env/resource_list.cpp:22-37). */
int lsjk_device_count(int* count);

/* Replaces IBoxInitShell(L, atomIdx, nprim, es, cs) per shell + IBoxUpdateCoors(threpair, xyz)
 * (reference tcintbox.cpp:99-110): a primitive-level basis (reference Basis, basis.hpp:425-609).
 * c, e: coefficients (normalisation included) and exponents; xyz: nprim x 3 (bohr); L: angular
 * momentum (0..2); cart: Cartesian offset of each primitive's shell (cartIdx); ncart: total
 * Cartesian functions.  Data are copied. */
int lsjk_basis_create(int nprim, const double* c, const double* e, const double* xyz, const int* L,
                      const int* cart, int ncart, lsjk_basis** out);
int lsjk_basis_destroy(lsjk_basis* b);
/* Atom index of each primitive (reference Primitive::atomIdx, basis.hpp:136-250), natom atoms:
 * needed by the gradient functions, which accumulate per-atom (natom x 3) results. */
int lsjk_basis_set_atoms(lsjk_basis* b, int natom, const int* atom);

/* A Schwarz pair list (reference PairList, pair_list.hpp:25-220) over (b1, b2): npair primitive
 * pairs (p1 index into b1, p2 into b2, float bound), grouped by (L1, L2) class in L1-major order
 * and sorted by bound descending inside each class (PairList::build_schwarz,
 * src/cpp/core/pair_list.cpp:152-173).  symmetric: b1 and b2 are the same basis and p2 <= p1.
 * Device-side Hermite and K-invert views are built once here (the role of TCIntBox::set_basis,
 * tcintbox.cpp:66-115, which caches the TeraChem shell-pair data). */
int lsjk_pairlist_create(const lsjk_basis* b1, const lsjk_basis* b2, int symmetric, int npair, const int* p1,
                         const int* p2, const float* bound, lsjk_pairlist** out);
int lsjk_pairlist_destroy(lsjk_pairlist* pl);

/* Replaces IBoxJFockGen(thresp, thredp, scalfr, scallr, omega, P, fock) per Ewald task
 * (reference tcintbox.cpp:261-278, call at :270, from tc/coulomb.cpp:14-63), with the semantics of
 * IntBox::coulomb (src/lightspeed/intbox.hpp:721-771):
 *   J12_dev (ncart1 x ncart2) += sum_34 (12|O|34) D34_dev (ncart3 x ncart4),
 *   O(r) = sum_i scales[i] erf(omegas[i] r)/r  (omegas[i] == -1: 1/r), nterm <= 8.
 * Precision (intbox.hpp:736-742): with thredp <= thresp every quartet with est >= thresp is FP64;
 * with thredp > thresp quartets with est >= thredp are FP64 and those with thresp <= est < thredp
 * are evaluated in FP32 -- except that quartets evaluated together with an FP64 one in the same
 * wave step are evaluated in FP64 as well, and so are all quartets of the launches that
 * lsjk_autotune found faster with the FP64 kernel (never less precision than asked for);
 * accumulation is FP64 throughout.  lsjk_last_counts_fp32 reports how many ran in FP32. */
int lsjk_coulomb(const lsjk_pairlist* pl12, const lsjk_pairlist* pl34, const double* D34_dev, int nterm,
                 const double* scales, const double* omegas, float thresp, float thredp, double* J12_dev);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): lsjk_coulomb with the
 * same arguments, kernels and results, but its per-call buffers (truncated ket bounds, ket Hermite
 * densities and their FP32 copy, bra accumulators: about 16 bytes per Hermite component of the
 * two lists) are taken from a fixed device scratch array of LSJK_NOALLOC_SCRATCH_BYTES that the
 * library holds on the device from module load until the process ends.  The call allocates and
 * frees nothing.  Limits: a call that needs more scratch fails with LSJK_E_UNSUPPORTED before any
 * launch; the J task list of (pl12, pl34) must exist (it is built when a list is created, for that
 * list on both sides, and by the first lsjk_coulomb call of any other pair), else LSJK_E_INVALID;
 * per-launch timing (lsjk_set_profiling) is not recorded.  Every call on the device shares the one
 * array, which keeps the last call's contents (each call clears what it accumulates into); calls
 * are serialised by the session lock, and calls that overlapped on one GPU would race on it. */
#define LSJK_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int lsjk_coulomb_noalloc(const lsjk_pairlist* pl12, const lsjk_pairlist* pl34, const double* D34_dev, int nterm,
                         const double* scales, const double* omegas, float thresp, float thredp, double* J12_dev);

/* Replaces IBoxKFockGen / IBoxKFockSym (reference tcintbox.cpp:299-334, called from
 * tc/exchange.cpp:14-75), with the semantics of IntBox::exchange (intbox.hpp:773-828):
 *   K13_dev (ncart1 x ncart3) += sum_24 (12|O|34) D24_dev (ncart2 x ncart4).
 * D24symm: the caller asserts D24 symmetric and bases 1==3, 2==4 (the effective flag of
 * exchange.cpp:73-81); with pl12 == pl34 the library then halves the work.  Precision as for
 * lsjk_coulomb (intbox.hpp:787-793). */
int lsjk_exchange(const lsjk_pairlist* pl12, const lsjk_pairlist* pl34, const double* D24_dev, int D24symm,
                  int nterm, const double* scales, const double* omegas, float thresp, float thredp,
                  double* K13_dev);

/* Batches of independent molecules.  The reference has no batched path: its callers run one
 * IntBox::coulomb + IntBox::exchange per molecule (an SCF per molecule, reference
 * est/rhf.py:616-643; MO integrals per pair density, src/cpp/intbox/mo_eri.cpp:10-250).  A batch
 * runs the J and K builds of many molecules together, one launch per angular-momentum class over
 * all of them, replayed from a captured hipGraph.
 *
 * lsjk_batch_create: nmol molecules, each given by a symmetric pair list over its own basis
 * (lsjk_pairlist_create with symmetric = 1 and b1 == b2).  The data are copied; the lists and
 * bases may be destroyed afterwards.
 * lsjk_batch_layout: the device matrix layout of the batch.  A batch matrix holds every molecule's
 * own ncart_m x ncart_m block: element (i, j) of molecule m at offsets[m] + i * ld + j in a buffer
 * of nelem doubles (ld = the largest ncart_m; the blocks do not overlap).  offsets: nmol entries,
 * or NULL.
 * lsjk_batch_jk: J_m += J[D_m] and K_m += K[D_m] (D24symm: each D_m must be symmetric) for every
 * molecule, with the semantics of lsjk_coulomb / lsjk_exchange on that molecule's list alone
 * (Ewald operator, thresp, thredp and the mixed route as there); J_dev or K_dev may be NULL to
 * skip that build.  The first call with a given (D_dev, J_dev, K_dev, thresp, thredp, operator)
 * runs the launches directly and captures them into a hipGraph; later calls with the same
 * arguments replay the graph (the graph keeps the device pointers: the buffers must stay
 * allocated while the batch lives or until other arguments have been used 4 times). */
typedef struct lsjk_batch lsjk_batch;
int lsjk_batch_create(int nmol, const lsjk_pairlist* const* lists, lsjk_batch** out);
int lsjk_batch_destroy(lsjk_batch* b);
int lsjk_batch_layout(const lsjk_batch* b, int* ld, size_t* nelem, size_t* offsets);
int lsjk_batch_jk(lsjk_batch* b, const double* D_dev, int nterm, const double* scales, const double* omegas,
                  float thresp, float thredp, double* J_dev, double* K_dev);

/* The learned screen.  The reference sieves exchange quartets by the Schwarz estimate alone,
 * bound12 * bound34 * max|D24 block| against thresp (src/cpp/intbox/cpu/exchange.cpp:186, contract
 * src/lightspeed/intbox.hpp:736-742 and 787-793).  The learned route tightens that estimate per
 * exchange work unit (a bra row primitive against a spatial group of ket rows) with a network
 * trained on the magnitudes of the unit's quartets, and routes each quartet by
 * est * tau_u: skipped below thresp, FP32 below thredp, FP64 above.  Schwarz stays the fallback:
 * tau_u <= 1, and classes the screen does not certify keep the plain FP64 route.
 *
 * Network (FP32): ni <= 32 inputs, 3 hidden layers of 64 ReLU units, one linear output; parameter
 * vector W1 (64 x ni, row-major), b1 (64), W2 (64 x 64), b2, W3 (64 x 64), b3, w4 (64), b4 (1).
 * lsjk_mlp_forward: out[s] = f(X[s]) for n rows of X (n x ni), device pointers.
 * lsjk_mlp_grad: grad += d/dparams and loss[0] += L of L = (1/nb) sum_s sw_s r_s^2 (times wu where
 * r_s < 0), r_s = f(X[idx[s]]) - y[idx[s]], over nb rows (idx NULL: rows 0..nb-1; sw NULL: 1).
 * lsjk_mlp_adam: one Adam step (bias-corrected, step t >= 1) of params with gradient grad and the
 * moment buffers m, v. */
int lsjk_mlp_nparam(int ni, int* nparam);
int lsjk_mlp_forward(int ni, const float* params_dev, int n, const float* X_dev, float* out_dev);
int lsjk_mlp_grad(int ni, const float* params_dev, int nb, const int* idx_dev, const float* X_dev, const float* y_dev,
                  const float* sw_dev, float wu, float* grad_dev, float* loss_dev);
int lsjk_mlp_adam(int nparam, float* params_dev, const float* grad_dev, float* m_dev, float* v_dev, float lr,
                  float beta1, float beta2, float eps, int t);

/* Unit features and training targets of the exchange class (L1 L2 | L3 L4) of one list on both
 * sides (all row pairs; the units lsjk_exchange forms at this thresp, in the same order).
 * lsjk_unit_features: *nunits, and (feat non-NULL) LSJK_SCREEN_NFEAT floats per unit: L1, L2, L3, L4;
 * ln of the bra row's exponent, of the smallest exponent among its partners in the unit, of the
 * smallest exponent among the ket group's rows and among the unit's ket partners; the distance of
 * the bra row to the ket group's centroid, the group's radius, the largest bra and ket pair extents
 * (bohr); ln of the unit's largest bra bound and of its first chunk's largest ket bound; ln of the
 * (ss|ss) ratio to its Schwarz product, sqrt(2) (pq)^(1/4) (p+q)^(-1/2) F0(rho R^2), for the unit's
 * most diffuse pairs at the closest approach the geometry allows and at the centroid distance.
 * lsjk_unit_magnitudes: per unit (host array of nunits floats), the largest
 * max_components |(ab|cd)| / (bound_ab bound_cd) over the unit's quartets with
 * bound_ab bound_cd >= smin, by the exact Rys evaluation on the GPU (1/r operator; classes with
 * s and p functions only). */
#define LSJK_SCREEN_NFEAT 16
int lsjk_unit_features(const lsjk_pairlist* pl, int L1, int L2, int L3, int L4, float thresp, int* nunits,
                       float* feat);
int lsjk_unit_magnitudes(const lsjk_pairlist* pl, int L1, int L2, int L3, int L4, float thresp, float smin,
                         float* mag);

/* A trained screen: the network (host params), the normalisation of its inputs
 * (x' = (x - mean) / scale, ni entries each) and, per exchange class index
 * c = ((L1*5 + L2)*5 + L3)*5 + L4, nbin bins of the network output (ascending edges
 * fedge[c * (nbin - 1) + b]) with a certified margin delta[c * nbin + b] each (NaN: not certified):
 * tau_u = min(1, exp(f(x'_u) + delta of f's bin)), applied while thresp >= smin * max|D24|.
 * lsjk_exchange_screened: lsjk_exchange on the learned route (certified classes with s and p
 * functions: est * tau_u against thresp and thredp; all other classes: plain FP64 route).  The
 * screen keeps the factors of each unit list it has seen until it is destroyed. */
typedef struct lsjk_screen lsjk_screen;
int lsjk_screen_create(int ni, const float* params, const float* mean, const float* scale, int nbin, const float* fedge,
                       const double* delta, float smin, lsjk_screen** out);
int lsjk_screen_destroy(lsjk_screen* screen);
int lsjk_exchange_screened(const lsjk_pairlist* pl12, const lsjk_pairlist* pl34, const double* D24_dev, int D24symm,
                           int nterm, const double* scales, const double* omegas, float thresp, float thredp,
                           lsjk_screen* screen, double* K13_dev);

/* Autotuner.  No counterpart in the reference, whose GPU tile choices live inside closed TeraChem
 * (tcintbox.cpp:261-334).  On a symmetric list over one basis, used on both sides with the density
 * D_dev (symmetric) and the route (thresp, thredp): times the J launches with the ket chunk of every
 * task scaled by 1/4, 1/2, 1, 2 or 4, and the K launches with the work-unit split target of every
 * class at 3072, 12288 or 49152 units (class launches serialised on one stream, HIP events around
 * each, best of two), and keeps the fastest choice per J group pair and per K class for later
 * builds of the list on both sides (the FP64 and mixed routes share these choices); before the K
 * splits it times the rows per exchange ket group (16, 32 or 64 per (row L, partner L) of the list,
 * where every class over that ket pair fits 64 KB of LDS) together with the orientations of the
 * symmetric mode's mirror classes: at each size every class is timed, those of a mirror pair in
 * both orientations; a ket pair keeps the size with the least summed time of its classes, and each
 * mirror pair then the faster orientation at the kept sizes.  Called with a
 * mixed route (thredp > thresp), it instead times each launch of that route with its mixed kernel
 * and with the FP64 kernel (the same quartets, all in FP64) and keeps the faster per launch, and
 * times the route's mirror-class orientations; call it after the FP64 route's tuning.  Each route
 * keeps its own orientations (a learned-route call uses those of its route).
 * lsjk_tuned_params: the choices, n entries: kind (0 J one-way, 1 J symmetric, 2 K, 3 mixed-route
 * kernel, 4 K ket group rows), counter index (J: L12*5 + L34, bit 24 for the symmetric kernel; K: as
 * lsjk_last_counts; kind 4: L3*5 + L4), value (chunk scale, split target, 1 = FP64 kernel / 0 =
 * mixed kernel, or rows per group), and the launch's ms (kind 4: the ket pair's launches summed)
 * with the default and with the choice.
 * lsjk_batch_autotune: the same on a batch's combined list (D_dev in the batch layout); drops the
 * batch's captured graphs.  lsjk_batch_pairlist: the batch's combined list (for lsjk_tuned_params). */
int lsjk_autotune(lsjk_pairlist* pl, const double* D_dev, int nterm, const double* scales, const double* omegas,
                  float thresp, float thredp);
int lsjk_tuned_params(const lsjk_pairlist* pl, int capacity, int* n, int* kind, int* idx, double* value,
                      double* ms_default, double* ms_tuned);
int lsjk_batch_autotune(lsjk_batch* b, const double* D_dev, int nterm, const double* scales, const double* omegas,
                        float thresp, float thredp);
int lsjk_batch_pairlist(const lsjk_batch* b, const lsjk_pairlist** out);
/* Sets the rows per exchange ket group of (row L, partner Lp) of the list, as lsjk_autotune would
 * (rows 16, 32 or 64; 0 restores the default); LSJK_E_INVALID when a class over that ket pair
 * would not fit 64 KB of LDS.  For reapplying saved tuning, or fixing it.  No counterpart in the
 * reference. */
int lsjk_set_kgroup_rows(lsjk_pairlist* pl, int L, int Lp, int rows);

/* Boys functions F_n(T) = int_0^1 t^(2n) exp(-T t^2) dt for n = 0..nboys (nboys <= 16) at npoint
 * arguments: F_dev (npoint x (nboys + 1), row-major) from T_dev, FP64.  Replaces the reference's
 * Boys::compute / Boys::interpolate_F (src/lightspeed/boys.hpp:49-114, 170-181;
 * src/cpp/quad/boys.cpp:9-49). */
int lsjk_boys(int nboys, int npoint, const double* T_dev, double* F_dev);

/* Coulomb gradient, bra-centre part (the reference's coulombGrad, intbox.hpp:954-1008 and
 * cpu/coulomb_grad.cpp:19-563, is the sum of this with the densities in both roles):
 *   G_dev (natom x 3) += sum_{ab,cd} Dbra_ab Dket_cd d/dR [bra centres a, b] (ab|O|cd),
 * over a symmetric list (both sides), Cartesian densities, screening as lsjk_coulomb with
 * Dket (thredp: FP64 route).  Needs lsjk_basis_set_atoms. */
int lsjk_coulomb_grad(const lsjk_pairlist* pl, const double* Dbra_dev, const double* Dket_dev, int nterm,
                      const double* scales, const double* omegas, float thresp, float thredp, double* G_dev);

/* Exchange gradient (the reference's exchangeGrad, intbox.hpp:1092-1160, cpu/exchange_grad.cpp):
 *   G_dev (natom x 3) += d/dR sum_{abcd} D13_ac D24_bd (ab|O|cd)
 * over a symmetric list (both sides), Cartesian densities, quartets with
 * float(b12 b34) max|D24 block| >= thresp (as lsjk_exchange without the D24symm half-sieve; FP64).
 * Needs lsjk_basis_set_atoms. */
int lsjk_exchange_grad(const lsjk_pairlist* pl, const double* D13_dev, const double* D24_dev, int nterm,
                       const double* scales, const double* omegas, float thresp, float thredp, double* G_dev);

/* One-electron integrals of a pair list, accumulated into ncart1 x ncart2 device matrices
 * (Cartesian, as lsjk_coulomb).  A symmetric list fills both triangles.
 * lsjk_overlap   replaces IntBox::overlap   (reference src/cpp/intbox/intbox.cpp:30-40,
 *                                            cpu/overlap.cpp:14-147): S += <a|b>;
 * lsjk_kinetic   replaces IntBox::kinetic   (intbox.cpp:183-193, cpu/kinetic.cpp:14-175):
 *                                            T += <a| -1/2 nabla^2 |b>;
 * lsjk_potential replaces IntBox::potential (intbox.cpp:196-210, cpu/potential.cpp:14-147):
 *                V += sum_A <a| O(r - R_A) |b> (-c_A) for ncharge point charges xyzc (host,
 *                ncharge x 4: x, y, z, c), O the Ewald operator as in lsjk_coulomb. */
int lsjk_overlap(const lsjk_pairlist* pl, double* S_dev);
int lsjk_kinetic(const lsjk_pairlist* pl, double* T_dev);
int lsjk_potential(const lsjk_pairlist* pl, int nterm, const double* scales, const double* omegas, int ncharge,
                   const double* xyzc, double* V_dev);

/* Device memory helpers for callers without their own HIP allocator. */
int lsjk_malloc(void** ptr, size_t bytes);
int lsjk_free(void* ptr);
int lsjk_memcpy_htod(void* dst, const void* src, size_t bytes);
int lsjk_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int lsjk_memset_zero(void* dst, size_t bytes);
/* Device-to-device copy by a plain copy kernel (16-byte loads and stores); also the reference
 * kernel for calibrating HBM byte counters. */
int lsjk_memcpy_dtod(void* dst, const void* src, size_t bytes);
int lsjk_synchronize(void);

/* Streams over which one build call spreads its per-class kernel launches (1..4, default 4; 1
 * runs them one after another on one stream, so each launch's duration is its own). */
int lsjk_set_streams(int n);

/* Instrumentation of the last build call: quartets evaluated per class (1024 counters: J at
 * [L12*5 + L34], K at [256 + ((L1*5+L2)*5+L3)*5+L4]; enabled by lsjk_set_counting(1)) and the
 * device time of each phase in milliseconds (HIP events on the library stream):
 * ms[0] = whole call, ms[1] = quartet kernels only. */
int lsjk_set_counting(int enable);
int lsjk_last_counts(uint64_t* counts1024);
/* The part of lsjk_last_counts evaluated on the FP32 route (mixed precision calls), same layout. */
int lsjk_last_counts_fp32(uint64_t* counts1024);
/* Lane occupancy of the quartet kernels in the last counted call, same layout: steps1024[i] =
 * wave steps of class i that evaluated at least one integral, lanes1024[i] = the lanes that
 * evaluated one, summed over those steps (lanes / (64 steps) = the share of the wave doing
 * integral work; the symmetric J kernel counts one lane per integral, serving both J blocks, at
 * the bra class's index).  No counterpart in the reference (instrumentation of this library). */
int lsjk_last_wave_steps(uint64_t* steps1024, uint64_t* lanes1024);
int lsjk_last_timing(double* ms2);

/* Quartet counts under the reference's screening conventions, no integrals evaluated (the
 * denominator of the quartets/s metric; reference coulomb.cpp:38-52, 115 for J and
 * exchange.cpp:54-71, 124-139, 186 for K, D24symm half-sieve included).  countsJ / countsK:
 * 625 entries each, indexed ((L1*5+L2)*5+L3)*5+L4.  Either density may be NULL to skip it. */
int lsjk_count_quartets(const lsjk_pairlist* pl12, const lsjk_pairlist* pl34, const double* D34_dev,
                        const double* D24_dev, int D24symm, float thresp, uint64_t* countsJ, uint64_t* countsK);

/* Per-launch timing of the quartet kernels of the last build call (HIP events around each
 * launch, enabled by lsjk_set_profiling(1)): ids[i] is the launch's counter index (as in
 * lsjk_last_counts), ms[i] its duration.  *n receives the number of launches. */
int lsjk_set_profiling(int enable);
int lsjk_last_launches(int capacity, int* n, int* ids, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* LSJK_H */

// Host C++ API of lightspeed_amd: the reference's Basis / PairList / Tensor / Ewald / IntBox
// interface for the direct J/K build (reference src/lightspeed/{basis,pair_list,tensor,ewald,
// intbox,resource_list}.hpp), with IntBox::coulomb / IntBox::exchange dispatching into the HIP
// library through its C-ABI (include/lsjk.h).
#pragma once

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

struct lsjk_basis;
struct lsjk_pairlist;
struct lsjk_batch;

namespace lightspeed {

// ------------------------------------------------------------------ Tensor
// Row-major FP64 n-d array (reference src/lightspeed/tensor.hpp:19-320; only what the J/K path
// and its callers use).
class Tensor {
public:
    explicit Tensor(const std::vector<size_t>& shape, const std::string& name = "T");
    const std::vector<size_t>& shape() const { return shape_; }
    size_t ndim() const { return shape_.size(); }
    size_t size() const { return data_.size(); }
    std::vector<double>& data() { return data_; }
    const std::vector<double>& data() const { return data_; }
    const std::string& name() const { return name_; }
    void set_name(const std::string& n) { name_ = n; }
    void shape_error(const std::vector<size_t>& shape) const;
    void ndim_error(size_t ndim) const;
    void zero();
    void scale(double a);
    // this = a * other + b * this
    void axpby(const std::shared_ptr<Tensor>& other, double a, double b);
    std::shared_ptr<Tensor> clone() const;

private:
    std::vector<size_t> shape_;
    std::vector<double> data_;
    std::string name_;
};

// ------------------------------------------------------------------ Primitive / Basis
// Reference src/lightspeed/basis.hpp:136-260 (Primitive) and 425-609 (Basis).
class Primitive {
public:
    Primitive(double c, double e, double x, double y, double z, double c0, int L, bool is_pure, int aoIdx,
              int cartIdx, int primIdx, int shellIdx, int atomIdx)
        : c_(c), e_(e), x_(x), y_(y), z_(z), c0_(c0), L_(L), is_pure_(is_pure), aoIdx_(aoIdx), cartIdx_(cartIdx),
          primIdx_(primIdx), shellIdx_(shellIdx), atomIdx_(atomIdx)
    {
    }
    double c() const { return c_; }
    double e() const { return e_; }
    double x() const { return x_; }
    double y() const { return y_; }
    double z() const { return z_; }
    double c0() const { return c0_; }
    int L() const { return L_; }
    bool is_pure() const { return is_pure_; }
    int aoIdx() const { return aoIdx_; }
    int cartIdx() const { return cartIdx_; }
    int primIdx() const { return primIdx_; }
    int shellIdx() const { return shellIdx_; }
    int atomIdx() const { return atomIdx_; }
    int nao() const { return is_pure_ ? npure() : ncart(); }
    int npure() const { return 2 * L_ + 1; }
    int ncart() const { return (L_ + 1) * (L_ + 2) / 2; }
    bool operator==(const Primitive& o) const;
    bool operator!=(const Primitive& o) const { return !(*this == o); }

private:
    double c_, e_, x_, y_, z_, c0_;
    int L_;
    bool is_pure_;
    int aoIdx_, cartIdx_, primIdx_, shellIdx_, atomIdx_;
};

class Basis {
public:
    Basis(const std::string& name, const std::vector<Primitive>& primitives);
    ~Basis();
    const std::string& name() const { return name_; }
    const std::vector<Primitive>& primitives() const { return primitives_; }
    size_t nao() const { return nao_; }
    size_t ncart() const { return ncart_; }
    size_t nprim() const { return nprim_; }
    size_t nshell() const { return nshell_; }
    size_t natom() const { return natom_; }
    int max_L() const;
    bool has_pure() const;
    static bool equivalent(const std::shared_ptr<Basis>& a, const std::shared_ptr<Basis>& b);
    // Device copy used by the HIP library (created on first use).
    lsjk_basis* device_handle() const;

private:
    std::string name_;
    std::vector<Primitive> primitives_;
    size_t nao_ = 0, ncart_ = 0, nprim_ = 0, nshell_ = 0, natom_ = 0;
    mutable lsjk_basis* dev_ = nullptr;
};

// ------------------------------------------------------------------ Ewald
// O(r) = sum_i scales[i] erf(omegas[i] r) / r, omegas[i] == -1 meaning 1/r
// (reference src/lightspeed/ewald.hpp:25-109).
class Ewald {
public:
    Ewald(const std::vector<double>& scales, const std::vector<double>& omegas);
    const std::vector<double>& scales() const { return scales_; }
    const std::vector<double>& omegas() const { return omegas_; }
    static std::shared_ptr<Ewald> coulomb();
    bool is_coulomb() const;

private:
    std::vector<double> scales_, omegas_;
};

// ------------------------------------------------------------------ pair lists
// Reference src/lightspeed/pair_list.hpp:25-220.
class Pair {
public:
    Pair(const Primitive* p1, const Primitive* p2, float bound) : p1_(p1), p2_(p2), bound_(bound) {}
    const Primitive& prim1() const { return *p1_; }
    const Primitive& prim2() const { return *p2_; }
    float bound() const { return bound_; }
    static Pair build_schwarz(const Primitive& p1, const Primitive& p2);

private:
    const Primitive* p1_;
    const Primitive* p2_;
    float bound_;
};

class PairListL {
public:
    PairListL(bool is_symmetric, int L1, int L2, const std::vector<Pair>& pairs);
    bool is_symmetric() const { return is_symmetric_; }
    int L1() const { return L1_; }
    int L2() const { return L2_; }
    const std::vector<Pair>& pairs() const { return pairs_; }

private:
    bool is_symmetric_;
    int L1_, L2_;
    std::vector<Pair> pairs_;
};

class PairList {
public:
    PairList(const std::shared_ptr<Basis>& basis1, const std::shared_ptr<Basis>& basis2, bool is_symmetric,
             float thre, const std::vector<PairListL>& pairlists);
    ~PairList();
    std::shared_ptr<Basis> basis1() const { return basis1_; }
    std::shared_ptr<Basis> basis2() const { return basis2_; }
    bool is_symmetric() const { return is_symmetric_; }
    float thre() const { return thre_; }
    const std::vector<PairListL>& pairlists() const { return pairlists_; }
    size_t npair() const;
    // PairList::build_schwarz, reference src/cpp/core/pair_list.cpp:29-177.
    static std::shared_ptr<PairList> build_schwarz(const std::shared_ptr<Basis>& basis1,
                                                   const std::shared_ptr<Basis>& basis2, bool is_symmetric,
                                                   float threpq);
    lsjk_pairlist* device_handle() const;

private:
    std::shared_ptr<Basis> basis1_, basis2_;
    bool is_symmetric_;
    float thre_;
    std::vector<PairListL> pairlists_;
    mutable lsjk_pairlist* dev_ = nullptr;
};

struct PairListUtil {
    // criteria: bound descending, prim1 primIdx ascending, prim2 primIdx ascending
    // (reference src/cpp/core/pair_list_util.cpp:12-20)
    static bool compare_prim_pairs(const Pair& a, const Pair& b);
    static double max_bound(const std::shared_ptr<PairList>& pairlist);
};

// ------------------------------------------------------------------ resources
// Reference src/lightspeed/resource_list.hpp:25-73: CPU threads plus logical GPUs.
class ResourceList {
public:
    ResourceList(int nthread, const std::vector<int>& gpus) : nthread_(nthread), gpus_(gpus) {}
    int nthread() const { return nthread_; }
    int ngpu() const { return (int)gpus_.size(); }
    const std::vector<int>& gpus() const { return gpus_; }
    static std::shared_ptr<ResourceList> build(size_t cpu_bytes = 0, size_t gpu_bytes = 0);
    static std::shared_ptr<ResourceList> build_cpu();

private:
    int nthread_;
    std::vector<int> gpus_;
};

// ------------------------------------------------------------------ pure transforms
// Real solid harmonics (reference src/cpp/core/am.cpp:74-203) and the two-index transforms
// (reference src/cpp/core/pure_transform.cpp:9-247).
struct PureTransform {
    static std::vector<double> solid_harmonics(int L);  // (2L+1) x ncart(L), row-major
    static std::shared_ptr<Tensor> pureToCart2(const std::shared_ptr<Basis>& b1, const std::shared_ptr<Basis>& b2,
                                               const std::shared_ptr<Tensor>& Tpure);
    // Tpure (allocated if null) += C1 Tcart C2^T
    static std::shared_ptr<Tensor> cartToPure2(const std::shared_ptr<Basis>& b1, const std::shared_ptr<Basis>& b2,
                                               const std::shared_ptr<Tensor>& Tcart,
                                               const std::shared_ptr<Tensor>& Tpure);
};

// ------------------------------------------------------------------ IntBox
// Reference src/lightspeed/intbox.hpp:721-828 (contract) and src/cpp/intbox/intbox.cpp:213-329
// (dispatch).  The GPU backend is the HIP library; there is no CPU fallback: calls that the HIP
// backend cannot serve throw std::runtime_error.
struct IntBox {
    static std::shared_ptr<Tensor> coulomb(const std::shared_ptr<ResourceList>& resources,
                                           const std::shared_ptr<Ewald>& ewald,
                                           const std::shared_ptr<PairList>& pairlist12,
                                           const std::shared_ptr<PairList>& pairlist34,
                                           const std::shared_ptr<Tensor>& D34, float thresp, float thredp,
                                           const std::shared_ptr<Tensor>& J12 = std::shared_ptr<Tensor>());
    static std::shared_ptr<Tensor> exchange(const std::shared_ptr<ResourceList>& resources,
                                            const std::shared_ptr<Ewald>& ewald,
                                            const std::shared_ptr<PairList>& pairlist12,
                                            const std::shared_ptr<PairList>& pairlist34,
                                            const std::shared_ptr<Tensor>& D24, bool D24symm, float thresp,
                                            float thredp,
                                            const std::shared_ptr<Tensor>& K13 = std::shared_ptr<Tensor>());
    // Device-resident variants: Cartesian D and outputs as device pointers (accumulated).
    static void coulomb_device(const std::shared_ptr<Ewald>& ewald, const std::shared_ptr<PairList>& pairlist12,
                               const std::shared_ptr<PairList>& pairlist34, const double* D34_dev, float thresp,
                               float thredp, double* J12_dev);
    static void exchange_device(const std::shared_ptr<Ewald>& ewald, const std::shared_ptr<PairList>& pairlist12,
                                const std::shared_ptr<PairList>& pairlist34, const double* D24_dev, bool D24symm,
                                float thresp, float thredp, double* K13_dev);
    // One-electron integrals (reference intbox.hpp:225-406, 620-652; dispatch intbox.cpp:30-40,
    // 183-210): S += <a|b>, T += <a|-1/2 nabla^2|b>, V += sum_A <a|O(r_1A)|b> (-c_A) with xyzc an
    // (nA, 4) Tensor of (x, y, z, c) rows.
    static std::shared_ptr<Tensor> overlap(const std::shared_ptr<ResourceList>& resources,
                                           const std::shared_ptr<PairList>& pairlist,
                                           const std::shared_ptr<Tensor>& S = std::shared_ptr<Tensor>());
    static std::shared_ptr<Tensor> kinetic(const std::shared_ptr<ResourceList>& resources,
                                           const std::shared_ptr<PairList>& pairlist,
                                           const std::shared_ptr<Tensor>& T = std::shared_ptr<Tensor>());
    static std::shared_ptr<Tensor> potential(const std::shared_ptr<ResourceList>& resources,
                                             const std::shared_ptr<Ewald>& ewald,
                                             const std::shared_ptr<PairList>& pairlist,
                                             const std::shared_ptr<Tensor>& xyzc,
                                             const std::shared_ptr<Tensor>& V = std::shared_ptr<Tensor>());
    // MO-basis ERIs I[i,j,k,l] = (C1_i C2_j | C3_k C4_l) through one J build per (i, j) pair
    // (reference intbox.hpp:1316-1363, mo_eri.cpp:10-157), with the reference's symmetrisations
    // when C1 == C2, C3 == C4 or (C1, C2) == (C3, C4) (the same Tensor objects).  I is overwritten.
    static std::shared_ptr<Tensor> eriJ(const std::shared_ptr<ResourceList>& resources,
                                        const std::shared_ptr<Ewald>& ewald,
                                        const std::shared_ptr<PairList>& pairlist12,
                                        const std::shared_ptr<PairList>& pairlist34,
                                        const std::shared_ptr<Tensor>& C1, const std::shared_ptr<Tensor>& C2,
                                        const std::shared_ptr<Tensor>& C3, const std::shared_ptr<Tensor>& C4,
                                        float thresp, float thredp,
                                        const std::shared_ptr<Tensor>& I = std::shared_ptr<Tensor>());
    // Coulomb gradient G (natom, 3) += d/dR sum D12_pq D34_rs (pq|rs) over a symmetric pair list
    // (reference intbox.hpp:954-1008; dispatch intbox.cpp:424-496).
    static std::shared_ptr<Tensor> coulombGrad(const std::shared_ptr<ResourceList>& resources,
                                               const std::shared_ptr<Ewald>& ewald,
                                               const std::shared_ptr<PairList>& pairlist,
                                               const std::shared_ptr<Tensor>& D12, const std::shared_ptr<Tensor>& D34,
                                               float thresp, float thredp,
                                               const std::shared_ptr<Tensor>& G = std::shared_ptr<Tensor>());
    // Exchange gradient G (natom, 3) += d/dR sum D13_pq D24_rs (pr|qs) over a symmetric pair list
    // (reference intbox.hpp:1092-1160; dispatch intbox.cpp:519-620).  The symmetry flags of the
    // reference only select faster CPU paths; they are checked as the reference does.
    static std::shared_ptr<Tensor> exchangeGrad(const std::shared_ptr<ResourceList>& resources,
                                                const std::shared_ptr<Ewald>& ewald,
                                                const std::shared_ptr<PairList>& pairlist,
                                                const std::shared_ptr<Tensor>& D13, const std::shared_ptr<Tensor>& D24,
                                                bool D13symm, bool D24symm, bool Dsame, float thresp, float thredp,
                                                const std::shared_ptr<Tensor>& G = std::shared_ptr<Tensor>());
    // Backend predicate, the counterpart of is_terachem_ready (reference intbox.cpp:12-26).
    static bool is_hip_ready(const std::shared_ptr<ResourceList>& resources,
                             const std::shared_ptr<PairList>& pairlist12, const std::shared_ptr<PairList>& pairlist34,
                             std::string* why = nullptr);
};

// ------------------------------------------------------------------ batches of molecules
// The J and K builds of many independent molecules in one set of class launches (the HIP library's
// lsjk_batch_*), replayed from a captured hipGraph after the first call.  No counterpart in the
// reference, whose callers run one IntBox::coulomb + IntBox::exchange per molecule (reference
// est/rhf.py:616-643, src/cpp/intbox/mo_eri.cpp:10-250).  Molecule m's J and K equal
// IntBox::coulomb(res, ewald, pl_m, pl_m, D_m, thresp, thredp) and IntBox::exchange(res, ewald,
// pl_m, pl_m, D_m, true, thresp, thredp) (D_m symmetric).
class JKBatch {
public:
    static std::shared_ptr<JKBatch> build(const std::shared_ptr<ResourceList>& resources,
                                          const std::vector<std::shared_ptr<PairList>>& pairlists);
    ~JKBatch();
    JKBatch(const JKBatch&) = delete;
    JKBatch& operator=(const JKBatch&) = delete;
    size_t nmol() const { return pairlists_.size(); }
    const std::vector<std::shared_ptr<PairList>>& pairlists() const { return pairlists_; }
    // Device layout: molecule m's Cartesian (ncart_m x ncart_m) block, element (i, j), at
    // offsets()[m] + i * ld() + j of an nelem()-double buffer.
    int ld() const { return ld_; }
    size_t nelem() const { return nelem_; }
    const std::vector<size_t>& offsets() const { return offsets_; }
    // J_dev += J[D], K_dev += K[D] for every molecule, device buffers in the layout above (either
    // output may be null).
    void jk_device(const std::shared_ptr<Ewald>& ewald, const double* D_dev, float thresp, float thredp,
                   double* J_dev, double* K_dev);
    // lsjk_batch_autotune on the batch's combined list with a density in the batch layout; returns
    // the tuned table (kind, counter index, value, ms default, ms tuned).
    std::vector<std::tuple<int, int, double, double, double>> autotune(const std::shared_ptr<Ewald>& ewald,
                                                                       const double* D_dev, float thresp,
                                                                       float thredp);
    // Host densities (nao_m x nao_m each) -> fresh J_m, K_m.
    std::pair<std::vector<std::shared_ptr<Tensor>>, std::vector<std::shared_ptr<Tensor>>> jk(
        const std::shared_ptr<Ewald>& ewald, const std::vector<std::shared_ptr<Tensor>>& Ds, float thresp,
        float thredp);

private:
    JKBatch() = default;
    std::vector<std::shared_ptr<PairList>> pairlists_;
    lsjk_batch* dev_ = nullptr;
    int ld_ = 0;
    size_t nelem_ = 0;
    std::vector<size_t> offsets_;
};

// ------------------------------------------------------------------ Boys function
// Reference src/lightspeed/boys.hpp:170-181 (Boys::compute): an (npoint, nboys + 1) Tensor of
// F_0..F_nboys at the npoint arguments of T (nboys <= 16), computed on the GPU.
struct Boys {
    static std::shared_ptr<Tensor> compute(int nboys, const std::shared_ptr<Tensor>& T);
};

// Session helpers (the HIP library is one process-wide session, like TCIntBox).
void hip_session_init(int device = 0);
void hip_session_finalize();
int hip_device_count();

}  // namespace lightspeed

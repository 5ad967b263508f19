// Host C++ implementation: Tensor, Basis, Ewald, pair lists, pure transforms, resources.
#include <algorithm>
#include <cmath>
#include <map>
#include <sstream>
#include <tuple>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../../../include/lsjk.h"
#include "lightspeed.hpp"

namespace lightspeed {

// ------------------------------------------------------------------ Tensor

Tensor::Tensor(const std::vector<size_t>& shape, const std::string& name) : shape_(shape), name_(name)
{
    size_t n = 1;
    for (size_t s : shape) n *= s;
    data_.assign(n, 0.0);
}

static std::string shape_str(const std::vector<size_t>& s)
{
    std::ostringstream o;
    o << "(";
    for (size_t i = 0; i < s.size(); i++) o << (i ? "," : "") << s[i];
    o << ")";
    return o.str();
}

void Tensor::shape_error(const std::vector<size_t>& shape) const
{
    if (shape != shape_)
        throw std::runtime_error("Tensor: " + name_ + " has shape " + shape_str(shape_) + ", should be " +
                                 shape_str(shape));
}

void Tensor::ndim_error(size_t ndim) const
{
    if (ndim != shape_.size())
        throw std::runtime_error("Tensor: " + name_ + " has " + std::to_string(shape_.size()) + " dims, should be " +
                                 std::to_string(ndim));
}

void Tensor::zero() { std::fill(data_.begin(), data_.end(), 0.0); }

void Tensor::scale(double a)
{
    for (auto& v : data_) v *= a;
}

void Tensor::axpby(const std::shared_ptr<Tensor>& other, double a, double b)
{
    other->shape_error(shape_);
    const double* o = other->data().data();
    for (size_t i = 0; i < data_.size(); i++) data_[i] = a * o[i] + b * data_[i];
}

std::shared_ptr<Tensor> Tensor::clone() const
{
    auto t = std::make_shared<Tensor>(shape_, name_);
    t->data() = data_;
    return t;
}

// ------------------------------------------------------------------ Primitive / Basis

bool Primitive::operator==(const Primitive& o) const
{
    return c_ == o.c_ && e_ == o.e_ && x_ == o.x_ && y_ == o.y_ && z_ == o.z_ && c0_ == o.c0_ && L_ == o.L_ &&
           is_pure_ == o.is_pure_ && aoIdx_ == o.aoIdx_ && cartIdx_ == o.cartIdx_ && primIdx_ == o.primIdx_ &&
           shellIdx_ == o.shellIdx_ && atomIdx_ == o.atomIdx_;
}

// Dense-index validation as reference src/lightspeed/basis.hpp:438-468.
Basis::Basis(const std::string& name, const std::vector<Primitive>& prims) : name_(name), primitives_(prims)
{
    for (size_t P = 0; P < prims.size(); P++) {
        const Primitive& p = prims[P];
        if (p.primIdx() != (int)P) throw std::runtime_error("Basis: primIdx field is not dense");
        if (P == 0) {
            if (p.aoIdx() != 0) throw std::runtime_error("Basis: aoIdx field is not dense");
            if (p.cartIdx() != 0) throw std::runtime_error("Basis: cartIdx field is not dense");
            if (p.shellIdx() != 0) throw std::runtime_error("Basis: shellIdx field is not dense");
            if (p.atomIdx() != 0) throw std::runtime_error("Basis: atomIdx field is not dense");
        } else {
            const Primitive& q = prims[P - 1];
            if (q.shellIdx() == p.shellIdx()) {
                if (q.aoIdx() != p.aoIdx()) throw std::runtime_error("Basis: aoIdx field is not dense");
                if (q.cartIdx() != p.cartIdx()) throw std::runtime_error("Basis: cartIdx field is not dense");
                if (q.atomIdx() != p.atomIdx()) throw std::runtime_error("Basis: atomIdx field is not dense");
            } else if (q.shellIdx() + 1 == p.shellIdx()) {
                if (q.aoIdx() + q.nao() != p.aoIdx()) throw std::runtime_error("Basis: aoIdx field is not dense");
                if (q.cartIdx() + q.ncart() != p.cartIdx()) throw std::runtime_error("Basis: cartIdx field is not dense");
                if (q.atomIdx() != p.atomIdx() && q.atomIdx() + 1 != p.atomIdx())
                    throw std::runtime_error("Basis: atomIdx field is not dense");
            } else {
                throw std::runtime_error("Basis: shellIdx field is not dense");
            }
        }
    }
    if (!prims.empty()) {
        const Primitive& p = prims.back();
        nao_ = p.aoIdx() + p.nao();
        ncart_ = p.cartIdx() + p.ncart();
        nshell_ = p.shellIdx() + 1;
        natom_ = p.atomIdx() + 1;
        nprim_ = prims.size();
    }
}

Basis::~Basis()
{
    if (dev_) lsjk_basis_destroy(dev_);
}

int Basis::max_L() const
{
    int m = 0;
    for (auto& p : primitives_) m = std::max(m, p.L());
    return m;
}

bool Basis::has_pure() const
{
    for (auto& p : primitives_)
        if (p.is_pure()) return true;
    return false;
}

bool Basis::equivalent(const std::shared_ptr<Basis>& a, const std::shared_ptr<Basis>& b)
{
    if (a->nao() != b->nao() || a->ncart() != b->ncart() || a->nprim() != b->nprim() || a->nshell() != b->nshell() ||
        a->natom() != b->natom())
        return false;
    return a->primitives() == b->primitives();
}

static void lsjk_check(int rc)
{
    if (rc != LSJK_OK) throw std::runtime_error(std::string("lightspeed_amd HIP backend: ") + lsjk_last_error());
}

lsjk_basis* Basis::device_handle() const
{
    if (!dev_) {
        hip_session_init(0);
        size_t n = primitives_.size();
        std::vector<double> c(n), e(n), xyz(3 * n);
        std::vector<int> L(n), cart(n);
        for (size_t i = 0; i < n; i++) {
            const Primitive& p = primitives_[i];
            c[i] = p.c();
            e[i] = p.e();
            xyz[3 * i] = p.x();
            xyz[3 * i + 1] = p.y();
            xyz[3 * i + 2] = p.z();
            L[i] = p.L();
            cart[i] = p.cartIdx();
        }
        lsjk_check(lsjk_basis_create((int)n, c.data(), e.data(), xyz.data(), L.data(), cart.data(), (int)ncart_, &dev_));
        std::vector<int> atom(n);
        for (size_t i = 0; i < n; i++) atom[i] = primitives_[i].atomIdx();
        lsjk_check(lsjk_basis_set_atoms(dev_, (int)natom(), atom.data()));
    }
    return dev_;
}

// ------------------------------------------------------------------ Ewald

Ewald::Ewald(const std::vector<double>& scales, const std::vector<double>& omegas) : scales_(scales), omegas_(omegas)
{
    if (scales.size() != omegas.size()) throw std::runtime_error("Ewald: scales and omegas must have the same size");
}

std::shared_ptr<Ewald> Ewald::coulomb() { return std::make_shared<Ewald>(std::vector<double>{1.0}, std::vector<double>{-1.0}); }

bool Ewald::is_coulomb() const { return scales_.size() == 1 && scales_[0] == 1.0 && omegas_[0] == -1.0; }

// ------------------------------------------------------------------ pair lists

// Pair::build_schwarz, reference src/cpp/core/pair_list.cpp:15-27 (same operation order, so the
// float bounds agree bit for bit).
Pair Pair::build_schwarz(const Primitive& p1, const Primitive& p2)
{
    double R2 = 0.0;
    R2 += std::pow(p1.x() - p2.x(), 2);
    R2 += std::pow(p1.y() - p2.y(), 2);
    R2 += std::pow(p1.z() - p2.z(), 2);
    float bound = (float)(std::fabs(p1.c() * p2.c()) * std::exp(-p1.e() * p2.e() / (p1.e() + p2.e()) * R2) *
                          std::pow(2.0, 0.25) * std::pow(M_PI / (p1.e() + p2.e()), 1.25));
    return Pair(&p1, &p2, bound);
}

bool PairListUtil::compare_prim_pairs(const Pair& a, const Pair& b)
{
    if (a.bound() == b.bound()) {
        if (a.prim1().primIdx() == b.prim1().primIdx()) return a.prim2().primIdx() < b.prim2().primIdx();
        return a.prim1().primIdx() < b.prim1().primIdx();
    }
    return a.bound() > b.bound();
}

double PairListUtil::max_bound(const std::shared_ptr<PairList>& pl)
{
    double v = 0.0;
    for (auto& l : pl->pairlists()) v = std::max(v, l.pairs().empty() ? 0.0 : (double)l.pairs()[0].bound());
    return v;
}

PairListL::PairListL(bool is_symmetric, int L1, int L2, const std::vector<Pair>& pairs)
    : is_symmetric_(is_symmetric), L1_(L1), L2_(L2), pairs_(pairs)
{
    for (size_t k = 1; k < pairs_.size(); k++)
        if (!PairListUtil::compare_prim_pairs(pairs_[k - 1], pairs_[k]))
            throw std::runtime_error(
                "PairListL: pairs not sorted according to descending bounds, prim1().primIdx() ascending, "
                "prim2().primIdx() ascending.");
}

PairList::PairList(const std::shared_ptr<Basis>& b1, const std::shared_ptr<Basis>& b2, bool sym, float thre,
                   const std::vector<PairListL>& pls)
    : basis1_(b1), basis2_(b2), is_symmetric_(sym), thre_(thre), pairlists_(pls)
{
}

PairList::~PairList()
{
    if (dev_) lsjk_pairlist_destroy(dev_);
}

size_t PairList::npair() const
{
    size_t n = 0;
    for (auto& l : pairlists_) n += l.pairs().size();
    return n;
}

// HashedGrid::significant_box, reference src/lightspeed/gridbox.hpp:132-150.
static bool significant_box(int kx, int ky, int kz, double R, double xA, double yA, double zA, double Rc)
{
    double C1x = kx * R, C2x = (kx + 1) * R, C1y = ky * R, C2y = (ky + 1) * R, C1z = kz * R, C2z = (kz + 1) * R;
    double D2 = std::pow(Rc, 2);
    D2 -= (xA < C1x) * std::pow(xA - C1x, 2);
    D2 -= (xA > C2x) * std::pow(xA - C2x, 2);
    D2 -= (yA < C1y) * std::pow(yA - C1y, 2);
    D2 -= (yA > C2y) * std::pow(yA - C2y, 2);
    D2 -= (zA < C1z) * std::pow(zA - C1z, 2);
    D2 -= (zA > C2z) * std::pow(zA - C2z, 2);
    return D2 >= 0.0;
}

// PairList::build_schwarz, reference src/cpp/core/pair_list.cpp:29-177: candidate partners are the
// primitives in 5-bohr boxes that a sphere of the largest reach Rmax touches; a pair is kept if its
// Schwarz bound exceeds threpq; the classes (L1, L2) are sorted by compare_prim_pairs.
std::shared_ptr<PairList> PairList::build_schwarz(const std::shared_ptr<Basis>& basis1,
                                                  const std::shared_ptr<Basis>& basis2, bool is_symmetric, float threpq)
{
    if (is_symmetric && (basis1->nprim() != basis2->nprim() || basis1->natom() != basis2->natom()))
        throw std::runtime_error("PairList::build_schwarz: basis sets are not symmetric");
    const auto& P1 = basis1->primitives();
    const auto& P2 = basis2->primitives();
    double Rmax = 0.0;
    for (const auto* set : {&P1, &P2})
        for (const auto& p : *set) {
            double R2 = -2.0 / p.e() * std::log(threpq / (p.c() * p.c() * (std::pow(2.0, 0.25) * std::pow(M_PI / (p.e() + p.e()), 1.25))));
            if (R2 < 0.0) continue;
            Rmax = std::max(Rmax, std::sqrt(R2));
        }
    const double Rval = 5.0;
    const double Rinv = 1.0 / Rval;
    std::map<std::tuple<int, int, int>, std::vector<int>> grid;
    for (size_t Q = 0; Q < P2.size(); Q++)
        grid[std::make_tuple((int)std::floor(P2[Q].x() * Rinv), (int)std::floor(P2[Q].y() * Rinv),
                             (int)std::floor(P2[Q].z() * Rinv))]
            .push_back((int)Q);
    const int nstep = (int)std::ceil(Rmax / Rval);
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    std::vector<std::vector<Pair>> tp(nthr);
#pragma omp parallel for schedule(dynamic)
    for (long P = 0; P < (long)P1.size(); P++) {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        const Primitive& p1 = P1[P];
        int kx = (int)std::floor(p1.x() * Rinv), ky = (int)std::floor(p1.y() * Rinv), kz = (int)std::floor(p1.z() * Rinv);
        for (int nx = -nstep; nx <= nstep; nx++)
            for (int ny = -nstep; ny <= nstep; ny++)
                for (int nz = -nstep; nz <= nstep; nz++) {
                    if (!significant_box(kx + nx, ky + ny, kz + nz, Rval, p1.x(), p1.y(), p1.z(), Rmax)) continue;
                    auto it = grid.find(std::make_tuple(kx + nx, ky + ny, kz + nz));
                    if (it == grid.end()) continue;
                    for (int Q : it->second) {
                        const Primitive& p2 = P2[Q];
                        if (is_symmetric && p2.primIdx() > p1.primIdx()) continue;
                        Pair pr = Pair::build_schwarz(p1, p2);
                        if (pr.bound() > threpq) tp[t].push_back(pr);
                    }
                }
    }
    const int ml1 = basis1->max_L(), ml2 = basis2->max_L();
    std::vector<std::vector<Pair>> byL((ml1 + 1) * (ml2 + 1));
    for (auto& v : tp)
        for (auto& pr : v) byL[pr.prim1().L() * (ml2 + 1) + pr.prim2().L()].push_back(pr);
    std::vector<PairListL> pls;
    for (int L1 = 0; L1 <= ml1; L1++)
        for (int L2 = 0; L2 <= ml2; L2++) {
            auto& v = byL[L1 * (ml2 + 1) + L2];
            if (v.empty()) continue;
            std::sort(v.begin(), v.end(), PairListUtil::compare_prim_pairs);
            pls.emplace_back(is_symmetric, L1, L2, v);
        }
    return std::make_shared<PairList>(basis1, basis2, is_symmetric, threpq, pls);
}

lsjk_pairlist* PairList::device_handle() const
{
    if (!dev_) {
        lsjk_basis* d1 = basis1_->device_handle();
        lsjk_basis* d2 = basis2_->device_handle();
        size_t n = npair();
        std::vector<int> p1, p2;
        std::vector<float> b;
        p1.reserve(n);
        p2.reserve(n);
        b.reserve(n);
        for (auto& l : pairlists_)
            for (auto& pr : l.pairs()) {
                p1.push_back(pr.prim1().primIdx());
                p2.push_back(pr.prim2().primIdx());
                b.push_back(pr.bound());
            }
        lsjk_check(lsjk_pairlist_create(d1, d2, is_symmetric_ ? 1 : 0, (int)n, p1.data(), p2.data(), b.data(), &dev_));
    }
    return dev_;
}

// ------------------------------------------------------------------ resources

std::shared_ptr<ResourceList> ResourceList::build(size_t, size_t)
{
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    std::vector<int> gpus;
    int n = hip_device_count();
    for (int i = 0; i < n; i++) gpus.push_back(i);
    return std::make_shared<ResourceList>(nthr, gpus);
}

std::shared_ptr<ResourceList> ResourceList::build_cpu()
{
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    return std::make_shared<ResourceList>(nthr, std::vector<int>{});
}

// ------------------------------------------------------------------ pure transforms

// Real solid harmonics by the recursions of Helgaker, Jorgensen & Olsen, "Molecular
// Electronic-Structure Theory", eqs. 6.4.70-6.4.73; rows ordered (0, 1c, 1s, 2c, 2s, ...) and
// Cartesian columns in lexical order (reference src/lightspeed/am.hpp:82-117).
std::vector<double> PureTransform::solid_harmonics(int L)
{
    typedef std::map<std::tuple<int, int, int>, double> Poly;
    auto add = [](Poly a, const Poly& b, double f) {
        for (auto& kv : b) a[kv.first] += f * kv.second;
        return a;
    };
    auto mul = [](const Poly& a, int dx, int dy, int dz, double f) {
        Poly o;
        for (auto& kv : a)
            o[std::make_tuple(std::get<0>(kv.first) + dx, std::get<1>(kv.first) + dy, std::get<2>(kv.first) + dz)] +=
                f * kv.second;
        return o;
    };
    std::vector<Poly> Sc(L + 1), Ss(L + 1);
    Sc[0][std::make_tuple(0, 0, 0)] = 1.0;
    for (int l = 0; l < L; l++) {
        double Nc = std::sqrt((l == 0 ? 2.0 : 1.0) * (2.0 * l + 1) / (2.0 * l + 2.0));
        double Ns = (l == 0) ? 0.0 : Nc;
        Sc[l + 1] = add(mul(Sc[l], 1, 0, 0, Nc), mul(Ss[l], 0, 1, 0, 1.0), -Ns);
        Ss[l + 1] = add(mul(Sc[l], 0, 1, 0, Nc), mul(Ss[l], 1, 0, 0, 1.0), Ns);
    }
    std::vector<Poly> rows(2 * L + 1);
    for (int m = 0; m <= L; m++) {
        Poly cc = Sc[m], cs = Ss[m], pc, ps;
        for (int l = m; l < L; l++) {
            double a = (2.0 * l + 1.0) / std::sqrt((l + m + 1.0) * (l - m + 1.0));
            double b = std::sqrt(((l + m) * (double)(l - m)) / ((l + m + 1.0) * (l - m + 1.0)));
            Poly r2c = add(add(mul(pc, 2, 0, 0, 1.0), mul(pc, 0, 2, 0, 1.0), 1.0), mul(pc, 0, 0, 2, 1.0), 1.0);
            Poly r2s = add(add(mul(ps, 2, 0, 0, 1.0), mul(ps, 0, 2, 0, 1.0), 1.0), mul(ps, 0, 0, 2, 1.0), 1.0);
            Poly nc = add(mul(cc, 0, 0, 1, a), r2c, -b);
            Poly ns = add(mul(cs, 0, 0, 1, a), r2s, -b);
            pc = cc;
            ps = cs;
            cc = nc;
            cs = ns;
        }
        if (m == 0)
            rows[0] = cc;
        else {
            rows[2 * m - 1] = cc;
            rows[2 * m] = cs;
        }
    }
    const int nc = (L + 1) * (L + 2) / 2;
    std::vector<double> C((2 * L + 1) * nc, 0.0);
    for (int r = 0; r < 2 * L + 1; r++)
        for (auto& kv : rows[r]) {
            if (std::fabs(kv.second) <= 1.0e-15) continue;
            int lx = std::get<0>(kv.first), lz = std::get<2>(kv.first);
            int i = L - lx, idx = i * (i + 1) / 2 + lz;
            C[r * nc + idx] += kv.second;
        }
    return C;
}

namespace {
// The shells of a basis with their transform blocks: C (npure x ncart, the real solid harmonics)
// for a pure shell, none (identity) for a Cartesian one.
struct ShellInfo {
    int ao, cart, L, nao, nc;
    const std::vector<double>* C;  // nullptr: Cartesian shell
};
std::vector<ShellInfo> shells_of(const Basis& b, std::map<int, std::vector<double>>& cache)
{
    std::vector<ShellInfo> s;
    int last = -1;
    for (auto& p : b.primitives())
        if (p.shellIdx() != last) {
            const int L = p.L(), nc = (L + 1) * (L + 2) / 2;
            const std::vector<double>* C = nullptr;
            if (p.is_pure()) {
                if (!cache.count(L)) cache[L] = PureTransform::solid_harmonics(L);
                C = &cache[L];
            }
            s.push_back({p.aoIdx(), p.cartIdx(), L, p.is_pure() ? 2 * L + 1 : nc, nc, C});
            last = p.shellIdx();
        }
    return s;
}
}  // namespace

// Both transforms work shell pair by shell pair, as the reference's (pure_transform.cpp:46-247):
// a block of one shell pair maps to the block of the same pair, so the cost is O(nshell^2) small
// products, not dense nao x ncart matrices.
std::shared_ptr<Tensor> PureTransform::pureToCart2(const std::shared_ptr<Basis>& b1, const std::shared_ptr<Basis>& b2,
                                                   const std::shared_ptr<Tensor>& Tpure)
{
    Tpure->shape_error({b1->nao(), b2->nao()});
    if (!b1->has_pure() && !b2->has_pure()) return Tpure;
    std::map<int, std::vector<double>> cache;
    const auto S1 = shells_of(*b1, cache), S2 = shells_of(*b2, cache);
    const size_t n2 = b2->nao(), c2 = b2->ncart();
    auto out = std::make_shared<Tensor>(std::vector<size_t>{b1->ncart(), c2});
    const double* P = Tpure->data().data();
    double* O = out->data().data();
    size_t m = 1;
    for (const auto& x : S1) m = std::max(m, (size_t)x.nc);
    for (const auto& x : S2) m = std::max(m, (size_t)x.nc);
    std::vector<double> X(m * m), Y(m * m);
    for (const auto& s1 : S1)
        for (const auto& s2 : S2) {
            // X (nao1 x nc2) = P_block C2, Y (nc1 x nc2) = C1^T X
            for (int i = 0; i < s1.nao; i++)
                for (int b = 0; b < s2.nc; b++) {
                    double v = 0.0;
                    if (s2.C) {
                        for (int k = 0; k < s2.nao; k++) v += P[(size_t)(s1.ao + i) * n2 + s2.ao + k] * (*s2.C)[k * s2.nc + b];
                    } else {
                        v = P[(size_t)(s1.ao + i) * n2 + s2.ao + b];
                    }
                    X[i * s2.nc + b] = v;
                }
            for (int a = 0; a < s1.nc; a++)
                for (int b = 0; b < s2.nc; b++) {
                    double v = 0.0;
                    if (s1.C) {
                        for (int i = 0; i < s1.nao; i++) v += (*s1.C)[i * s1.nc + a] * X[i * s2.nc + b];
                    } else {
                        v = X[a * s2.nc + b];
                    }
                    Y[a * s2.nc + b] = v;
                }
            for (int a = 0; a < s1.nc; a++)
                for (int b = 0; b < s2.nc; b++) O[(size_t)(s1.cart + a) * c2 + s2.cart + b] = Y[a * s2.nc + b];
        }
    return out;
}

std::shared_ptr<Tensor> PureTransform::cartToPure2(const std::shared_ptr<Basis>& b1, const std::shared_ptr<Basis>& b2,
                                                   const std::shared_ptr<Tensor>& Tcart,
                                                   const std::shared_ptr<Tensor>& Tpure)
{
    if (Tpure) Tpure->shape_error({b1->nao(), b2->nao()});
    Tcart->shape_error({b1->ncart(), b2->ncart()});
    if (!b1->has_pure() && !b2->has_pure()) return Tcart;
    auto out = Tpure ? Tpure : std::make_shared<Tensor>(std::vector<size_t>{b1->nao(), b2->nao()});
    std::map<int, std::vector<double>> cache;
    const auto S1 = shells_of(*b1, cache), S2 = shells_of(*b2, cache);
    const size_t n2 = b2->nao(), c2 = b2->ncart();
    const double* C = Tcart->data().data();
    double* O = out->data().data();
    size_t m = 1;
    for (const auto& x : S1) m = std::max(m, (size_t)x.nc);
    for (const auto& x : S2) m = std::max(m, (size_t)x.nc);
    std::vector<double> X(m * m);
    for (const auto& s1 : S1)
        for (const auto& s2 : S2) {
            // X (nc1 x nao2) = T_block C2^T, out_block += C1 X
            for (int a = 0; a < s1.nc; a++)
                for (int k = 0; k < s2.nao; k++) {
                    double v = 0.0;
                    if (s2.C) {
                        for (int b = 0; b < s2.nc; b++) v += C[(size_t)(s1.cart + a) * c2 + s2.cart + b] * (*s2.C)[k * s2.nc + b];
                    } else {
                        v = C[(size_t)(s1.cart + a) * c2 + s2.cart + k];
                    }
                    X[a * s2.nao + k] = v;
                }
            for (int r = 0; r < s1.nao; r++)
                for (int k = 0; k < s2.nao; k++) {
                    double v = 0.0;
                    if (s1.C) {
                        for (int a = 0; a < s1.nc; a++) v += (*s1.C)[r * s1.nc + a] * X[a * s2.nao + k];
                    } else {
                        v = X[r * s2.nao + k];
                    }
                    O[(size_t)(s1.ao + r) * n2 + s2.ao + k] += v;
                }
        }
    return out;
}

// ------------------------------------------------------------------ session

void hip_session_init(int device) { lsjk_check(lsjk_init(device)); }

void hip_session_finalize() { lsjk_check(lsjk_finalize()); }

int hip_device_count()
{
    int n = 0;
    if (lsjk_device_count(&n) != LSJK_OK) return 0;
    return n;
}

}  // namespace lightspeed

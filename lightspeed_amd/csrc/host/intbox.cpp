// IntBox::coulomb / IntBox::exchange: the reference's J/K entry points (reference
// src/lightspeed/intbox.hpp:721-828, dispatch src/cpp/intbox/intbox.cpp:213-329) routed to the HIP
// library.  Pure <-> Cartesian transforms happen here on the host, as reference coulombCPU does
// around its kernels (src/cpp/intbox/cpu/coulomb.cpp:26-35, 186-193).
#include <algorithm>
#include <cstring>

#include "../../../include/lsjk.h"
#include "lightspeed.hpp"

namespace lightspeed {

namespace {

void check(int rc)
{
    if (rc != LSJK_OK) throw std::runtime_error(std::string("lightspeed_amd HIP backend: ") + lsjk_last_error());
}

// RAII device buffer
struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) { check(lsjk_malloc(&p, bytes)); }
    ~DevBuf()
    {
        if (p) lsjk_free(p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    double* d() const { return (double*)p; }
};

void ewald_arrays(const std::shared_ptr<Ewald>& ewald, std::vector<double>& s, std::vector<double>& o)
{
    s = ewald->scales();
    o = ewald->omegas();
}

void require_ready(const char* who, const std::shared_ptr<ResourceList>& res, const std::shared_ptr<PairList>& pl12,
                   const std::shared_ptr<PairList>& pl34)
{
    std::string why;
    if (!IntBox::is_hip_ready(res, pl12, pl34, &why))
        throw std::runtime_error(std::string(who) + ": HIP backend unavailable (" + why + ")");
    hip_session_init(res->gpus()[0]);
}

// accumulate a freshly computed Cartesian result into the caller's tensor (reference semantics:
// allocCart2 aliases J12 for Cartesian bases, cartToPure2 accumulates for pure ones)
std::shared_ptr<Tensor> finish(const std::shared_ptr<Basis>& b1, const std::shared_ptr<Basis>& b2,
                               const std::shared_ptr<Tensor>& Tcart, const std::shared_ptr<Tensor>& Tout)
{
    if (!b1->has_pure() && !b2->has_pure()) {
        if (!Tout) return Tcart;
        Tout->axpby(Tcart, 1.0, 1.0);
        return Tout;
    }
    return PureTransform::cartToPure2(b1, b2, Tcart, Tout);
}

}  // namespace

bool IntBox::is_hip_ready(const std::shared_ptr<ResourceList>& res, const std::shared_ptr<PairList>& pl12,
                          const std::shared_ptr<PairList>& pl34, std::string* why)
{
    auto fail = [&](const char* m) {
        if (why) *why = m;
        return false;
    };
    if (!res || res->ngpu() == 0) return fail("the ResourceList has no GPU");
    if (pl12->basis1()->max_L() > 2 || pl12->basis2()->max_L() > 2 || pl34->basis1()->max_L() > 2 ||
        pl34->basis2()->max_L() > 2)
        return fail("angular momentum above d is not supported");
    return true;
}

std::shared_ptr<Tensor> IntBox::coulomb(const std::shared_ptr<ResourceList>& resources,
                                        const std::shared_ptr<Ewald>& ewald, const std::shared_ptr<PairList>& pl12,
                                        const std::shared_ptr<PairList>& pl34, const std::shared_ptr<Tensor>& D34,
                                        float thresp, float thredp, const std::shared_ptr<Tensor>& J12)
{
    auto b1 = pl12->basis1(), b2 = pl12->basis2(), b3 = pl34->basis1(), b4 = pl34->basis2();
    D34->shape_error({b3->nao(), b4->nao()});
    if (J12) J12->shape_error({b1->nao(), b2->nao()});
    require_ready("IntBox::coulomb", resources, pl12, pl34);
    auto Dc = PureTransform::pureToCart2(b3, b4, D34);
    const size_t nD = b3->ncart() * b4->ncart(), nJ = b1->ncart() * b2->ncart();
    DevBuf dD(nD * sizeof(double)), dJ(nJ * sizeof(double));
    check(lsjk_memcpy_htod(dD.p, Dc->data().data(), nD * sizeof(double)));
    check(lsjk_memset_zero(dJ.p, nJ * sizeof(double)));
    coulomb_device(ewald, pl12, pl34, dD.d(), thresp, thredp, dJ.d());
    auto Jc = std::make_shared<Tensor>(std::vector<size_t>{b1->ncart(), b2->ncart()}, "J");
    check(lsjk_memcpy_dtoh(Jc->data().data(), dJ.p, nJ * sizeof(double)));
    return finish(b1, b2, Jc, J12);
}

std::shared_ptr<Tensor> IntBox::exchange(const std::shared_ptr<ResourceList>& resources,
                                         const std::shared_ptr<Ewald>& ewald, const std::shared_ptr<PairList>& pl12,
                                         const std::shared_ptr<PairList>& pl34, const std::shared_ptr<Tensor>& D24,
                                         bool D24symm, float thresp, float thredp, const std::shared_ptr<Tensor>& K13)
{
    auto b1 = pl12->basis1(), b2 = pl12->basis2(), b3 = pl34->basis1(), b4 = pl34->basis2();
    D24->shape_error({b2->nao(), b4->nao()});
    if (K13) K13->shape_error({b1->nao(), b3->nao()});
    // reference exchange.cpp:73-81
    if (D24symm && !Basis::equivalent(b2, b4))
        throw std::runtime_error("IntBox::exchange: D24symm implied, but basis sets 2 and 4 not the same");
    if (D24symm) D24symm = Basis::equivalent(b1, b3);
    require_ready("IntBox::exchange", resources, pl12, pl34);
    auto Dc = PureTransform::pureToCart2(b2, b4, D24);
    const size_t nD = b2->ncart() * b4->ncart(), nK = b1->ncart() * b3->ncart();
    DevBuf dD(nD * sizeof(double)), dK(nK * sizeof(double));
    check(lsjk_memcpy_htod(dD.p, Dc->data().data(), nD * sizeof(double)));
    check(lsjk_memset_zero(dK.p, nK * sizeof(double)));
    exchange_device(ewald, pl12, pl34, dD.d(), D24symm, thresp, thredp, dK.d());
    auto Kc = std::make_shared<Tensor>(std::vector<size_t>{b1->ncart(), b3->ncart()}, "K");
    check(lsjk_memcpy_dtoh(Kc->data().data(), dK.p, nK * sizeof(double)));
    return finish(b1, b3, Kc, K13);
}

namespace {
enum class OneE { S, T, V };

std::shared_ptr<Tensor> one_electron(OneE kind, const char* who, const std::shared_ptr<ResourceList>& resources,
                                     const std::shared_ptr<Ewald>& ewald, const std::shared_ptr<PairList>& pl,
                                     const std::shared_ptr<Tensor>& xyzc, const std::shared_ptr<Tensor>& Out)
{
    auto b1 = pl->basis1(), b2 = pl->basis2();
    if (Out) Out->shape_error({b1->nao(), b2->nao()});
    if (xyzc && (xyzc->shape().size() != 2 || xyzc->shape()[1] != 4))
        throw std::runtime_error(std::string(who) + ": xyzc must be an (nA, 4) Tensor");
    require_ready(who, resources, pl, pl);
    const size_t n = b1->ncart() * b2->ncart();
    DevBuf dM(std::max<size_t>(n, 1) * sizeof(double));
    check(lsjk_memset_zero(dM.p, std::max<size_t>(n, 1) * sizeof(double)));
    lsjk_pairlist* d = pl->device_handle();
    if (kind == OneE::S) {
        check(lsjk_overlap(d, dM.d()));
    } else if (kind == OneE::T) {
        check(lsjk_kinetic(d, dM.d()));
    } else {
        std::vector<double> s, o;
        ewald_arrays(ewald, s, o);
        check(lsjk_potential(d, (int)s.size(), s.data(), o.data(), (int)xyzc->shape()[0], xyzc->data().data(),
                             dM.d()));
    }
    auto Mc = std::make_shared<Tensor>(std::vector<size_t>{b1->ncart(), b2->ncart()}, "M");
    check(lsjk_memcpy_dtoh(Mc->data().data(), dM.p, n * sizeof(double)));
    return finish(b1, b2, Mc, Out);
}
}  // namespace

std::shared_ptr<Tensor> IntBox::overlap(const std::shared_ptr<ResourceList>& resources,
                                        const std::shared_ptr<PairList>& pairlist, const std::shared_ptr<Tensor>& S)
{
    return one_electron(OneE::S, "IntBox::overlap", resources, nullptr, pairlist, nullptr, S);
}

std::shared_ptr<Tensor> IntBox::kinetic(const std::shared_ptr<ResourceList>& resources,
                                        const std::shared_ptr<PairList>& pairlist, const std::shared_ptr<Tensor>& T)
{
    return one_electron(OneE::T, "IntBox::kinetic", resources, nullptr, pairlist, nullptr, T);
}

std::shared_ptr<Tensor> IntBox::potential(const std::shared_ptr<ResourceList>& resources,
                                          const std::shared_ptr<Ewald>& ewald,
                                          const std::shared_ptr<PairList>& pairlist,
                                          const std::shared_ptr<Tensor>& xyzc, const std::shared_ptr<Tensor>& V)
{
    if (!xyzc) throw std::runtime_error("IntBox::potential: xyzc is required");
    return one_electron(OneE::V, "IntBox::potential", resources, ewald, pairlist, xyzc, V);
}

std::shared_ptr<Tensor> IntBox::coulombGrad(const std::shared_ptr<ResourceList>& resources,
                                            const std::shared_ptr<Ewald>& ewald, const std::shared_ptr<PairList>& pl,
                                            const std::shared_ptr<Tensor>& D12, const std::shared_ptr<Tensor>& D34,
                                            float thresp, float thredp, const std::shared_ptr<Tensor>& G)
{
    if (!pl->is_symmetric()) throw std::runtime_error("IntBox::coulombGrad: pairlist must be symmetric");
    auto b = pl->basis1();
    D12->shape_error({b->nao(), b->nao()});
    D34->shape_error({b->nao(), b->nao()});
    const size_t natom = b->natom();
    if (G) G->shape_error({natom, 3});
    require_ready("IntBox::coulombGrad", resources, pl, pl);
    auto Dc12 = PureTransform::pureToCart2(b, b, D12);
    auto Dc34 = PureTransform::pureToCart2(b, b, D34);
    const size_t nD = b->ncart() * b->ncart();
    DevBuf d12(nD * sizeof(double)), d34(nD * sizeof(double)), dG(std::max<size_t>(1, 3 * natom) * sizeof(double));
    check(lsjk_memcpy_htod(d12.p, Dc12->data().data(), nD * sizeof(double)));
    check(lsjk_memcpy_htod(d34.p, Dc34->data().data(), nD * sizeof(double)));
    check(lsjk_memset_zero(dG.p, std::max<size_t>(1, 3 * natom) * sizeof(double)));
    std::vector<double> s, o;
    ewald_arrays(ewald, s, o);
    lsjk_pairlist* d = pl->device_handle();
    // bra-centre derivatives with each density in the bra role (d/dR of (12|34) = bra part + ket
    // part, and the ket part of (12|34) is the bra part of (34|12))
    check(lsjk_coulomb_grad(d, d12.d(), d34.d(), (int)s.size(), s.data(), o.data(), thresp, thredp, dG.d()));
    check(lsjk_coulomb_grad(d, d34.d(), d12.d(), (int)s.size(), s.data(), o.data(), thresp, thredp, dG.d()));
    auto Gn = std::make_shared<Tensor>(std::vector<size_t>{natom, 3}, "G");
    check(lsjk_memcpy_dtoh(Gn->data().data(), dG.p, 3 * natom * sizeof(double)));
    if (!G) return Gn;
    G->axpby(Gn, 1.0, 1.0);
    return G;
}

std::shared_ptr<Tensor> IntBox::exchangeGrad(const std::shared_ptr<ResourceList>& resources,
                                             const std::shared_ptr<Ewald>& ewald, const std::shared_ptr<PairList>& pl,
                                             const std::shared_ptr<Tensor>& D13, const std::shared_ptr<Tensor>& D24,
                                             bool D13symm, bool D24symm, bool Dsame, float thresp, float thredp,
                                             const std::shared_ptr<Tensor>& G)
{
    (void)D13symm;
    (void)D24symm;
    (void)Dsame;
    if (!pl->is_symmetric()) throw std::runtime_error("IntBox::exchangeGrad: pairlist must be symmetric");
    auto b = pl->basis1();
    D13->shape_error({b->nao(), b->nao()});
    D24->shape_error({b->nao(), b->nao()});
    const size_t natom = b->natom();
    if (G) G->shape_error({natom, 3});
    require_ready("IntBox::exchangeGrad", resources, pl, pl);
    auto Dc13 = PureTransform::pureToCart2(b, b, D13);
    auto Dc24 = PureTransform::pureToCart2(b, b, D24);
    const size_t nD = b->ncart() * b->ncart();
    DevBuf d13(nD * sizeof(double)), d24(nD * sizeof(double)), dG(std::max<size_t>(1, 3 * natom) * sizeof(double));
    check(lsjk_memcpy_htod(d13.p, Dc13->data().data(), nD * sizeof(double)));
    check(lsjk_memcpy_htod(d24.p, Dc24->data().data(), nD * sizeof(double)));
    check(lsjk_memset_zero(dG.p, std::max<size_t>(1, 3 * natom) * sizeof(double)));
    std::vector<double> s, o;
    ewald_arrays(ewald, s, o);
    check(lsjk_exchange_grad(pl->device_handle(), d13.d(), d24.d(), (int)s.size(), s.data(), o.data(), thresp, thredp,
                             dG.d()));
    auto Gn = std::make_shared<Tensor>(std::vector<size_t>{natom, 3}, "G");
    check(lsjk_memcpy_dtoh(Gn->data().data(), dG.p, 3 * natom * sizeof(double)));
    if (!G) return Gn;
    G->axpby(Gn, 1.0, 1.0);
    return G;
}

std::shared_ptr<Tensor> IntBox::eriJ(const std::shared_ptr<ResourceList>& resources,
                                     const std::shared_ptr<Ewald>& ewald, const std::shared_ptr<PairList>& pl12,
                                     const std::shared_ptr<PairList>& pl34, const std::shared_ptr<Tensor>& C1,
                                     const std::shared_ptr<Tensor>& C2, const std::shared_ptr<Tensor>& C3,
                                     const std::shared_ptr<Tensor>& C4, float thresp, float thredp,
                                     const std::shared_ptr<Tensor>& I)
{
    C1->ndim_error(2);
    C2->ndim_error(2);
    C3->ndim_error(2);
    C4->ndim_error(2);
    const size_t nao1 = pl12->basis1()->nao(), nao2 = pl12->basis2()->nao();
    const size_t nao3 = pl34->basis1()->nao(), nao4 = pl34->basis2()->nao();
    const size_t n1 = C1->shape()[1], n2 = C2->shape()[1], n3 = C3->shape()[1], n4 = C4->shape()[1];
    C1->shape_error({nao1, n1});
    C2->shape_error({nao2, n2});
    C3->shape_error({nao3, n3});
    C4->shape_error({nao4, n4});
    std::shared_ptr<Tensor> I2 = I ? I : std::make_shared<Tensor>(std::vector<size_t>{n1, n2, n3, n4}, "I");
    I2->shape_error({n1, n2, n3, n4});
    double* Ip = I2->data().data();
    const double *c1 = C1->data().data(), *c2 = C2->data().data(), *c3 = C3->data().data(), *c4 = C4->data().data();
    const bool symm12 = (C1 == C2);
    auto D = std::make_shared<Tensor>(std::vector<size_t>{nao1, nao2}, "D");
    std::vector<double> T(nao3 * n4);
    for (size_t i = 0; i < n1; i++) {
        for (size_t j = 0; j < n2; j++) {
            if (symm12 && i < j) continue;
            double* Dp = D->data().data();
            for (size_t p = 0; p < nao1; p++)
                for (size_t q = 0; q < nao2; q++) Dp[p * nao2 + q] = c1[p * n1 + i] * c2[q * n2 + j];
            // J over the (3,4) pairs from the (1,2) density: (pl34 | pl12) as in mo_eri.cpp:89-97
            auto J = IntBox::coulomb(resources, ewald, pl34, pl12, D, thresp, thredp);
            const double* Jp = J->data().data();
            // L = C3^T J C4 (n3 x n4)
            std::fill(T.begin(), T.end(), 0.0);
            for (size_t r = 0; r < nao3; r++)
                for (size_t s = 0; s < nao4; s++) {
                    const double v = Jp[r * nao4 + s];
                    if (v == 0.0) continue;
                    for (size_t l = 0; l < n4; l++) T[r * n4 + l] += v * c4[s * n4 + l];
                }
            double* dst = Ip + i * (n2 * n3 * n4) + j * (n3 * n4);
            for (size_t k = 0; k < n3; k++)
                for (size_t l = 0; l < n4; l++) {
                    double v = 0.0;
                    for (size_t r = 0; r < nao3; r++) v += c3[r * n3 + k] * T[r * n4 + l];
                    dst[k * n4 + l] = v;
                }
            if (symm12 && i != j) std::memcpy(Ip + j * (n2 * n3 * n4) + i * (n3 * n4), dst, n3 * n4 * sizeof(double));
        }
    }
    // symmetrisations of mo_eri.cpp:120-155
    if (C1 == C2)
        for (size_t rs = 0; rs < n3 * n4; rs++)
            for (size_t p = 0; p < n1; p++)
                for (size_t q = 0; q < p; q++) {
                    double& a = Ip[p * n2 * n3 * n4 + q * n3 * n4 + rs];
                    double& b = Ip[q * n2 * n3 * n4 + p * n3 * n4 + rs];
                    a = b = 0.5 * (a + b);
                }
    if (C3 == C4)
        for (size_t pq = 0; pq < n1 * n2; pq++) {
            double* I2p = Ip + pq * n3 * n4;
            for (size_t r = 0; r < n3; r++)
                for (size_t s = 0; s < r; s++) {
                    double& a = I2p[r * n4 + s];
                    double& b = I2p[s * n4 + r];
                    a = b = 0.5 * (a + b);
                }
        }
    if (C1 == C3 && C2 == C4)
        for (size_t pq = 0; pq < n1 * n2; pq++)
            for (size_t rs = 0; rs < pq; rs++) {
                double& a = Ip[pq * n3 * n4 + rs];
                double& b = Ip[rs * n3 * n4 + pq];
                a = b = 0.5 * (a + b);
            }
    return I2;
}

void IntBox::coulomb_device(const std::shared_ptr<Ewald>& ewald, const std::shared_ptr<PairList>& pl12,
                            const std::shared_ptr<PairList>& pl34, const double* D34_dev, float thresp, float thredp,
                            double* J12_dev)
{
    std::vector<double> s, o;
    ewald_arrays(ewald, s, o);
    lsjk_pairlist* d12 = pl12->device_handle();
    lsjk_pairlist* d34 = (pl34 == pl12) ? d12 : pl34->device_handle();
    check(lsjk_coulomb(d12, d34, D34_dev, (int)s.size(), s.data(), o.data(), thresp, thredp, J12_dev));
}

void IntBox::exchange_device(const std::shared_ptr<Ewald>& ewald, const std::shared_ptr<PairList>& pl12,
                             const std::shared_ptr<PairList>& pl34, const double* D24_dev, bool D24symm, float thresp,
                             float thredp, double* K13_dev)
{
    std::vector<double> s, o;
    ewald_arrays(ewald, s, o);
    lsjk_pairlist* d12 = pl12->device_handle();
    lsjk_pairlist* d34 = (pl34 == pl12) ? d12 : pl34->device_handle();
    check(lsjk_exchange(d12, d34, D24_dev, D24symm ? 1 : 0, (int)s.size(), s.data(), o.data(), thresp, thredp,
                        K13_dev));
}

}  // namespace lightspeed

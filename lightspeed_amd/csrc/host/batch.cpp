// JKBatch: the J and K builds of many independent molecules together, through the HIP library's
// batch entry points (lsjk_batch_*).  The reference has no batched path: its callers run one
// IntBox::coulomb + IntBox::exchange per molecule (reference est/rhf.py:616-643 per SCF iteration;
// src/cpp/intbox/mo_eri.cpp:10-250 per pair density).  Per molecule the results are those of
// IntBox::coulomb(res, ewald, pl, pl, D, thresp, thredp) and IntBox::exchange(res, ewald, pl, pl,
// D, true, thresp, thredp) on that molecule's own symmetric pair list; pure bases are transformed
// on the host around the call, as IntBox::coulomb does.
#include <cstring>

#include "../../../include/lsjk.h"
#include "lightspeed.hpp"

namespace lightspeed {

namespace {
void check(int rc)
{
    if (rc != LSJK_OK) throw std::runtime_error(std::string("lightspeed_amd HIP backend: ") + lsjk_last_error());
}
}  // namespace

std::shared_ptr<JKBatch> JKBatch::build(const std::shared_ptr<ResourceList>& resources,
                                        const std::vector<std::shared_ptr<PairList>>& pairlists)
{
    if (pairlists.empty()) throw std::runtime_error("JKBatch::build: no molecules");
    for (const auto& pl : pairlists) {
        if (!pl || !pl->is_symmetric() || pl->basis1() != pl->basis2())
            throw std::runtime_error("JKBatch::build: each molecule needs a symmetric pair list over one basis");
        std::string why;
        if (!IntBox::is_hip_ready(resources, pl, pl, &why))
            throw std::runtime_error("JKBatch::build: HIP backend unavailable (" + why + ")");
    }
    hip_session_init(resources->gpus()[0]);
    std::vector<const lsjk_pairlist*> h;
    for (const auto& pl : pairlists) h.push_back(pl->device_handle());
    auto b = std::shared_ptr<JKBatch>(new JKBatch());
    b->pairlists_ = pairlists;
    check(lsjk_batch_create((int)h.size(), h.data(), &b->dev_));
    b->offsets_.assign(h.size(), 0);
    check(lsjk_batch_layout(b->dev_, &b->ld_, &b->nelem_, b->offsets_.data()));
    return b;
}

JKBatch::~JKBatch()
{
    if (dev_) lsjk_batch_destroy(dev_);
}

void JKBatch::jk_device(const std::shared_ptr<Ewald>& ewald, const double* D_dev, float thresp, float thredp,
                        double* J_dev, double* K_dev)
{
    const auto& s = ewald->scales();
    const auto& o = ewald->omegas();
    check(lsjk_batch_jk(dev_, D_dev, (int)s.size(), s.data(), o.data(), thresp, thredp, J_dev, K_dev));
}

std::vector<std::tuple<int, int, double, double, double>> JKBatch::autotune(const std::shared_ptr<Ewald>& ewald,
                                                                            const double* D_dev, float thresp,
                                                                            float thredp)
{
    const auto& s = ewald->scales();
    const auto& o = ewald->omegas();
    check(lsjk_batch_autotune(dev_, D_dev, (int)s.size(), s.data(), o.data(), thresp, thredp));
    const lsjk_pairlist* pl = nullptr;
    check(lsjk_batch_pairlist(dev_, &pl));
    int n = 0;
    check(lsjk_tuned_params(pl, 0, &n, nullptr, nullptr, nullptr, nullptr, nullptr));
    std::vector<int> kind(std::max(n, 1)), idx(std::max(n, 1));
    std::vector<double> val(std::max(n, 1)), m0(std::max(n, 1)), m1(std::max(n, 1));
    check(lsjk_tuned_params(pl, n, &n, kind.data(), idx.data(), val.data(), m0.data(), m1.data()));
    std::vector<std::tuple<int, int, double, double, double>> out;
    for (int i = 0; i < n; i++) out.emplace_back(kind[i], idx[i], val[i], m0[i], m1[i]);
    return out;
}

std::pair<std::vector<std::shared_ptr<Tensor>>, std::vector<std::shared_ptr<Tensor>>> JKBatch::jk(
    const std::shared_ptr<Ewald>& ewald, const std::vector<std::shared_ptr<Tensor>>& Ds, float thresp, float thredp)
{
    const size_t nmol = pairlists_.size();
    if (Ds.size() != nmol) throw std::runtime_error("JKBatch::jk: one density per molecule is needed");
    std::vector<double> h(nelem_, 0.0);
    for (size_t m = 0; m < nmol; m++) {
        auto b = pairlists_[m]->basis1();
        Ds[m]->shape_error({b->nao(), b->nao()});
        auto Dc = PureTransform::pureToCart2(b, b, Ds[m]);
        const size_t n = b->ncart();
        for (size_t i = 0; i < n; i++)
            std::memcpy(&h[offsets_[m] + i * ld_], Dc->data().data() + i * n, n * sizeof(double));
    }
    void *dD = nullptr, *dJ = nullptr, *dK = nullptr;
    std::vector<double> hj(nelem_), hk(nelem_);
    try {
        check(lsjk_malloc(&dD, nelem_ * sizeof(double)));
        check(lsjk_malloc(&dJ, nelem_ * sizeof(double)));
        check(lsjk_malloc(&dK, nelem_ * sizeof(double)));
        check(lsjk_memcpy_htod(dD, h.data(), nelem_ * sizeof(double)));
        check(lsjk_memset_zero(dJ, nelem_ * sizeof(double)));
        check(lsjk_memset_zero(dK, nelem_ * sizeof(double)));
        jk_device(ewald, (const double*)dD, thresp, thredp, (double*)dJ, (double*)dK);
        check(lsjk_memcpy_dtoh(hj.data(), dJ, nelem_ * sizeof(double)));
        check(lsjk_memcpy_dtoh(hk.data(), dK, nelem_ * sizeof(double)));
    } catch (...) {
        for (void* p : {dD, dJ, dK})
            if (p) lsjk_free(p);
        throw;
    }
    for (void* p : {dD, dJ, dK}) lsjk_free(p);
    std::vector<std::shared_ptr<Tensor>> Js, Ks;
    for (size_t m = 0; m < nmol; m++) {
        auto b = pairlists_[m]->basis1();
        const size_t n = b->ncart();
        auto Jc = std::make_shared<Tensor>(std::vector<size_t>{n, n}, "J");
        auto Kc = std::make_shared<Tensor>(std::vector<size_t>{n, n}, "K");
        for (size_t i = 0; i < n; i++) {
            std::memcpy(Jc->data().data() + i * n, &hj[offsets_[m] + i * ld_], n * sizeof(double));
            std::memcpy(Kc->data().data() + i * n, &hk[offsets_[m] + i * ld_], n * sizeof(double));
        }
        if (b->has_pure()) {
            Js.push_back(PureTransform::cartToPure2(b, b, Jc, std::shared_ptr<Tensor>()));
            Ks.push_back(PureTransform::cartToPure2(b, b, Kc, std::shared_ptr<Tensor>()));
        } else {
            Js.push_back(Jc);
            Ks.push_back(Kc);
        }
    }
    return {Js, Ks};
}

std::shared_ptr<Tensor> Boys::compute(int nboys, const std::shared_ptr<Tensor>& T)
{
    if (nboys < 0 || nboys > 16) throw std::runtime_error("Boys::compute: nboys must be 0..16");
    const size_t n = T->size();
    hip_session_init(0);
    auto F = std::make_shared<Tensor>(std::vector<size_t>{n, (size_t)nboys + 1}, "F");
    void *dT = nullptr, *dF = nullptr;
    try {
        check(lsjk_malloc(&dT, std::max<size_t>(n, 1) * sizeof(double)));
        check(lsjk_malloc(&dF, std::max<size_t>(n * (nboys + 1), 1) * sizeof(double)));
        if (n) check(lsjk_memcpy_htod(dT, T->data().data(), n * sizeof(double)));
        check(lsjk_boys(nboys, (int)n, (const double*)dT, (double*)dF));
        if (n) check(lsjk_memcpy_dtoh(F->data().data(), dF, n * (nboys + 1) * sizeof(double)));
    } catch (...) {
        if (dT) lsjk_free(dT);
        if (dF) lsjk_free(dF);
        throw;
    }
    lsjk_free(dT);
    lsjk_free(dF);
    return F;
}

}  // namespace lightspeed

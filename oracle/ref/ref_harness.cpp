// TEST INFRASTRUCTURE ONLY -- a driver (this project's code) around the reference's own CPU
// J/K path, compiled from the reference sources where they lie (oracle/ref/Makefile).  It reads a
// binary problem file, builds the reference's Basis and PairList, calls IntBox::coulombCPU and
// IntBox::exchangeCPU (reference src/cpp/intbox/cpu/coulomb.cpp:14-196, exchange.cpp:17-355) and
// writes J, K and timings.  Used to pin the oracle and as the "reference" CPU baseline.
//
// Problem file (little endian): int32 nprim; nprim x {f64 c, e, x, y, z, c0; i32 L, is_pure, aoIdx,
// cartIdx, primIdx, shellIdx, atomIdx}; int32 nao; f64 D[nao*nao]; f32 threpq, thresp;
// int32 symmetric, D24symm, nrep, do_J, do_K, stride.  With stride > 1 the bra list (12) keeps
// every stride-th pair of each angular-momentum class (a bounded sample of the workload); the ket
// list (34) is the whole list.
// Output file: f64 J[nao*nao], K[nao*nao], tJ, tK (median seconds), tPL (pair list seconds).
#include <lightspeed/basis.hpp>
#include <lightspeed/ewald.hpp>
#include <lightspeed/intbox.hpp>
#include <lightspeed/pair_list.hpp>
#include <lightspeed/resource_list.hpp>
#include <lightspeed/tensor.hpp>
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

using namespace lightspeed;

template <typename T>
static T rd(FILE* f)
{
    T v;
    if (fread(&v, sizeof(T), 1, f) != 1) throw std::runtime_error("ref_harness: short read");
    return v;
}

static double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char** argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: ref_harness problem.bin out.bin\n");
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 3;
    int nprim = rd<int>(f);
    std::vector<Primitive> prims;
    for (int i = 0; i < nprim; i++) {
        double c = rd<double>(f), e = rd<double>(f), x = rd<double>(f), y = rd<double>(f), z = rd<double>(f),
               c0 = rd<double>(f);
        int L = rd<int>(f), pure = rd<int>(f), ao = rd<int>(f), cart = rd<int>(f), pi = rd<int>(f), si = rd<int>(f),
            ai = rd<int>(f);
        prims.push_back(Primitive(c, e, x, y, z, c0, L, pure != 0, ao, cart, pi, si, ai));
    }
    int nao = rd<int>(f);
    std::shared_ptr<Tensor> D(new Tensor({(size_t)nao, (size_t)nao}));
    if (fread(D->data().data(), sizeof(double), (size_t)nao * nao, f) != (size_t)nao * nao) return 4;
    float threpq = rd<float>(f), thresp = rd<float>(f);
    int symmetric = rd<int>(f), D24symm = rd<int>(f), nrep = rd<int>(f), doJ = rd<int>(f), doK = rd<int>(f);
    int stride = rd<int>(f);
    fclose(f);

    std::shared_ptr<Basis> bas(new Basis("basis", prims));
    std::shared_ptr<ResourceList> res(new ResourceList(omp_get_max_threads(), {}));
    std::shared_ptr<Ewald> ew(new Ewald({1.0}, {-1.0}));
    double t0 = now();
    std::shared_ptr<PairList> pl = PairList::build_schwarz(bas, bas, symmetric != 0, threpq);
    double tPL = now() - t0;
    std::shared_ptr<PairList> bra = pl;
    if (stride > 1) {
        std::vector<PairListL> sub;
        for (const PairListL& l : pl->pairlists()) {
            std::vector<Pair> keep;
            for (size_t i = 0; i < l.pairs().size(); i += stride) keep.push_back(l.pairs()[i]);
            sub.push_back(PairListL(l.is_symmetric(), l.L1(), l.L2(), keep));
        }
        bra = std::make_shared<PairList>(bas, bas, pl->is_symmetric(), pl->thre(), sub);
    }
    std::shared_ptr<Tensor> J, K;
    std::vector<double> tj, tk;
    for (int r = 0; r < std::max(1, nrep); r++) {
        if (doJ) {
            double a = now();
            J = IntBox::coulombCPU(res, ew, bra, pl, D, thresp, thresp);
            tj.push_back(now() - a);
        }
        if (doK) {
            double a = now();
            K = IntBox::exchangeCPU(res, ew, bra, pl, D, D24symm != 0, thresp, thresp);
            tk.push_back(now() - a);
        }
    }
    auto med = [](std::vector<double> v) {
        if (v.empty()) return 0.0;
        std::sort(v.begin(), v.end());
        return v[v.size() / 2];
    };
    FILE* o = fopen(argv[2], "wb");
    if (!o) return 5;
    std::vector<double> zero((size_t)nao * nao, 0.0);
    fwrite(J ? J->data().data() : zero.data(), sizeof(double), (size_t)nao * nao, o);
    fwrite(K ? K->data().data() : zero.data(), sizeof(double), (size_t)nao * nao, o);
    double tt[3] = {med(tj), med(tk), tPL};
    fwrite(tt, sizeof(double), 3, o);
    fclose(o);
    return 0;
}

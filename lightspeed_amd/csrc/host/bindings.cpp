// Python bindings (pybind11) of the host API: the call signatures of the reference's Boost.Python
// module for the J/K path (reference src/python/export_intbox.cpp:36-51, 127-148;
// src/python/export_core.cpp:256-264; src/python/export_tensor.cpp:8-23, 46-57).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>

#include "../../../include/lsjk.h"
#include "lightspeed.hpp"

namespace py = pybind11;
using namespace lightspeed;

static std::shared_ptr<Tensor> tensor_from_array(py::array arr)
{
    py::array_t<double, py::array::c_style | py::array::forcecast> a(arr);
    std::vector<size_t> shape(a.ndim());
    for (py::ssize_t i = 0; i < a.ndim(); i++) shape[i] = (size_t)a.shape(i);
    auto t = std::make_shared<Tensor>(shape);
    if (t->size()) std::memcpy(t->data().data(), a.data(), t->size() * sizeof(double));
    return t;
}

static py::array tensor_view(const std::shared_ptr<Tensor>& t)
{
    std::vector<py::ssize_t> shape(t->shape().begin(), t->shape().end());
    std::vector<py::ssize_t> strides(shape.size());
    py::ssize_t s = sizeof(double);
    for (int i = (int)shape.size() - 1; i >= 0; i--) {
        strides[i] = s;
        s *= shape[i];
    }
    return py::array(py::dtype::of<double>(), shape, strides, t->data().data(), py::cast(t));
}

static void check(int rc)
{
    if (rc != LSJK_OK) throw std::runtime_error(std::string("lightspeed_amd HIP backend: ") + lsjk_last_error());
}

// FP64 device (HBM) array owned through the library's allocator; the HIP runtime of the library is
// the only one in the process, so HBM-resident inputs are made here rather than by another framework.
struct DeviceArray {
    std::vector<size_t> shape;
    size_t n = 0;
    void* p = nullptr;
    explicit DeviceArray(std::vector<size_t> s) : shape(std::move(s))
    {
        hip_session_init(0);
        n = 1;
        for (size_t v : shape) n *= v;
        check(lsjk_malloc(&p, std::max<size_t>(n, 1) * sizeof(double)));
        check(lsjk_memset_zero(p, std::max<size_t>(n, 1) * sizeof(double)));
    }
    ~DeviceArray()
    {
        if (p) lsjk_free(p);
    }
    void upload(py::array a)
    {
        py::array_t<double, py::array::c_style | py::array::forcecast> b(a);
        if ((size_t)b.size() != n) throw std::runtime_error("DeviceArray.upload: size mismatch");
        check(lsjk_memcpy_htod(p, b.data(), n * sizeof(double)));
    }
    py::array_t<double> download() const
    {
        std::vector<py::ssize_t> s(shape.begin(), shape.end());
        py::array_t<double> out(s);
        check(lsjk_memcpy_dtoh(out.mutable_data(), p, n * sizeof(double)));
        return out;
    }
    void zero() { check(lsjk_memset_zero(p, std::max<size_t>(n, 1) * sizeof(double))); }
    void copy_from(const DeviceArray& o)
    {
        if (o.n != n) throw std::runtime_error("DeviceArray.copy_from: size mismatch");
        check(lsjk_memcpy_dtod(p, o.p, n * sizeof(double)));
    }
};

PYBIND11_MODULE(_lightspeed, m)
{
    m.doc() = "lightspeed_amd native module: host API over the MI355X J/K library";

    py::class_<DeviceArray, std::shared_ptr<DeviceArray>>(m, "DeviceArray")
        .def(py::init<std::vector<size_t>>(), py::arg("shape"))
        .def_property_readonly("shape", [](const DeviceArray& d) { return d.shape; })
        .def_property_readonly("ptr", [](const DeviceArray& d) { return (uintptr_t)d.p; })
        .def("upload", &DeviceArray::upload)
        .def("download", &DeviceArray::download)
        .def("zero", &DeviceArray::zero)
        .def("copy_from", &DeviceArray::copy_from);

    py::class_<Tensor, std::shared_ptr<Tensor>>(m, "Tensor", py::buffer_protocol())
        .def(py::init([](std::vector<size_t> shape, std::string name) { return std::make_shared<Tensor>(shape, name); }),
             py::arg("shape"), py::arg("name") = "T")
        .def(py::init(&tensor_from_array))
        .def_static("array", &tensor_from_array)
        .def_static("zeros", [](std::vector<size_t> shape) { return std::make_shared<Tensor>(shape); })
        .def_property_readonly("shape", [](const Tensor& t) { return t.shape(); })
        .def_property_readonly("ndim", &Tensor::ndim)
        .def_property_readonly("size", &Tensor::size)
        .def_property("name", &Tensor::name, &Tensor::set_name)
        .def_property_readonly("np", &tensor_view)
        .def("zero", &Tensor::zero)
        .def("scale", &Tensor::scale)
        .def("axpby", &Tensor::axpby)
        .def("clone", &Tensor::clone)
        .def("shape_error", &Tensor::shape_error)
        .def("__getitem__", [](const std::shared_ptr<Tensor>& t, py::object key) { return tensor_view(t).attr("__getitem__")(key); })
        .def("__setitem__", [](const std::shared_ptr<Tensor>& t, py::object key, py::object v) { tensor_view(t).attr("__setitem__")(key, v); })
        .def_buffer([](Tensor& t) -> py::buffer_info {
            std::vector<py::ssize_t> shape(t.shape().begin(), t.shape().end());
            std::vector<py::ssize_t> strides(shape.size());
            py::ssize_t s = sizeof(double);
            for (int i = (int)shape.size() - 1; i >= 0; i--) {
                strides[i] = s;
                s *= shape[i];
            }
            return py::buffer_info(t.data().data(), sizeof(double), py::format_descriptor<double>::format(),
                                   (py::ssize_t)shape.size(), shape, strides);
        });
    py::implicitly_convertible<py::array, Tensor>();

    py::class_<Primitive>(m, "Primitive")
        .def(py::init<double, double, double, double, double, double, int, bool, int, int, int, int, int>(),
             py::arg("c"), py::arg("e"), py::arg("x"), py::arg("y"), py::arg("z"), py::arg("c0"), py::arg("L"),
             py::arg("is_pure"), py::arg("aoIdx"), py::arg("cartIdx"), py::arg("primIdx"), py::arg("shellIdx"),
             py::arg("atomIdx"))
        .def_property_readonly("c", &Primitive::c)
        .def_property_readonly("e", &Primitive::e)
        .def_property_readonly("x", &Primitive::x)
        .def_property_readonly("y", &Primitive::y)
        .def_property_readonly("z", &Primitive::z)
        .def_property_readonly("c0", &Primitive::c0)
        .def_property_readonly("L", &Primitive::L)
        .def_property_readonly("is_pure", &Primitive::is_pure)
        .def_property_readonly("aoIdx", &Primitive::aoIdx)
        .def_property_readonly("cartIdx", &Primitive::cartIdx)
        .def_property_readonly("primIdx", &Primitive::primIdx)
        .def_property_readonly("shellIdx", &Primitive::shellIdx)
        .def_property_readonly("atomIdx", &Primitive::atomIdx)
        .def_property_readonly("nao", &Primitive::nao)
        .def_property_readonly("npure", &Primitive::npure)
        .def_property_readonly("ncart", &Primitive::ncart)
        .def("__eq__", &Primitive::operator==);

    py::class_<Basis, std::shared_ptr<Basis>>(m, "Basis")
        .def(py::init<const std::string&, const std::vector<Primitive>&>(), py::arg("name"), py::arg("primitives"))
        .def_property_readonly("name", &Basis::name)
        .def_property_readonly("primitives", &Basis::primitives)
        .def_property_readonly("nao", &Basis::nao)
        .def_property_readonly("ncart", &Basis::ncart)
        .def_property_readonly("nprim", &Basis::nprim)
        .def_property_readonly("nshell", &Basis::nshell)
        .def_property_readonly("natom", &Basis::natom)
        .def_property_readonly("max_L", &Basis::max_L)
        .def_property_readonly("has_pure", &Basis::has_pure)
        .def_static("equivalent", &Basis::equivalent);

    py::class_<Ewald, std::shared_ptr<Ewald>>(m, "Ewald")
        .def(py::init<const std::vector<double>&, const std::vector<double>&>(), py::arg("scales"), py::arg("omegas"))
        .def_property_readonly("scales", &Ewald::scales)
        .def_property_readonly("omegas", &Ewald::omegas)
        .def_property_readonly("is_coulomb", &Ewald::is_coulomb)
        .def_static("coulomb", &Ewald::coulomb);

    py::class_<PairListL>(m, "PairListL")
        .def_property_readonly("is_symmetric", &PairListL::is_symmetric)
        .def_property_readonly("L1", &PairListL::L1)
        .def_property_readonly("L2", &PairListL::L2)
        .def_property_readonly("npair", [](const PairListL& l) { return l.pairs().size(); })
        .def("arrays", [](const PairListL& l) {
            size_t n = l.pairs().size();
            py::array_t<int> p1(n), p2(n);
            py::array_t<float> b(n);
            auto P1 = p1.mutable_unchecked<1>();
            auto P2 = p2.mutable_unchecked<1>();
            auto B = b.mutable_unchecked<1>();
            for (size_t k = 0; k < n; k++) {
                P1(k) = l.pairs()[k].prim1().primIdx();
                P2(k) = l.pairs()[k].prim2().primIdx();
                B(k) = l.pairs()[k].bound();
            }
            return py::make_tuple(p1, p2, b);
        });

    py::class_<PairList, std::shared_ptr<PairList>>(m, "PairList")
        .def_static("build_schwarz", &PairList::build_schwarz, py::arg("basis1"), py::arg("basis2"),
                    py::arg("is_symmetric"), py::arg("thre"))
        .def_property_readonly("basis1", &PairList::basis1)
        .def_property_readonly("basis2", &PairList::basis2)
        .def_property_readonly("is_symmetric", &PairList::is_symmetric)
        .def_property_readonly("thre", &PairList::thre)
        .def_property_readonly("pairlists", &PairList::pairlists)
        .def_property_readonly("npair", &PairList::npair)
        .def("max_bound", [](const std::shared_ptr<PairList>& p) { return PairListUtil::max_bound(p); })
        .def_property_readonly("device_ptr",
                               [](const std::shared_ptr<PairList>& p) { return (uintptr_t)p->device_handle(); })
        // every stride-th pair of each class, order kept (a bounded sample of a list for tests
        // against the CPU oracle at full-size workloads)
        .def("subsample", [](const std::shared_ptr<PairList>& p, int stride) {
            if (stride < 1) throw std::runtime_error("PairList.subsample: stride must be >= 1");
            std::vector<PairListL> ls;
            for (const auto& l : p->pairlists()) {
                std::vector<Pair> v;
                for (size_t k = 0; k < l.pairs().size(); k += (size_t)stride) v.push_back(l.pairs()[k]);
                ls.emplace_back(l.is_symmetric(), l.L1(), l.L2(), v);
            }
            return std::make_shared<PairList>(p->basis1(), p->basis2(), p->is_symmetric(), p->thre(), ls);
        }, py::arg("stride"));

    py::class_<ResourceList, std::shared_ptr<ResourceList>>(m, "ResourceList")
        .def_static("build", &ResourceList::build, py::arg("cpu_bytes") = 0, py::arg("gpu_bytes") = 0)
        .def_static("build_cpu", &ResourceList::build_cpu)
        .def_property_readonly("nthread", &ResourceList::nthread)
        .def_property_readonly("ngpu", &ResourceList::ngpu)
        .def("__str__", [](const ResourceList& r) {
            return "ResourceList: nthread = " + std::to_string(r.nthread()) + ", ngpu = " + std::to_string(r.ngpu());
        });

    py::class_<PureTransform>(m, "PureTransform")
        .def_static("solid_harmonics", [](int L) {
            auto v = PureTransform::solid_harmonics(L);
            py::array_t<double> a({2 * L + 1, (L + 1) * (L + 2) / 2});
            std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(double));
            return a;
        })
        .def_static("pureToCart2", &PureTransform::pureToCart2)
        .def_static("cartToPure2", &PureTransform::cartToPure2, py::arg("basis1"), py::arg("basis2"),
                    py::arg("Tcart"), py::arg("Tpure") = std::shared_ptr<Tensor>());

    py::class_<IntBox>(m, "IntBox")
        .def_static("coulomb", &IntBox::coulomb, py::arg("resources"), py::arg("ewald"), py::arg("pairlist12"),
                    py::arg("pairlist34"), py::arg("D34"), py::arg("thresp"), py::arg("thredp"),
                    py::arg("J12") = std::shared_ptr<Tensor>(), py::call_guard<py::gil_scoped_release>())
        .def_static("exchange", &IntBox::exchange, py::arg("resources"), py::arg("ewald"), py::arg("pairlist12"),
                    py::arg("pairlist34"), py::arg("D24"), py::arg("D24symm"), py::arg("thresp"), py::arg("thredp"),
                    py::arg("K13") = std::shared_ptr<Tensor>(), py::call_guard<py::gil_scoped_release>())
        .def_static(
            "coulomb_device",
            [](std::shared_ptr<Ewald> ew, std::shared_ptr<PairList> a, std::shared_ptr<PairList> b, uintptr_t D,
               float thresp, float thredp, uintptr_t J) {
                IntBox::coulomb_device(ew, a, b, (const double*)D, thresp, thredp, (double*)J);
            },
            py::call_guard<py::gil_scoped_release>())
        .def_static(
            "exchange_device",
            [](std::shared_ptr<Ewald> ew, std::shared_ptr<PairList> a, std::shared_ptr<PairList> b, uintptr_t D,
               bool D24symm, float thresp, float thredp, uintptr_t K) {
                IntBox::exchange_device(ew, a, b, (const double*)D, D24symm, thresp, thredp, (double*)K);
            },
            py::call_guard<py::gil_scoped_release>())
        .def_static("coulombGrad", &IntBox::coulombGrad, py::arg("resources"), py::arg("ewald"), py::arg("pairlist"),
                    py::arg("D12"), py::arg("D34"), py::arg("thresp"), py::arg("thredp"),
                    py::arg("G") = std::shared_ptr<Tensor>(), py::call_guard<py::gil_scoped_release>())
        .def_static("exchangeGrad", &IntBox::exchangeGrad, py::arg("resources"), py::arg("ewald"),
                    py::arg("pairlist"), py::arg("D13"), py::arg("D24"), py::arg("D13symm"), py::arg("D24symm"),
                    py::arg("Dsame"), py::arg("thresp"), py::arg("thredp"), py::arg("G") = std::shared_ptr<Tensor>(),
                    py::call_guard<py::gil_scoped_release>())
        .def_static("eriJ", &IntBox::eriJ, py::arg("resources"), py::arg("ewald"), py::arg("pairlist12"),
                    py::arg("pairlist34"), py::arg("C1"), py::arg("C2"), py::arg("C3"), py::arg("C4"),
                    py::arg("thresp"), py::arg("thredp"), py::arg("I") = std::shared_ptr<Tensor>(),
                    py::call_guard<py::gil_scoped_release>())
        .def_static("overlap", &IntBox::overlap, py::arg("resources"), py::arg("pairlist"),
                    py::arg("S") = std::shared_ptr<Tensor>(), py::call_guard<py::gil_scoped_release>())
        .def_static("kinetic", &IntBox::kinetic, py::arg("resources"), py::arg("pairlist"),
                    py::arg("T") = std::shared_ptr<Tensor>(), py::call_guard<py::gil_scoped_release>())
        .def_static("potential", &IntBox::potential, py::arg("resources"), py::arg("ewald"), py::arg("pairlist"),
                    py::arg("xyzc"), py::arg("V") = std::shared_ptr<Tensor>(), py::call_guard<py::gil_scoped_release>())
        .def_static("is_hip_ready", [](std::shared_ptr<ResourceList> r, std::shared_ptr<PairList> a,
                                       std::shared_ptr<PairList> b) { return IntBox::is_hip_ready(r, a, b); });

    py::class_<JKBatch, std::shared_ptr<JKBatch>>(m, "JKBatch")
        .def_static("build", &JKBatch::build, py::arg("resources"), py::arg("pairlists"))
        .def_property_readonly("nmol", &JKBatch::nmol)
        .def_property_readonly("ld", &JKBatch::ld)
        .def_property_readonly("nelem", &JKBatch::nelem)
        .def_property_readonly("offsets", &JKBatch::offsets)
        .def_property_readonly("pairlists", &JKBatch::pairlists)
        .def("jk", &JKBatch::jk, py::arg("ewald"), py::arg("Ds"), py::arg("thresp"), py::arg("thredp"),
             py::call_guard<py::gil_scoped_release>())
        .def(
            "jk_device",
            [](JKBatch& b, std::shared_ptr<Ewald> ew, uintptr_t D, float thresp, float thredp, uintptr_t J,
               uintptr_t K) { b.jk_device(ew, (const double*)D, thresp, thredp, (double*)J, (double*)K); },
            py::arg("ewald"), py::arg("D_ptr"), py::arg("thresp"), py::arg("thredp"), py::arg("J_ptr"),
            py::arg("K_ptr"), py::call_guard<py::gil_scoped_release>())
        .def(
            "autotune",
            [](JKBatch& b, std::shared_ptr<Ewald> ew, uintptr_t D, float thresp, float thredp) {
                return b.autotune(ew, (const double*)D, thresp, thredp);
            },
            py::arg("ewald"), py::arg("D_ptr"), py::arg("thresp"), py::arg("thredp"));

    py::class_<Boys>(m, "Boys").def_static("compute", &Boys::compute, py::arg("nboys"), py::arg("T"));
    m.def(
        "set_kgroup_rows",
        [](std::shared_ptr<PairList> pl, int L, int Lp, int rows) {
            check(lsjk_set_kgroup_rows(pl->device_handle(), L, Lp, rows));
        },
        py::arg("pl"), py::arg("L"), py::arg("Lp"), py::arg("rows"));

    auto tuned_list = [](const lsjk_pairlist* d) {
        int n = 0;
        check(lsjk_tuned_params(d, 0, &n, nullptr, nullptr, nullptr, nullptr, nullptr));
        std::vector<int> kind(std::max(n, 1)), idx(std::max(n, 1));
        std::vector<double> val(std::max(n, 1)), m0(std::max(n, 1)), m1(std::max(n, 1));
        check(lsjk_tuned_params(d, n, &n, kind.data(), idx.data(), val.data(), m0.data(), m1.data()));
        py::list out;
        for (int i = 0; i < n; i++) out.append(py::make_tuple(kind[i], idx[i], val[i], m0[i], m1[i]));
        return out;
    };
    m.def(
        "autotune",
        [tuned_list](std::shared_ptr<PairList> pl, std::shared_ptr<Ewald> ew, uintptr_t D, float thresp, float thredp) {
            lsjk_pairlist* d = pl->device_handle();
            const auto& sc = ew->scales();
            const auto& om = ew->omegas();
            {
                py::gil_scoped_release rel;
                check(lsjk_autotune(d, (const double*)D, (int)sc.size(), sc.data(), om.data(), thresp, thredp));
            }
            return tuned_list(d);
        },
        py::arg("pairlist"), py::arg("ewald"), py::arg("D_ptr"), py::arg("thresp"), py::arg("thredp"),
        "Autotune the J chunk and K split of a symmetric list on both sides; returns "
        "[(kind, index, value, ms_default, ms_tuned), ...]");

    m.def("hip_init", &hip_session_init, py::arg("device") = 0);
    m.def("hip_finalize", &hip_session_finalize);
    m.def("device_count", &hip_device_count);
    m.def("set_counting", [](bool on) { check(lsjk_set_counting(on ? 1 : 0)); });
    m.def("set_streams", [](int n) {
        hip_session_init(0);
        check(lsjk_set_streams(n));
    }, py::arg("n"));
    m.def("last_counts", []() {
        py::array_t<uint64_t> a(1024);
        check(lsjk_last_counts((uint64_t*)a.mutable_data()));
        return a;
    });
    m.def("last_counts_fp32", []() {
        py::array_t<uint64_t> a(1024);
        check(lsjk_last_counts_fp32((uint64_t*)a.mutable_data()));
        return a;
    });
    m.def("last_wave_steps", []() {
        py::array_t<uint64_t> s(1024), l(1024);
        check(lsjk_last_wave_steps((uint64_t*)s.mutable_data(), (uint64_t*)l.mutable_data()));
        return py::make_tuple(s, l);
    });
    m.def("last_timing", []() {
        double ms[2];
        check(lsjk_last_timing(ms));
        return py::make_tuple(ms[0], ms[1]);
    });
    m.def("synchronize", []() { check(lsjk_synchronize()); });
    m.def("set_profiling", [](bool on) { check(lsjk_set_profiling(on ? 1 : 0)); });
    m.def("last_launches", []() {
        int n = 0;
        check(lsjk_last_launches(0, &n, nullptr, nullptr));
        std::vector<int> ids(std::max(n, 1));
        std::vector<double> ms(std::max(n, 1));
        check(lsjk_last_launches(n, &n, ids.data(), ms.data()));
        py::list out;
        for (int i = 0; i < n; i++) out.append(py::make_tuple(ids[i], ms[i]));
        return out;
    });
    m.def(
        "count_quartets",
        [](std::shared_ptr<PairList> pl12, std::shared_ptr<PairList> pl34, uintptr_t D34, uintptr_t D24,
           bool D24symm, float thresp) {
            lsjk_pairlist* d12 = pl12->device_handle();
            lsjk_pairlist* d34 = (pl34 == pl12) ? d12 : pl34->device_handle();
            py::array_t<uint64_t> cj(625), ck(625);
            check(lsjk_count_quartets(d12, d34, (const double*)D34, (const double*)D24, D24symm ? 1 : 0, thresp,
                                      (uint64_t*)cj.mutable_data(), (uint64_t*)ck.mutable_data()));
            return py::make_tuple(cj, ck);
        },
        py::arg("pairlist12"), py::arg("pairlist34"), py::arg("D34_ptr"), py::arg("D24_ptr"), py::arg("D24symm"),
        py::arg("thresp"));
}

// pymodule.cpp -- pybind11 modules with the reference's names and class signatures:
//   pyElastic_iso_{float,double}_nl   : spaceInterpGpu, nonlinearPropElasticShotsGpu
//                                       (python_float/pyElastic_iso_float_nl.cpp:15-57)
//   pyElastic_iso_{float,double}_born : spaceInterpGpu, BornElasticShotsGpu
//                                       (python_float/pyElastic_iso_float_born.cpp:72-97)
// Compiled once per (EIO_REAL, EIO_KIND_*) with -DEIO_MODULE=<name>.
//
// Vectors are the package's SEP-style vectors (sep/SepVector.py): any object with
// getNdArray() (C-contiguous numpy array, slowest axis first) and getHyper().axes
// (objects with n, o, d).  Arrays are used in place (no copies); dtype must match the
// module's real type.  Parameters are a mapping of key -> string (genericIO shim `.param`).
#include <pybind11/iostream.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "host_ops.h"

namespace py = pybind11;
using real = EIO_REAL;
using namespace eioh;

namespace {

std::vector<Axis> axes_of(const py::object& hyper) {
    std::vector<Axis> out;
    py::object axes = py::hasattr(hyper, "axes") ? hyper.attr("axes") : hyper;
    for (auto a : axes) {
        py::object ao = py::reinterpret_borrow<py::object>(a);
        Axis x;
        x.n = ao.attr("n").cast<int>();
        x.o = ao.attr("o").cast<double>();
        x.d = ao.attr("d").cast<double>();
        out.push_back(x);
    }
    return out;
}

py::object cpp_of(const py::object& v) {
    return py::hasattr(v, "getCpp") ? v.attr("getCpp")() : v;
}

// numpy view of a vector, checked for dtype and contiguity; never copies.
py::array_t<real> array_of(const py::object& vec, bool writable) {
    py::object v = cpp_of(vec);
    py::object a = py::hasattr(v, "getNdArray") ? v.attr("getNdArray")() : v;
    if (!py::isinstance<py::array>(a)) throw std::runtime_error("vector has no numpy storage");
    py::array arr = py::reinterpret_borrow<py::array>(a);
    if (!arr.dtype().is(py::dtype::of<real>()))
        throw std::runtime_error(std::string("vector dtype must be ") + (sizeof(real) == 4 ? "float32" : "float64"));
    if (!(arr.flags() & py::array::c_style)) throw std::runtime_error("vector storage must be C-contiguous");
    if (writable && !arr.writeable()) throw std::runtime_error("output vector is read-only");
    return py::reinterpret_borrow<py::array_t<real>>(arr);
}

std::vector<Axis> hyper_of_vec(const py::object& vec) {
    py::object v = cpp_of(vec);
    return axes_of(v.attr("getHyper")());
}

Params params_of(const py::object& p) {
    Params out;
    py::object m = py::hasattr(p, "param") ? p.attr("param") : p;
    if (py::hasattr(m, "items")) {
        for (auto kv : m.attr("items")()) {
            py::tuple t = py::reinterpret_borrow<py::tuple>(kv);
            out.kv[py::str(t[0]).cast<std::string>()] = py::str(t[1]).cast<std::string>();
        }
        return out;
    }
    throw std::runtime_error("parameter object must provide a mapping of key -> value");
}

using SpaceInterp = SpaceInterpT<real>;

std::vector<real> coords_of(const py::object& vec) {
    py::object v = cpp_of(vec);
    py::object a = py::hasattr(v, "getNdArray") ? v.attr("getNdArray")() : v;
    py::array_t<real, py::array::c_style | py::array::forcecast> arr(a);
    return std::vector<real>(arr.data(), arr.data() + arr.size());
}

std::vector<InterpVec<real>> interp_lists(const std::vector<std::shared_ptr<SpaceInterp>>& c,
                                          const std::vector<std::shared_ptr<SpaceInterp>>& x,
                                          const std::vector<std::shared_ptr<SpaceInterp>>& z,
                                          const std::vector<std::shared_ptr<SpaceInterp>>& xz) {
    if (c.size() != x.size() || c.size() != z.size() || c.size() != xz.size())
        throw std::runtime_error("geometry lists of the four grids differ in length");
    std::vector<InterpVec<real>> out(c.size());
    for (size_t i = 0; i < c.size(); ++i) out[i] = {c[i], x[i], z[i], xz[i]};
    return out;
}

// [nts, nTrace, 5, nShot] SEP vector -> dims (nShot, nTrace)
void trace_dims(const py::array_t<real>& a, int nts, int& nshot, int& ntr) {
    if (a.ndim() == 4) {
        nshot = int(a.shape(0));
        if (a.shape(1) != 5) throw std::runtime_error("component axis must have 5 entries");
        ntr = int(a.shape(2));
        if (a.shape(3) != nts) throw std::runtime_error("time axis not consistent with nts");
    } else if (a.ndim() == 3) {
        nshot = 1;
        if (a.shape(0) != 5) throw std::runtime_error("component axis must have 5 entries");
        ntr = int(a.shape(1));
        if (a.shape(2) != nts) throw std::runtime_error("time axis not consistent with nts");
    } else {
        throw std::runtime_error("traces must be a 3-D or 4-D vector");
    }
}

template <class Op>
bool dot_test(Op& op, bool verbose, float maxError,
              const std::function<void(bool, std::vector<real>&, std::vector<real>&)>& fwd,
              const std::function<void(bool, std::vector<real>&, std::vector<real>&)>& adj, size_t nm, size_t nd) {
    // Operator::dotTest (prop/include/operator.h:50-89)
    std::mt19937 gen(1234);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::vector<real> d1(nm), d2(nm), r1(nd), r2(nd);
    for (auto& x : d1) x = real(U(gen));
    for (auto& x : r1) x = real(U(gen));
    auto dot = [](const std::vector<real>& a, const std::vector<real>& b) {
        double s = 0;
        for (size_t i = 0; i < a.size(); ++i) s += double(a[i]) * double(b[i]);
        return s;
    };
    fwd(false, d1, r2);
    adj(false, d2, r1);
    double dot1 = dot(r1, r2), dot2 = dot(d1, d2);
    const double eNoAdd = std::abs((dot1 - dot2) / dot2);
    fwd(true, d1, r2);
    adj(true, d2, r1);
    double dot3 = dot(r1, r2), dot4 = dot(d1, d2);
    const double eAdd = std::abs((dot3 - dot4) / dot4);
    if (verbose) {
        py::print("WITHOUT ADD: dot1:", dot1, "dot2:", dot2, "error:", eNoAdd);
        py::print("WITH ADD: dot3:", dot3, "dot4:", dot4, "error:", eAdd);
    }
    (void)op;
    return eNoAdd <= maxError && eAdd <= maxError;
}

}  // namespace

PYBIND11_MODULE(EIO_MODULE, m) {
    py::add_ostream_redirect(m, "ostream_redirect");

#if defined(EIO_DEFINE_INTERP)
    py::class_<SpaceInterp, std::shared_ptr<SpaceInterp>>(m, "spaceInterpGpu")
        .def(py::init([](py::object zCoord, py::object xCoord, py::object hyper, int nt, std::string method,
                         int nFilt, int dipole, real zShift, real xShift) {
                 return std::make_shared<SpaceInterp>(coords_of(zCoord), coords_of(xCoord),
                                                      axes_of(cpp_of(hyper)), nt, method, nFilt, dipole, zShift,
                                                      xShift);
             }),
             "Initialize a spaceInterpGpu object using zcoord, xcoord, hypercube and nt")
        .def("getInfo", &SpaceInterp::getInfo)
        .def("getRegPosUnique", [](const SpaceInterp& s) { return s.uniq; })
        .def("getRegPos", [](const SpaceInterp& s) { return s.gridPointIndex; })
        .def("getWeights", [](const SpaceInterp& s) { return s.weight; })
        .def("getNDeviceReg", &SpaceInterp::nReg)
        .def("getNDeviceIrreg", [](const SpaceInterp& s) { return s.nDeviceIrreg; })
        .def("getNt", [](const SpaceInterp& s) { return s.nt; })
        .def("getSizePosUnique", [](const SpaceInterp& s) { return int(s.uniq.size()); });
#else
    // one registration of the geometry class per precision (in its NL module), shared with the
    // Born module of that precision
    m.attr("spaceInterpGpu") =
        py::module_::import(sizeof(real) == 4 ? "elastic_isotropic_operators_amd.pyElastic_iso_float_nl"
                                              : "elastic_isotropic_operators_amd.pyElastic_iso_double_nl")
            .attr("spaceInterpGpu");
#endif

#if defined(EIO_KIND_NL)
    // fdParamElastic (fdParamElastic.cpp): parameter checks and scaled media, host only
    using FD = FdParam<real>;
    py::class_<FD, std::shared_ptr<FD>>(m, "fdParamElastic")
        .def(py::init([](py::object elasticParam, py::object par) {
            py::array_t<real> el = array_of(elasticParam, false);
            return std::make_shared<FD>(el.data(), hyper_of_vec(elasticParam), params_of(par));
        }))
        .def("getInfo", &FD::getInfo)
        .def_readonly("nts", &FD::nts)
        .def_readonly("sub", &FD::sub)
        .def_readonly("ntw", &FD::ntw)
        .def_readonly("dtw", &FD::dtw)
        .def_readonly("minPad", &FD::minPad)
        .def_readonly("Courant", &FD::Courant)
        .def_readonly("dispersionRatio", &FD::dispersionRatio)
        .def_readonly("minVpVs", &FD::minVpVs)
        .def_readonly("maxVpVs", &FD::maxVpVs)
        .def("getMedia", [](const FD& f) {
            py::array_t<real> out({size_t(5), size_t(f.nx), size_t(f.nz)});
            std::copy(f.media.begin(), f.media.end(), out.mutable_data());
            return out;
        });
#else
    m.attr("fdParamElastic") =
        py::module_::import(sizeof(real) == 4 ? "elastic_isotropic_operators_amd.pyElastic_iso_float_nl"
                                              : "elastic_isotropic_operators_amd.pyElastic_iso_double_nl")
            .attr("fdParamElastic");
#endif

#if defined(EIO_KIND_NL)
    using NL = NonlinearShots<real>;
    py::class_<NL, std::shared_ptr<NL>>(m, "nonlinearPropElasticShotsGpu")
        .def(py::init([](py::object elasticParam, py::object par, std::vector<std::shared_ptr<SpaceInterp>> sc,
                         std::vector<std::shared_ptr<SpaceInterp>> sx, std::vector<std::shared_ptr<SpaceInterp>> sz,
                         std::vector<std::shared_ptr<SpaceInterp>> sxz,
                         std::vector<std::shared_ptr<SpaceInterp>> rc, std::vector<std::shared_ptr<SpaceInterp>> rx,
                         std::vector<std::shared_ptr<SpaceInterp>> rz,
                         std::vector<std::shared_ptr<SpaceInterp>> rxz) {
                 py::array_t<real> el = array_of(elasticParam, false);
                 return std::make_shared<NL>(el.data(), hyper_of_vec(elasticParam), params_of(par),
                                             interp_lists(sc, sx, sz, sxz), interp_lists(rc, rx, rz, rxz));
             }),
             "Initialize a nonlinearPropElasticShotsGpu")
        .def("forward",
             [](NL& op, bool add, py::object model, py::object data) {
                 py::array_t<real> mo = array_of(model, false), da = array_of(data, true);
                 int ms, ns, ds, nr;
                 trace_dims(mo, op.fd().nts, ms, ns);
                 trace_dims(da, op.fd().nts, ds, nr);
                 if (ds != op.nExp()) throw std::runtime_error("data shot axis must equal nExp");
                 py::gil_scoped_release rel;
                 op.forward(add, mo.data(), ms, ns, da.mutable_data(), nr);
             })
        .def("adjoint",
             [](NL& op, bool add, py::object model, py::object data) {
                 py::array_t<real> mo = array_of(model, true), da = array_of(data, false);
                 int ms, ns, ds, nr;
                 trace_dims(mo, op.fd().nts, ms, ns);
                 trace_dims(da, op.fd().nts, ds, nr);
                 if (ds != op.nExp()) throw std::runtime_error("data shot axis must equal nExp");
                 py::gil_scoped_release rel;
                 op.adjoint(add, mo.mutable_data(), ms, ns, da.data(), nr);
             })
        .def("setBackground",
             [](NL& op, py::object elasticParam) {
                 py::array_t<real> el = array_of(elasticParam, false);
                 op.setBackground(el.data());
             })
        .def("forwardWavefield",
             [](NL& op, bool add, py::object model, py::object data) {
                 py::array_t<real> mo = array_of(model, false), da = array_of(data, true);
                 int ms, ns, ds, nr;
                 trace_dims(mo, op.fd().nts, ms, ns);
                 trace_dims(da, op.fd().nts, ds, nr);
                 if (ds != op.nExp()) throw std::runtime_error("data shot axis must equal nExp");
                 py::gil_scoped_release rel;
                 op.forward(add, mo.data(), ms, ns, da.mutable_data(), nr, true);
             })
        .def("adjointWavefield",
             [](NL& op, bool add, py::object model, py::object data) {
                 py::array_t<real> mo = array_of(model, true), da = array_of(data, false);
                 int ms, ns, ds, nr;
                 trace_dims(mo, op.fd().nts, ms, ns);
                 trace_dims(da, op.fd().nts, ds, nr);
                 if (ds != op.nExp()) throw std::runtime_error("data shot axis must equal nExp");
                 py::gil_scoped_release rel;
                 op.adjoint(add, mo.mutable_data(), ms, ns, da.data(), nr, true);
             })
        .def("getWavefield",
             // [nts][5][nx][nz] (SEP axes z, x, component, time), a copy
             [](NL& op) {
                 const auto& w = op.wavefield();
                 if (w.empty()) throw std::runtime_error("no wavefield saved (saveWavefield=1 and a *Wavefield call)");
                 const auto& fd = op.fd();
                 py::array_t<real> out({py::ssize_t(fd.nts), py::ssize_t(5), py::ssize_t(fd.nx), py::ssize_t(fd.nz)});
                 std::memcpy(out.mutable_data(), w.data(), w.size() * sizeof(real));
                 return out;
             })
        .def("lastPropagationMs", &NL::last_ms)
        .def("dotTest",
             [](NL& op, bool verbose, float maxError) {
                 const int nts = op.fd().nts;
                 // domain: one source signal [5][nSrc][nts]; range: [nExp][5][nRec][nts]
                 const int nSrc = op.nSrcIrreg(), nRec = op.nRecIrreg();
                 const size_t nm = size_t(5) * nSrc * nts, nd = size_t(op.nExp()) * 5 * nRec * nts;
                 auto fwd = [&](bool add, std::vector<real>& x, std::vector<real>& y) {
                     op.forward(add, x.data(), 1, nSrc, y.data(), nRec);
                 };
                 auto adj = [&](bool add, std::vector<real>& x, std::vector<real>& y) {
                     op.adjoint(add, x.data(), 1, nSrc, y.data(), nRec);
                 };
                 return dot_test(op, verbose, maxError, fwd, adj, nm, nd);
             },
             py::arg("verbose") = false, py::arg("maxError") = 1e-5f);
#endif

#if defined(EIO_KIND_BORN)
    using BO = BornShots<real>;
    py::class_<BO, std::shared_ptr<BO>>(m, "BornElasticShotsGpu")
        .def(py::init([](py::object elasticParam, py::object par, py::list sourcesSignals,
                         std::vector<std::shared_ptr<SpaceInterp>> sc, std::vector<std::shared_ptr<SpaceInterp>> sx,
                         std::vector<std::shared_ptr<SpaceInterp>> sz, std::vector<std::shared_ptr<SpaceInterp>> sxz,
                         std::vector<std::shared_ptr<SpaceInterp>> rc, std::vector<std::shared_ptr<SpaceInterp>> rx,
                         std::vector<std::shared_ptr<SpaceInterp>> rz,
                         std::vector<std::shared_ptr<SpaceInterp>> rxz) {
                 py::array_t<real> el = array_of(elasticParam, false);
                 std::vector<std::vector<real>> sig;
                 for (auto s : sourcesSignals) {
                     py::array_t<real> a = array_of(py::reinterpret_borrow<py::object>(s), false);
                     sig.emplace_back(a.data(), a.data() + a.size());
                 }
                 return std::make_shared<BO>(el.data(), hyper_of_vec(elasticParam), params_of(par), sig,
                                             interp_lists(sc, sx, sz, sxz), interp_lists(rc, rx, rz, rxz));
             }),
             "Initialize a BornElasticShotsGpu")
        .def("forward",
             [](BO& op, bool add, py::object model, py::object data) {
                 py::array_t<real> mo = array_of(model, false), da = array_of(data, true);
                 if (size_t(mo.size()) != 3 * op.fd().ncell()) throw std::runtime_error("Born model must be [3][nx][nz]");
                 int ds, nr;
                 trace_dims(da, op.fd().nts, ds, nr);
                 if (ds != op.nExp()) throw std::runtime_error("data shot axis must equal nExp");
                 py::gil_scoped_release rel;
                 op.forward(add, mo.data(), da.mutable_data(), nr);
             })
        .def("adjoint",
             [](BO& op, bool add, py::object model, py::object data) {
                 py::array_t<real> mo = array_of(model, true), da = array_of(data, false);
                 if (size_t(mo.size()) != 3 * op.fd().ncell()) throw std::runtime_error("Born model must be [3][nx][nz]");
                 int ds, nr;
                 trace_dims(da, op.fd().nts, ds, nr);
                 if (ds != op.nExp()) throw std::runtime_error("data shot axis must equal nExp");
                 py::gil_scoped_release rel;
                 op.adjoint(add, mo.mutable_data(), da.data(), nr);
             })
        .def("lastPropagationMs", &BO::last_ms)
        .def("setBackground", [](BO& op, py::object elasticParam) {
            py::array_t<real> el = array_of(elasticParam, false);
            op.setBackground(el.data());
        });
#endif
}

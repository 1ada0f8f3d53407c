// scene_pybind: reference renderer/pybindFiles/scene_pybind.cpp:8-89.  The
// three constructors are the Bessel, spline (Bessel + VOL file) and
// point-source builds of the reference, distinguished by arity.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "simpler_compat.h"

namespace py = pybind11;
using S = scn::Scene<tvec::TVector3>;
using V3 = tvec::TVector3<Float>;
using V2 = tvec::Vec2f;

PYBIND11_MODULE(scene_pybind, m) {
    m.doc() = "scene pybind module";
    py::class_<S>(m, "Scene")
        .def(py::init<Float, const V3 &, const V3 &, const V3 &, const V3 &, const Float &, const std::string &,
                      const V2 &, const Float, const V3 &, const V3 &, const V3 &, const V2 &, const V2 &,
                      const bool &, const std::string &, const Float &, const V3 &, const Float &, const Float &,
                      const bool &, const V3 &, const Float &, const Float &, const bool &, const Float &,
                      const Float &, const Float &, const Float &, const Float &, const Float &, const Float &,
                      const int &, const V3 &, const V3 &, const V3 &, const Float &, const Float &, const Float &,
                      const int &, const Float &, const Float &, const bool &>())
        .def(py::init<Float, const V3 &, const V3 &, const V3 &, const V3 &, const Float &, const std::string &,
                      const V2 &, const Float, const V3 &, const V3 &, const V3 &, const V2 &, const V2 &,
                      const bool &, const std::string &, const Float &, const V3 &, const Float &, const Float &,
                      const bool &, const V3 &, const Float &, const Float &, const bool &, const Float &,
                      const Float &, const Float &, const Float &, const Float &, const Float &, const Float &,
                      const int &, const V3 &, const V3 &, const V3 &, const Float &, const Float &, const Float &,
                      const int &, const Float &, const Float &, const bool &, const std::string &>())
        .def(py::init<Float, const V3 &, const V3 &, const V3 &, const V3 &, const Float &, const std::string &,
                      const V2 &, const Float, const V3 &, const V3 &, const V3 &, const V2 &, const V2 &,
                      const bool &, const std::string &, const Float &, const V3 &, const Float &, const Float &,
                      const bool &, const V3 &, const Float &, const Float &, const bool &, const Float &,
                      const Float &, const Float &, const Float &, const int &, const Float &, const V3 &,
                      const V3 &, const bool &, const bool &, const Float &, const Float &, const Float &,
                      const Float &, const int &, const Float &, const Float &, const int &, const V3 &, const V3 &,
                      const V3 &, const Float &, const Float &, const Float &, const int &, const Float &,
                      const Float &, const bool &>())
        .def("set_f_u", &S::set_f_u)
        .def("set_n_max", &S::set_n_max)
        .def("set_n_scaling", &S::set_n_scaling)
        .def("set_phase1", &S::set_phase1)
        .def("set_phase2", &S::set_phase2)
        .def("field_kind", [](const S &s) { return s.desc().field_kind; });
    // VOL grids and the B-spline prefilter (reference spline.h:115-162,686-764)
    m.def("prefilter3d", [](py::array_t<float, py::array::c_style | py::array::forcecast> data) {
        if (data.ndim() != 3) throw std::invalid_argument("prefilter3d: expected a (z, y, x) array");
        const int N[3] = {(int)data.shape(2), (int)data.shape(1), (int)data.shape(0)};
        std::vector<float> v(data.data(), data.data() + data.size());
        std::vector<float> c = spline::prefilter3d(v, N);
        py::array_t<float> out({data.shape(0), data.shape(1), data.shape(2)});
        std::copy(c.begin(), c.end(), out.mutable_data());
        return out;
    });
    m.def("read_vol", [](const std::string &f) {
        spline::VolGrid g = spline::readVol(f);
        py::array_t<float> a({g.N[2], g.N[1], g.N[0]});
        std::copy(g.data.begin(), g.data.end(), a.mutable_data());
        return py::make_tuple(a, std::vector<float>(g.xmin, g.xmin + 3), std::vector<float>(g.xmax, g.xmax + 3));
    });
    m.def("write_vol", [](const std::string &f, py::array_t<float, py::array::c_style | py::array::forcecast> a,
                          std::vector<float> xmin, std::vector<float> xmax) {
        if (a.ndim() != 3 || xmin.size() != 3 || xmax.size() != 3) throw std::invalid_argument("write_vol: bad shape");
        spline::VolGrid g;
        g.N[0] = (int)a.shape(2);
        g.N[1] = (int)a.shape(1);
        g.N[2] = (int)a.shape(0);
        for (int i = 0; i < 3; ++i) {
            g.xmin[i] = xmin[i];
            g.xmax[i] = xmax[i];
        }
        g.data.assign(a.data(), a.data() + a.size());
        spline::writeVol(f, g);
    });
}

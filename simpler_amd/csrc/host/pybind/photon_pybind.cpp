// photon_pybind: reference renderer/pybindFiles/photon_pybind.cpp:9-15.
// renderImage releases the GIL while the GPU renders.
#include <pybind11/pybind11.h>

#include "simpler_compat.h"

namespace py = pybind11;
using R = photon::Renderer<tvec::TVector3>;

PYBIND11_MODULE(photon_pybind, m) {
    m.doc() = "photon pybind module";
    py::class_<R>(m, "Renderer")
        .def(py::init<const int, const Float, const bool, const bool, const int64_t>())
        .def("renderImage",
             [](const R &r, image::Image3<double> &img, const med::Medium &medium,
                const scn::Scene<tvec::TVector3> &scene, int64 n) {
                 py::gil_scoped_release nogil;
                 r.renderImage(img, medium, scene, n);
             })
        .def("setSeed", &R::setSeed)
        .def("getSeed", &R::getSeed);
}

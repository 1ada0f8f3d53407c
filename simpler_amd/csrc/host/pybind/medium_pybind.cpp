// medium_pybind: reference renderer/pybindFiles/medium_pybind.cpp:10-15
#include <pybind11/pybind11.h>

#include "simpler_compat.h"

namespace py = pybind11;

PYBIND11_MODULE(medium_pybind, m) {
    m.doc() = "medium pybind module";
    py::class_<med::Medium>(m, "Medium")
        .def(py::init<Float, Float, pfunc::henyey_greenstein *>(), py::keep_alive<1, 4>())
        .def("getSigmaT", &med::Medium::getSigmaT)
        .def("getAlbedo", &med::Medium::getAlbedo)
        .def("getMfp", &med::Medium::getMfp);
}

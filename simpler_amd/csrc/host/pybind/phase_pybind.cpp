// phase_pybind: reference renderer/pybindFiles/phase_pybind.cpp:8-11
#include <pybind11/pybind11.h>

#include "simpler_compat.h"

namespace py = pybind11;

PYBIND11_MODULE(phase_pybind, m) {
    m.doc() = "phase pybind module";
    py::class_<pfunc::henyey_greenstein>(m, "henyey_greenstein")
        .def(py::init<Float>())
        .def("getG", &pfunc::henyey_greenstein::getG);
}

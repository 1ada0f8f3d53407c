// image_pybind: reference renderer/pybindFiles/image_pybind.cpp:10-30
// (Image3<double> as SmallImage), plus whole-image access as a numpy array
// (z, y, x) and PFM3D I/O.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include "simpler_compat.h"

namespace py = pybind11;
using Img = image::Image3<double>;

PYBIND11_MODULE(image_pybind, m) {
    m.doc() = "image pybind module";
    py::class_<Img>(m, "SmallImage", py::buffer_protocol())
        .def(py::init<int, int, int>())
        .def("get_x_res", &Img::get_x_res)
        .def("get_y_res", &Img::get_y_res)
        .def("getZRes", &Img::getZRes)
        .def("get_pixel", &Img::get_pixel)
        .def("setPixel", &Img::setPixel)
        .def("zero", &Img::zero)
        .def("writePFM3D", &Img::writePFM3D)
        .def("readPFM3D", &Img::readPFM3D)
        .def("numpy", [](Img &im) {
            return py::array_t<double>({im.getZRes(), im.get_y_res(), im.get_x_res()},
                                       {sizeof(double) * im.get_x_res() * im.get_y_res(),
                                        sizeof(double) * im.get_x_res(), sizeof(double)},
                                       im.data(), py::cast(im));
        })
        .def_buffer([](Img &im) -> py::buffer_info {
            return py::buffer_info(im.data(), sizeof(double), py::format_descriptor<double>::format(), 3,
                                   {im.getZRes(), im.get_y_res(), im.get_x_res()},
                                   {sizeof(double) * im.get_x_res() * im.get_y_res(),
                                    sizeof(double) * im.get_x_res(), sizeof(double)});
        });
}
